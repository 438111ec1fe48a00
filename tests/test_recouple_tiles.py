"""Host tables of the recouple kernel, on the CPU: every destination element of every job is
covered by exactly one tile, and no tile is empty, under the kernel's tile -> element map (``include/tenet_amd.h``:
linear tiles of 2048 / 512 elements in destination C order; 64 x 64 / 32 x 32 transpose tiles
over dims (rank-1, J) with the remaining dims as batch), for programs large and small enough
to take every tile kind.  Also the packed upload of a program's device tables."""

import numpy as np
import pytest
import torch

from tenet.engine import JOB_WORDS, RecoupleProgram, _DeviceCache

LIN = {-1: 2048, -2: 512}


def _coverage(prog):
    h = prog._tables.host
    jobs = np.asarray(h["jobs"]).reshape(-1, JOB_WORDS)
    tiles = np.asarray(h["tiles"]).reshape(-1, 2)
    cover = [np.zeros(int(j[23]), dtype=np.int64) for j in jobs]
    for job, tile in tiles:
        jw = jobs[job]
        rank, mode, nelem = int(jw[3]), int(jw[22]), int(jw[23])
        shape = [int(x) for x in jw[4:4 + rank]]
        if mode < 0:
            lo = tile * LIN[mode]
            assert lo < nelem, "a tile past the job's end"
            cover[job][lo:min(lo + LIN[mode], nelem)] += 1
            continue
        E = 32 if mode >= 8 else 64
        J = mode - 8 if mode >= 8 else mode
        I = rank - 1
        nI, nJ = -(-shape[I] // E), -(-shape[J] // E)
        ti, tj, b = tile % nI, (tile // nI) % nJ, tile // (nI * nJ)
        rest = [d for d in range(rank) if d not in (I, J)]
        idx = [0] * rank
        for d in reversed(rest):  # batch index over the other dims, last fastest
            idx[d] = b % shape[d]
            b //= shape[d]
        assert b == 0 and ti * E < shape[I] and tj * E < shape[J], "a tile past the job's end"
        ii, jj = np.meshgrid(np.arange(ti * E, min(ti * E + E, shape[I])),
                             np.arange(tj * E, min(tj * E + E, shape[J])), indexing="ij")
        full = [np.full(ii.size, idx[d]) for d in range(rank)]
        full[I], full[J] = ii.ravel(), jj.ravel()
        np.add.at(cover[job], np.ravel_multi_index(full, shape), 1)
    return cover


def _regions(rng, n, big):
    regions, src, dst = [], 0, 0
    for _ in range(n):
        if rng.random() < 0.3:  # a transpose whose destination-contiguous dim is strided in the source
            a, b = (int(x) for x in rng.integers(16, 150, 2))
            c = int(rng.integers(1, 4))
            shape = (c, a, b)  # destination order; the source holds (c, b, a)
            regions.append((dst, shape, (a * b, b, 1), (a * b, 1, a), [(src, 1.0)]))
        else:  # a contiguous run, two terms
            m = int(rng.integers(1, 400_000 if big else 700))
            shape = (m,)
            regions.append((dst, shape, (1,), (1,), [(src, 1.0), (src, -0.5)]))
        size = int(np.prod(shape))
        src += size
        dst += size
    return regions, src, dst


@pytest.mark.parametrize("big,seed", [(True, 0), (True, 1), (False, 2), (False, 3)])
def test_every_element_is_covered_once(big, seed):
    rng = np.random.default_rng(seed)
    regions, ssz, dsz = _regions(rng, 30 if big else 6, big)
    prog = RecoupleProgram(regions, ssz, dsz)
    modes = set(int(m) for m in np.asarray(prog._tables.host["jobs"]).reshape(-1, JOB_WORDS)[:, 22])
    # large programs keep the large tiles, small ones switch to the small tiles
    assert modes <= (set(range(-1, 6)) if big else {-2, *range(8, 14)}), modes
    assert (prog.ntiles >= 2048) == big
    for c in _coverage(prog):
        assert (c == 1).all()


def test_regions_past_the_buffers_are_refused():
    with pytest.raises(ValueError):
        RecoupleProgram([(0, (10,), (1,), (1,), [(5, 1.0)])], 12, 10)
    with pytest.raises(ValueError):
        RecoupleProgram([(3, (10,), (1,), (1,), [(0, 1.0)])], 12, 12)


def test_packed_device_tables_keep_values_and_types():
    arrays = dict(a=np.arange(7, dtype=np.int64), b=np.array([0.5, -2.0, 3.25]),
                  c=np.arange(5, dtype=np.int32), d=np.zeros(0, dtype=np.int64),
                  e=np.array([1.5], dtype=np.float32))
    got = _DeviceCache(**arrays).get(torch.device("cpu"))
    offs = []
    for k, v in arrays.items():
        t = got[k]
        assert t.dtype == {np.int64: torch.int64, np.float64: torch.float64, np.int32: torch.int32,
                           np.float32: torch.float32}[v.dtype.type]
        assert t.numpy().tolist() == v.tolist()
        offs.append(t.storage_offset() * t.element_size())
    assert all(o % _DeviceCache._ALIGN == 0 for o in offs)
    with pytest.raises(TypeError):
        _DeviceCache(x=np.zeros(3, dtype=np.int16)).get(torch.device("cpu"))
