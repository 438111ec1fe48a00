"""Host logic of the sector-stacked two-site H_eff (``tenet.network.heff._SectorChain``):
its compiled tables -- entry gather, FR/FL stacking, the two GEMM problem lists, the MPO
slab jobs and the exit scatter -- are executed here by a NumPy interpreter of the kernels'
documented semantics and compared with the FP64 oracle of the reference's ``_heff2_full``
(``network/env.py:416-432``).  No device is needed: any wrong offset, stride, coefficient or
missing term in the tables fails these tests on the CPU."""

import numpy as np
import pytest

from oracle import u1_network as N
from tests.helpers import rel_err, to_oracle
from tests.interp import run_gemm, run_mpo, run_recouple


def emulate(ch, fl, fr, aa, rows: bool = False):
    """The chain's tables on the host; ``rows``: step 1 reads aa's blocks through its row
    table (the FP32 / BF16x3 form) instead of the entry-gathered x."""
    x = aa.data.double().numpy()
    frs = np.zeros(max(ch.frs_size, 1))
    run_recouple(ch.fr_stack, fr.data.double().numpy(), frs)
    fls = np.zeros(max(ch.fls_size, 1))
    run_recouple(ch.fl_stack, fl.data.double().numpy(), fls)
    t1 = np.zeros(max(ch.t1_size, 1))
    if rows:
        run_gemm(ch.gemm1_rows, x, frs, t1)
    else:
        xs = np.zeros(max(ch.xs_size, 1))
        run_recouple(ch.entry, x, xs)
        run_gemm(ch.gemm1, xs, frs, t1)
    t3 = np.zeros(max(ch.t3_size, 1))
    run_mpo(*ch.mpo_host, t1, t3)
    ws = run_gemm(ch.gemm4, fls, t3, None)
    out = np.zeros(aa.structure.size)
    run_recouple(ch.exit, ws, out)
    return out


def _problem(left, right, n_sites, dim, sigma, seed):
    from tenet.network.models import MPO_BOND, PHYS, gaussian_bond, heisenberg_mpo
    from tenet.space import IN, OUT, Leg
    from tenet.tensor import SymmetricTensor

    bl = gaussian_bond(left, n_sites, dim, sigma)
    br = gaussian_bond(right, n_sites, dim, sigma)
    h = heisenberg_mpo(n_sites, 0.8, device="cpu")
    fl = SymmetricTensor.random((Leg(bl, IN), Leg(MPO_BOND, OUT), Leg(bl, OUT)), seed=seed, device="cpu")
    fr = SymmetricTensor.random((Leg(br, OUT), Leg(MPO_BOND, IN), Leg(br, IN)), seed=seed + 1, device="cpu")
    aa = SymmetricTensor.random((Leg(bl, OUT), Leg(PHYS, OUT), Leg(PHYS, OUT), Leg(br, IN)),
                                seed=seed + 2, device="cpu")
    return fl, h[left], h[left + 1], fr, aa


def _chain(fl, w1, w2, fr, aa):
    from tenet.network.heff import _SectorChain, _w_host

    return _SectorChain(aa.structure, fr.structure, fl.structure, _w_host(w1), _w_host(w2))


@pytest.mark.parametrize("left,right,n_sites,dim,sigma,seed", [
    (5, 7, 12, 40, 1.5, 3),      # centre bond, symmetric sector sets
    (2, 4, 12, 64, 3.0, 11),     # left bond 3 sectors, right bond 5
    (9, 11, 20, 512, 4.0, 5),    # step 3 split into K slices
])
def test_sector_chain_tables_reproduce_the_oracle(left, right, n_sites, dim, sigma, seed):
    fl, w1, w2, fr, aa = _problem(left, right, n_sites, dim, sigma, seed)
    ch = _chain(fl, w1, w2, fr, aa)
    ref = N.heff2(*(to_oracle(t) for t in (fl, w1, w2, fr, aa)))
    for rows in (False, True):
        got = emulate(ch, fl, fr, aa, rows=rows)
        assert rel_err(got, ref.flat()) < 1e-12


def test_row_table_addresses_every_row_of_aa_once():
    """The step-1 row table of the FP32 chain: each row of every aa block appears exactly
    once over the problems, in the entry gather's row order, and a row table shifted by one
    row in a single block is caught by the oracle comparison."""
    fl, w1, w2, fr, aa = _problem(9, 11, 20, 512, 4.0, 5)
    ch = _chain(fl, w1, w2, fr, aa)
    g = ch.gemm1_rows
    starts = np.sort(g._arows)
    L = aa.structure.layout
    want = np.sort(np.concatenate([int(L.offsets[i]) + int(L.shapes[i][3]) * np.arange(int(np.prod(L.shapes[i][:3])))
                                   for i in range(len(L.offsets))]))
    assert np.array_equal(starts, want)
    ref = N.heff2(*(to_oracle(t) for t in (fl, w1, w2, fr, aa))).flat()
    g._arows = g._arows.copy()
    g._arows[3] += int(L.shapes[0][3])  # a planted off-by-one-row
    assert rel_err(emulate(ch, fl, fr, aa, rows=True), ref) > 1e-6


def test_sector_chain_counts_the_same_work_as_the_block_chain():
    from tenet.network.heff import _FusedChain, _w_host

    fl, w1, w2, fr, aa = _problem(9, 11, 20, 512, 4.0, 5)
    sec = _chain(fl, w1, w2, fr, aa)
    blk = _FusedChain(aa.structure, fr.structure, fl.structure, _w_host(w1), _w_host(w2), "rys")
    assert sec.flops == blk.flops
    assert sec.gemm1.flops == blk.ex1.flops and sec.gemm4.flops == blk.gemm4.flops
    # one GEMM problem per charge sector instead of one per block pair
    assert len(sec.gemm1.problems) < len(blk.ex1.gemm.problems)
    assert len(sec.gemm4.problems) < len(blk.gemm4.problems)
    # step 3 is split into K slices and the exit pass reads every slice once
    assert sec.gemm4.ws_size > aa.structure.size
    assert sec.exit.moved_terms == sec.gemm4.ws_size
    assert sec.exit.elements == aa.structure.size


def test_planted_mpo_coefficient_error_is_caught():
    """The interpreter is not vacuous: one wrong MPO coefficient moves the result."""
    fl, w1, w2, fr, aa = _problem(5, 7, 12, 40, 1.5, 3)
    ch = _chain(fl, w1, w2, fr, aa)
    jobs, ins, outs, coef = ch.mpo_host
    coef = list(coef)
    k = next(i for i, c in enumerate(coef) if c != 0.0)
    coef[k] *= -1.0
    ch.mpo_host = (jobs, ins, outs, coef)
    got = emulate(ch, fl, fr, aa)
    ref = N.heff2(*(to_oracle(t) for t in (fl, w1, w2, fr, aa)))
    # far outside the 1e-12 agreement of the correct tables (the flipped slab is small)
    assert rel_err(got, ref.flat()) > 1e-9


def _region_cover(regions, size):
    cover = np.zeros(size, dtype=np.int64)
    for doff, shape, dstr, _, _ in regions:
        idx = np.full(1, doff, dtype=np.int64)
        for n, d in zip(shape, dstr):
            idx = (idx[:, None] + np.arange(n, dtype=np.int64)[None, :] * d).ravel()
        np.add.at(cover, idx, 1)
    return cover


@pytest.mark.parametrize("left,right,n_sites,dim,sigma,seed", [
    (5, 7, 12, 40, 1.5, 3), (2, 4, 12, 64, 3.0, 11), (9, 11, 20, 512, 4.0, 5)])
def test_every_intermediate_element_is_written_once(left, right, n_sites, dim, sigma, seed):
    """The stacked x, t3 and the output are allocated uninitialised: the entry gather, the
    MPO pass and the exit scatter must each write every element exactly once."""
    fl, w1, w2, fr, aa = _problem(left, right, n_sites, dim, sigma, seed)
    ch = _chain(fl, w1, w2, fr, aa)
    assert (_region_cover(ch.xs_regions, ch.xs_size) == 1).all()
    jobs, ins, outs, coef = ch.mpo_host
    t3 = np.zeros(ch.t3_size, dtype=np.int64)
    for R, S, _, _, ob, oe, _, _ in jobs:
        for off, rs in outs[ob:oe]:
            np.add.at(t3, (off + np.arange(R)[:, None] * rs + np.arange(S)[None, :]).ravel(), 1)
    assert (t3 == 1).all()
    h = ch.exit._tables.host
    from tenet.engine import JOB_WORDS
    ex = []
    for j in np.asarray(h["jobs"]).reshape(-1, JOB_WORDS):
        r = int(j[3])
        ex.append((int(j[0]), tuple(int(x) for x in j[4:4 + r]), tuple(int(x) for x in j[10:10 + r]), None, None))
    assert (_region_cover(ex, aa.structure.size) == 1).all()
