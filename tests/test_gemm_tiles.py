"""Host tile tables of the grouped GEMM, on the CPU: under the kernel's tile -> output map
(tile (p, row block, column block) of a BM x BN launch writes rows [row*BM, +BM) and columns
[col*BN, +BN) of problem p, clipped to M x N) every output element of every problem -- the
K slices of a split problem included -- is written by exactly one tile, no tile of a real
problem is empty, and the XCD layout's padding tiles point at the all-zero problem row past
the real ones.  The persistent layout covers the same tiles."""

import numpy as np
import pytest
import torch

from tenet.engine import PROB_WORDS, GemmProgram


def _problems(rng, n, kmax):
    probs, a, b, c = [], 0, 0, 0
    for _ in range(n):
        M, N = (int(x) for x in rng.integers(1, 300, 2))
        K = int(rng.integers(1, kmax))
        probs.append((M, N, c, N, [(a, b, K, K, N)]))
        a += M * K
        b += K * N
        c += M * N
    return probs, a, b, c


def _check(g, tile, grid=0):
    t, ntiles = g._tables(tile, torch.device("cpu"), grid)
    probs = t["probs"].numpy().reshape(-1, PROB_WORDS)
    tiles = t["tiles"].numpy().reshape(-1, 4)
    assert len(tiles) == ntiles
    nreal = g.nproblems
    assert (probs[nreal] == 0).all()  # the padding row
    bm, bn = tile
    cover = [np.zeros((int(p[0]), int(p[1])), dtype=np.int64) for p in probs[:nreal]]
    real = 0
    for p, tr, tc, _ in tiles:
        if p == nreal:
            continue
        real += 1
        M, N = cover[p].shape
        assert tr * bm < M and tc * bn < N, "an empty tile of a real problem"
        cover[p][tr * bm:(tr + 1) * bm, tc * bn:(tc + 1) * bn] += 1
    assert real == sum(-(-M // bm) * -(-N // bn) for M, N in (c.shape for c in cover))
    for c in cover:
        assert (c == 1).all()


@pytest.mark.parametrize("tile", [(64, 64), (128, 64), (64, 128), (128, 128), (32, 32)])
def test_every_output_element_written_once(tile):
    rng = np.random.default_rng(sum(tile))
    probs, a, b, c = _problems(rng, 40, 900)
    g = GemmProgram(probs, a, b, c, tile=tile)
    _check(g, tile)


def test_split_k_slices_and_persistent_layout():
    rng = np.random.default_rng(5)
    probs, a, b, c = _problems(rng, 6, 4000)  # few tiles, long K: the launch splits K
    g = GemmProgram(probs, a, b, c, tile=(64, 64))
    assert g.nproblems > len(probs) and g.ws_size > 0
    _check(g, (64, 64))
    _check(g, (64, 64), grid=512)
