"""The batched one-sided Jacobi SVD against LAPACK on the host: singular values to FP64
rounding, exact reconstruction, orthonormal factors, for square, tall, wide, rank-deficient
and multi-block sector matrices in one batch."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


class _Layout:
    """A packed set of sector matrices, shaped like the lowering's matrix layout."""

    def __init__(self, shapes):
        self.sectors = list(range(len(shapes)))
        self.shapes = dict(enumerate(shapes))
        self.offsets, acc = {}, 0
        for c, (m, n) in enumerate(shapes):
            self.offsets[c] = acc
            acc += m * n
        self.size = acc


def _check(mats, parts, tol=1e-13):
    for c, a in mats.items():
        u, s, vh = (t.cpu().numpy() for t in parts[c])
        if a.size == 0:
            continue
        ref = np.linalg.svd(a, compute_uv=False)
        scale = max(ref.max(), 1.0 if not ref.any() else 1e-300)
        assert np.abs(s - ref).max() <= tol * scale, c
        assert np.all(np.diff(s) <= 0)
        assert np.abs(u @ np.diag(s) @ vh - a).max() <= 10 * tol * scale
        k = len(s)
        # both factors are isometries on every column, zero singular values included
        # (the reference asserts per-sector isometry of U and Vh, tests/ops/test_linalg.py:199)
        assert np.abs(u.T @ u - np.eye(k)).max() < 1e-12
        assert np.abs(vh @ vh.T - np.eye(k)).max() < 1e-12
        assert u.shape == (a.shape[0], k) and vh.shape == (k, a.shape[1])


def _run(shapes, seed=0, make=None):
    from tenet import linalg

    rng = np.random.default_rng(seed)
    ml = _Layout(shapes)
    mats = {}
    flat = np.empty(ml.size)
    for c, (m, n) in enumerate(shapes):
        a = make(rng, m, n) if make else rng.standard_normal((m, n))
        mats[c] = a
        flat[ml.offsets[c] : ml.offsets[c] + m * n] = a.reshape(-1)
    buf = torch.from_numpy(flat).cuda()
    return mats, linalg._svd_parts((buf, ml))


def test_jacobi_svd_mixed_batch():
    # up to (90, 90), (101, 99): one workgroup holds the whole problem (odd k: one zero
    # vector); (110, 110), (130, 200), (300, 280): block tournaments
    shapes = [(1, 1), (5, 7), (7, 5), (40, 9), (9, 40), (64, 64), (70, 66), (65, 67), (77, 75), (33, 33),
              (90, 90), (101, 99), (110, 110), (130, 200), (300, 280), (3, 0)]
    mats, parts = _run(shapes)
    _check(mats, parts)


def test_jacobi_svd_graded_and_rank_deficient():
    def make(rng, m, n):
        k = min(m, n)
        u, _ = np.linalg.qr(rng.standard_normal((m, k)))
        v, _ = np.linalg.qr(rng.standard_normal((n, k)))
        s = 10.0 ** -np.linspace(0, 14, k)
        s[k // 2 :: 3] = 0.0
        return (u * s) @ v.T

    mats, parts = _run([(48, 48), (96, 80), (33, 120), (540, 540)], seed=3, make=make)
    _check(mats, parts)


def test_jacobi_svd_exact_zero_singular_values():
    """Exact zeros (zero rows or columns, an all-zero sector) leave W rows that the kernel
    must complete to an orthonormal set; tall and wide shapes exercise both factor sides."""
    def make(rng, m, n):
        a = rng.standard_normal((m, n))
        if m == n == 6:
            return np.zeros((m, n))
        if m <= n:
            a[1::3, :] = 0.0
        else:
            a[:, ::2] = 0.0
        return a

    mats, parts = _run([(6, 6), (9, 14), (14, 9), (40, 70), (70, 40), (5, 5)], seed=4, make=make)
    _check(mats, parts)
    assert float(parts[0][1].abs().max()) == 0.0


def test_jacobi_svd_two_calls_share_nothing():
    a1, p1 = _run([(20, 30), (100, 90)], seed=5)
    a2, p2 = _run([(20, 30), (100, 90)], seed=5)
    for c in a1:
        assert torch.equal(p1[c][1], p2[c][1])
