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


def test_null_rows_are_completed_on_demand():
    """A DMRG-like spectrum decaying into rounding noise, through the truncating callers'
    form: the sweeps stop once the vectors above the numerical-null floor (k eps |A|_F) are
    orthogonal and the length-L factor rows below it stay zero; every value matches LAPACK
    (the null ones to the floor), the live factor rows are orthonormal and reproduce A's
    live part, and the on-demand completion makes both factors isometries."""
    from tenet import linalg

    rng = np.random.default_rng(7)
    shapes = [(96, 100), (130, 128), (40, 44)]
    ml = _Layout(shapes)
    mats, flat = {}, np.empty(ml.size)
    for c, (m, n) in enumerate(shapes):
        k = min(m, n)
        u, _ = np.linalg.qr(rng.standard_normal((m, k)))
        v, _ = np.linalg.qr(rng.standard_normal((n, k)))
        a = (u * np.exp(-40.0 / k * np.arange(k))) @ v.T  # 1 down to ~1e-17
        mats[c] = a
        flat[ml.offsets[c] : ml.offsets[c] + m * n] = a.reshape(-1)
    buf = torch.from_numpy(flat).cuda()
    parts, complete = linalg._svd_run((buf, ml), False)
    for c, a in mats.items():
        u, s, vh = (t.cpu().numpy() for t in parts[c])
        ref = np.linalg.svd(a, compute_uv=False)
        assert np.abs(s - ref).max() <= 1e-13 * ref.max()
        floor = min(a.shape) * 2.220446049250313e-16 * np.linalg.norm(a)
        live = s > floor
        assert 0 < live.sum() < len(s)  # the spectrum crosses the floor
        w = vh if a.shape[0] <= a.shape[1] else u.T
        assert np.all(w[~live] == 0.0)
        assert np.abs(w[live] @ w[live].T - np.eye(live.sum())).max() < 1e-12
        assert np.abs(u[:, live] * s[live] @ vh[live] - a).max() <= 1e-12 * ref.max()
    complete()
    _check(mats, parts)


def test_truncating_form_keeps_the_values_of_the_exact_one():
    """Graded and rank-deficient sectors through both forms: the same singular values to
    FP64 rounding, and the live factor rows agree up to sign."""
    from tenet import linalg

    rng = np.random.default_rng(11)
    shapes = [(48, 48), (96, 80), (33, 120), (200, 190)]
    ml = _Layout(shapes)
    flat = np.empty(ml.size)
    for c, (m, n) in enumerate(shapes):
        k = min(m, n)
        u, _ = np.linalg.qr(rng.standard_normal((m, k)))
        v, _ = np.linalg.qr(rng.standard_normal((n, k)))
        s = 10.0 ** -np.linspace(0, 18, k)
        s[k // 2 :: 5] = 0.0
        flat[ml.offsets[c] : ml.offsets[c] + m * n] = ((u * s) @ v.T).reshape(-1)
    buf = torch.from_numpy(flat).cuda()
    exact = linalg._svd_parts((buf, ml), complete=True)
    trunc = linalg._svd_parts((buf, ml), complete=False)
    for c in exact:
        se, st = exact[c][1].cpu().numpy(), trunc[c][1].cpu().numpy()
        assert np.abs(se - st).max() <= 1e-13 * se.max()
        live = st > 1e-6 * st.max()  # well separated values: vectors fixed up to sign
        ue, ut = exact[c][0].cpu().numpy()[:, live], trunc[c][0].cpu().numpy()[:, live]
        assert np.abs(np.abs(np.sum(ue * ut, axis=0)) - 1.0).max() < 1e-9


@pytest.mark.parametrize("complete", [True, False])
def test_sweeps_queued_past_convergence_change_nothing(complete):
    """The first batch of queued Jacobi sweeps follows the previous SVD of the same launch
    shape; sweeps queued after the convergence flag is set must leave every factor
    bit-identical (a block tournament and single-workgroup problems in one call)."""
    from tenet import linalg

    rng = np.random.default_rng(5)
    shapes = [(130, 128), (40, 60), (96, 100)]
    ml = _Layout(shapes)
    flat = rng.standard_normal(ml.size)
    buf = torch.from_numpy(flat).cuda()
    saved = dict(linalg._SVD_SWEEP_HINT)
    try:
        linalg._SVD_SWEEP_HINT.clear()
        base = linalg._svd_parts((buf, ml), complete=complete)
        queued_small = linalg._SVD_LAST_SWEEPS
        for key in list(linalg._SVD_SWEEP_HINT):
            linalg._SVD_SWEEP_HINT[key] = 50
        long = linalg._svd_parts((buf, ml), complete=complete)
        assert linalg._SVD_LAST_SWEEPS >= 50 > queued_small
    finally:
        linalg._SVD_SWEEP_HINT.clear()
        linalg._SVD_SWEEP_HINT.update(saved)
    for c in base:
        for x, y in zip(base[c], long[c]):
            assert torch.equal(x, y), c
    if complete:
        _check({c: flat[ml.offsets[c]: ml.offsets[c] + m * n].reshape(m, n) for c, (m, n) in enumerate(shapes)},
               long)
