"""The device tridiagonal eigensolver of the Lanczos step, and the FP64 residual
refinement.

``tn_tridiag_ground`` is compiled for the host too (``tn_tridiag_ground_host``): the CPU
tests check that routine against LAPACK (``numpy.linalg.eigh``) on random, graded,
clustered and broken-down tridiagonals; the GPU tests check that the device launch returns
the host routine's answer and that the Lanczos step built on it keeps its guarantees."""

import ctypes

import numpy as np
import pytest


def _lib():
    from tenet import _native

    return _native.load_library()


def ground_host(a, b, tol=0.0):
    n = len(a)
    val = ctypes.c_double(0.0)
    coef = np.zeros(n)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b if len(b) else np.zeros(1), dtype=np.float64)
    rc = _lib().tn_tridiag_ground_host(a.ctypes.data, b.ctypes.data, n, tol, ctypes.byref(val), coef.ctypes.data)
    assert rc == 0
    return val.value, coef


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 17, 40, 64])
@pytest.mark.parametrize("scale", [1e-6, 1.0, 1e4])
def test_ground_pair_matches_lapack(n, scale):
    rng = np.random.default_rng(n)
    for _ in range(20):
        a = rng.standard_normal(n) * scale
        b = rng.standard_normal(n - 1) * scale
        t = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        w, v = np.linalg.eigh(t)
        val, x = ground_host(a, b)
        span = max(abs(w).max(), 1e-300)
        assert abs(val - w[0]) <= 4e-15 * span
        if n > 1 and w[1] - w[0] > 1e-8 * span:  # a simple eigenvalue: the vector is fixed
            assert 1.0 - abs(x @ v[:, 0]) < 1e-12
        assert abs(np.linalg.norm(x) - 1.0) < 1e-14
        assert x[np.flatnonzero(x)[0]] > 0


def test_happy_breakdown_truncates_the_matrix():
    a = np.array([1.0, 2.0, 3.0, -7.0])
    b = np.array([0.5, 1e-15, 0.3])
    val, x = ground_host(a, b, tol=1e-13)
    w, v = np.linalg.eigh(np.array([[1.0, 0.5], [0.5, 2.0]]))
    assert abs(val - w[0]) < 1e-15
    np.testing.assert_allclose(np.abs(x[:2]), np.abs(v[:, 0]), atol=1e-14)
    assert x[2] == 0.0 and x[3] == 0.0  # the -7 past the breakdown is not reachable


def test_lanczos_tridiagonal_of_a_real_operator():
    """The recurrence's own matrix: Lanczos on a dense symmetric operator, the tridiagonal
    ground value equal to LAPACK's on the same tridiagonal, and converging to the operator's
    ground value as the Krylov space grows."""
    rng = np.random.default_rng(3)
    h = rng.standard_normal((200, 200))
    h = h + h.T
    v = rng.standard_normal(200)
    v /= np.linalg.norm(v)
    vs, al, be = [v], [], []
    for j in range(60):
        w = h @ vs[j]
        al.append(vs[j] @ w)
        w -= al[-1] * vs[j] + (be[-1] * vs[j - 1] if j else 0)
        for u in vs:
            w -= (u @ w) * u
        be.append(np.linalg.norm(w))
        vs.append(w / be[-1])
    val, x = ground_host(np.array(al), np.array(be[:-1]))
    t = np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1)
    assert abs(val - np.linalg.eigh(t)[0][0]) < 1e-13 * abs(val)
    assert abs(val - np.linalg.eigh(h)[0][0]) < 1e-8 * abs(val)


def test_refuses_sizes_outside_the_table():
    from tenet._native import TRIDIAG_MAX

    val = ctypes.c_double(0.0)
    buf = np.zeros(TRIDIAG_MAX + 1)
    assert _lib().tn_tridiag_ground_host(buf.ctypes.data, buf.ctypes.data, TRIDIAG_MAX + 1, 0.0,
                                         ctypes.byref(val), buf.ctypes.data) != 0


@pytest.mark.gpu
def test_device_launch_returns_the_host_routine_answer():
    import torch

    from tenet import _native

    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 8, 33, 64):
        a = rng.standard_normal(n)
        b = rng.standard_normal(n)
        b[n // 2] = 1e-16 if n > 2 else b[n // 2]
        want_v, want_x = ground_host(a, b[: n - 1], tol=1e-13)
        da = torch.tensor(a, device="cuda")
        db = torch.tensor(b, device="cuda")
        val = torch.zeros((), dtype=torch.float64, device="cuda")
        coef = torch.zeros(n, dtype=torch.float64, device="cuda")
        _native.check(_native.lib().tn_tridiag_ground(da.data_ptr(), db.data_ptr(), n, 1e-13, val.data_ptr(),
                                                      coef.data_ptr(), _native.stream_ptr()))
        torch.cuda.synchronize()
        assert abs(float(val) - want_v) <= 1e-14 * max(1.0, abs(want_v))
        np.testing.assert_allclose(coef.cpu().numpy(), want_x, atol=1e-13)


@pytest.mark.gpu
def test_residual_refinement_of_an_fp32_solve():
    """FP32 Krylov plus the FP64 residual step lands on the FP64 solve's energy, never above
    the Rayleigh quotient it started from, and leaves a smaller FP64 residual."""
    from tenet import ops
    from tenet.network import Heff2, lanczos
    from tenet.network.dmrg import residual_refine
    from tenet.network.workloads import heff2_problem

    fl, w1, w2, fr, aa = heff2_problem(12, 40, 2.0, seed=6, device="cuda")
    op64, op32 = Heff2(fl, w1, w2, fr, "fp64"), Heff2(fl, w1, w2, fr, "fp32")
    v, e64 = aa, None
    for _ in range(40):
        e64, v = lanczos(op64, v, ncv=8)
    u = aa
    for _ in range(40):
        _, u = lanczos(op32, u, ncv=8)
    rq = float(ops.inner(u, op64(u)))
    e_ref, u_ref, rho = residual_refine(op64, u)
    assert float(e_ref) <= rq + 1e-15 * abs(rq)
    r2 = ops.subtract(op64(u_ref), ops.multiply(u_ref, float(e_ref)))
    assert float(ops.norm(r2)) < float(rho)
    assert abs(float(e_ref) - e64) < 1e-11 * abs(e64)
