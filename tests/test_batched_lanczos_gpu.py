"""The batched Lanczos step (many XXZ H_eff problems, one batched matvec per iteration,
segmented FP64 Krylov scalars, one batched tridiagonal solve) against the single-problem
solver and the oracle's Lanczos; graph replays against eager runs; FP32 with the closing
FP64 refinement against FP64 at convergence."""

import numpy as np
import pytest
import torch

from oracle import u1_network as N
from tests.helpers import rel_err, to_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def insts():
    from tenet.network.batched import batch_workload

    out, dims, _ = batch_workload(count=12, d_min=48, d_max=256, seed=3, device="cuda")
    print(f"[batched-lanczos] D = {sorted(int(d) for d in dims)}", flush=True)
    return out


def test_segmented_kernels_match_torch(insts):
    from tenet.network.batched_lanczos import _Segments

    offs = np.array([0, 5, 5, 40000, 40001, 90000])  # an empty and a one-element segment
    S = _Segments(offs, torch.device("cuda"))
    g = torch.Generator(device="cuda").manual_seed(1)
    x = torch.randn(90000, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(90000, dtype=torch.float64, device="cuda", generator=g)
    out = torch.full((10,), 7.0, dtype=torch.float64, device="cuda")
    S.dot(x, y, out, stride=2)
    want = [float(torch.dot(x[a:b], y[a:b])) for a, b in zip(offs[:-1], offs[1:])]
    got = out.cpu().numpy()
    assert np.allclose(got[0::2], want, rtol=1e-13, atol=1e-13) and np.all(got[1::2] == 7.0)
    a = torch.tensor([2.0, 0.0, -1.0, 4.0, 0.5], dtype=torch.float64, device="cuda")
    z = y.clone()
    S.axpby(-1.0, a, 1, False, x, 1.0, None, 1, z)
    ref = y.clone()
    for s, (lo, hi) in enumerate(zip(offs[:-1], offs[1:])):
        ref[lo:hi] -= float(a[s]) * x[lo:hi]
    assert torch.equal(z, ref)
    z.fill_(float("nan"))  # b = 0: the old y is never read
    S.axpby(1.0, a, 1, True, x, 0.0, None, 1, z)  # inverted per-segment scalars
    for s, (lo, hi) in enumerate(zip(offs[:-1], offs[1:])):
        if hi > lo:
            assert torch.allclose(z[lo:hi], x[lo:hi] / float(a[s]), rtol=1e-15, atol=0)
    xn = x.clone()
    xn[5:40000] = float("nan")  # a zero factor drops its operand unread
    a0 = torch.tensor([1.0, 1.0, 0.0, 1.0, 1.0], dtype=torch.float64, device="cuda")
    z = y.clone()
    S.axpby(1.0, a0, 1, False, xn, 1.0, None, 1, z)
    assert torch.equal(z[5:40000], y[5:40000]) and torch.equal(z[:5], y[:5] + x[:5])


@pytest.mark.parametrize("layout", ["sector", "blocks"])
def test_batched_lanczos_matches_single_problem_solver(insts, layout):
    from tenet.network import Heff2, lanczos
    from tenet.network.batched import BatchedHeff2
    from tenet.network.batched_lanczos import BatchedLanczos

    b = BatchedHeff2(insts, "fp64", layout=layout)
    bl = BatchedLanczos(b, ncv=4)
    vals, vecs = bl.solve([i[4] for i in insts])
    for inst, e, v in zip(insts, vals, vecs):
        e1, v1 = lanczos(Heff2(*inst[:4], "fp64"), inst[4], ncv=4)
        assert abs(e - e1) <= 1e-12 * abs(e1)
        assert rel_err(v.data.cpu().numpy(), v1.data.cpu().numpy()) < 1e-9


def test_batched_lanczos_matches_the_oracle(insts):
    from tenet.network.batched import BatchedHeff2
    from tenet.network.batched_lanczos import BatchedLanczos

    b = BatchedHeff2(insts, "fp64")
    vals, vecs = BatchedLanczos(b, ncv=3).solve([i[4] for i in insts])
    for k in (0, 5, 11):
        fl, w1, w2, fr, aa = (to_oracle(t) for t in insts[k])
        eo, vo = N.lanczos(lambda x: N.heff2(fl, w1, w2, fr, x), aa, 3)
        assert abs(vals[k] - eo) <= 1e-10 * abs(eo)
        assert rel_err(to_oracle(vecs[k]).flat(), vo.flat()) < 1e-7


def test_graph_replays_equal_eager(insts):
    from tenet.network.batched import BatchedHeff2
    from tenet.network.batched_lanczos import BatchedLanczos

    b = BatchedHeff2(insts, "fp32")
    g = BatchedLanczos(b, ncv=4, graphs=True)
    e = BatchedLanczos(b, ncv=4, graphs=False)
    start = [i[4] for i in insts]
    ve, xe = e.solve(start)
    for _ in range(3):
        vg, xg = g.solve(start)
        torch.cuda.synchronize()
        assert vg == ve
        for a, c in zip(xg, xe):
            assert torch.equal(a.data, c.data)


def test_restarted_fp32_with_fp64_refinement_reaches_the_ground_energy(insts):
    """Restarted from its own Ritz vectors, the FP32 batch with the closing FP64
    refinement converges to each problem's ground energy (from ARPACK on the oracle's FP64
    operator) as the FP64 batch does."""
    from scipy.sparse.linalg import LinearOperator, eigsh

    from tenet.network.batched import BatchedHeff2
    from tenet.network.batched_lanczos import BatchedLanczos

    f64 = BatchedHeff2(insts, "fp64")
    ref = BatchedLanczos(f64, ncv=8)
    low = BatchedLanczos(BatchedHeff2(insts, "fp32"), ncv=8, refine=f64)
    v64 = v32 = [i[4] for i in insts]
    for _ in range(60):
        e64, v64 = ref.solve(v64)
        e32, v32 = low.solve(v32)
    sizes = [i[4].structure.size for i in insts]
    for k in np.argsort(sizes)[:3]:
        fl, w1, w2, fr, aa = (to_oracle(t) for t in insts[k])

        def mv(x, aa=aa, fl=fl, w1=w1, w2=w2, fr=fr):
            return N.heff2(fl, w1, w2, fr, type(aa).from_flat(aa.legs, np.asarray(x).ravel())).flat()

        op = LinearOperator((aa.flat().size,) * 2, matvec=mv, dtype=np.float64)
        e0 = float(eigsh(op, k=1, which="SA", tol=1e-14, v0=aa.flat())[0][0])
        assert e32[k] >= e0 - 1e-12 * abs(e0)  # a Rayleigh quotient of the exact operator
        assert abs(e32[k] - e0) <= 1e-9 * abs(e0), (k, e32[k], e0)
        assert abs(e64[k] - e0) <= 1e-9 * abs(e0), (k, e64[k], e0)
