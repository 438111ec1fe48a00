"""Parity at the benchmarked sizes.

* The D = 2048 U(1) Heisenberg two-site H_eff (L = 64, centre bond, 22 odd-charge sectors of
  1...269, MPO bond 5): the FP64 engine against the oracle's NumPy FP64 ``heff2`` to 1e-12
  relative; FP32 (2e-6) and split-BF16x3 (5e-5) against FP64; the operator is symmetric.
* A batch of XXZ H_eff problems (Delta ~ U[0, 2], D log-uniform in [64, 1024]) run by
  ``BatchedHeff2`` in both layouts, every instance against the oracle, with the merged
  step-3 GEMM splitting its K across slices; and the full 256-problem batch through its
  size-independent properties (symmetric operator, batch = per-instance operators,
  a sample against the oracle).
* An SU(2) two-site H_eff whose bonds carry spins 2j = 11, 13 (inner fusion lines reach
  2j = 14, past the reference's F fixture) against the dense Clebsch-Gordan contraction.
"""

import numpy as np
import pytest
import torch

from oracle import u1_network as N
from tests.helpers import rel_err, to_oracle

pytestmark = pytest.mark.gpu


def _oracle_heff(inst):
    fl, w1, w2, fr, aa = (to_oracle(t) for t in inst)
    return N.heff2(fl, w1, w2, fr, aa).flat()


@pytest.fixture(scope="module")
def cfg2():
    from tenet.network.workloads import heff2_problem

    inst = heff2_problem(64, 2048, 6.0, seed=0, device="cuda")
    print(f"[fullsize] cfg2 aa {inst[4].structure.size} elements", flush=True)
    return inst, _oracle_heff(inst)


def test_cfg2_fp64_matches_the_oracle(cfg2):
    from tenet.network import Heff2

    inst, ref = cfg2
    aa = inst[4]
    assert aa.structure.size == 1554890
    got = Heff2(*inst[:4], "fp64")(aa)
    assert got.structure == aa.structure
    assert rel_err(got.data.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 5e-5)])
def test_cfg2_reduced_precision_against_fp64(cfg2, prec, tol):
    from tenet.network import Heff2

    inst, ref = cfg2
    got = Heff2(*inst[:4], prec)(inst[4]).data.cpu().numpy()
    assert rel_err(got, ref) < tol


def test_cfg2_operator_is_symmetric(cfg2):
    import tenet
    from tenet.network import Heff2

    inst, _ = cfg2
    op = Heff2(*inst[:4], "fp64")
    x = tenet.SymmetricTensor.random(inst[4].legs, seed=3, device="cuda")
    y = tenet.SymmetricTensor.random(inst[4].legs, seed=4, device="cuda")
    a, b = float(tenet.inner(x, op(y))), float(tenet.inner(op(x), y))
    assert abs(a - b) <= 1e-12 * max(abs(a), 1e-300)


# --------------------------------------------------------------------------- batched


@pytest.fixture(scope="module")
def batch16():
    from tenet.network.batched import batch_workload

    insts, dims, _ = batch_workload(count=16, seed=7, device="cuda")
    print(f"[fullsize] batch of 16, D = {sorted(int(d) for d in dims)}", flush=True)
    return insts, dims, [_oracle_heff(i) for i in insts]


@pytest.mark.parametrize("layout", ["sector", "blocks"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 5e-5)])
def test_batched_instances_match_the_oracle(batch16, layout, prec, tol):
    from tenet.network.batched import BatchedHeff2

    insts, dims, refs = batch16
    assert dims.min() < 128 and dims.max() > 700
    b = BatchedHeff2(insts, prec, layout=layout)
    if layout == "sector":
        assert b.gemm4.ws_size > b.x_size  # step 3 runs in K slices
    out = b.unpack(b.apply(b.pack([i[4] for i in insts])))
    for inst, got, ref in zip(insts, out, refs):
        assert got.structure == inst[4].structure
        assert rel_err(got.data.double().cpu().numpy(), ref) < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_batched_blocks_layout_split_k_matches_the_oracle(prec, tol):
    """Few instances leave the per-block-pair GEMMs too few tiles: both launches split K
    and sum the slices in a reduction pass."""
    from tenet.network.batched import BatchedHeff2, synthetic_operator

    insts = [synthetic_operator(n, d, sg, dl, s, device="cuda")
             for n, d, sg, dl, s in ((24, 600, 4.0, 1.3, 2), (16, 90, 3.0, 0.4, 5))]
    b = BatchedHeff2(insts, prec, layout="blocks")
    assert b.gemm4.reduce is not None
    out = b.unpack(b.apply(b.pack([i[4] for i in insts])))
    for inst, got in zip(insts, out):
        assert rel_err(got.data.double().cpu().numpy(), _oracle_heff(inst)) < tol


def test_full_256_batch_properties():
    import tenet
    from tenet.network.batched import BatchedHeff2, batch_workload

    insts, dims, _ = batch_workload(count=256, seed=0, device="cuda")
    print(f"[fullsize] batch of 256 built, D {int(dims.min())}..{int(dims.max())}", flush=True)
    b = BatchedHeff2(insts, "fp64")
    xs = [tenet.SymmetricTensor.random(i[4].legs, seed=100 + k, device="cuda") for k, i in enumerate(insts)]
    ys = [tenet.SymmetricTensor.random(i[4].legs, seed=900 + k, device="cuda") for k, i in enumerate(insts)]
    hx = b.unpack(b.apply(b.pack(xs)))
    hy = b.unpack(b.apply(b.pack(ys)))
    for k in range(len(insts)):
        a, c = float(tenet.inner(xs[k], hy[k])), float(tenet.inner(hx[k], ys[k]))
        assert abs(a - c) <= 1e-11 * max(abs(a), 1e-300), k
    # every instance equals its own operator, a sample equals the oracle
    for k in range(len(insts)):
        ref = b.ops[k](xs[k])
        assert rel_err(hx[k].data.cpu().numpy(), ref.data.cpu().numpy()) < 1e-13, k
    for k in (0, 37, 101, 255):
        want = _oracle_heff((*insts[k][:4], xs[k]))
        assert rel_err(hx[k].data.cpu().numpy(), want) < 1e-12, k
    f32 = BatchedHeff2(insts, "fp32")
    out32 = f32.apply(f32.pack(xs))
    assert rel_err(out32.double().cpu().numpy(), b.apply(b.pack(xs)).cpu().numpy()) < 2e-6


# --------------------------------------------------------------------------- SU(2)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_su2_heff2_at_high_spin_matches_dense(precision, tol):
    import tenet
    from tenet import IN, OUT, GradedSpace, Leg
    from tenet.network import Heff2
    from tenet.network.models import SU2_MPO_BOND, SU2_PHYS, su2_heisenberg_mpo
    from tenet.symmetry import SU2, SU2Sector
    from tests.test_su2_gpu import dense

    BL = GradedSpace.new(SU2, {SU2Sector(11): 1, SU2Sector(13): 2})
    BR = GradedSpace.new(SU2, {SU2Sector(11): 1, SU2Sector(13): 1})
    fl = tenet.SymmetricTensor.random((Leg(BL, IN), Leg(SU2_MPO_BOND, OUT), Leg(BL, OUT)), seed=1, device="cuda")
    fr = tenet.SymmetricTensor.random((Leg(BR, OUT), Leg(SU2_MPO_BOND, IN), Leg(BR, IN)), seed=2, device="cuda")
    aa = tenet.SymmetricTensor.random((Leg(BL, OUT), Leg(SU2_PHYS, OUT), Leg(SU2_PHYS, OUT), Leg(BR, IN)),
                                      seed=3, device="cuda")
    h = su2_heisenberg_mpo(4, device="cuda")
    w1, w2 = h[1], h[2]
    spins = {x.two_j for l in aa.legs for x, _ in l.space.sectors}
    assert max(spins) == 13
    want = np.einsum("axB,apqr,xPpm,mQqy,rys->BPQs", dense(fl), dense(aa), dense(w1), dense(w2), dense(fr),
                     optimize=True)
    got = dense(Heff2(fl, w1, w2, fr, precision)(aa))
    assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)
