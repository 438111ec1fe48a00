"""GPU parity of tenet.einsum / tensordot / transpose / repartition against the oracle's
NumPy float64 restatement of the reference lowering.  Sector maps (block order, shapes,
offsets) must be identical; FP64 values agree to 1e-12 relative; FP32/BF16x3 GEMMs to the
tolerance written beside each case."""

import numpy as np
import pytest
import torch

from oracle import u1 as O
from tests.helpers import gaussian_degs, leg, oleg, rel_err, to_oracle

pytestmark = pytest.mark.gpu

CFG1 = gaussian_degs(range(-2, 3), 32, 1.0)  # {4, 19, 32, 19, 4}
SMALL = {-1: 2, 0: 3, 1: 2, 2: 1}


def _pair(d, sides, duals, seed):
    import tenet

    legs = tuple(leg(d, o, du) for o, du in zip(sides, duals))
    olegs = tuple(oleg(d, o, du) for o, du in zip(sides, duals))
    t = tenet.SymmetricTensor.random(legs, seed=seed, device="cuda")
    ot = O.OTensor.random(olegs, seed)
    return t, ot


EQUATIONS = [
    # (equation, a sides, a duals, b sides, b duals)
    ("abc,cde->abde", (1, 1, 0), (0, 0, 0), (1, 0, 0), (0, 0, 0)),  # bend-free layout
    ("abc,cde->abde", (1, 1, 1), (0, 0, 0), (1, 0, 0), (1, 0, 0)),  # README layout: a bend
    ("abc,cbd->ad", (1, 1, 0), (0, 0, 0), (1, 0, 0), (0, 0, 0)),
    ("abc,dca->bd", (1, 0, 0), (0, 0, 0), (1, 1, 0), (0, 0, 0)),
    ("abcd,dc->ab", (1, 1, 0, 0), (0, 0, 0, 0), (1, 1), (0, 0)),
    ("ab,bc->ca", (1, 0), (0, 0), (1, 0), (0, 0)),
]


@pytest.mark.parametrize("case", EQUATIONS, ids=[e[0] + str(i) for i, e in enumerate(EQUATIONS)])
def test_einsum_fp64_matches_oracle(case):
    import tenet

    eq, sa, da, sb, db = case
    a, oa = _pair(SMALL, sa, da, 1)
    b, ob = _pair(SMALL, sb, db, 2)
    out = tenet.einsum(eq, a, b)
    ref = O.einsum(eq, oa, ob)
    # sector map bit-exact: same legs, same block order
    got = to_oracle(out)
    assert [tuple(l.sectors) for l in got.legs] == [tuple(l.sectors) for l in ref.legs]
    assert [(l.out, l.dual) for l in got.legs] == [(l.out, l.dual) for l in ref.legs]
    assert got.keys() == ref.keys()
    assert rel_err(got.flat(), ref.flat()) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 5e-5)])
def test_cfg1_einsum(prec, tol):
    """abc,cde->abde at q in {-2..2}, D_q = {4,19,32,19,4}, A seed 0, B seed 1."""
    import tenet

    a, oa = _pair(CFG1, (1, 1, 0), (0, 0, 0), 0)
    b, ob = _pair(CFG1, (1, 0, 0), (0, 0, 0), 1)
    ref = O.einsum("abc,cde->abde", oa, ob)
    if prec != "fp64":
        tenet.set_gemm_precision(prec)
        a, b = a.to(dtype=torch.float32), b.to(dtype=torch.float32)
    out = tenet.einsum("abc,cde->abde", a, b)
    torch.cuda.synchronize()
    got = to_oracle(out)
    assert got.keys() == ref.keys()
    # 85 structurally allowed blocks; the 75 with |c| <= 2 are the ones a GEMM fills
    assert out.structure.num_blocks == 85
    nonzero = sum(1 for b in got.blocks.values() if np.any(b))
    assert nonzero == 75
    assert rel_err(got.flat(), ref.flat()) < tol
    tenet.set_gemm_precision("fp32")


def test_transpose_and_repartition_bit_exact():
    import tenet

    t, ot = _pair(SMALL, (1, 1, 0, 0), (0, 1, 0, 1), 5)
    for axes in [(2, 0, 3, 1), (3, 2, 1, 0), (0, 1, 2, 3), (1, 0, 2, 3)]:
        got = to_oracle(tenet.transpose(t, axes))
        ref = O.transpose(ot, axes)
        assert got.keys() == ref.keys()
        assert np.array_equal(got.flat(), ref.flat())
    for outs, ins in [((0,), (1, 2, 3)), ((0, 1, 2), (3,)), ((3, 1), (0, 2)), ((), (0, 1, 2, 3))]:
        got = to_oracle(tenet.repartition(t, outs, ins))
        ref = O.repartition(ot, outs, ins)
        assert got.keys() == ref.keys()
        assert np.array_equal(got.flat(), ref.flat())


def test_vector_space_ops():
    import tenet

    a, oa = _pair(CFG1, (1, 1, 0), (0, 0, 0), 3)
    b, ob = _pair(CFG1, (1, 1, 0), (0, 0, 0), 4)
    assert abs(float(tenet.inner(a, b)) - O.inner(oa, ob)) < 1e-9 * O.norm(oa) * O.norm(ob)
    assert abs(float(tenet.norm(a)) - O.norm(oa)) < 1e-12 * O.norm(oa)
    c = tenet.subtract(tenet.add(a, b * 2.5), a / 4.0)
    oc = O.axpy(-0.25, oa, O.axpy(2.5, ob, oa))
    assert rel_err(to_oracle(c).flat(), oc.flat()) < 1e-14
    d = tenet.divide(a, tenet.norm(a))
    assert abs(float(tenet.norm(d)) - 1.0) < 1e-14


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 5e-5)])
def test_split_k_long_contraction(prec, tol):
    """Few output tiles and a long K: the GEMM splits K into slices summed in a fixed order."""
    import tenet
    from tenet.engine import contraction_exec

    small, big = {0: 3, 1: 2, -1: 5}, {0: 700, 1: 500, -1: 300}
    import tenet as tn
    a = tn.SymmetricTensor.random((leg(small, 1), leg(big, 0)), seed=21, device="cuda")
    b = tn.SymmetricTensor.random((leg(big, 1), leg(small, 0)), seed=22, device="cuda")
    oa = O.OTensor.random((oleg(small, 1), oleg(big, 0)), 21)
    ob = O.OTensor.random((oleg(big, 1), oleg(small, 0)), 22)
    ex = contraction_exec(a.structure, b.structure, ((1,), (0,)), (0, 1))
    assert ex.gemm.reduce is not None and ex.gemm.ws_size > 0  # the split path is the one tested
    ref = O.einsum("ab,bc->ac", oa, ob)
    if prec != "fp64":
        tenet.set_gemm_precision(prec)
        a, b = a.to(dtype=torch.float32), b.to(dtype=torch.float32)
    out = tenet.einsum("ab,bc->ac", a, b)
    tenet.set_gemm_precision("fp32")
    assert rel_err(to_oracle(out).flat(), ref.flat()) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
@pytest.mark.parametrize("shape", ["rows", "cols"])
def test_skinny_contractions(prec, tol, shape):
    """Outputs at most 8 wide in one dimension with a short K run as one streaming launch
    (one thread per output line) instead of MFMA tiles."""
    import tenet
    from tenet.engine import contraction_exec

    small, big = {0: 2, 1: 1, -1: 3}, {0: 900, 1: 700, -1: 500}
    if shape == "rows":
        la, lb = (big, small), (small, small)
    else:
        la, lb = (small, small), (small, big)
    a = tenet.SymmetricTensor.random((leg(la[0], 1), leg(la[1], 0)), seed=31, device="cuda")
    b = tenet.SymmetricTensor.random((leg(lb[0], 1), leg(lb[1], 0)), seed=32, device="cuda")
    oa = O.OTensor.random((oleg(la[0], 1), oleg(la[1], 0)), 31)
    ob = O.OTensor.random((oleg(lb[0], 1), oleg(lb[1], 0)), 32)
    ex = contraction_exec(a.structure, b.structure, ((1,), (0,)), (0, 1))
    assert ex.gemm.skinny == (1 if shape == "rows" else 0)
    ref = O.einsum("ab,bc->ac", oa, ob)
    if prec != "fp64":
        a, b = a.to(dtype=torch.float32), b.to(dtype=torch.float32)
    out = tenet.einsum("ab,bc->ac", a, b)
    assert rel_err(to_oracle(out).flat(), ref.flat()) < tol
