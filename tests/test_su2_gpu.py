"""SU(2) and fermion-parity contractions on the GPU against the dense oracle.

Block values that pass through multi-term fusion-tree recoupling (SU(2) F/R/B moves,
Frobenius-Schur signs on dual legs) or through Koszul signs (fZ2) are checked the way the
reference checks them (``tests/ops/test_einsum.py:106-131``,
``tests/ops/test_contraction.py:198-232``, ``tests/symmetry/test_fz2.py:304-437``): the
oracle's dense expansion of the GPU result against ``np.tensordot`` of the dense
expansions of the operands (graded with the per-side Koszul sign for fZ2).  The oracle's
dense expansion (``oracle/su2.py``) is pinned to sympy's Clebsch-Gordan values and to the
reference's F fixture on the CPU.
"""

import numpy as np
import pytest
import torch

from oracle import su2 as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _spaces():
    from tenet import GradedSpace
    from tenet.symmetry import SU2, FZ2Sector, SU2Sector, fZ2

    su2 = {k: GradedSpace.new(SU2, {SU2Sector(a): m for a, m in v})
           for k, v in {"V": ((0, 2), (1, 2), (2, 1)), "W": ((0, 1), (1, 2)), "U": ((1, 1), (2, 2))}.items()}
    fz2 = {k: GradedSpace.new(fZ2, {FZ2Sector(a): m for a, m in v})
           for k, v in {"P": ((0, 2), (1, 3)), "Q": ((0, 1), (1, 2)), "R": ((0, 2), (1, 1)),
                        "S": ((0, 1), (1, 1))}.items()}
    return su2, fz2


def _rules(t):
    return O.SU2Rules if t.provider.name == "SU2" else O.FZ2Rules


def _label(a):
    return a.two_j if hasattr(a, "two_j") else a.parity


def _olegs(legs):
    from tenet import OUT

    return [O.SLeg([(_label(a), m) for a, m in l.space.sectors], l.side is OUT, l.dual) for l in legs]


def dense(t):
    """The oracle's dense expansion of a product tensor, after checking that its block
    order is the oracle's (bit-exact sector map)."""
    rules = _rules(t)
    ol = _olegs(t.legs)
    keys = O.block_keys(rules, ol)
    got = [tuple(_label(x) for x in t.structure.axis_sectors(k)) for k in t.structure.block_order]
    assert got == [k[2] for k in keys]
    blocks = [np.asarray(b, dtype=np.float64) for b in t.numpy_blocks()]
    return O.to_dense(rules, ol, blocks)


def _pair(a_legs, b_legs, seeds=(0, 1)):
    from tenet import SymmetricTensor

    return (SymmetricTensor.random(a_legs, seed=seeds[0], device="cuda"),
            SymmetricTensor.random(b_legs, seed=seeds[1], device="cuda"))


def _layout(x, y, z, w):
    """Operand legs in the bend-free layout of the reference's einsum tests."""
    from tenet import IN, OUT, Leg

    return ((Leg(x, OUT), Leg(y, OUT), Leg(z, IN)), (Leg(z, OUT), Leg(w, IN), Leg(x, IN)),
            (Leg(x, OUT), Leg(y, OUT)), (Leg(z, IN), Leg(w, IN)))


def _dense_einsum(a, b, equation, graded):
    """``np.tensordot`` of the dense expansions, with each operand graded-transposed into
    (free, contracted) / (contracted, free) order and the result into the output order."""
    lhs, out = equation.split("->")
    ta, tb = lhs.split(",")
    shared = [c for c in ta if c in tb]
    fa = [c for c in ta if c not in shared]
    fb = [c for c in tb if c not in shared]
    pa = tuple(ta.index(c) for c in fa + shared)
    pb = tuple(tb.index(c) for c in shared + fb)
    da, db = np.transpose(dense(a), pa), np.transpose(dense(b), pb)
    if graded:
        da = O.supersign(_olegs(a.legs), pa) * da
        db = O.supersign(_olegs(b.legs), pb) * db
    c = np.tensordot(da, db, axes=len(shared))
    free = [a.legs[ta.index(x)] for x in fa] + [b.legs[tb.index(x)] for x in fb]
    perm = tuple((fa + fb).index(x) for x in out)
    got = np.transpose(c, perm)
    return O.supersign(_olegs(free), perm) * got if graded else got


EQUATIONS = ["abc,cde->abde", "abc,cde->bade", "abc,cde->adbe", "ab,cd->abcd"]


@pytest.mark.parametrize("provider", ["su2", "fz2"])
@pytest.mark.parametrize("equation", EQUATIONS)
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_einsum_matches_dense_oracle(provider, equation, dtype, tol):
    import tenet

    su2, fz2 = _spaces()
    lay = _layout(su2["V"], su2["W"], su2["U"], su2["W"]) if provider == "su2" else \
        _layout(fz2["P"], fz2["Q"], fz2["R"], fz2["S"])
    legs = (lay[0], lay[1]) if len(equation.split(",")[0]) == 3 else (lay[2], lay[3])
    a, b = _pair(*legs)
    want = _dense_einsum(a, b, equation, provider == "fz2")
    c = tenet.einsum(equation, a.to(dtype=dtype), b.to(dtype=dtype))
    got = dense(c.to(dtype=torch.float64))
    err = np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300)
    assert err < tol, f"{provider} {equation}: rel err {err:.3e}"


def _su2_cases():
    from tenet import IN, OUT, Leg

    su2, _ = _spaces()
    V, W, U = su2["V"], su2["W"], su2["U"]
    return {
        "within-side-transpose": ((Leg(V, OUT), Leg(W, IN), Leg(U, IN)),
                                  (Leg(U, OUT), Leg(W, OUT), Leg(V, IN)), ((2, 1), (0, 1))),
        "bends-on-both-operands": ((Leg(V, OUT), Leg(W, OUT), Leg(U, IN)),
                                   (Leg(W, IN), Leg(U, OUT), Leg(V, IN)), ((1,), (0,))),
        "out-against-out-dual": ((Leg(V, OUT), Leg(W, OUT)),
                                 (Leg(W, OUT, True), Leg(U, IN)), ((1,), (0,))),
    }


@pytest.mark.parametrize("case", ["within-side-transpose", "bends-on-both-operands", "out-against-out-dual"])
def test_su2_tensordot_cases(case):
    import tenet

    a_legs, b_legs, axes = _su2_cases()[case]
    a, b = _pair(a_legs, b_legs)
    c = tenet.tensordot(a, b, axes)
    want = np.tensordot(dense(a), dense(b), axes)
    got = dense(c)
    assert np.linalg.norm(got - want) <= TOL * np.linalg.norm(want)


def test_fz2_tensordot_with_a_nontrivial_supersign():
    """Contracted axes paired in reversed order: the lowering transposes two IN lines past
    each other, a genuine Koszul sign (reference ``test_contraction.py:235-254``)."""
    import tenet
    from tenet import IN, OUT, Leg

    _, fz2 = _spaces()
    a_legs = (Leg(fz2["P"], OUT), Leg(fz2["Q"], IN), Leg(fz2["R"], IN))
    b_legs = (Leg(fz2["R"], OUT), Leg(fz2["Q"], OUT), Leg(fz2["S"], IN))
    a, b = _pair(a_legs, b_legs, seeds=(3, 4))
    perm = (0, 2, 1)
    sign = O.supersign(_olegs(a_legs), perm)
    assert np.any(sign != 1.0)
    want = np.tensordot(sign * np.transpose(dense(a), perm), dense(b), axes=((1, 2), (0, 1)))
    got = dense(tenet.tensordot(a, b, ((2, 1), (0, 1))))
    assert np.linalg.norm(got - want) <= TOL * np.linalg.norm(want)
    plain = np.tensordot(dense(a), dense(b), ((2, 1), (0, 1)))
    assert not np.allclose(got, plain, atol=1e-8)


ALL_PERMS = [(0, 1, 2, 3), (2, 1, 0, 3), (1, 0, 3, 2), (3, 1, 2, 0), (0, 3, 2, 1), (2, 3, 0, 1), (3, 2, 1, 0)]


@pytest.mark.parametrize("p", ALL_PERMS)
def test_su2_transpose_and_repartition_match_dense(p):
    from tenet import IN, OUT, Leg, SymmetricTensor, ops

    su2, _ = _spaces()
    legs = (Leg(su2["V"], OUT), Leg(su2["W"], OUT, True), Leg(su2["U"], IN), Leg(su2["W"], IN))
    t = SymmetricTensor.random(legs, seed=9, device="cuda")
    d = dense(t)
    got = dense(ops.transpose(t, p))
    assert np.linalg.norm(got - np.transpose(d, p)) <= TOL * np.linalg.norm(d)
    for outs, ins in (((0, 1, 2), (3,)), ((0,), (1, 2, 3)), ((3, 0), (2, 1)), ((), (0, 1, 2, 3))):
        r = ops.repartition(t, outs, ins)
        got = dense(r)
        assert np.linalg.norm(got - np.transpose(d, (*outs, *ins))) <= TOL * np.linalg.norm(d)


@pytest.mark.parametrize("p", ALL_PERMS)
def test_fz2_transpose_carries_the_koszul_sign(p):
    from tenet import IN, OUT, Leg, SymmetricTensor, ops

    _, fz2 = _spaces()
    legs = (Leg(fz2["P"], OUT), Leg(fz2["Q"], IN), Leg(fz2["R"], OUT), Leg(fz2["S"], IN))
    t = SymmetricTensor.random(legs, seed=17, device="cuda")
    want = O.supersign(_olegs(legs), p) * np.transpose(dense(t), p)
    got = dense(ops.transpose(t, p))
    np.testing.assert_allclose(got, want, atol=1e-14)
    back = ops.transpose(ops.transpose(t, p), tuple(np.argsort(p)))
    assert torch.equal(back.data, t.data)


def test_su2_inner_and_norm_are_qdim_weighted():
    from tenet import IN, OUT, Leg, SymmetricTensor, ops

    su2, _ = _spaces()
    legs = (Leg(su2["V"], OUT), Leg(su2["U"], OUT), Leg(su2["W"], IN))
    a = SymmetricTensor.random(legs, seed=1, device="cuda")
    b = SymmetricTensor.random(legs, seed=2, device="cuda")
    da, db = dense(a), dense(b)
    assert abs(float(ops.norm(a)) - np.linalg.norm(da)) < 1e-12 * np.linalg.norm(da)
    assert abs(float(ops.inner(a, b)) - float(np.sum(da * db))) < 1e-12 * np.linalg.norm(da) * np.linalg.norm(db)


# ------------------------------------------------------------------ SU(2) network path


def _su2_heff_problem(n_sites, mult, seed=0):
    from tenet import ops
    from tenet.network.env import Env
    from tenet.network.models import su2_heisenberg_mpo, su2_random_chain

    psi = su2_random_chain(n_sites, mult, 3.0, seed=seed, device="cuda")
    psi.canonize_()
    h = su2_heisenberg_mpo(n_sites, device="cuda")
    env = Env(psi, h)
    c = n_sites // 2 - 1
    for n in range(n_sites - 1, c + 1, -1):
        env.update_(n, to="first")
    for n in range(0, c):
        env.update_(n, to="last")
    aa = ops.einsum("apx,xqr->apqr", psi[c], psi[c + 1])
    return env, c, aa, h


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 5e-5)])
def test_su2_heff2_matches_dense_oracle(precision, tol):
    """The compiled SU(2) H_eff (generic path: four contractions with their bends) against
    the dense contraction of the dense environments, MPO tensors and two-site tensor."""
    env, c, aa, h = _su2_heff_problem(10, 6)
    fl, fr = env.F[c - 1, c], env.F[c + 2, c + 1]
    want = np.einsum("axB,apqr,xPpm,mQqy,rys->BPQs", dense(fl), dense(aa), dense(h[c]),
                     dense(h[c + 1]), dense(fr), optimize=True)
    got_t = env.operator(c, precision)(aa)
    got = dense(got_t)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < tol, f"{precision}: {err:.3e}"
    ref = dense(env.heff2(c, aa))
    assert np.linalg.norm(ref - want) <= 1e-12 * np.linalg.norm(want)


def test_su2_env_update_matches_dense_oracle():
    env, c, aa, h = _su2_heff_problem(8, 5, seed=3)
    psi = env.psi
    n = c
    want = np.einsum("axB,apr,xPpm,BPs->rms", dense(env.F[n - 1, n]), dense(psi[n]), dense(h[n]),
                     dense(psi[n]), optimize=True)
    env.update_(n, to="last")
    got = dense(env.F[n, n + 1])
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-8)])
def test_su2_dmrg_heisenberg_12_matches_ed(precision, tol):
    """SU(2) two-site DMRG reaches the reference's exact-diagonalisation literal
    ``E_OBC_12 = -5.142090632840532`` (``tests/integration/test_dmrg.py:82``)."""
    from tenet.network import dmrg_
    from tenet.network.models import su2_heisenberg_mpo, su2_random_chain

    psi = su2_random_chain(12, 8, 3.0, seed=1, device="cuda")
    out = dmrg_(psi, su2_heisenberg_mpo(12, device="cuda"), chi=64, max_sweeps=20, precision=precision)
    assert abs(out.energy - (-5.142090632840532)) < tol * 5.2


def test_su2_dmrg_reproduces_the_u1_energy_at_l20():
    """SU(2) and U(1) runs of the same chain agree (reference ``test_dmrg.py:462-490``) and
    reach the reference's L=20 literal ``-8.682473334398956`` (``test_dmrg.py:83``)."""
    from tenet.network import dmrg_
    from tenet.network.models import su2_heisenberg_mpo, su2_random_chain

    psi = su2_random_chain(20, 16, 3.0, seed=2, device="cuda")
    out = dmrg_(psi, su2_heisenberg_mpo(20, device="cuda"), chi=160, max_sweeps=20, precision="fp64")
    assert abs(out.energy - (-8.682473334398956)) < 1e-9
