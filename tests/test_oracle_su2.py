"""The SU(2) / fermion-parity dense oracle (CPU), pinned before it is trusted:

* Clebsch-Gordan tensors against sympy's exact values (an independent library routine);
* CG orthonormality (the defining identity of a unitary change of basis);
* F-symbols recomputed from CG tensors against the reference fixture ``su2_f.txt``
  (TensorKitSectors' values): this fixes the oracle's CG gauge to the reference's F gauge;
* the norm identity ``|to_dense(T)| = sqrt(sum qdim |A|^2)`` (reference ``basic.py:340-346``)
  and the ``from_dense`` round trip;
* planted mistakes (a transposed CG, a dropped Z on a dual leg, a flipped Koszul sign)
  that one of the checks above must catch.

Then the product's host-side pieces are checked against it: block order (bit-exact sector
maps), its own Clebsch-Gordan construction, ``to_dense``/``from_dense`` and the invariant
SU(2) Heisenberg MPO.
"""

import itertools
from pathlib import Path

import numpy as np
import pytest

from oracle import su2 as O

FIX = Path(__file__).resolve().parent / "golden"


def test_cg_matches_sympy_exact_values():
    from sympy import Rational, N
    from sympy.physics.quantum.cg import CG

    worst = 0.0
    for j1, j2 in itertools.product(range(0, 5), range(0, 5)):
        for j3 in O.SU2Rules.fuse(j1, j2):
            c = O.cg(j1, j2, j3)
            for a, b, k in itertools.product(range(j1 + 1), range(j2 + 1), range(j3 + 1)):
                m1, m2, m3 = j1 - 2 * a, j2 - 2 * b, j3 - 2 * k
                v = CG(Rational(j1, 2), Rational(m1, 2), Rational(j2, 2), Rational(m2, 2),
                       Rational(j3, 2), Rational(m3, 2)).doit()
                worst = max(worst, abs(float(N(v, 20)) - c[a, b, k]))
    assert worst < 1e-15


def test_cg_orthonormal():
    for j1, j2 in itertools.product(range(0, 9), range(0, 7)):
        stack = np.concatenate([O.cg(j1, j2, j3).reshape((j1 + 1) * (j2 + 1), -1)
                                for j3 in O.SU2Rules.fuse(j1, j2)], axis=1)
        np.testing.assert_allclose(stack.T @ stack, np.eye(stack.shape[1]), atol=1e-13)


def _f_rows():
    path = FIX / "su2_f.txt"
    if not path.exists():
        pytest.skip("su2_f.txt fixture missing")
    return np.loadtxt(path, comments="#")


def test_f_from_cg_matches_reference_fixture():
    rows = _f_rows()
    rng = np.random.default_rng(7)
    small = rows[np.max(rows[:, :6], axis=1) <= 4]
    pick = np.concatenate([small, rows[rng.choice(len(rows), 3000, replace=False)]])
    worst = max(abs(O.f_symbol_from_cg(*map(int, r[:6])) - r[6]) for r in pick)
    assert worst < 1e-13


@pytest.mark.parametrize("mistake", ["sign of the coupled line", "dropped sqrt(2j+1)"])
def test_planted_cg_mistakes_are_caught_by_the_fixture(monkeypatch, mistake):
    """A CG tensor whose phase depends on the coupled spin in another gauge, or that lost
    its ``sqrt(2j3+1)`` normalisation, changes some F: the fixture comparison must fail."""
    rows = _f_rows()
    orig = O.cg

    def wrong(j1, j2, j3):
        if mistake == "dropped sqrt(2j+1)":
            return orig(j1, j2, j3) / np.sqrt(j3 + 1)
        return orig(j1, j2, j3) * (-1.0) ** ((j3 // 2) % 2)

    monkeypatch.setattr(O, "cg", wrong)
    small = rows[np.max(rows[:, :6], axis=1) <= 3]
    worst = max(abs(O.f_symbol_from_cg(*map(int, r[:6])) - r[6]) for r in small)
    assert worst > 0.1


def test_z_matrix_closed_form():
    for j in range(0, 9):
        z = O.z_matrix(j)
        want = np.zeros((j + 1, j + 1))
        for i in range(j + 1):
            want[i, j - i] = (-1.0) ** i
        np.testing.assert_allclose(z, want, atol=1e-15)


def _legs(*specs):
    return [O.SLeg(sectors, out, dual) for sectors, out, dual in specs]


V = ((0, 2), (1, 2), (2, 1))
W = ((0, 1), (1, 2))
U = ((1, 1), (2, 2))


@pytest.mark.parametrize("specs", [
    ((V, True, False), (W, False, False), (U, False, False)),
    ((V, True, False), (W, True, True), (U, False, False), (V, False, True)),
    ((U, False, True), (W, False, False)),
])
def test_norm_identity_and_round_trip(specs):
    legs = _legs(*specs)
    blocks = O.random_blocks(O.SU2Rules, legs, 5)
    dense = O.to_dense(O.SU2Rules, legs, blocks)
    assert abs(np.linalg.norm(dense) - O.norm(O.SU2Rules, legs, blocks)) < 1e-12 * np.linalg.norm(dense)
    back, resid = O.from_dense(O.SU2Rules, legs, dense)
    assert resid < 1e-12
    for a, b in zip(back, blocks):
        np.testing.assert_allclose(a, b, atol=1e-13)


def test_planted_dropped_z_breaks_invariance(monkeypatch):
    """Without Z on a dual leg the dense array of a tensor contracted against its dual
    partner is no longer the dense contraction: the bilinear pairing of a dual and a
    plain leg only closes with Z."""
    legs_a = _legs((W, True, False), (W, True, False))
    legs_b = _legs((W, True, True), (U, False, False))
    a = O.to_dense(O.SU2Rules, legs_a, O.random_blocks(O.SU2Rules, legs_a, 1))
    b = O.to_dense(O.SU2Rules, legs_b, O.random_blocks(O.SU2Rules, legs_b, 2))
    full = np.tensordot(a, b, ([1], [0]))
    out = _legs((W, True, False), (U, False, False))
    _, resid = O.from_dense(O.SU2Rules, out, full)
    assert resid < 1e-12 * max(1.0, np.linalg.norm(full))
    monkeypatch.setattr(O, "z_matrix", lambda j: np.eye(j + 1))
    b2 = O.to_dense(O.SU2Rules, legs_b, O.random_blocks(O.SU2Rules, legs_b, 2))
    _, resid2 = O.from_dense(O.SU2Rules, out, np.tensordot(a, b2, ([1], [0])))
    assert resid2 > 1e-3


def test_supersign_counts_odd_odd_inversions():
    par = ((0, 1), (1, 1))
    legs = _legs((par, True, False), (par, True, False))
    s = O.supersign(legs, (1, 0))
    np.testing.assert_array_equal(s, [[1, 1], [1, -1]])
    legs2 = _legs((par, True, False), (par, False, False))
    np.testing.assert_array_equal(O.supersign(legs2, (1, 0), per_side=True), np.ones((2, 2)))


# ------------------------------------------------------------------ product host side


def _product_legs(specs):
    from tenet import IN, OUT, GradedSpace, Leg
    from tenet.symmetry import SU2, SU2Sector

    return tuple(Leg(GradedSpace.new(SU2, {SU2Sector(a): m for a, m in sec}), OUT if out else IN, dual)
                 for sec, out, dual in specs)


def _tree_tuple(t):
    return (tuple(x.two_j for x in t.uncoupled), tuple(x.two_j for x in t.inner), t.coupled.two_j)


SPECS = [
    ((V, True, False), (W, False, False), (U, False, False)),
    ((V, True, False), (W, True, True), (U, False, False), (V, False, True)),
    ((V, True, False), (W, True, False), (U, True, False), (W, False, False)),
    ((U, False, True), (W, False, False)),
    ((V, True, False), (V, True, False)),
]


@pytest.mark.parametrize("specs", SPECS)
def test_product_block_order_matches_oracle(specs):
    from tenet.structure import TensorStructure

    s = TensorStructure(_product_legs(specs))
    keys = O.block_keys(O.SU2Rules, _legs(*specs))
    assert [(_tree_tuple(k.output_tree), _tree_tuple(k.input_tree)) for k in s.block_order] == \
        [(k[0], k[1]) for k in keys]
    assert [tuple(x.two_j for x in s.axis_sectors(k)) for k in s.block_order] == [k[2] for k in keys]


def test_product_cg_matches_oracle():
    from tenet import _su2coeff as C

    worst = max(float(np.abs(C.cg(a, b, c) - O.cg(a, b, c)).max())
                for a, b in itertools.product(range(0, 9), range(0, 9)) for c in O.SU2Rules.fuse(a, b))
    assert worst < 1e-12


@pytest.mark.parametrize("specs", SPECS)
def test_product_dense_matches_oracle(specs):
    from tenet import SymmetricTensor
    from tenet.dense import from_dense, to_dense

    legs = _product_legs(specs)
    t = SymmetricTensor.random(legs, seed=11, device="cpu")
    ol = _legs(*specs)
    want = O.to_dense(O.SU2Rules, ol, O.random_blocks(O.SU2Rules, ol, 11))
    np.testing.assert_allclose(to_dense(t), want, atol=1e-13)
    back = from_dense(want, legs, device="cpu")
    np.testing.assert_allclose(back.data.numpy(), t.data.numpy(), atol=1e-12)


def test_product_from_dense_refuses_a_non_invariant_array():
    from tenet.dense import from_dense

    legs = _product_legs(((W, True, False), (W, False, False)))
    rng = np.random.default_rng(0)
    with pytest.raises(ValueError, match="not symmetric"):
        from_dense(rng.standard_normal((5, 5)), legs, device="cpu")


def test_su2_heisenberg_mpo_is_invariant_and_reproduces_s_dot_s():
    from tenet.dense import to_dense
    from tenet.network.models import SU2_END, SU2_START, su2_heisenberg_mpo_w, su2_heisenberg_w

    w = su2_heisenberg_w()
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    ss = 0.25 * np.kron(np.diag([1.0, -1.0]), np.diag([1.0, -1.0])) + 0.5 * (np.kron(sp, sp.T) + np.kron(sp.T, sp))
    two = np.einsum("Ppa,aQq->PQpq", w[SU2_START, :, :, 2:], w[2:, :, :, SU2_END]).reshape(4, 4)
    np.testing.assert_allclose(two, ss, atol=1e-15)
    mpo = su2_heisenberg_mpo_w(4, device="cpu")  # from_dense refuses a non-invariant W
    np.testing.assert_allclose(to_dense(mpo[1]), w, atol=1e-14)


def test_su2_bond_profile_of_the_512_multiplet_chain():
    from tenet.network.models import su2_gaussian_bond

    even = su2_gaussian_bond(32, 64, 512, 6.0, 12)
    assert [(a.two_j, m) for a, m in even.sectors] == [(0, 122), (2, 117), (4, 99), (6, 75), (8, 51), (10, 31), (12, 17)]
    odd = su2_gaussian_bond(31, 64, 512, 6.0, 13)
    assert [a.two_j for a, _ in odd.sectors] == list(range(1, 14, 2))
    assert sum(m for _, m in odd.sectors) == 512
