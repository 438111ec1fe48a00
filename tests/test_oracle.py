"""The oracle pinned against what the reference's own tests hold and what mathematics fixes
(CPU only): dense einsum, brute-force sector enumeration, exact diagonalization, the
reference's ED and MPSKit energy literals, the XX closed form, Eckart-Young."""

import itertools

import numpy as np
import pytest

from oracle import ed
from oracle import u1 as O
from oracle import u1_network as N

SMALL = {-1: 2, 0: 3, 1: 2, 2: 1}


def _t(sides, duals, seed, d=SMALL):
    return O.OTensor.random(tuple(O.OLeg(O.space(d), s, du) for s, du in zip(sides, duals)), seed)


def test_block_keys_match_brute_force_conservation():
    legs = tuple(O.OLeg(O.space(SMALL), s, du) for s, du in [(True, False), (True, True), (False, False)])
    keys = O.block_keys(legs)
    brute = [k for k in itertools.product(*([q for q, _ in l.sectors] for l in legs))
             if legs[0].fused(k[0]) + legs[1].fused(k[1]) == legs[2].fused(k[2])]
    assert sorted(keys) == sorted(brute)
    # sorted by (fused OUT tuple, fused IN tuple): independent of enumeration order
    assert keys == sorted(keys, key=lambda k: ((legs[0].fused(k[0]), legs[1].fused(k[1])), (legs[2].fused(k[2]),)))


@pytest.mark.parametrize("eq,sa,da,sb,db", [
    ("abc,cde->abde", (1, 1, 0), (0, 0, 0), (1, 0, 0), (0, 0, 0)),
    ("abc,cde->abde", (1, 1, 1), (0, 0, 0), (1, 0, 0), (1, 0, 0)),
    ("abc,dca->bd", (1, 0, 0), (0, 0, 0), (1, 1, 0), (0, 0, 0)),
    ("abcd,dc->ab", (1, 1, 0, 0), (0, 0, 0, 0), (1, 1), (0, 0)),
])
def test_oracle_einsum_equals_dense_einsum(eq, sa, da, sb, db):
    a, b = _t(sa, da, 1), _t(sb, db, 2)
    got = O.densify(O.einsum(eq, a, b))
    ref = np.einsum(eq, O.densify(a), O.densify(b))
    assert np.allclose(got, ref, rtol=0, atol=1e-12)


def test_oracle_planted_transpose_mistake_is_caught():
    """A dropped transpose in the oracle's repartition would be caught by the dense pin."""
    a, b = _t((1, 1, 0), (0, 0, 0), 1), _t((1, 0, 0), (0, 0, 0), 2)
    c = O.einsum("abc,cde->abde", a, b)
    wrong = O.OTensor(c.legs, {k: v.swapaxes(0, 1) if v.shape[0] == v.shape[1] else v
                               for k, v in c.blocks.items()})
    ref = np.einsum("abc,cde->abde", O.densify(a), O.densify(b))
    assert not np.allclose(O.densify(wrong), ref)


def test_ed_against_reference_literals_and_closed_form():
    assert abs(ed.xxz_sz0_energy(12) - ed.E_OBC_12) < 1e-10
    for L in (6, 8, 10):
        assert abs(ed.xxz_sz0_energy(L, 0.0) - ed.xx_energy(L)) < 1e-10


def test_ed_levels_against_closed_forms():
    """The two lowest S^z = 0 levels: the free-fermion ground state and its lowest
    particle-number-conserving excitation at delta = 0, and the two-site singlet/triplet
    (-3/4, +1/4) and four-site ground state -3/4 - sqrt(3)/2 at delta = 1."""
    for L in (6, 8, 10):
        lv = ed.xxz_sz0_levels(L, 0.0, 2)
        assert abs(lv[0] - ed.xx_energy(L)) < 1e-10
        assert abs(lv[1] - ed.xx_first_excited(L)) < 1e-10
        assert lv[1] - lv[0] > 1e-3  # a non-degenerate first level
    assert np.allclose(ed.xxz_sz0_levels(2, 1.0, 2), [-0.75, 0.25], atol=1e-14)
    assert abs(ed.xxz_sz0_levels(4, 1.0, 1)[0] - (-0.75 - 3 ** 0.5 / 2)) < 1e-13
    assert abs(ed.xxz_sz0_levels(12, 1.0, 1)[0] - ed.E_OBC_12) < 1e-10


def test_singlet_levels_against_closed_forms():
    """Two sites: the singlet -3/4 alone.  Four sites (S1.S2 + S2.S3 + S3.S4): the two
    singlets -3/4 -+ sqrt(3)/2 (the S^z = 0 block also holds the triplets -1/4 -+ 1/sqrt(2)
    and -1/4 and the quintet 3/4, all of which reappear at S^z = 1).  Twelve sites: the
    ground state is the singlet E_OBC_12."""
    sz0 = np.linalg.eigvalsh(ed._xxz_block(4, 1.0, 2).toarray())
    assert np.allclose(sz0, sorted([-0.75 - 3 ** 0.5 / 2, -0.75 + 3 ** 0.5 / 2, -0.25 - 0.5 ** 0.5,
                                    -0.25, -0.25 + 0.5 ** 0.5, 0.75]), atol=1e-13)
    assert np.allclose(ed.heisenberg_singlet_levels(2, 2), [-0.75], atol=1e-14)
    assert np.allclose(ed.heisenberg_singlet_levels(4, 2), [-0.75 - 3 ** 0.5 / 2, -0.75 + 3 ** 0.5 / 2],
                       atol=1e-13)
    assert abs(ed.heisenberg_singlet_levels(12, 1)[0] - ed.E_OBC_12) < 1e-10


def _heis(L, delta=1.0):
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "tenet-py_amd"))
    from tenet.network.models import BOUNDARY, MPO_BOND, PHYS, END, START, spin_half_w

    conv = lambda s: tuple((a.charge, m) for a, m in s.sectors)
    return N.mpo_from_w(spin_half_w(delta), L, conv(PHYS), conv(MPO_BOND), conv(BOUNDARY), START, END)


def test_oracle_dmrg_reaches_exact_energy():
    L = 10
    mpo = _heis(L)
    phys = ((-1, 1), (1, 1))
    bonds = [((0, 1),)]
    for i in range(1, L):
        right = L - i
        qs = [q for q in range(-i, i + 1, 2) if abs(q) <= right]
        bonds.append(tuple((q, 2) for q in qs))
    bonds.append(((0, 1),))
    sites = [O.OTensor.random((O.OLeg(bonds[i], True), O.OLeg(phys, True), O.OLeg(bonds[i + 1], False)), i)
             for i in range(L)]
    e, hist = N.dmrg(sites, mpo, chi=32, max_sweeps=8)
    assert abs(e - ed.xxz_sz0_energy(L)) < 1e-9


def test_truncated_svd_eckart_young_and_pythagoras():
    t = _t((1, 1, 0, 0), (0, 0, 0, 0), 7, {-1: 3, 0: 4, 1: 3})
    u, s, vh, spec, dw = N.svd_truncated(t, (0, 1), (2, 3), max_bond=6)
    approx = O.compose(O.compose(u, s), vh)
    m = O.repartition(t, (0, 1), (2, 3))
    err2 = O.norm(O.axpy(-1.0, approx, m)) ** 2
    assert abs(err2 - dw) < 1e-10 * O.norm(m) ** 2
    assert abs(O.norm(m) ** 2 - O.norm(s) ** 2 - dw) < 1e-10 * O.norm(m) ** 2
