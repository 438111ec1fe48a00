"""Local operators, the spin-1/2 site and ``MPO.from_terms`` on the host (no GPU): every MPO
built from terms, contracted densely along a short chain, equals the Hamiltonian written
with Kronecker products (reference behaviour ``network/mps.py:960-1041,3135-3342``,
``models/sites.py:121-194``), and the refusals fire."""

import numpy as np
import pytest

L = 6
SZ = np.diag([-0.5, 0.5])
SP = np.array([[0.0, 0.0], [1.0, 0.0]])


def _kron_at(o, i, n=L):
    m = np.eye(1)
    for k in range(n):
        m = np.kron(m, o if k == i else np.eye(2))
    return m


def _xxz(delta, n=L, field=0.0):
    h = sum(delta * _kron_at(SZ, i, n) @ _kron_at(SZ, i + 1, n)
            + 0.5 * (_kron_at(SP, i, n) @ _kron_at(SP.T, i + 1, n) + _kron_at(SP.T, i, n) @ _kron_at(SP, i + 1, n))
            for i in range(n - 1))
    return h + field * sum(_kron_at(SZ, i, n) for i in range(n))


def _dense_chain(mpo):
    from tenet.dense import to_dense

    ws = [to_dense(w) for w in mpo.sites]
    acc = ws[0][0]
    for w in ws[1:]:
        d = w.shape[1]
        acc = np.einsum("abx,xcdy->acbdy", acc, w).reshape(acc.shape[0] * d, acc.shape[1] * d, w.shape[3])
    return acc[:, :, 0]


def test_spin_half_sites():
    from tenet.network import spin_half
    from tenet.symmetry import SU2, U1

    s = spin_half(U1, device="cpu")
    assert sorted(s.ops) == ["S+", "S-", "Sz"] and sorted(s.matrices) == ["S+", "S-", "S.S", "Sz"]
    assert s.phys.dim == 2 and all(op.ndim == 3 for op in s.ops.values())
    su = spin_half(SU2, device="cpu")
    assert sorted(su.ops) == ["S.S"] and su.ops["S.S"].ndim == 4
    with pytest.raises(ValueError, match="gradings"):
        spin_half(object())


def test_local_op_refusals():
    from tenet.network import local_op, spin_half
    from tenet.symmetry import U1, U1Sector

    phys = spin_half(U1, device="cpu").phys
    with pytest.raises(ValueError):  # S+ does not conserve the charge as a rank-2 operator
        local_op(SP, phys=phys, device="cpu")
    with pytest.raises(ValueError):  # S+ emits -2, not +2
        local_op(SP, phys=phys, charge=U1Sector(2), device="cpu")
    with pytest.raises(ValueError, match="invariant k-site term"):
        local_op(np.zeros((3, 3)), phys=phys, device="cpu")


@pytest.mark.parametrize("delta,field", [(1.0, 0.0), (0.3, 0.7)])
def test_u1_chain_terms_build_the_exact_hamiltonian(delta, field):
    from tenet.network import MPO, spin_half

    o = spin_half(device="cpu").ops
    terms = [(field, [(o["Sz"], i)]) for i in range(L)]
    for i in range(L - 1):
        terms += [(delta, [(o["Sz"], i), (o["Sz"], i + 1)]), (0.5, [(o["S+"], i), (o["S-"], i + 1)]),
                  (0.5, [(o["S-"], i), (o["S+"], i + 1)])]
    h = MPO.from_terms(L, terms, device="cpu")
    assert h.edges is not None and h.materialize().edges is None
    assert h[2].legs[3].space.dim == 5  # the two corner channels + three open terms
    assert np.abs(_dense_chain(h) - _xxz(delta, field=field)).max() < 1e-13


def test_longer_range_term_carries_its_charge_through_identities():
    from tenet.network import MPO, spin_half

    o = spin_half(device="cpu").ops
    h = MPO.from_terms(L, [(0.8, [(o["S+"], 1), (o["S-"], 4)])], device="cpu")
    want = 0.8 * _kron_at(SP, 1) @ _kron_at(SP.T, 4)
    assert np.abs(_dense_chain(h) - want).max() < 1e-14


def test_invariant_two_site_terms_u1_and_su2():
    from tenet.network import MPO, local_op, spin_half
    from tenet.symmetry import SU2

    s = spin_half(device="cpu")
    ss = local_op(s.matrices["S.S"], phys=s.phys, device="cpu")
    h = MPO.from_terms(L, [(1.0, [(ss, (i, i + 1))]) for i in range(L - 1)], device="cpu")
    assert np.abs(_dense_chain(h) - _xxz(1.0)).max() < 1e-13
    su = spin_half(SU2, device="cpu")
    h2 = MPO.from_terms(L, [(1.0, [(su.ops["S.S"], (i, i + 1))]) for i in range(L - 1)], device="cpu")
    bond = h2[2].legs[3].space
    assert [(a.two_j, m) for a, m in bond.sectors] == [(0, 2), (2, 1)]  # corners + one triplet
    # the SU(2) dense basis orders m descending: same spectrum as the U(1) chain
    ev = np.linalg.eigvalsh(_dense_chain(h2))
    assert np.abs(ev - np.linalg.eigvalsh(_xxz(1.0))).max() < 1e-12


def test_from_terms_refusals():
    from tenet.network import MPO, local_op, spin_half

    o = spin_half(device="cpu").ops
    with pytest.raises(ValueError, match="sum to the unit"):
        MPO.from_terms(L, [(1.0, [(o["S+"], 0), (o["Sz"], 1)])], device="cpu")
    with pytest.raises(ValueError, match="increasing"):
        MPO.from_terms(L, [(1.0, [(o["Sz"], 2), (o["Sz"], 1)])], device="cpu")
    with pytest.raises(ValueError, match="outside"):
        MPO.from_terms(L, [(1.0, [(o["Sz"], L)])], device="cpu")


def test_su2_heisenberg_from_terms_equals_the_hand_written_w():
    from tenet.network.models import su2_heisenberg_mpo, su2_heisenberg_mpo_w

    a, b = su2_heisenberg_mpo(L, device="cpu"), su2_heisenberg_mpo_w(L, device="cpu")
    assert all(x.legs == y.legs for x, y in zip(a.sites, b.sites))
    assert np.abs(_dense_chain(a) - _dense_chain(b)).max() < 1e-13
