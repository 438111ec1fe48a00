"""Perturbative noise: the density-matrix decimation and the perturbation built from the
operator's own action on the two-site tensor.

With nothing folded in, the density-matrix split is the SVD split (same bond space, same
Schmidt values, same discarded weight, same truncated state).  With a perturbation the
kept space is the top of ``M M^T + P P^T``, checked against NumPy's ``eigh`` of the
dense matrices.  The sweeps reach the exact ground energy with the mixer on,
from a random state and from a Neel product state, with U(1) and SU(2) symmetry."""

import numpy as np
import pytest
import torch

from oracle import ed

gpu = pytest.mark.gpu


def _at_bond(psi, h, n, chi=16):
    """``(env, aa)`` at bond ``n`` of a mixed-canonical ``psi``: the state a sweep holds
    there (noiseless sweep steps up to the bond, then the bond's Lanczos step)."""
    from tenet.network import Env, lanczos
    from tenet.network.dmrg import _pair, sweep_steps

    psi.canonize_(0)
    env = Env(psi, h).setup_(0)
    steps = sweep_steps(psi, h, env, {}, chi=chi, cutoff=1e-14)
    for _ in range(n):
        next(steps)
    steps.close()
    aa = _pair(psi[n], psi[n + 1])
    _, aa = lanczos(env.operator(n), aa)
    return env, aa


def _u1_case():
    from tenet.network import dmrg_
    from tenet.network.models import heisenberg_mpo, random_chain

    psi, h = random_chain(10, 16, 3.0, seed=4, device="cuda"), heisenberg_mpo(10, device="cuda")
    dmrg_(psi, h, chi=16, max_sweeps=2)
    return psi, h


def _su2_case():
    from tenet.network import dmrg_
    from tenet.network.models import su2_heisenberg_mpo, su2_random_chain

    psi, h = su2_random_chain(10, 6, 3.0, seed=4, device="cuda"), su2_heisenberg_mpo(10, device="cuda")
    dmrg_(psi, h, chi=16, max_sweeps=2)
    return psi, h


def _ramp(chi, noise=1e-4):
    from tenet.network import Sweep

    return [Sweep(chi, noise=noise, noise_type="perturbative")] * 3 + [Sweep(chi)] * 3


@gpu
@pytest.mark.parametrize("case", [_u1_case, _su2_case], ids=["u1", "su2"])
@pytest.mark.parametrize("chi", [4, 8, 16])
@pytest.mark.parametrize("forward", [True, False])
def test_noiseless_density_matrix_split_is_the_svd_split(case, chi, forward):
    """rho's spectrum is aa's squared and the keep rule is monotone in the value, so the
    kept bond space is the same; the values then agree to what squaring leaves."""
    import tenet
    from tenet.network.dmrg import _pair, _split_dm, spectrum

    psi, h = case()
    _, aa = _at_bond(psi, h, 3)
    u, s, vh = tenet.linalg.svd_truncated(aa, ((0, 1), (2, 3)), max_bond=chi, cutoff=1e-14)
    left, right, s2 = _split_dm(aa, forward, chi=chi, cutoff=1e-14)
    assert s2.legs[0].space == s.legs[0].space
    assert np.allclose(spectrum(s2), spectrum(s), rtol=0, atol=1e-10)
    carrier = right if forward else left
    norm_aa = float(tenet.norm(aa))
    dw_svd = 1.0 - (float(tenet.norm(s)) / norm_aa) ** 2
    dw_dm = 1.0 - (float(tenet.norm(carrier)) / norm_aa) ** 2
    assert abs(dw_dm - dw_svd) < 1e-12
    # both on the MPS site partition (bond, phys | bond), as the sweep writes them
    site = lambda t: tenet.repartition(t, (0, 1), (2,))  # noqa: E731
    mine = _pair(site(left), site(right))
    theirs = _pair(site(u), site(tenet.einsum("xy,yqr->xqr", s, vh)))
    assert float(tenet.norm(tenet.subtract(mine, theirs))) < 1e-7


@gpu
def test_perturbation_is_the_operator_action_scaled_to_the_noise():
    """One vector (the site-tensor operator has no term families to resolve): ``H_eff aa``
    normalised, carrying the noise as its squared norm."""
    import tenet
    from tenet.network.dmrg import _perturbations

    psi, h = _u1_case()
    env, aa = _at_bond(psi, h, 3)
    assert abs(float(tenet.norm(aa)) - 1.0) < 1e-12
    hv = env.heff2(3, aa)
    for noise in (1e-4, 1e-5):
        (p,) = _perturbations(env, 3, aa, noise)
        assert abs(float(tenet.norm(p)) ** 2 - noise) <= 1e-10 * noise
        unit = tenet.divide(hv, tenet.norm(hv))
        assert float(tenet.norm(tenet.subtract(tenet.divide(p, noise ** 0.5), unit))) < 1e-12
    assert len(env.heff2_families(3, aa)) == 1


@gpu
@pytest.mark.parametrize("forward", [True, False])
def test_perturbed_split_keeps_the_top_of_the_perturbed_density_matrix(forward):
    """The kept values squared are the top eigenvalues of ``M M^T + P P^T`` (forward;
    ``M^T M + P^T P`` coming back) from NumPy on the dense matrices, the isometry spans
    their eigenvectors, and the carrier is the state projected on that span."""
    from tenet.dense import to_dense
    from tenet.network.dmrg import _perturbations, _split_dm, spectrum

    psi, h = _u1_case()
    env, aa = _at_bond(psi, h, 3)
    (p,) = _perturbations(env, 3, aa, 1e-2)
    d = to_dense(aa)
    nl, nr = d.shape[0] * d.shape[1], d.shape[2] * d.shape[3]
    m, pm = d.reshape(nl, nr), to_dense(p).reshape(nl, nr)
    rho = m @ m.T + pm @ pm.T if forward else m.T @ m + pm.T @ pm
    w, v = np.linalg.eigh(rho)
    order = np.argsort(w)[::-1]
    w, v = w[order], v[:, order]
    # a bond dimension at a gap of the spectrum, so the kept span is unique
    chi = next(k for k in range(8, len(w)) if w[k - 1] - w[k] > 1e-3 * w[k - 1])
    left, right, s = _split_dm(aa, forward, chi=chi, cutoff=1e-14, perturbations=(p,))
    got = np.asarray(spectrum(s)) ** 2
    assert len(got) == chi
    assert np.allclose(got, w[:chi], rtol=1e-9, atol=1e-13)
    if forward:
        q = to_dense(left).reshape(nl, chi)
        c = to_dense(right).reshape(chi, nr)
    else:
        q = to_dense(right).reshape(chi, nr).T
        c = to_dense(left).reshape(nl, chi)
    top = v[:, :chi]
    assert np.allclose(q.T @ q, np.eye(chi), atol=1e-12)
    assert np.allclose(q @ q.T, top @ top.T, atol=1e-8)
    assert np.allclose(c, q.T @ m if forward else m @ q, atol=1e-12)


@gpu
def test_perturbative_sweep_keeps_the_state_normalised():
    from tenet.network import Env, sweep_
    from tenet.network.models import heisenberg_mpo, random_chain

    psi = random_chain(8, 8, 3.0, seed=0, device="cuda").canonize_(0)
    h = heisenberg_mpo(8, device="cuda")
    env = Env(psi, h).setup_(0)
    _, dw = sweep_(psi, h, env, {}, chi=8, cutoff=1e-14, noise=1e-3, noise_type="perturbative")
    assert abs(psi.norm() - 1.0) < 1e-12
    assert 0.0 <= dw < 1.0


@gpu
def test_perturbative_noise_fills_sectors_from_a_product_state():
    """From a Neel product state (every bond one-dimensional) a noiseless sweep grows the
    bonds only as far as the Lanczos vector reaches; the perturbation adds what the
    Hamiltonian couples to, so the bond comes back a strict superset."""
    from tenet.network import MPS, Env, sweep_
    from tenet.network.models import PHYS, heisenberg_mpo
    from tenet.symmetry import U1Sector

    h = heisenberg_mpo(6, device="cuda")
    spaces = {}
    for noise in (0.0, 1e-2):
        psi = MPS.product(PHYS, [U1Sector(1), U1Sector(-1)] * 3, device="cuda").canonize_(0)
        env = Env(psi, h).setup_(0)
        sweep_(psi, h, env, {}, chi=8, cutoff=1e-14, noise=noise, noise_type="perturbative")
        spaces[noise] = psi[3].legs[0].space
    plain, mixed = spaces[0.0], spaces[1e-2]
    assert set(plain) <= set(mixed)
    assert mixed.dim > plain.dim


@gpu
def test_the_mixer_does_not_move_the_ground_energy_u1():
    from tenet.network import MPS, dmrg_
    from tenet.network.models import PHYS, heisenberg_mpo, random_chain
    from tenet.symmetry import U1Sector

    out = dmrg_(random_chain(12, 16, 3.0, seed=0, device="cuda"), heisenberg_mpo(12, device="cuda"),
                schedule=_ramp(64), max_sweeps=8)
    assert abs(out.energy - ed.E_OBC_12) < 1e-9
    neel = MPS.product(PHYS, [U1Sector(1), U1Sector(-1)] * 6, device="cuda")
    out = dmrg_(neel, heisenberg_mpo(12, device="cuda"), schedule=_ramp(64), max_sweeps=10)
    assert abs(out.energy - ed.E_OBC_12) < 1e-9


@gpu
def test_the_mixer_does_not_move_the_ground_energy_su2():
    from tenet.network import dmrg_
    from tenet.network.models import su2_heisenberg_mpo, su2_random_chain

    psi = su2_random_chain(12, 6, 3.0, seed=0, device="cuda")
    out = dmrg_(psi, su2_heisenberg_mpo(12, device="cuda"), schedule=_ramp(32), max_sweeps=8)
    assert abs(psi.norm() - 1.0) < 1e-12
    assert abs(out.energy - ed.E_OBC_12) < 1e-9


@gpu
def test_noise_type_on_a_noiseless_entry_changes_nothing():
    from tenet.network import Sweep, dmrg_
    from tenet.network.models import heisenberg_mpo, random_chain

    energies = []
    for kind in ("wavefunction", "perturbative"):
        psi = random_chain(8, 8, 3.0, seed=0, device="cuda")
        out = dmrg_(psi, heisenberg_mpo(8, device="cuda"), schedule=[Sweep(16, noise_type=kind)] * 3,
                    max_sweeps=3)
        energies.append(out.energy)
    assert energies[0] == energies[1]


def test_an_unknown_noise_type_is_refused():
    from tenet.network import sweep_

    with pytest.raises(ValueError, match="noise_type"):
        sweep_(None, None, None, {}, chi=8, cutoff=1e-14, noise=1e-4, noise_type="density_matrix")


def test_product_state_refusals():
    from tenet.network import MPS
    from tenet.network.models import PHYS, SU2_PHYS
    from tenet.space import GradedSpace
    from tenet.symmetry import U1, SU2Sector, U1Sector

    with pytest.raises(ValueError, match="degeneracy 1"):
        MPS.product(PHYS, [U1Sector(3)])
    with pytest.raises(ValueError, match="degeneracy 1"):
        MPS.product(GradedSpace.new(U1, {U1Sector(1): 2}), [U1Sector(1)])
    with pytest.raises(ValueError, match="fusion channels"):
        MPS.product(SU2_PHYS, [SU2Sector(1), SU2Sector(1)])


@gpu
def test_product_state_bonds_carry_the_charges():
    """Bonds derived right to left: the last is the unit, bond 0 shows the total charge
    (on a dual leg), every site holds one amplitude 1."""
    from tenet.dense import to_dense
    from tenet.network import MPS
    from tenet.network.models import PHYS
    from tenet.symmetry import U1Sector

    psi = MPS.product(PHYS, [U1Sector(1), U1Sector(1), U1Sector(-1)], device="cuda")
    assert dict(psi[0].legs[0].space.sectors) == {U1Sector(1): 1} and psi[0].legs[0].dual
    assert [dict(t.legs[2].space.sectors) for t in psi] == [{U1Sector(0): 1}, {U1Sector(1): 1}, {U1Sector(0): 1}]
    for t, k in zip(psi, (1, 1, 0)):  # PHYS's dense order is (-1, +1)
        d = to_dense(t)
        assert d.shape == (1, 2, 1) and d[0, k, 0] == 1.0 and d.sum() == 1.0
    assert abs(psi.norm() - 1.0) < 1e-15
