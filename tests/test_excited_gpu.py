"""Excited states on the GPU (reference ``network/dmrg.py:620-843``, ``orthogonal_to``;
``env.py:843-977``, ``project2``/``measure``; ``mps.py:2731-2778``, ``MPO.identity``):
overlaps of the two-state environment over the identity MPO against the oracle's dense
contraction, and the first excited S^z = 0 level of open XX and Heisenberg chains against
the free-fermion closed form and exact diagonalization."""

import numpy as np
import pytest

from oracle import ed
from oracle import u1 as O
from tests.helpers import to_oracle

pytestmark = pytest.mark.gpu


def _dense_overlap(bra, ket) -> float:
    """``<bra|ket>`` by contracting the oracle's dense site arrays left to right."""
    env = np.ones((1, 1))
    for b, k in zip(bra, ket):
        db, dk = O.densify(to_oracle(b)), O.densify(to_oracle(k))
        env = np.einsum("ab,apx,bpy->xy", env, dk, db)
    return float(env.sum())


def test_identity_mpo_overlaps_match_dense():
    from tenet.network import MPO, Env
    from tenet.network.models import PHYS, random_chain

    psi = random_chain(8, 12, 2.0, seed=11, device="cuda")
    phi = random_chain(8, 12, 2.0, seed=12, device="cuda")
    ident = MPO.identity(8, PHYS, device="cuda")
    got = Env(psi, ident, bra=phi).measure()
    ref = _dense_overlap(phi, psi)
    assert abs(got - ref) <= 1e-12 * max(abs(ref), 1.0)
    assert abs(Env(psi, ident).measure() - psi.norm() ** 2) <= 1e-12 * psi.norm() ** 2


def test_su2_identity_overlap_is_the_norm():
    from tenet.network import MPO, Env
    from tenet.network.models import SU2_PHYS, su2_random_chain

    psi = su2_random_chain(8, 6, 2.0, seed=3, device="cuda", max_two_j=4)
    got = Env(psi, MPO.identity(8, SU2_PHYS, device="cuda")).measure()
    assert abs(got - psi.norm() ** 2) <= 1e-12 * psi.norm() ** 2


@pytest.mark.parametrize("delta,prec,tol", [(0.0, "fp64", 1e-10), (1.0, "fp64", 1e-10), (1.0, "fp32", 1e-8)])
def test_first_excited_level(delta, prec, tol):
    from tenet.network import MPO, Env, dmrg_
    from tenet.network.models import PHYS, heisenberg_mpo, random_chain

    L = 10
    h = heisenberg_mpo(L, delta=delta, device="cuda")
    psi0 = random_chain(L, 16, 3.0, seed=1, device="cuda")
    out0 = dmrg_(psi0, h, chi=32, max_sweeps=12)
    psi1 = random_chain(L, 16, 3.0, seed=2, device="cuda")
    out1 = dmrg_(psi1, h, chi=32, max_sweeps=16, orthogonal_to=[psi0], precision=prec)
    e0, e1 = ed.xxz_sz0_levels(L, delta, 2)
    if delta == 0.0:
        e0, e1 = ed.xx_energy(L), ed.xx_first_excited(L)
    assert abs(out0.energy - e0) < 1e-10 * abs(e0)
    assert abs(out1.energy - e1) < tol * abs(e1), (out1.energy, e1)
    # the excited state is orthogonal to the ground state and normalised
    ident = MPO.identity(L, PHYS, device="cuda")
    assert abs(Env(psi1, ident, bra=psi0).measure()) < 1e-6
    assert abs(psi1.norm() - 1.0) < 1e-10


def test_external_solves_refuse_an_excited_state_sweep():
    from tenet.network.dmrg import dmrg_steps
    from tenet.network.models import heisenberg_mpo, random_chain

    h = heisenberg_mpo(6, device="cuda")
    phi = random_chain(6, 8, 3.0, seed=1, device="cuda")
    psi = random_chain(6, 8, 3.0, seed=2, device="cuda")
    with pytest.raises(ValueError):
        next(dmrg_steps(psi, h, chi=8, external=True, orthogonal_to=[phi]))


def test_su2_first_excited_singlet():
    """SU(2) chains hold total spin 0 only: the excited-state sweep finds the second singlet
    level of the L=10 Heisenberg chain (ED: the S^z = 0 spectrum less the S^z = 1 one)."""
    from tenet.network import MPO, Env, dmrg_
    from tenet.network.models import SU2_PHYS, su2_heisenberg_mpo, su2_random_chain

    L = 10
    h = su2_heisenberg_mpo(L, device="cuda")
    psi0 = su2_random_chain(L, 8, 3.0, seed=1, device="cuda")
    out0 = dmrg_(psi0, h, chi=48, max_sweeps=12)
    psi1 = su2_random_chain(L, 8, 3.0, seed=2, device="cuda")
    out1 = dmrg_(psi1, h, chi=48, max_sweeps=16, orthogonal_to=[psi0])
    e0, e1 = ed.heisenberg_singlet_levels(L, 2)
    assert abs(out0.energy - e0) < 1e-10 * abs(e0)
    assert abs(out1.energy - e1) < 1e-10 * abs(e1), (out1.energy, e1)
    assert abs(Env(psi1, MPO.identity(L, SU2_PHYS, device="cuda"), bra=psi0).measure()) < 1e-6


def test_overlap_and_measure_mpo():
    """``overlap`` against the oracle's dense contraction; ``measure_mpo(psi, h, psi)`` of a
    converged ground state is its energy times ``<psi|psi>``."""
    from tenet.network import dmrg_, measure_mpo, overlap
    from tenet.network.models import heisenberg_mpo, random_chain

    psi = random_chain(8, 12, 2.0, seed=21, device="cuda")
    phi = random_chain(8, 12, 2.0, seed=22, device="cuda")
    ref = _dense_overlap(phi, psi)
    assert abs(overlap(phi, psi) - ref) <= 1e-12 * max(abs(ref), 1.0)
    h = heisenberg_mpo(10, device="cuda")
    g = random_chain(10, 16, 3.0, seed=1, device="cuda")
    out = dmrg_(g, h, chi=32, max_sweeps=12)
    assert abs(measure_mpo(g, h, g) - out.energy * overlap(g, g)) <= 1e-10 * abs(out.energy)
