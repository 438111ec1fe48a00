"""Synthetic workloads of the benchmark and the full-size parity tests.

``heff2_problem(L, D, sigma, seed)`` -- the two-site H_eff of a U(1) Heisenberg chain at
its centre bond: a random N(0,1) MPS whose bonds follow ``models.gaussian_bond`` (odd
charges at the centre, a Gaussian profile of width ``sigma`` capped by reachability),
each site normalised, and environments built from it by ``Env.update_`` (each normalised
after its fold, which leaves the operator Hermitian).  Returns ``(fl, w_left, w_right,
fr, aa)`` on the device in float64.
"""

from __future__ import annotations

from tenet import ops
from tenet.network.env import Env
from tenet.network.models import chain_bonds, heisenberg_mpo, PHYS
from tenet.network.mps import MPS

__all__ = ["heff2_problem", "su2_heff2_problem"]


def heff2_problem(n_sites: int = 64, max_dim: int = 2048, sigma: float = 6.0, seed: int = 0,
                  delta: float = 1.0, device="cuda"):
    psi = MPS.random(PHYS, chain_bonds(n_sites, max_dim, sigma), seed=seed, device=device)
    for i in range(n_sites):
        psi.sites[i] = ops.divide(psi[i], ops.norm(psi[i]))
    h = heisenberg_mpo(n_sites, delta, device=device)
    env = Env(psi, h)
    c = n_sites // 2 - 1
    for n in range(0, c):
        env.update_(n, to="last")
        f = env.F[n, n + 1]
        env.F[n, n + 1] = ops.divide(f, ops.norm(f))
    for n in range(n_sites - 1, c + 1, -1):
        env.update_(n, to="first")
        f = env.F[n, n - 1]
        env.F[n, n - 1] = ops.divide(f, ops.norm(f))
    aa = ops.einsum("apx,xqr->apqr", psi[c], psi[c + 1])
    aa = ops.divide(aa, ops.norm(aa))
    return env.F[c - 1, c], h[c], h[c + 1], env.F[c + 2, c + 1], aa


def su2_heff2_problem(n_sites: int = 64, max_mult: int = 512, sigma: float = 6.0, seed: int = 0,
                      device="cuda"):
    """The two-site H_eff of the SU(2) Heisenberg chain at its centre bond: bonds of
    ``max_mult`` multiplets with a Gaussian profile in ``2j`` (at most ``2j = 12`` on even
    bonds, ``13`` on odd ones), random N(0,1) sites each normalised, environments folded
    from them and normalised.  Returns ``(fl, w_left, w_right, fr, aa)`` in float64."""
    from tenet.network.models import SU2_PHYS, su2_chain_bonds, su2_heisenberg_mpo

    bonds = su2_chain_bonds(n_sites, max_mult, sigma, max_two_j=12)
    psi = MPS.random(SU2_PHYS, bonds, seed=seed, device=device)
    for i in range(n_sites):
        psi.sites[i] = ops.divide(psi[i], ops.norm(psi[i]))
    h = su2_heisenberg_mpo(n_sites, device=device)
    env = Env(psi, h)
    c = n_sites // 2 - 1
    for n in range(0, c):
        env.update_(n, to="last")
        f = env.F[n, n + 1]
        env.F[n, n + 1] = ops.divide(f, ops.norm(f))
    for n in range(n_sites - 1, c + 1, -1):
        env.update_(n, to="first")
        f = env.F[n, n - 1]
        env.F[n, n - 1] = ops.divide(f, ops.norm(f))
    aa = ops.einsum("apx,xqr->apqr", psi[c], psi[c + 1])
    aa = ops.divide(aa, ops.norm(aa))
    return env.F[c - 1, c], h[c], h[c + 1], env.F[c + 2, c + 1], aa

