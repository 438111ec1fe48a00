"""Lattice models and bond spaces used by the workloads.

* ``spin_half_w`` -- the 5x5 U(1) Heisenberg/XXZ ``W[wl, p_out, p_in, wr]`` of reference
  ``examples/heisenberg_walkthrough.py:101-176`` (physical charge ``t = 2 S^z`` in
  ``{-1, +1}``, MPO bond ``{-2: 1, 0: 3, +2: 1}``), with the ``S^z S^z`` term scaled by
  ``delta`` for XXZ.
* ``heisenberg_mpo`` -- ``MPO.from_w`` on that ``W``.
* ``gaussian_bond`` -- the U(1) bond space with a Gaussian charge profile capped by what
  the chain can reach, used to synthesize MPS of a prescribed bond dimension.
"""

from __future__ import annotations

import math

import numpy as np

from tenet.network.mps import MPO, MPS
from tenet.space import GradedSpace
from tenet.symmetry import U1, U1Sector

__all__ = ["PHYS", "BOUNDARY", "MPO_BOND", "spin_half_w", "heisenberg_mpo", "gaussian_bond",
           "chain_bonds", "random_chain"]

PHYS = GradedSpace.new(U1, {U1Sector(-1): 1, U1Sector(1): 1})
BOUNDARY = GradedSpace.new(U1, {U1Sector(0): 1})
MPO_BOND = GradedSpace.new(U1, {U1Sector(0): 3, U1Sector(2): 1, U1Sector(-2): 1})
# dense channel positions of MPO_BOND (sectors sorted: -2, 0, 0, 0, +2)
SP_CH, END, SZ_CH, START, SM_CH = 0, 1, 2, 3, 4


def spin_half_w(delta: float = 1.0, j: float = 1.0) -> np.ndarray:
    """``H = J sum_i (delta S^z_i S^z_{i+1} + (S^+_i S^-_{i+1} + S^-_i S^+_{i+1}) / 2)``."""
    eye = np.eye(2)
    sz = np.diag([-0.5, 0.5])
    sp = np.array([[0.0, 0.0], [1.0, 0.0]])
    sm = sp.T
    w = np.zeros((5, 2, 2, 5))
    w[START, :, :, START] = eye
    w[END, :, :, END] = eye
    w[START, :, :, SM_CH] = 0.5 * j * sm
    w[SM_CH, :, :, END] = sp
    w[START, :, :, SP_CH] = 0.5 * j * sp
    w[SP_CH, :, :, END] = sm
    w[START, :, :, SZ_CH] = delta * j * sz
    w[SZ_CH, :, :, END] = sz
    return w


def heisenberg_mpo(n_sites: int, delta: float = 1.0, device=None) -> MPO:
    return MPO.from_w(spin_half_w(delta), n_sites, phys=PHYS, bond=MPO_BOND, boundary=BOUNDARY,
                      start=START, end=END, device=device)


def gaussian_bond(i: int, n_sites: int, max_dim: int, sigma: float, total: int = 0) -> GradedSpace:
    """Bond ``i`` (left of site ``i``) of a spin-1/2 chain at total charge ``total``.

    Allowed charges have the parity of ``i`` and are reachable from both ends; each sector
    is capped by its binomial count ``min(C(i, (i+q)/2), C(L-i, (L-i-(total-q))/2))``.  The
    dimension ``min(max_dim, sum of caps)`` is spread with weights ``exp(-q^2 / 2 sigma^2)``
    (capped sectors' excess redistributed), rounded to integers summing exactly to it;
    sectors rounding to zero are dropped.
    """
    right = n_sites - i
    caps = {}
    for q in range(-i, i + 1, 2):
        qr = total - q
        if abs(qr) > right or (right + qr) % 2:
            continue
        c = min(math.comb(i, (i + q) // 2), math.comb(right, (right + qr) // 2))
        if c > 0:
            caps[q] = c
    dim = min(max_dim, sum(caps.values()))
    alloc = {q: 0.0 for q in caps}
    free = set(caps)
    remaining = float(dim)
    while free and remaining > 1e-9:
        wsum = sum(math.exp(-q * q / (2 * sigma * sigma)) for q in free)
        capped = []
        for q in sorted(free):
            share = remaining * math.exp(-q * q / (2 * sigma * sigma)) / wsum
            if alloc[q] + share >= caps[q]:
                capped.append(q)
        if not capped:
            for q in free:
                alloc[q] += remaining * math.exp(-q * q / (2 * sigma * sigma)) / wsum
            remaining = 0.0
            break
        for q in capped:
            remaining -= caps[q] - alloc[q]
            alloc[q] = float(caps[q])
            free.discard(q)
    m = {q: min(caps[q], int(math.floor(x + 0.5))) for q, x in alloc.items()}
    diff = dim - sum(m.values())
    order = sorted(alloc, key=lambda q: (-(alloc[q] - math.floor(alloc[q])), abs(q), q))
    k = 0
    while diff != 0 and k < 10 * len(order) + 10:
        q = order[k % len(order)]
        if diff > 0 and m[q] < caps[q]:
            m[q] += 1
            diff -= 1
        elif diff < 0 and m[q] > 0:
            m[q] -= 1
            diff += 1
        k += 1
    return GradedSpace.new(U1, {U1Sector(q): v for q, v in m.items() if v > 0})


def chain_bonds(n_sites: int, max_dim: int, sigma: float = 6.0) -> list:
    bonds = [BOUNDARY]
    bonds += [gaussian_bond(i, n_sites, max_dim, sigma) for i in range(1, n_sites)]
    bonds.append(BOUNDARY)
    return bonds


def random_chain(n_sites: int, max_dim: int, sigma: float = 6.0, seed: int = 0, device=None) -> MPS:
    return MPS.random(PHYS, chain_bonds(n_sites, max_dim, sigma), seed=seed, device=device)
