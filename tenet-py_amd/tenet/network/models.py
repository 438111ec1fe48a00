"""Lattice models and bond spaces used by the workloads.

* ``spin_half_w`` -- the 5x5 U(1) Heisenberg/XXZ ``W[wl, p_out, p_in, wr]`` of reference
  ``examples/heisenberg_walkthrough.py:101-176`` (physical charge ``t = 2 S^z`` in
  ``{-1, +1}``, MPO bond ``{-2: 1, 0: 3, +2: 1}``), with the ``S^z S^z`` term scaled by
  ``delta`` for XXZ.
* ``heisenberg_mpo`` -- ``MPO.from_w`` on that ``W``.
* ``gaussian_bond`` -- the U(1) bond space with a Gaussian charge profile capped by what
  the chain can reach, used to synthesize MPS of a prescribed bond dimension.
* ``su2_heisenberg_w``/``su2_heisenberg_mpo`` -- the same chain with SU(2) symmetry: the
  invariant ``S.S`` MPO on the bond ``{singlet: 2, spin-1: 1}`` (start/end channels and
  the spherical components of ``S``), projected onto the symmetric blocks by
  ``from_dense`` (the reference builds it from ``spin_half(SU2)`` terms and
  ``materialize()``, ``examples/su2_heisenberg.py:27-31``).
* ``su2_gaussian_bond``/``su2_random_chain`` -- SU(2) bond spaces: spin labels ``2j`` of
  the parity of the site count, multiplicities with a Gaussian profile in ``2j`` capped by
  how many spin-``j`` multiplets the chain's two halves can form.
"""

from __future__ import annotations

import math

import numpy as np

from tenet.network.mps import MPO, MPS
from tenet.space import GradedSpace
from tenet.symmetry import SU2, U1, SU2Sector, U1Sector

__all__ = ["PHYS", "BOUNDARY", "MPO_BOND", "spin_half_w", "heisenberg_mpo", "gaussian_bond",
           "chain_bonds", "random_chain", "SU2_PHYS", "SU2_BOUNDARY", "SU2_MPO_BOND",
           "su2_heisenberg_w", "su2_heisenberg_mpo", "su2_heisenberg_mpo_w", "su2_gaussian_bond", "su2_chain_bonds",
           "su2_random_chain"]

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


SU2_PHYS = GradedSpace.new(SU2, {SU2Sector(1): 1})
SU2_BOUNDARY = GradedSpace.new(SU2, {SU2Sector(0): 1})
SU2_MPO_BOND = GradedSpace.new(SU2, {SU2Sector(0): 2, SU2Sector(2): 1})
# dense channels of SU2_MPO_BOND: two singlets, then the spin-1 triplet m = +1, 0, -1
SU2_START, SU2_END = 0, 1


def su2_heisenberg_w(j: float = 1.0) -> np.ndarray:
    """Dense ``W[wl, p_out, p_in, wr]`` of ``H = J sum_i S_i . S_{i+1}`` in the SU(2)
    dense basis (magnetic indices descending).

    ``S.S = sum_q (-1)^q T_{-q} T_q`` with the spherical components ``T_{+1} = -S^+/sqrt2``,
    ``T_0 = S^z``, ``T_{-1} = S^-/sqrt2``: on the OUT leg ``wr`` the triplet channel ``q``
    carries ``(-1)^q T_{-q}``, on the IN leg ``wl`` it carries ``T_q`` -- the two
    placements under which ``W`` is invariant."""
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T
    sz = np.diag([0.5, -0.5])
    t = {1: -sp / math.sqrt(2.0), 0: sz, -1: sm / math.sqrt(2.0)}
    w = np.zeros((5, 2, 2, 5))
    w[SU2_START, :, :, SU2_START] = np.eye(2)
    w[SU2_END, :, :, SU2_END] = np.eye(2)
    for k, q in enumerate((1, 0, -1)):
        w[SU2_START, :, :, 2 + k] = j * (-1) ** q * t[-q]
        w[2 + k, :, :, SU2_END] = t[q]
    return w


def su2_heisenberg_mpo(n_sites: int, j: float = 1.0, device=None) -> MPO:
    """``H = J sum_i S_i . S_{i+1}`` as one invariant two-site term per bond through
    ``MPO.from_terms`` and ``materialize`` (reference ``examples/su2_heisenberg.py:27-31``):
    the MPO bond is the two corner singlets plus the spin-1 channel of ``S . S``."""
    from tenet.network.terms import spin_half

    ss = spin_half(SU2, device=device).ops["S.S"]
    return MPO.from_terms(n_sites, [(j, [(ss, (i, i + 1))]) for i in range(n_sites - 1)],
                          device=device).materialize()


def su2_heisenberg_mpo_w(n_sites: int, j: float = 1.0, device=None) -> MPO:
    """The same operator from the hand-written invariant ``W`` of ``su2_heisenberg_w``."""
    return MPO.from_w(su2_heisenberg_w(j), n_sites, phys=SU2_PHYS, bond=SU2_MPO_BOND,
                      boundary=SU2_BOUNDARY, start=SU2_START, end=SU2_END, device=device)


def _multiplets(n: int, two_j: int) -> int:
    """How many spin-``j`` multiplets ``n`` spin-1/2 sites carry."""
    if two_j > n or (n - two_j) % 2:
        return 0
    k = (n - two_j) // 2
    return math.comb(n, k) - (math.comb(n, k - 1) if k >= 1 else 0)


def su2_gaussian_bond(i: int, n_sites: int, max_mult: int, sigma: float,
                      max_two_j: int | None = None) -> GradedSpace:
    """Bond ``i`` (left of site ``i``) of a singlet spin-1/2 chain: labels ``2j`` with the
    parity of ``i``, at most ``min(i, L-i)`` (and ``max_two_j``), multiplicity caps from
    both halves, ``max_mult`` multiplets spread with weights ``exp(-(2j)^2 / 2 sigma^2)``."""
    top = min(i, n_sites - i) if max_two_j is None else min(i, n_sites - i, max_two_j)
    caps = {}
    for tj in range(i % 2, top + 1, 2):
        c = min(_multiplets(i, tj), _multiplets(n_sites - i, tj))
        if c > 0:
            caps[tj] = c
    total = min(max_mult, sum(caps.values()))
    alloc = {tj: 0 for tj in caps}
    free = set(caps)
    left = total
    while free and left > 0:
        wsum = sum(math.exp(-tj * tj / (2 * sigma * sigma)) for tj in free)
        share = {tj: left * math.exp(-tj * tj / (2 * sigma * sigma)) / wsum for tj in free}
        capped = [tj for tj in free if alloc[tj] + share[tj] >= caps[tj]]
        if not capped:
            for tj in sorted(free):
                alloc[tj] += int(math.floor(share[tj] + 0.5))
            break
        for tj in capped:
            left -= caps[tj] - alloc[tj]
            alloc[tj] = caps[tj]
            free.discard(tj)
    diff = total - sum(alloc.values())
    order = sorted(caps)
    k = 0
    while diff != 0 and k < 10 * len(order) + 10:
        tj = order[k % len(order)]
        if diff > 0 and alloc[tj] < caps[tj]:
            alloc[tj] += 1
            diff -= 1
        elif diff < 0 and alloc[tj] > 0:
            alloc[tj] -= 1
            diff += 1
        k += 1
    return GradedSpace.new(SU2, {SU2Sector(tj): m for tj, m in alloc.items() if m > 0})


def su2_chain_bonds(n_sites: int, max_mult: int, sigma: float = 6.0, max_two_j: int | None = None) -> list:
    """Bond spaces of an ``n_sites`` chain; ``max_two_j`` caps even bonds at ``2j = max_two_j``
    and odd bonds at ``max_two_j + 1``."""
    cap = (lambda i: None) if max_two_j is None else (lambda i: max_two_j + i % 2)
    return [SU2_BOUNDARY] + [su2_gaussian_bond(i, n_sites, max_mult, sigma, cap(i))
                             for i in range(1, n_sites)] + [SU2_BOUNDARY]


def su2_random_chain(n_sites: int, max_mult: int, sigma: float = 6.0, seed: int = 0,
                     max_two_j: int | None = None, device=None) -> MPS:
    return MPS.random(SU2_PHYS, su2_chain_bonds(n_sites, max_mult, sigma, max_two_j), seed=seed, device=device)
