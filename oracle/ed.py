"""ORACLE -- test infrastructure only.  Exact ground-state energies of open spin-1/2 XXZ
chains, the independent pins for the DMRG path.

* ``xxz_sz0_energy`` -- sparse exact diagonalization in the S^z_tot = 0 block, the
  generalization to XXZ of the reference's ``heisenberg_sz0`` oracle
  (``tests/integration/test_dmrg.py:120-139``), ``H = sum_i (delta Sz Sz + (S+S- + S-S+)/2)``.
* ``xxz_sz0_levels`` -- the ``k`` lowest levels of the same block (dense ``eigvalsh``), the
  pins of the excited-state sweep (reference ``dmrg.py:620-843``, ``orthogonal_to``).
* ``heisenberg_singlet_levels`` -- the total-spin-0 levels (the S^z = 0 spectrum less the
  S^z = 1 spectrum), the pins of an SU(2) excited-state sweep.
* ``xx_energy`` -- closed form at delta = 0 (free fermions by Jordan-Wigner):
  ``E0 = sum_{k: eps_k < 0} eps_k``, ``eps_k = cos(pi k / (L+1))``.
* ``xx_first_excited`` -- the lowest excitation at fixed particle number in the same closed
  form: the highest occupied mode ``k = L/2 + 1`` moved to the lowest empty one ``k = L/2``,
  ``E1 = E0 + eps_{L/2} - eps_{L/2+1} = E0 + 2 cos(pi L / (2 (L+1)))``.
* Reference literals: ``E_OBC_12``, ``E_OBC_20`` (``test_dmrg.py:82-83``) and the MPSKit
  ``N32``/``N64`` energies of ``tests/fixtures/mpskit_heisenberg.json``.
"""

from __future__ import annotations

from itertools import combinations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

E_OBC_12 = -5.142090632840532
E_OBC_20 = -8.682473334398956
E_MPSKIT_N32 = -13.997315618224897
E_MPSKIT_N64 = -28.175424859745323


def _xxz_sz0_matrix(n_sites: int, delta: float):
    if n_sites % 2:
        raise ValueError("S^z_tot = 0 needs an even chain")
    return _xxz_block(n_sites, delta, n_sites // 2)


def _xxz_block(n_sites: int, delta: float, n_up: int):
    states = []
    for ups in combinations(range(n_sites), n_up):
        s = 0
        for u in ups:
            s |= 1 << u
        states.append(s)
    index = {s: i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for i, s in enumerate(states):
        diag = 0.0
        for b in range(n_sites - 1):
            x = (s >> b) & 1
            y = (s >> (b + 1)) & 1
            diag += delta * (0.25 if x == y else -0.25)
            if x != y:
                t = s ^ (1 << b) ^ (1 << (b + 1))
                rows.append(index[t])
                cols.append(i)
                vals.append(0.5)
        rows.append(i)
        cols.append(i)
        vals.append(diag)
    n = len(states)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def xxz_sz0_levels(n_sites: int, delta: float = 1.0, k: int = 2) -> np.ndarray:
    return np.linalg.eigvalsh(_xxz_sz0_matrix(n_sites, delta).toarray())[:k]


def heisenberg_singlet_levels(n_sites: int, k: int = 2) -> np.ndarray:
    """The ``k`` lowest total-spin-0 levels of the open Heisenberg chain: every multiplet
    of spin S >= 1 has one state in the S^z = 0 block and one in the S^z = 1 block, a
    singlet only the first, so the singlet levels are the S^z = 0 spectrum with the
    S^z = 1 spectrum removed level by level."""
    e0 = np.linalg.eigvalsh(_xxz_block(n_sites, 1.0, n_sites // 2).toarray())
    e1 = list(np.linalg.eigvalsh(_xxz_block(n_sites, 1.0, n_sites // 2 + 1).toarray()))
    out = []
    for e in e0:
        j = next((i for i, x in enumerate(e1) if abs(x - e) < 1e-9), None)
        if j is None:
            out.append(float(e))
            if len(out) == k:
                break
        else:
            e1.pop(j)
    return np.asarray(out)


def xxz_sz0_energy(n_sites: int, delta: float = 1.0) -> float:
    h = _xxz_sz0_matrix(n_sites, delta)
    n = h.shape[0]
    if n < 200:
        return float(np.linalg.eigvalsh(h.toarray())[0])
    return float(spla.eigsh(h, k=1, which="SA", tol=1e-14)[0][0])


def xx_energy(n_sites: int) -> float:
    eps = np.cos(np.pi * np.arange(1, n_sites + 1) / (n_sites + 1))
    return float(eps[eps < 0].sum())


def xx_first_excited(n_sites: int) -> float:
    if n_sites % 2:
        raise ValueError("S^z_tot = 0 needs an even chain")
    return xx_energy(n_sites) + 2.0 * float(np.cos(np.pi * n_sites / (2 * (n_sites + 1))))
