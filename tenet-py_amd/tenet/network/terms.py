"""Local operators and Hamiltonians written as sums of terms.

* ``local_op(dense, phys=, charge=)`` -- a dense operator as a term operator (reference
  ``network/mps.py:960-1041``): rank 3 on ``(phys OUT, phys IN, charge OUT)`` when the
  operator changes an abelian charge (the charge it emits onto the MPO bond), or the
  invariant rank-2k form ``(phys OUT)*k + (phys IN)*k`` of one whole k-site term.
* ``spin_half(U1 | SU2)`` -- the spin-1/2 site (reference ``models/sites.py:121-194``):
  under U(1) (charge ``2 S^z``) ``Sz``, ``S+``, ``S-`` in the charge-leg form and the two-site
  ``S.S`` as a matrix; under SU(2) the invariant two-site ``S.S`` alone.
* ``build_mpo(n_sites, terms)`` -- the MPO of ``sum_t coef_t prod_f op_f`` behind
  ``MPO.from_terms`` (reference ``mps.py:3135-3342``), as a finite-state machine: at every
  cut the MPO bond carries the two corner channels ("nothing applied yet", "everything
  applied") plus one channel per open term, each channel a graded space of its own; every
  site tensor is assembled densely from the pieces that act on it and projected onto its
  symmetric blocks with ``from_dense`` (which refuses anything not symmetric).

A term is ``(coef, [(op, site), ...])`` with rank-3 charge-leg factors on distinct sites
(their charges summing to zero), or ``(coef, [(op, (i, i + 1))])`` with one invariant
two-site operator.  An invariant two-site operator is split into per-site pieces by its
symmetry: for an abelian symmetry by charge transfer, each transfer block by a dense SVD
(the channel space holds its nonzero singular values); for SU(2) by the spherical tensor
operators of the site, ``T = sum_k c_k sum_q (-1)^q T^k_{-q} (x) T^k_q`` (one spin-k
multiplet channel per nonzero ``c_k``).  Channels are not compressed across terms.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from tenet.space import IN, OUT, GradedSpace, Leg
from tenet.symmetry import SU2, U1, SU2Sector, U1Sector

__all__ = ["Site", "spin_half", "local_op", "build_mpo", "Channel"]


@dataclass(frozen=True)
class Site:
    """A lattice site: its physical space, its term operators and their dense matrices."""

    phys: GradedSpace
    ops: dict
    matrices: dict


# spin-1/2 in charge order (|down>, |up>) -- the sorted U(1) sectors -1, +1
_SZ = np.diag([-0.5, 0.5])
_SP = np.array([[0.0, 0.0], [1.0, 0.0]])
_SS = np.kron(_SZ, _SZ) + (np.kron(_SP, _SP.T) + np.kron(_SP.T, _SP)) / 2


def local_op(dense, *, phys: GradedSpace, charge=None, device=None):
    """A dense operator as a term operator (reference ``network/mps.py:960-1041``).

    With ``charge``: a ``(d, d)`` array as rank 3 on ``(phys OUT, phys IN, charge OUT)``;
    without: an invariant ``k``-site term, ``(d**k, d**k)`` or ``(d,)*2k``, as rank
    ``2k`` on ``(phys OUT)*k + (phys IN)*k``.  A symmetry-forbidden array raises in
    ``from_dense``."""
    from tenet.dense import from_dense

    d = phys.dim
    a = np.asarray(dense, dtype=np.float64)
    if charge is not None:
        if a.shape != (d, d):
            raise ValueError(f"local_op: expected a ({d}, {d}) array on this phys, got {a.shape}")
        emitted = Leg(GradedSpace.new(phys.provider, {charge: 1}), OUT)
        return from_dense(a.reshape(d, d, 1), (Leg(phys, OUT), Leg(phys, IN), emitted), device=device)
    k = next((k for k in range(1, 9) if a.shape in ((d**k, d**k), (d,) * (2 * k))), None)
    if k is None:
        raise ValueError(f"local_op: with charge=None the array is an invariant k-site term, so it is "
                         f"({d}**k, {d}**k) or ({d},)*2k on this phys (d={d}); got {a.shape}")
    legs = (Leg(phys, OUT),) * k + (Leg(phys, IN),) * k
    return from_dense(a.reshape((d,) * (2 * k)), legs, device=device)


def spin_half(symmetry=U1, device=None) -> Site:
    """The spin-1/2 site graded by ``U1`` (charge ``2 S^z``) or ``SU2``."""
    if symmetry is U1:
        phys = GradedSpace.new(U1, {U1Sector(-1): 1, U1Sector(1): 1})
        table = {"Sz": (_SZ, U1Sector(0)), "S+": (_SP, U1Sector(-2)), "S-": (_SP.T, U1Sector(2))}
        ops_ = {n: local_op(m, phys=phys, charge=q, device=device) for n, (m, q) in table.items()}
        return Site(phys, ops_, {**{n: m for n, (m, _) in table.items()}, "S.S": _SS})
    if symmetry is SU2:
        phys = GradedSpace.new(SU2, {SU2Sector(1): 1})
        # the SU(2) dense basis runs m descending, (|up>, |down>): reverse the charge order
        flip = np.ix_([3, 2, 1, 0], [3, 2, 1, 0])
        ss = _SS[flip]
        return Site(phys, {"S.S": local_op(ss, phys=phys, device=device)}, {"S.S": ss})
    raise ValueError(f"spin_half: the shipped gradings are U1 (charge 2 S^z) and SU2, got {symmetry!r}")


# ----------------------------------------------------------------------------- channels


@dataclass
class Channel:
    """One open term's MPO-bond channel at one cut: a graded space (``sector -> slots``)."""

    sectors: dict


def _dense(t) -> np.ndarray:
    from tenet.dense import to_dense

    return to_dense(t)


def _slab(provider, space: GradedSpace) -> dict:
    """Dense offset of every sector's slab in ``space`` (slots x irrep dimension)."""
    o, acc = {}, 0
    for a, m in space.sectors:
        o[a] = acc
        acc += m * provider.irrep_dim(a)
    return o


def _charge_factor(op):
    """``(dense d x d, emitted sector)`` of a rank-3 charge-leg operator."""
    if op.ndim != 3 or len(op.legs[2].space.sectors) != 1 or op.legs[2].space.dim != 1:
        raise ValueError("a single-site term factor must be a rank-3 charge-leg operator from local_op")
    return _dense(op)[:, :, 0], op.legs[2].space.sectors[0][0]


def _split_abelian(provider, phys, t4: np.ndarray, tol: float):
    """Invariant two-site ``t4[p1o, p1i, p2o, p2i]`` (abelian) as
    ``[(sector, L[p1o, p1i, r], R[r, p2o, p2i])]`` by charge transfer and SVD."""
    q = np.concatenate([[a] * m for a, m in phys.sectors])  # sector of every dense index
    d = len(q)
    pieces = []
    seen = {}
    for i in range(d):
        for j in range(d):
            # the emitted charge: q(p_out) + c = q(p_in)
            c = _transfer(provider, q[i], q[j])
            seen.setdefault(c, []).append((i, j))
    total = float(np.linalg.norm(t4))
    covered = 0.0
    for c, pairs in sorted(seen.items()):
        back = [(i, j) for i in range(d) for j in range(d) if _transfer(provider, q[i], q[j]) == provider.dual(c)]
        if not back:
            continue
        m = np.array([[t4[i, j, k, l] for k, l in back] for i, j in pairs])
        covered += float(np.sum(m * m))
        u, s, vt = np.linalg.svd(m, full_matrices=False)
        keep = s > tol * max(total, 1e-300)
        if not keep.any():
            continue
        L = np.zeros((d, d, int(keep.sum())))
        R = np.zeros((int(keep.sum()), d, d))
        for r, (uc, sc, vc) in enumerate(zip(u.T[keep], s[keep], vt[keep])):
            for (i, j), x in zip(pairs, uc * sc):
                L[i, j, r] = x
            for (k, l), x in zip(back, vc):
                R[r, k, l] = x
        pieces.append((c, L, R))
    if abs(covered - total * total) > 1e-20 + 1e-12 * total * total:
        raise ValueError("from_terms: the two-site operator is not charge conserving")
    return pieces


def _transfer(provider, q_out, q_in):
    """The sector ``c`` with ``q_out (x) c = q_in`` for an abelian provider."""
    cs = provider.fusion(provider.dual(q_out), q_in)
    (c,) = cs
    return c


def _spherical(provider, j) -> dict:
    """``{k: [T^k_q for q = k .. -k]}`` for a spin-``j`` site: ``T^k_q[m, m'] = <j m | j m'; k q>``
    (the provider's Clebsch-Gordan tensor, m descending)."""
    out = {}
    for k2 in range(0, 2 * j.two_j + 1, 2):
        if not provider.n_symbol(j, SU2Sector(k2), j):
            continue
        cg = provider.cgc(j, SU2Sector(k2), j)  # [m', q, m]
        out[k2] = [cg[:, iq, :].T.copy() for iq in range(k2 + 1)]
    return out


def _split_su2(provider, phys, t4: np.ndarray, tol: float):
    """Invariant two-site ``t4[p1o, p1i, p2o, p2i]`` on one spin-j multiplet per site as
    ``[(SU2Sector(k), L[p1o, p1i, m], R[m, p2o, p2i])]``: the channel of spin k carries
    ``(-1)^q T^k_{-q}`` on the left piece (the bond leg leaves it) and ``T^k_q`` on the right,
    with ``c_k`` the projection of ``t4`` onto ``sum_q (-1)^q T^k_{-q} (x) T^k_q``."""
    if len(phys.sectors) != 1 or phys.sectors[0][1] != 1:
        raise ValueError("from_terms: SU(2) two-site terms need one multiplet per site")
    j = phys.sectors[0][0]
    rest = t4.copy()
    pieces = []
    for k2, T in _spherical(provider, j).items():
        n = k2 + 1
        qs = [k2 - 2 * i for i in range(n)]  # 2q, descending
        left = np.stack([(-1) ** ((k2 - q2) // 2 % 2) * T[n - 1 - i] for i, q2 in enumerate(qs)], axis=-1)
        right = np.stack(T, axis=0)
        # (-1)^q with q = q2 / 2 integer; T[n-1-i] is T^k_{-q}
        basis = np.einsum("abm,mcd->abcd", left, right)
        nb = float(np.sum(basis * basis))
        ck = float(np.sum(basis * t4)) / nb
        if abs(ck) <= tol * max(float(np.linalg.norm(t4)), 1e-300):
            continue
        rest -= ck * basis
        pieces.append((SU2Sector(k2), ck * left, right))
    if float(np.linalg.norm(rest)) > 1e-10 * max(float(np.linalg.norm(t4)), 1.0):
        raise ValueError("from_terms: the two-site operator is not SU(2) invariant")
    return pieces


# ----------------------------------------------------------------------------- assembly


def _normalize_terms(n_sites, terms):
    """Each term as ``(first site, [dense piece per site], [channel graded dict per cut])``;
    a piece is ``(p_out, p_in)`` for a one-slot chain or ``(slots_in, p_out, p_in, slots_out)``
    arrays in the channel's own dense basis."""
    out = []
    for t in terms:
        coef, factors = t
        factors = list(factors)
        if not factors:
            raise ValueError("from_terms: a term needs at least one factor")
        op0, where0 = factors[0]
        if isinstance(where0, tuple):
            if len(factors) != 1 or len(where0) != 2 or where0[1] != where0[0] + 1:
                raise ValueError("from_terms: an invariant operator spans two neighbouring sites (i, i + 1) "
                                 "and is the term's only factor")
            out.append(("pair", float(coef), op0, where0[0]))
            continue
        sites = [w for _, w in factors]
        if sorted(sites) != sites or len(set(sites)) != len(sites):
            raise ValueError("from_terms: the factors of a term act on distinct sites in increasing order")
        if sites[0] < 0 or sites[-1] >= n_sites:
            raise ValueError(f"from_terms: a factor acts outside sites 0..{n_sites - 1}")
        out.append(("chain", float(coef), factors, sites[0]))
    return out


def build_mpo(n_sites: int, terms, *, tol: float = 1e-13, device=None):
    """The MPO site tensors of ``sum_t coef_t prod_f op_f`` and the channel table
    ``[(cut, term index, Channel)]`` (see the module docstring)."""
    from tenet.dense import from_dense

    if n_sites < 2:
        raise ValueError("from_terms: an MPO needs at least two sites")
    parsed = _normalize_terms(n_sites, terms)
    phys = None
    for kind, _, f, _ in parsed:
        op = f if kind == "pair" else f[0][0]
        phys = op.legs[0].space if phys is None else phys
        if op.legs[0].space != phys:
            raise ValueError("from_terms: every operator must act on the same physical space")
    provider = phys.provider
    unit = provider.unit
    d = phys.dim
    eye = np.eye(d)
    # per term: pieces[site] = (dense block, in-channel, out-channel); channels per cut
    chans = {}   # (cut, term) -> {sector: slots}
    pieces = []  # (site, term, in-key, out-key, array[slots_in_dense?, d, d, ...])
    for ti, (kind, coef, f, s0) in enumerate(parsed):
        if kind == "pair":
            dense = _dense(f).transpose(0, 2, 1, 3)  # (p1o, p2o, p1i, p2i) -> (p1o, p1i, p2o, p2i)
            split = (_split_su2 if provider is SU2 else _split_abelian)(provider, phys, dense, tol)
            if s0 < 0 or s0 + 1 >= n_sites:
                raise ValueError(f"from_terms: a two-site term acts outside sites 0..{n_sites - 1}")
            secs = {}
            for sec, L, R in split:
                secs[sec] = secs.get(sec, 0) + L.shape[-1] // provider.irrep_dim(sec)
            if not secs:
                continue
            chans[(s0, ti)] = secs
            pieces.append((s0, ti, "L", ("term", s0, ti), coef, split))
            pieces.append((s0 + 1, ti, ("term", s0, ti), "R", 1.0, split))
            continue
        mats = []
        charge = unit
        sites = [w for _, w in f]
        for op, _ in f:
            m, c = _charge_factor(op)
            mats.append((m, c))
        # channel charges at every cut the term spans
        at = dict(zip(sites, mats))
        q = unit
        for n in range(sites[0], sites[-1] + 1):
            m_c = at.get(n)
            prev = q
            if m_c is not None:
                (q,) = provider.fusion(q, m_c[1])
            if n < sites[-1]:
                chans[(n, ti)] = {q: 1}
            block = (m_c[0] * (coef if n == sites[0] else 1.0)) if m_c is not None else eye
            kin = "L" if n == sites[0] else ("term", n - 1, ti)
            kout = "R" if n == sites[-1] else ("term", n, ti)
            pieces.append((n, ti, kin, kout, 1.0, ("chain", block, prev, q)))
        if q != unit:
            raise ValueError("from_terms: a term's factor charges do not sum to the unit sector")
        charge = q
        _ = charge
    # graded space of every cut: the corner channels in the unit sector, then open terms
    cut_sp, cut_slot = [], []
    for n in range(-1, n_sites):
        counts = {}
        slot = {}
        if n < n_sites - 1:  # "nothing applied yet" continues to the right
            slot["L"] = (unit, counts.get(unit, 0))
            counts[unit] = counts.get(unit, 0) + 1
        if n >= 0:  # "everything applied" came from the left
            slot["R"] = (unit, counts.get(unit, 0))
            counts[unit] = counts.get(unit, 0) + 1
        for (c, ti), secs in sorted(chans.items(), key=lambda kv: kv[0][1]):
            if c != n:
                continue
            slot[("term", n, ti)] = {a: counts.get(a, 0) for a in secs}
            for a, mm in secs.items():
                counts[a] = counts.get(a, 0) + mm
        cut_sp.append(GradedSpace.new(provider, counts))
        cut_slot.append(slot)

    def index(n_cut, key, sector, alpha, m):
        sp = cut_sp[n_cut + 1]
        off = _slab(provider, sp)
        s = cut_slot[n_cut + 1][key]
        base = s[1] if isinstance(s, tuple) else s[sector]
        return off[sector] + (base + alpha) * provider.irrep_dim(sector) + m

    sites = []
    for n in range(n_sites):
        wl, wr = cut_sp[n], cut_sp[n + 1]
        w = np.zeros((wl.dim, d, d, wr.dim))
        if "L" in cut_slot[n] and "L" in cut_slot[n + 1]:
            w[index(n - 1, "L", unit, 0, 0), :, :, index(n, "L", unit, 0, 0)] += eye
        if "R" in cut_slot[n] and "R" in cut_slot[n + 1]:
            w[index(n - 1, "R", unit, 0, 0), :, :, index(n, "R", unit, 0, 0)] += eye
        for site, ti, kin, kout, coef, spec in pieces:
            if site != n:
                continue
            if spec[0] == "chain":
                _, block, qin, qout = spec
                i = index(n - 1, kin, qin, 0, 0)
                o = index(n, kout, qout, 0, 0)
                w[i, :, :, o] += block
                continue
            split = spec
            alpha = {}
            for sec, L, R in split:
                dim = provider.irrep_dim(sec)
                nslots = L.shape[-1] // dim
                for r in range(L.shape[-1]):
                    a0 = alpha.get(sec, 0) + r // dim
                    m = r % dim
                    if kin == "L":  # left piece: nothing applied -> open channel
                        w[index(n - 1, "L", unit, 0, 0), :, :, index(n, kout, sec, a0, m)] += coef * L[:, :, r]
                    else:  # right piece: open channel -> everything applied
                        w[index(n - 1, kin, sec, a0, m), :, :, index(n, "R", unit, 0, 0)] += R[r]
                alpha[sec] = alpha.get(sec, 0) + nslots
        legs = (Leg(wl, IN), Leg(phys, OUT), Leg(phys, IN), Leg(wr, OUT))
        sites.append(from_dense(w, legs, device=device))
    table = [(c, ti, Channel(dict(secs))) for (c, ti), secs in sorted(chans.items())]
    return sites, table
