"""ORACLE -- test infrastructure only.  NumPy float64 restatement of the reference's U(1)
DMRG path, on ``oracle.u1`` tensors:

* ``heff2``       -- ``_heff2_full`` (reference ``network/env.py:416-432``)
* ``env_update``  -- ``Env.update_`` site-tensor branch (``env.py:593-633``)
* ``lanczos``     -- ``network/dmrg.py:106-227`` (no re-orthogonalization, as there)
* ``svd_truncated`` with ``_spectrum``/``_admissible``/``keep_rule`` (``ops/linalg.py:1079-1324,
  1441-1612``); the keep rule takes a ``qdim`` weight, so it also states the SU(2) rule
* ``lq``/``canonize`` (``linalg.py:549-620``, ``network/mps.py:312-352``)
* ``sweep``/``dmrg`` (``network/dmrg.py:321-843``, noiseless)
* ``mpo_from_w``  -- ``MPO.from_w`` (``mps.py:2780-2845``) for U(1).
"""

from __future__ import annotations

import numpy as np

from oracle import u1 as O


def heff2(fl, w1, w2, fr, aa):
    t = O.einsum("apqr,rys->apqys", aa, fr)
    t = O.composed("apqys,mQqy->apQms", t, w2, bend="q")
    t = O.composed("apQms,xPpm->aPQxs", t, w1, bend="p")
    return O.composed("aPQxs,axB->BPQs", t, fl, bend="a")


def env_update(F, a, w, bra_site, to: str):
    bra = O.adjoint(bra_site)
    if to == "last":
        t = O.einsum("axB,apr->xBpr", F, a)
        t = O.einsum("xPpm,xBpr->BrPm", w, t)
        return O.einsum("BPs,BrPm->rms", bra, t)
    t = O.einsum("apr,rys->apys", a, F)
    t = O.composed("apys,xPpy->axPs", t, w, bend="p")
    return O.composed("axPs,BPs->axB", t, bra, bend="P")


def lanczos(matvec, v, ncv=3, tol=1e-13):
    vecs = [O.scale(v, 1.0 / O.norm(v))]
    alphas, betas = [], []
    for j in range(ncv):
        w = matvec(vecs[j])
        alphas.append(O.inner(vecs[j], w))
        w = O.axpy(-alphas[j], vecs[j], w)
        if j:
            w = O.axpy(-betas[j - 1], vecs[j - 1], w)
        beta = O.norm(w)
        if j + 1 == ncv or beta < tol:
            break
        betas.append(beta)
        vecs.append(O.scale(w, 1.0 / beta))
    tri = np.diag(alphas) + np.diag(betas, 1) + np.diag(betas, -1)
    vals, states = np.linalg.eigh(tri)
    g = states[:, 0]
    out = O.scale(vecs[0], float(g[0]))
    for k in range(1, len(vecs)):
        out = O.axpy(float(g[k]), vecs[k], out)
    return float(vals[0]), O.scale(out, 1.0 / O.norm(out))


def _spectrum(values: dict):
    e = [(float(s), c, i) for c, arr in values.items() for i, s in enumerate(arr)]
    e.sort(key=lambda x: (-x[0], x[1], x[2]))
    return e


def _admissible(spec, cutoff, mode, qdim=None):
    """Length of the prefix of ``spec`` that the cutoff admits (reference
    ``ops/linalg.py:1100-1129``): bare sigma for abs/rel, qdim-weighted powers of sigma for
    the cumulative modes, whose relative forms compare against the same power."""
    qdim = qdim or (lambda c: 1.0)
    if cutoff is None:
        return len(spec)
    if mode == "abs":
        return sum(1 for s, _, _ in spec if s > cutoff)
    if mode == "rel":
        return sum(1 for s, _, _ in spec if s > cutoff * spec[0][0])
    p = 2 if mode in ("sum2", "rsum2") else 1
    w = [qdim(c) * s**p for s, c, _ in spec]
    thr = cutoff * sum(w) if mode.startswith("r") else cutoff
    kept, dropped = len(spec), 0.0
    for x in reversed(w):
        if dropped + x >= thr:
            break
        dropped += x
        kept -= 1
    return kept


def keep_rule(spec, max_bond, cutoff, mode="rsum2", qdim=None):
    """The truncation keep rule (reference ``ops/linalg.py:1274-1324``): walk the admitted
    prefix, charging ``qdim(c)`` of the dense budget per value, and stop at the first value
    that would overflow ``max_bond``.  Returns ``(kept entries, dense dimension, discarded
    weight sum qdim sigma^2)``."""
    qdim = qdim or (lambda c: 1.0)
    adm = _admissible(spec, cutoff, mode, qdim)
    kept, dense = [], 0.0
    for e in spec[:adm]:
        if max_bond is not None and dense + qdim(e[1]) > max_bond:
            break
        dense += qdim(e[1])
        kept.append(e)
    dw = sum(qdim(c) * s * s for s, c, _ in spec[len(kept):])
    return kept, dense, dw


def decide(spec, max_bond, cutoff, mode="rsum2", qdim=None):
    """The keep rule as ``({sector: kept count}, discarded weight)``."""
    kept, _, dw = keep_rule(spec, max_bond, cutoff, mode, qdim)
    keep = {}
    for _, c, _ in kept:
        keep[c] = keep.get(c, 0) + 1
    return keep, dw


def svd_truncated(t, left, right, max_bond=None, cutoff=None, mode="rsum2"):
    m = O.repartition(t, left, right)
    mats = O.to_matrices(m)
    parts = {c: np.linalg.svd(b, full_matrices=False) for c, b in mats.items()}
    spec = _spectrum({c: p[1] for c, p in parts.items()})
    keep, dw = decide(spec, max_bond, cutoff, mode)
    bond = O.space(keep)
    nl = len(left)
    ulegs = tuple(m.legs[:nl]) + (O.OLeg(bond, False, False),)
    slegs = (O.OLeg(bond, True, False), O.OLeg(bond, False, False))
    vlegs = (O.OLeg(bond, True, False),) + tuple(m.legs[nl:])
    u = O.from_matrices(ulegs, {c: parts[c][0][:, :k] for c, k in keep.items()})
    s = O.from_matrices(slegs, {c: np.diag(parts[c][1][:k]) for c, k in keep.items()})
    vh = O.from_matrices(vlegs, {c: parts[c][2][:k, :] for c, k in keep.items()})
    return u, s, vh, spec, dw


def lq(t, left, right):
    m = O.repartition(t, left, right)
    mats = O.to_matrices(m)
    keep = {c: min(b.shape) for c, b in mats.items()}
    bond = O.space(keep)
    nl = len(left)
    ls, qs = {}, {}
    for c, b in mats.items():
        q, r = np.linalg.qr(b.T)
        ls[c], qs[c] = r.T, q.T
    L = O.from_matrices(tuple(m.legs[:nl]) + (O.OLeg(bond, False, False),), ls)
    Q = O.from_matrices((O.OLeg(bond, True, False),) + tuple(m.legs[nl:]), qs)
    return L, Q


def as_site(t):
    return O.repartition(t, (0, 1), (2,))


def canonize(sites):
    sites = list(sites)
    for n in range(len(sites) - 1, 0, -1):
        L, Q = lq(sites[n], (0,), (1, 2))
        sites[n] = as_site(Q)
        sites[n - 1] = as_site(O.einsum("apx,xy->apy", sites[n - 1], L))
    sites[0] = O.scale(sites[0], 1.0 / O.norm(sites[0]))
    return sites


def mpo_from_w(w, n_sites, phys, bond, boundary, start, end):
    """U(1) ``from_dense``: every block is a slice of the dense W at its sectors' offsets."""

    def make(arr, left, right):
        legs = (O.OLeg(left, False), O.OLeg(phys, True), O.OLeg(phys, False), O.OLeg(right, True))
        offs = []
        for l in legs:
            o, acc = {}, 0
            for q, m in l.sectors:
                o[q] = acc
                acc += m
            offs.append(o)
        blocks = {}
        for k in O.block_keys(legs):
            sl = tuple(slice(o[q], o[q] + l.deg(q)) for o, l, q in zip(offs, legs, k))
            blocks[k] = np.array(arr[sl], dtype=np.float64)
        return O.OTensor(legs, blocks)

    w = np.asarray(w, dtype=np.float64)
    first = make(w[start : start + 1], boundary, bond)
    bulk = make(w, bond, bond)
    last = make(w[:, :, :, end : end + 1], bond, boundary)
    return [first] + [bulk] * (n_sites - 2) + [last]


def boundary_envs(sites, mpo):
    n = len(sites)
    kl, kr = sites[0].legs[0], sites[-1].legs[2]
    wl, wr = mpo[0].legs[0], mpo[-1].legs[3]
    left = O.OTensor.ones((O.OLeg(kl.sectors, False, kl.dual), O.OLeg(wl.sectors, True, wl.dual),
                           O.OLeg(kl.sectors, True, kl.dual)))
    right = O.OTensor.ones((O.OLeg(kr.sectors, True, kr.dual), O.OLeg(wr.sectors, False, wr.dual),
                            O.OLeg(kr.sectors, False, kr.dual)))
    return {(-1, 0): left, (n, n - 1): right}


def spectrum(s):
    vals = []
    for b in s.blocks.values():
        vals.extend(np.diag(b).tolist())
    return sorted(vals, reverse=True)


def dmrg(sites, mpo, chi, cutoff=1e-14, max_sweeps=10, energy_tol=1e-12, schmidt_tol=1e-8, ncv=3):
    """Noiseless two-site DMRG; returns (energy, history)."""
    sites = canonize(sites)
    n = len(sites)
    F = boundary_envs(sites, mpo)
    for k in range(n - 1, 0, -1):
        F[k, k - 1] = env_update(F[k + 1, k], sites[k], mpo[k], sites[k], "first")
    schmidt, energy, history = {}, float("inf"), []
    for _ in range(max_sweeps):
        old_e, old_s = energy, dict(schmidt)
        for direction in ("right", "left"):
            fwd = direction == "right"
            for b in (range(n - 1) if fwd else range(n - 2, -1, -1)):
                aa = O.einsum("apx,xqr->apqr", sites[b], sites[b + 1])
                fl, fr = F[b - 1, b], F[b + 2, b + 1]
                energy, aa = lanczos(lambda v: heff2(fl, mpo[b], mpo[b + 1], fr, v), aa, ncv)
                u, s, vh, _, _ = svd_truncated(aa, (0, 1), (2, 3), chi, cutoff)
                s = O.scale(s, 1.0 / O.norm(s))
                vh = as_site(vh)
                if fwd:
                    sites[b] = as_site(u)
                    sites[b + 1] = as_site(O.einsum("xy,yqr->xqr", s, vh))
                else:
                    sites[b] = as_site(O.einsum("apx,xy->apy", u, s))
                    sites[b + 1] = vh
                schmidt[b] = spectrum(s)
                for k in ((b, b - 1), (b, b + 1), (b + 1, b), (b + 1, b + 2)):
                    F.pop(k, None)
                if fwd:
                    F[b, b + 1] = env_update(F[b - 1, b], sites[b], mpo[b], sites[b], "last")
                else:
                    F[b + 1, b] = env_update(F[b + 2, b + 1], sites[b + 1], mpo[b + 1], sites[b + 1], "first")
        de = abs(old_e - energy)
        ds = float("inf")
        if old_s:
            ds = max(
                float(np.linalg.norm(np.pad(schmidt[k], (0, max(0, len(old_s.get(k, [])) - len(schmidt[k]))))
                                     - np.pad(old_s.get(k, []), (0, max(0, len(schmidt[k]) - len(old_s.get(k, [])))))))
                for k in schmidt
            )
        history.append((energy, de, ds))
        if de < energy_tol and ds < schmidt_tol:
            break
    return energy, history
