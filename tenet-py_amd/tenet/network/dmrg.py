"""The Krylov step, the two-site sweep and the DMRG loop, with FP64 Krylov scalars on the
device.

``lanczos`` follows reference ``network/dmrg.py:106-227`` (three-term recurrence, happy
breakdown below ``tol``, the ground pair of the small tridiagonal, one recombination)
with MI355X-side changes that do not change the mathematics: every Krylov scalar
(``alpha_j = <v_j|w>``, ``beta_j = |w|``) stays in device memory, the tridiagonal
eigenproblem and the recombination run on the device too (``ritz``), so the host reads
one number -- the energy -- per Lanczos step; and each new vector is fully
re-orthogonalized in FP64 against the previous ones (the reference's ncv=3 recurrence is
orthogonal in exact arithmetic; this keeps it so when the matvec is FP32).  ``refine``
takes an FP64 matvec and closes the step with ``residual_refine``: a Rayleigh-Ritz
correction in the span of the Ritz vector and its FP64 residual, which removes the FP32
operator's error from the energy and the vector.

``sweep_``/``dmrg_`` follow reference ``dmrg.py:321-843``: the noiseless SVD split, the
wavefunction-noise split, and perturbative noise (the density-matrix decimation with the
operator's own action on the two-site tensor folded into ``rho``).
"""

from __future__ import annotations

import gc
from collections.abc import Callable, Sequence
from typing import Any, NamedTuple

import numpy as np
import torch

from tenet import linalg, ops
from tenet.tensor import SymmetricTensor

__all__ = ["Sweep", "DMRG_out", "LanczosRequest", "lanczos", "sweep_", "sweep_steps", "dmrg_", "dmrg_steps",
           "spectrum"]


class Sweep(NamedTuple):
    chi: int = 64
    cutoff: float = 1e-14
    noise: float = 0.0
    noise_type: str = "wavefunction"


class DMRG_out(NamedTuple):
    sweeps: int
    energy: float
    denergy: float
    max_dSchmidt: float
    max_discarded_weight: float
    history: list
    schedule: list
    psi: Any


def lanczos(matvec: Callable[[SymmetricTensor], SymmetricTensor], v: SymmetricTensor, *,
            ncv: int = 3, tol: float = 1e-13, orthogonal_to: Sequence[SymmetricTensor] = (),
            refine: Callable[[SymmetricTensor], SymmetricTensor] | None = None,
            reorthogonalize: bool = True, host_value: bool = True) -> tuple[float, SymmetricTensor]:
    """Ground eigenpair ``(value, vector)`` of a Hermitian ``matvec``.  ``host_value=False``
    returns the value as a 0-d device tensor (no host synchronisation)."""
    basis = []
    for t in orthogonal_to:
        for b in basis:
            t = ops.subtract(t, ops.multiply(b, float(ops.inner(b, t))))
        if float(ops.norm(t)) > 1e-12:
            basis.append(ops.divide(t, ops.norm(t)))
    if basis:
        for b in basis:
            v = ops.subtract(v, ops.multiply(b, float(ops.inner(b, v))))
        if float(ops.norm(v)) < 1e-12:
            raise ValueError(
                "the start vector lies in the span of orthogonal_to, so there is no vector "
                "left to iterate on at this bond"
            )
        def projected(op):
            def apply(x):
                y = op(x)
                for b in basis:
                    y = ops.subtract(y, ops.multiply(b, ops.inner(b, y)))
                return y
            return apply

        # the FP64 refinement step runs on the same projected operator
        matvec = projected(matvec)
        if refine is not None:
            refine = projected(refine)

    dev = v.device
    alphas = torch.zeros(ncv, dtype=torch.float64, device=dev)
    betas = torch.zeros(ncv, dtype=torch.float64, device=dev)
    vecs = [ops.divide(v, ops.norm(v))]
    for j in range(ncv):
        w = matvec(vecs[j])
        if reorthogonalize and j + 1 <= ops.ORTH_MAX and w.dtype == vecs[j].dtype:
            # the three-term step and its re-orthogonalisation as two classical
            # Gram-Schmidt passes against the whole basis (one native call): the first
            # pass's coefficient of v_j is alpha_j, the norm after the second is beta_j
            if any(w.data.data_ptr() == u.data.data_ptr() for u in vecs):
                w = ops.multiply(w, 1.0)
            ops.orthogonalize_(w, vecs, passes=2, last_coef=alphas[j], norm=betas[j])
        else:
            a = ops.inner_device(vecs[j], w, alphas[j])
            w = ops.subtract(w, ops.multiply(vecs[j], a))
            if j:
                w = ops.subtract(w, ops.multiply(vecs[j - 1], betas[j - 1]))
            if reorthogonalize:
                for u in vecs:
                    w = ops.subtract(w, ops.multiply(u, ops.inner(u, w)))
            betas[j] = ops.norm(w)
        if j + 1 == ncv:
            break
        vecs.append(ops.divide(w, betas[j]))
    value, out = ritz(alphas, betas, [x.data for x in vecs], tol)
    out = SymmetricTensor(v.structure, out)
    out = ops.divide(out, ops.norm(out))
    if refine is not None:
        value, out, _ = residual_refine(refine, out)
    return (float(value) if host_value else value), out


def ritz(alphas: torch.Tensor, betas: torch.Tensor, basis, tol: float, coef=None, value=None, out=None):
    """Ground Ritz pair of a Lanczos recurrence, on the device: the tridiagonal eigenproblem
    (``tn_tridiag_ground``, happy breakdown at the first ``beta < tol``) and the
    recombination ``sum_k g_k v_k``, normalised.  Returns ``(value, vector)`` as device
    tensors; nothing is read back to the host.  ``coef``/``value``/``out`` are optional
    preallocated buffers (the captured Krylov graphs pass theirs)."""
    from tenet import _native

    n = len(basis)
    dev = alphas.device
    coef = torch.empty(n, dtype=torch.float64, device=dev) if coef is None else coef
    value = torch.empty((), dtype=torch.float64, device=dev) if value is None else value
    _native.check(_native.lib().tn_tridiag_ground(alphas.data_ptr(), betas.data_ptr(), n, tol,
                                                  value.data_ptr(), coef.data_ptr(), _native.stream_ptr()))
    out = torch.empty_like(basis[0]) if out is None else out
    lib = _native.lib()
    for k, x in enumerate(basis):
        # out = g_k x + (k ? out : 0); entries of g past the breakdown are exactly zero
        _native.check(lib.tn_axpby(_native.TN_F64, x.numel(), 1.0, coef[k:].data_ptr(), 0, x.data_ptr(),
                                   1.0 if k else 0.0, None, 0, out.data_ptr(), _native.stream_ptr()))
    return value, out


def residual_refine(op64: Callable[[SymmetricTensor], SymmetricTensor], psi: SymmetricTensor):
    """The closing FP64 residual-refinement step of a reduced-precision Krylov solve.

    With ``w = H psi`` (FP64 operator), ``E = <psi|H|psi>``, the residual ``r = w - E psi``
    is orthogonal to ``psi``; the Rayleigh-Ritz problem of ``H`` on ``span{psi, r/|r|}`` is
    the 2x2 tridiagonal ``[[E, |r|], [|r|, <u|H|u>]]`` whose ground pair (solved on the
    device like the Lanczos one) gives the refined energy and vector.  The energy is never
    above the Rayleigh quotient of ``psi``, and an FP32 operator's first-order error is
    gone from it.  Returns ``(energy, vector, |r|)`` with device scalars.  Costs two FP64
    matvecs."""
    psi = ops.divide(psi, ops.norm(psi))
    w = op64(psi)
    e = ops.inner_device(psi, w)
    r = ops.subtract(w, ops.multiply(psi, e))
    rho = ops.norm(r)
    u = ops.divide(r, torch.clamp(rho, min=1e-300))  # r == 0 leaves u == 0 and g = (1, 0)
    hu = op64(u)
    euu = ops.inner_device(u, hu)
    alphas = torch.stack((e, euu))
    betas = torch.stack((rho, torch.zeros_like(rho)))
    value, out = ritz(alphas, betas, [psi.data, u.data], 1e-300)
    out = SymmetricTensor(psi.structure, out)
    return value, ops.divide(out, ops.norm(out)), rho


def spectrum(s: SymmetricTensor) -> list:
    """Every diagonal value of a singular-value tensor, ``sqrt(qdim)``-weighted, descending."""
    return _spectrum_of(s.structure, s.data.detach().to("cpu", torch.float64).numpy())


def _spectrum_of(structure, flat: np.ndarray) -> list:
    """``spectrum`` of a diagonal tensor of ``structure`` whose flat blocks are on the host."""
    prov = structure.provider
    lay = structure.layout
    vals = []
    for key, off, shp in zip(structure.block_order, lay.offsets.tolist(), lay.shapes.tolist()):
        k = shp[0]
        diag = flat[off: off + k * k: k + 1]
        vals.extend((diag * prov.qdim(key.coupled) ** 0.5).tolist())
    return sorted(vals, reverse=True)


def _pair(left, right):
    return ops.einsum("apx,xqr->apqr", left, right)


class LanczosRequest(NamedTuple):
    """One bond's Krylov step, handed to whoever solves it: the compiled operator, the FP64
    refinement operator (or None), the start vector, the Krylov dimension and the vectors the
    solve stays orthogonal to (an excited-state sweep's projection vectors)."""

    op: Any
    refine: Any
    aa: SymmetricTensor
    ncv: int
    ortho: tuple = ()


def _rho(m: SymmetricTensor, forward: bool) -> SymmetricTensor:
    """``m m^dag`` going forward, ``m^dag m`` coming back, for ``m`` on ``(l, p | q, r)``
    (reference ``network/dmrg.py:245-253``)."""
    return ops.compose(m, ops.adjoint(m)) if forward else ops.compose(ops.adjoint(m), m)


def _perturbations(env, n: int, aa: SymmetricTensor, noise: float) -> tuple:
    """The perturbation vectors of perturbative noise (reference ``dmrg.py:256-278``): the
    operator's term families on ``aa``, each normalised, the set scaled to total squared
    norm ``noise``; a numerically zero family is dropped."""
    parts = [(p, float(ops.norm(p))) for p in env.heff2_families(n, aa)]
    live = [(p, w) for p, w in parts if w > 0.0]
    if not live:
        return ()
    scale = (noise / len(live)) ** 0.5
    return tuple(ops.multiply(p, scale / w) for p, w in live)


def _split_dm(aa: SymmetricTensor, forward: bool, *, chi: int, cutoff: float, perturbations=()):
    """The density-matrix decimation (reference ``dmrg.py:281-322``): ``svd_truncated`` of
    ``rho = m m^dag`` (forward) or ``m^dag m`` with every perturbation's ``p p^dag`` added,
    in cutoff mode ``rsum1`` (on rho's spectrum, the same rule as ``rsum2`` on ``aa``'s).
    Returns ``(left, right, s)``: the isometry on the side the sweep leaves, the weighted
    factor on the other, and the singular values ``sqrt(spectrum)``."""
    m = ops.repartition(aa, (0, 1), (2, 3))
    rho = _rho(m, forward)
    for p in perturbations:
        rho = ops.add(rho, _rho(ops.repartition(p, (0, 1), (2, 3)), forward))
    v, w, _ = linalg.svd_truncated(rho, max_bond=chi, cutoff=cutoff, cutoff_mode="rsum1")
    s = ops.block_sqrt(w)
    if forward:
        return v, ops.compose(ops.adjoint(v), m), s
    # the adjoint keeps v's axis order, so the new bond comes last; the next site wants it
    # first
    return ops.compose(m, v), ops.transpose(ops.adjoint(v), (2, 0, 1)), s


def _bond_gen(psi, env, schmidt, n, forward, *, chi, cutoff, ncv, noise, seed, precision,
              refine, orthogonal_to=(), noise_type: str = "wavefunction"):
    """One bond of a two-site sweep as a generator: pair, hand the Lanczos step out
    (``yield LanczosRequest``; the caller sends back ``(energy, ground vector)``), split,
    write back, update environments.  Returns (energy, discarded weight)."""
    aa = _pair(psi[n], psi[n + 1])
    op = env.operator(n, precision)
    ref = env.operator(n, "fp64") if (precision != "fp64" and refine) else None
    ortho = tuple(o.project2(n, _pair(o.psi[n], o.psi[n + 1])) for o in orthogonal_to)
    energy, aa = yield LanczosRequest(op, ref, aa, ncv, ortho)
    if noise and noise_type == "perturbative":
        left, right, s = _split_dm(aa, forward, chi=chi, cutoff=cutoff,
                                   perturbations=_perturbations(env, n, aa, noise))
        # the factor that carries the weight is the truncated state, the other an
        # isometry: its norm is the kept fraction (the SVD split's Pythagoras)
        carrier = right if forward else left
        norm_c = ops.norm(carrier)
        host = torch.cat([norm_c.view(1), ops.norm(aa).view(1)]).cpu().numpy()
        dw = 1.0 - (host[0] / host[1]) ** 2
        psi[n] = left if forward else ops.divide(left, norm_c)
        psi[n + 1] = ops.divide(right, norm_c) if forward else right
        schmidt[n] = spectrum(ops.divide(s, ops.norm(s)))
        psi.center = n + 1 if forward else n
        for e in (env, *orthogonal_to):
            e.clear_(n, n + 1)
            if forward:
                e.update_(n, to="last")
            else:
                e.update_(n + 1, to="first")
        return energy, dw
    if noise:
        r = SymmetricTensor.random(aa.legs, seed=seed + n, device=aa.device)
        aa = ops.add(aa, ops.multiply(r, float(noise * ops.norm(aa) / ops.norm(r))))
        aa = ops.divide(aa, ops.norm(aa))
    u, s, vh = linalg.svd_truncated(aa, ((0, 1), (2, 3)), max_bond=chi, cutoff=cutoff)
    norm_s = ops.norm(s)
    # one readback per bond for the norms and the Schmidt values
    host = torch.cat([norm_s.view(1), ops.norm(aa).view(1), s.data.to(torch.float64)]).cpu().numpy()
    dw = 1.0 - (host[0] / host[1]) ** 2
    s = ops.divide(s, norm_s)
    psi[n + 1] = vh
    if forward:
        psi[n] = u
        psi[n + 1] = ops.einsum("xy,yqr->xqr", s, psi[n + 1])
    else:
        psi[n] = ops.einsum("apx,xy->apy", u, s)
    schmidt[n] = _spectrum_of(s.structure, host[2:] / host[0])
    psi.center = n + 1 if forward else n
    for e in (env, *orthogonal_to):
        e.clear_(n, n + 1)
        if forward:
            e.update_(n, to="last")
        else:
            e.update_(n + 1, to="first")
    return energy, dw


def _bond_step(psi, env, schmidt, n, forward, **kw):
    """``_bond_gen`` with its Lanczos step solved here, one problem at a time."""
    g = _bond_gen(psi, env, schmidt, n, forward, **kw)
    req = next(g)
    try:
        g.send(lanczos(req.op, req.aa, ncv=req.ncv, refine=req.refine, orthogonal_to=req.ortho,
                       host_value=False))
    except StopIteration as stop:
        return stop.value
    raise RuntimeError("a bond step asked for more than one Lanczos solve")


def sweep_steps(psi, h, env, schmidt: dict, *, chi: int, cutoff: float, ncv: int = 3,
                noise: float = 0.0, noise_type: str = "wavefunction", orthogonal_to=(),
                seed: int = 0, precision: str = "fp64", refine: bool = True, external: bool = False):
    """``sweep_`` as a generator that yields ``(energy, discarded)`` after every bond, so a
    scheduler can interleave bonds of independent solves.  With ``external`` it also yields
    each bond's ``LanczosRequest`` first and expects ``(energy, vector)`` sent back, so a
    scheduler can solve the Krylov steps of many solves together."""
    if noise_type not in ("wavefunction", "perturbative"):
        raise ValueError(f"noise_type={noise_type!r} is not one of ('wavefunction', 'perturbative')")
    if external and orthogonal_to:
        raise ValueError("an excited-state sweep (orthogonal_to) solves its own projected Krylov "
                         "steps; it cannot hand them out (external=True)")
    n_sites = len(psi)
    kw = dict(chi=chi, cutoff=cutoff, ncv=ncv, noise=noise, seed=seed, precision=precision,
              refine=refine, orthogonal_to=orthogonal_to, noise_type=noise_type)
    for direction in ("right", "left"):
        forward = direction == "right"
        bonds = range(n_sites - 1) if forward else range(n_sites - 2, -1, -1)
        for n in bonds:
            if external:
                yield (yield from _bond_gen(psi, env, schmidt, n, forward, **kw))
            else:
                yield _bond_step(psi, env, schmidt, n, forward, **kw)


def sweep_(psi, h, env, schmidt: dict, *, chi: int, cutoff: float, ncv: int = 3,
           noise: float = 0.0, noise_type: str = "wavefunction", orthogonal_to=(), seed: int = 0,
           precision: str = "fp64", refine: bool = True):
    """One left-to-right then right-to-left two-site sweep; ``psi`` and ``env`` mutate.
    Returns ``(last energy, max discarded weight)`` (reference ``dmrg.py:321-542``).

    ``precision`` selects the H_eff arithmetic (``fp64``, ``fp32``, ``bf16x3``); with a
    reduced precision and ``refine`` the energy of every bond is the FP64 Rayleigh quotient.
    """
    energy, max_dw = 0.0, 0.0
    for e, dw in sweep_steps(psi, h, env, schmidt, chi=chi, cutoff=cutoff, ncv=ncv, noise=noise,
                             noise_type=noise_type, orthogonal_to=orthogonal_to, seed=seed,
                             precision=precision, refine=refine):
        energy, max_dw = e, max(max_dw, dw)
    return float(energy), max_dw


def _schmidt_change(old: dict, new: dict) -> float:
    if not old:
        return float("inf")
    worst = 0.0
    for n, cur in new.items():
        prev = old.get(n, [])
        m = max(len(prev), len(cur))
        a = prev + [0.0] * (m - len(prev))
        b = cur + [0.0] * (m - len(cur))
        worst = max(worst, sum((x - y) ** 2 for x, y in zip(a, b)) ** 0.5)
    return worst


def dmrg_steps(psi, h, *, schedule=None, chi=None, cutoff=None, energy_tol: float = 1e-12,
               schmidt_tol: float = 1e-8, max_sweeps: int = 40, ncv: int = 3, seed: int = 0,
               callback=None, compile=None, precision: str = "fp64", refine: bool = True,
               polish_tol: float | None = None, external: bool = False, gc_freeze: bool = True,
               orthogonal_to=None):
    """``dmrg_`` as a generator: yields ``None`` after every bond and returns the
    ``DMRG_out`` (``StopIteration.value``) when the loop ends.  With ``external`` every
    yield is instead one bond's ``LanczosRequest`` and the caller sends back
    ``(energy, ground vector)``: the Krylov steps of many solves can then run as one batch.

    ``gc_freeze``: after every sweep the objects alive at that point (compiled plans and
    device tables, mostly long-lived) are moved out of the cyclic collector's generations
    (``gc.freeze``), so its full collections do not re-traverse the plan caches on every
    bond (they cost ~15% of a steady sweep otherwise); reference counting still frees
    everything that dies.

    With a reduced ``precision`` the Schmidt values carry the operator's rounding
    (|dSchmidt| ~ 1e-6 for FP32), so the Schmidt criterion cannot be met by reduced
    sweeps alone: once a reduced sweep changes the energy by less than ``polish_tol``
    (default ``1000 * energy_tol``) the remaining sweeps run the FP64 operator."""
    if polish_tol is None:
        polish_tol = 1e3 * energy_tol
    from tenet.network.env import Env

    if schedule is not None:
        if chi is not None or cutoff is not None:
            raise ValueError("pass either schedule=[Sweep(...)] or the flat chi=/cutoff= kwargs, not both")
        plan = list(schedule)
        if not plan:
            raise ValueError("schedule is empty; a run needs at least one Sweep entry")
    else:
        plan = [Sweep(64 if chi is None else chi, 1e-14 if cutoff is None else cutoff)]
    psi.canonize_(0)
    env = Env(psi, h, compile=compile).setup_(0)
    # one two-state environment <psi|1|phi> per converged state, set up after psi's
    # canonize_ since psi is its bra and it follows psi through the sweep (reference
    # dmrg.py:787-806); a state whose boundary legs put it in another sector is orthogonal
    # to psi by the symmetry and is skipped
    orthos = ()
    if orthogonal_to:
        from tenet.network.mps import MPO

        ident = MPO.identity(len(psi), psi[0].legs[1].space, device=psi[0].device)
        ends = (psi[0].legs[0], psi[len(psi) - 1].legs[2])

        def same_ends(phi):
            a, b = phi[0].legs[0], phi[len(phi) - 1].legs[2]
            return ((a.space, a.dual) == (ends[0].space, ends[0].dual)
                    and (b.space, b.dual) == (ends[1].space, ends[1].dual))

        orthos = tuple(Env(phi, ident, bra=psi).setup_(0) for phi in orthogonal_to if same_ends(phi))
    schmidt: dict = {}
    energy = float("inf")
    history, realized = [], []
    out = None
    for it in range(1, max_sweeps + 1):
        entry = plan[it - 1] if it <= len(plan) else plan[-1]
        old_e, old_s = energy, dict(schmidt)
        max_dw = 0.0
        steps = sweep_steps(psi, h, env, schmidt, chi=entry.chi, cutoff=entry.cutoff, ncv=ncv,
                            noise=entry.noise, noise_type=entry.noise_type, orthogonal_to=orthos,
                            seed=seed + 977 * it, precision=precision, refine=refine,
                            external=external)
        answer = None
        while True:
            try:
                item = steps.send(answer)
            except StopIteration:
                break
            answer = None
            if isinstance(item, LanczosRequest):
                answer = yield item
                continue
            e, dw = item
            energy, max_dw = e, max(max_dw, dw)
            if not external:
                yield None
        energy = float(energy)  # the sweep's last bond energy (a device scalar until here)
        if gc_freeze:
            gc.freeze()
        de = abs(old_e - energy)
        if precision != "fp64" and de < polish_tol:
            precision = "fp64"
        ds = _schmidt_change(old_s, schmidt)
        history.append((energy, de, ds, max_dw))
        realized.append(entry)
        out = DMRG_out(it, energy, de, ds, max_dw, history, realized, psi)
        if callback:
            callback(out)
        if de < energy_tol and ds < schmidt_tol and it >= len(plan) and entry.noise == 0.0:
            break
    return out


def dmrg_(psi, h, *, schedule=None, chi=None, cutoff=None, energy_tol: float = 1e-12,
          schmidt_tol: float = 1e-8, max_sweeps: int = 40, ncv: int = 3, orthogonal_to=None,
          seed: int = 0, callback=None, compile=None, precision: str = "fp64",
          refine: bool = True, polish_tol: float | None = None) -> DMRG_out:
    """Sweep ``psi`` to the ground state of ``h`` in place (reference ``dmrg.py:620-843``).

    ``orthogonal_to``: converged states ``phi`` (MPS, held fixed) the result is kept
    orthogonal to -- every bond's Krylov solve runs on ``H`` with the projection vectors
    ``<psi env|1|phi>`` of the two-site space projected out, so ``psi`` converges to the
    lowest state of ``h`` orthogonal to them (an excited state)."""
    gen = dmrg_steps(psi, h, schedule=schedule, chi=chi, cutoff=cutoff, energy_tol=energy_tol,
                     schmidt_tol=schmidt_tol, max_sweeps=max_sweeps, ncv=ncv, seed=seed,
                     callback=callback, compile=compile, precision=precision, refine=refine,
                     polish_tol=polish_tol, orthogonal_to=orthogonal_to)
    while True:
        try:
            next(gen)
        except StopIteration as stop:
            return stop.value
