"""The Krylov step, the two-site sweep and the DMRG loop, with FP64 Krylov scalars on the
device.

``lanczos`` follows reference ``network/dmrg.py:106-227`` (three-term recurrence, happy
breakdown below ``tol``, ``eigh`` of the small tridiagonal, one recombination) with two
MI355X-side changes that do not change the mathematics: every Krylov scalar
(``alpha_j = <v_j|w>``, ``beta_j = |w|``) stays in device memory and the host reads them
once, after the recurrence; and each new vector is fully re-orthogonalized in FP64
against the previous ones (the reference's ncv=3 recurrence is orthogonal in exact
arithmetic; this keeps it so when the matvec is FP32).  ``refine`` takes an FP64 matvec:
the returned value is then the FP64 Rayleigh quotient of the Ritz vector, which removes
the first-order error of an FP32 operator from the energy.

``sweep_``/``dmrg_`` follow reference ``dmrg.py:321-843`` (noiseless and wavefunction-noise
splits; perturbative noise is not part of this build).
"""

from __future__ import annotations

from collections.abc import Callable, Sequence
from typing import Any, NamedTuple

import numpy as np
import torch

from tenet import linalg, ops
from tenet.tensor import SymmetricTensor

__all__ = ["Sweep", "DMRG_out", "lanczos", "sweep_", "sweep_steps", "dmrg_", "dmrg_steps", "spectrum"]


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
            reorthogonalize: bool = True) -> tuple[float, SymmetricTensor]:
    """Ground eigenpair ``(value, vector)`` of a Hermitian ``matvec``."""
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
        full = matvec

        def matvec(x):
            y = full(x)
            for b in basis:
                y = ops.subtract(y, ops.multiply(b, ops.inner(b, y)))
            return y

    dev = v.device
    alphas = torch.zeros(ncv, dtype=torch.float64, device=dev)
    betas = torch.zeros(ncv, dtype=torch.float64, device=dev)
    vecs = [ops.divide(v, ops.norm(v))]
    for j in range(ncv):
        w = matvec(vecs[j])
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
    a_h = alphas.cpu().numpy()
    b_h = betas.cpu().numpy()
    m = ncv
    for j in range(ncv - 1):
        if b_h[j] < tol:  # happy breakdown: drop the rest, keep the space
            m = j + 1
            break
    tri = np.diag(a_h[:m]) + np.diag(b_h[: m - 1], 1) + np.diag(b_h[: m - 1], -1)
    values, states = np.linalg.eigh(tri)
    g = states[:, 0]
    out = ops.multiply(vecs[0], float(g[0]))
    for k in range(1, m):
        out = ops.add(out, ops.multiply(vecs[k], float(g[k])))
    out = ops.divide(out, ops.norm(out))
    value = float(values[0])
    if refine is not None:
        value = float(ops.inner(out, refine(out)))
    return value, out


def spectrum(s: SymmetricTensor) -> list:
    """Every diagonal value of a singular-value tensor, ``sqrt(qdim)``-weighted, descending."""
    vals = []
    prov = s.provider
    for key, blk in zip(s.structure.block_order, s.numpy_blocks()):
        w = prov.qdim(key.coupled) ** 0.5
        vals.extend((np.diag(blk) * w).tolist())
    return sorted(vals, reverse=True)


def _pair(left, right):
    return ops.einsum("apx,xqr->apqr", left, right)


def _bond_step(psi, env, schmidt, n, forward, *, chi, cutoff, ncv, noise, seed, precision,
               refine, orthogonal_to=()):
    """One bond of a two-site sweep: pair, Lanczos, split, write back, update environments.
    Returns (energy, discarded weight)."""
    aa = _pair(psi[n], psi[n + 1])
    op = env.operator(n, precision)
    ref = env.operator(n, "fp64") if (precision != "fp64" and refine) else None
    energy, aa = lanczos(op, aa, ncv=ncv, refine=ref)
    if noise:
        r = SymmetricTensor.random(aa.legs, seed=seed + n, device=aa.device)
        aa = ops.add(aa, ops.multiply(r, float(noise * ops.norm(aa) / ops.norm(r))))
        aa = ops.divide(aa, ops.norm(aa))
    u, s, vh = linalg.svd_truncated(aa, ((0, 1), (2, 3)), max_bond=chi, cutoff=cutoff)
    norm_s = float(ops.norm(s))
    dw = 1.0 - (norm_s / float(ops.norm(aa))) ** 2
    s = ops.divide(s, norm_s)
    psi[n + 1] = vh
    if forward:
        psi[n] = u
        psi[n + 1] = ops.einsum("xy,yqr->xqr", s, psi[n + 1])
    else:
        psi[n] = ops.einsum("apx,xy->apy", u, s)
    schmidt[n] = spectrum(s)
    psi.center = n + 1 if forward else n
    for e in (env, *orthogonal_to):
        e.clear_(n, n + 1)
        if forward:
            e.update_(n, to="last")
        else:
            e.update_(n + 1, to="first")
    return energy, dw


def sweep_steps(psi, h, env, schmidt: dict, *, chi: int, cutoff: float, ncv: int = 3,
                noise: float = 0.0, noise_type: str = "wavefunction", orthogonal_to=(),
                seed: int = 0, precision: str = "fp64", refine: bool = True):
    """``sweep_`` as a generator that yields ``(energy, discarded)`` after every bond, so a
    scheduler can interleave bonds of independent solves."""
    if noise_type not in ("wavefunction", "perturbative"):
        raise ValueError(f"noise_type={noise_type!r} is not one of ('wavefunction', 'perturbative')")
    if noise and noise_type == "perturbative":
        raise NotImplementedError("perturbative noise is not part of this build")
    n_sites = len(psi)
    for direction in ("right", "left"):
        forward = direction == "right"
        bonds = range(n_sites - 1) if forward else range(n_sites - 2, -1, -1)
        for n in bonds:
            yield _bond_step(psi, env, schmidt, n, forward, chi=chi, cutoff=cutoff, ncv=ncv,
                             noise=noise, seed=seed, precision=precision, refine=refine,
                             orthogonal_to=orthogonal_to)


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
    return energy, max_dw


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
               polish_tol: float | None = None):
    """``dmrg_`` as a generator: yields ``None`` after every bond and returns the
    ``DMRG_out`` (``StopIteration.value``) when the loop ends.

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
    schmidt: dict = {}
    energy = float("inf")
    history, realized = [], []
    out = None
    for it in range(1, max_sweeps + 1):
        entry = plan[it - 1] if it <= len(plan) else plan[-1]
        old_e, old_s = energy, dict(schmidt)
        max_dw = 0.0
        for e, dw in sweep_steps(psi, h, env, schmidt, chi=entry.chi, cutoff=entry.cutoff, ncv=ncv,
                                 noise=entry.noise, noise_type=entry.noise_type,
                                 seed=seed + 977 * it, precision=precision, refine=refine):
            energy, max_dw = e, max(max_dw, dw)
            yield None
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
    """Sweep ``psi`` to the ground state of ``h`` in place (reference ``dmrg.py:620-843``)."""
    if orthogonal_to:
        raise NotImplementedError("excited-state projection is not part of this build")
    gen = dmrg_steps(psi, h, schedule=schedule, chi=chi, cutoff=cutoff, energy_tol=energy_tol,
                     schmidt_tol=schmidt_tol, max_sweeps=max_sweeps, ncv=ncv, seed=seed,
                     callback=callback, compile=compile, precision=precision, refine=refine,
                     polish_tol=polish_tol)
    while True:
        try:
            next(gen)
        except StopIteration as stop:
            return stop.value
