"""Many independent two-site H_eff problems in one launch per step.

``BatchedHeff2`` takes the compiled abelian chains of any number of bonds/chains, packs
every instance's operands into shared arena buffers and merges their tables with offset
shifts, so one matvec of the whole batch is a fixed handful of launches whatever the
batch size.  ``layout="sector"`` (default) merges ``_SectorChain`` tables: entry gather,
one grouped GEMM over every instance's per-sector step-1 problems, one MPO pass, one
grouped GEMM over every per-sector step-3 problem (K slices to a workspace) and the exit
pass summing the slices into every instance's blocks.  ``layout="blocks"`` merges
``_FusedChain`` tables (one problem per block pair, three launches).  Uneven sector sizes of all instances share one tile table, sorted
longest-K first; the tile shape of each GEMM is chosen by the autotuner
(``tenet.autotune``) from the histogram of its problem sizes.

``synthetic_operator`` builds one XXZ instance directly: random environments made
Hermitian under the MPO's channel conjugation (``S^+ <-> S^-``), so ``H_eff`` is a real
symmetric operator without folding a whole chain.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from tenet._native import MPO_TILE

from tenet import _native, ops
from tenet.engine import GemmProgram, _DeviceCache
from tenet.network.heff import Heff2
from tenet.network.models import (
    BOUNDARY,
    END,
    MPO_BOND,
    PHYS,
    SM_CH,
    SP_CH,
    gaussian_bond,
    heisenberg_mpo,
)
from tenet.network.mps import from_dense
from tenet.space import IN, OUT, Leg
from tenet.structure import TensorStructure
from tenet.tensor import SymmetricTensor

__all__ = ["BatchedHeff2", "synthetic_operator", "batch_workload"]

_BAR = np.array([4, 1, 2, 3, 0])  # MPO channel conjugation: SP (0) <-> SM (4)


def _charges(space):
    return np.concatenate([np.full(m, a.charge) for a, m in space.sectors])


def _symmetric_env(bond, left: bool, rng, device):
    """A Hermitian environment on ``bond``: F[a, x, B] = F[B, bar(x), a] (left) or
    F[r, y, s] = F[s, bar(y), r] (right), zero outside the U(1) block set."""
    qb = _charges(bond)
    qw = _charges(MPO_BOND)
    D = len(qb)
    f = rng.standard_normal((D, 5, D))
    if left:  # (a IN, x OUT, B OUT): x + B = a
        mask = (qw[None, :, None] + qb[None, None, :]) == qb[:, None, None]
    else:     # (r OUT, y IN, s IN): r = y + s
        mask = qb[:, None, None] == (qw[None, :, None] + qb[None, None, :])
    f *= mask
    f = 0.5 * (f + np.transpose(f[:, _BAR, :], (2, 1, 0)))
    legs = ((Leg(bond, IN), Leg(MPO_BOND, OUT), Leg(bond, OUT)) if left
            else (Leg(bond, OUT), Leg(MPO_BOND, IN), Leg(bond, IN)))
    t = from_dense(f, legs, device=device)
    return ops.divide(t, ops.norm(t))


def synthetic_operator(n_sites: int, max_dim: int, sigma: float, delta: float, seed: int,
                       device="cuda"):
    """``(fl, w_left, w_right, fr, aa)`` of one XXZ two-site H_eff at the centre bond."""
    rng = np.random.default_rng(seed)
    c = n_sites // 2 - 1
    bl = gaussian_bond(c, n_sites, max_dim, sigma)
    br = gaussian_bond(c + 2, n_sites, max_dim, sigma)
    h = heisenberg_mpo(n_sites, delta, device=device)
    fl = _symmetric_env(bl, True, rng, device)
    fr = _symmetric_env(br, False, rng, device)
    aa = SymmetricTensor.random((Leg(bl, OUT), Leg(PHYS, OUT), Leg(PHYS, OUT), Leg(br, IN)),
                                seed=seed, device=device)
    aa = ops.divide(aa, ops.norm(aa))
    return fl, h[c], h[c + 1], fr, aa


def batch_workload(count: int = 256, n_sites: int = 32, d_min: int = 64, d_max: int = 1024,
                   sigma: float = 6.0, seed: int = 0, device="cuda"):
    """Delta ~ U[0, 2], D log-uniform in [d_min, d_max], from one seeded generator."""
    rng = np.random.default_rng(seed)
    deltas = rng.uniform(0.0, 2.0, count)
    dims = np.exp(rng.uniform(math.log(d_min), math.log(d_max), count)).astype(int)
    return [synthetic_operator(n_sites, int(d), sigma, float(dl), seed + 1 + i, device)
            for i, (d, dl) in enumerate(zip(dims, deltas))], dims, deltas


def _shift_problems(problems, c_shift, a_shift, b_shift):
    return [(M, N, c_off + c_shift, ldc, [(a + a_shift, b + b_shift, K, lda, ldb) for a, b, K, lda, ldb in segs])
            for M, N, c_off, ldc, segs in problems]


class BatchedHeff2:
    """One matvec over a batch of abelian two-site H_eff instances."""

    def __init__(self, instances, precision: str = "fp32", tiles: dict | None = None,
                 layout: str = "sector", operators=None) -> None:
        """``instances``: ``(fl, w_left, w_right, fr, aa)`` per problem; or ``operators``:
        ``(Heff2, structure)`` per problem, compiled operators reused as they are (their
        FP32 copies and stacked environments are not rebuilt) -- the form a scheduler uses
        to batch the current bonds of many running solves."""
        dt = torch.float64 if precision == "fp64" else torch.float32
        self.precision, self.dtype, self.layout = precision, dt, layout
        if operators is not None:
            if layout != "sector":
                raise ValueError("batching compiled operators needs the sector layout")
            for op, _ in operators:
                if op.precision != precision or op.layout != "sector" or not op.fused:
                    raise ValueError("every operator must be an abelian sector-layout Heff2 at the batch's precision")
            self._init_sector(None, tiles or {}, operators)
            return
        if layout == "sector":
            self._init_sector(instances, tiles or {})
            return
        if layout != "blocks":
            raise ValueError(f"unknown layout {layout!r}")
        self.ops = [Heff2(fl, w1, w2, fr, precision, layout="blocks") for fl, w1, w2, fr, _ in instances]
        self.structures = [inst[4].structure for inst in instances]
        chains = [op.chain(s) for op, s in zip(self.ops, self.structures)]
        for ch in chains:
            if not hasattr(ch, "mpo_host"):
                raise ValueError("BatchedHeff2 needs abelian (fused-path) operators")
            if ch.ex1.a.program is not None or ch.ex1.b.program is not None or ch.ex1.c_program is not None:
                raise ValueError("BatchedHeff2 needs in-place step-1 operands")
        dev = instances[0][4].device
        # arenas
        def offsets(sizes):
            o = np.zeros(len(sizes) + 1, dtype=np.int64)
            np.cumsum(sizes, out=o[1:])
            return o
        self.x_off = offsets([s.size for s in self.structures])
        fr_off = offsets([op._fr.structure.size for op in self.ops])
        fl_off = offsets([op._fl.structure.size for op in self.ops])
        t1_off = offsets([ch.t1_size for ch in chains])
        t3_off = offsets([ch.t3_size for ch in chains])
        self.fr = torch.cat([op._fr.data for op in self.ops]).to(dev)
        self.fl = torch.cat([op._fl.data for op in self.ops]).to(dev)
        self.t1_size, self.t3_size = int(t1_off[-1]), int(t3_off[-1])
        self.x_size = int(self.x_off[-1])
        p1, p4 = [], []
        jobs, ins, outs, coef, tl = [], [], [], [], []
        for i, ch in enumerate(chains):
            p1 += _shift_problems(ch.ex1.gemm.problems, int(t1_off[i]), int(self.x_off[i]), int(fr_off[i]))
            p4 += _shift_problems(ch.gemm4.problems, int(self.x_off[i]), int(fl_off[i]), int(t3_off[i]))
            j0, i0, o0, c0 = len(jobs), len(ins), len(outs), len(coef)
            cj, ci, co, cc = ch.mpo_host
            for R, S, ib, ie, ob, oe, cf, _ in cj:
                jobs.append((R, S, ib + i0, ie + i0, ob + o0, oe + o0, cf + c0, 0))
                nt = -(-(R * S) // MPO_TILE)
                tl.append(np.stack([np.full(nt, len(jobs) - 1, np.int64), np.arange(nt, dtype=np.int64)], 1))
            ins += [(off + int(t1_off[i]), rs) for off, rs in ci]
            outs += [(off + int(t3_off[i]), rs) for off, rs in co]
            coef += cc
        tune = tiles or {}
        self.gemm1 = GemmProgram(p1, self.x_size, int(fr_off[-1]), self.t1_size, tile=tune.get("gemm1"))
        self.gemm4 = GemmProgram(p4, int(fl_off[-1]), self.t3_size, self.x_size, trans_a=True,
                                 tile=tune.get("gemm4"))
        tl = np.concatenate(tl) if tl else np.zeros((0, 2), np.int64)
        self.mpo_njobs, self.mpo_ntiles = len(jobs), len(tl)
        self.mpo = _DeviceCache(jobs=np.asarray(jobs, np.int64).reshape(-1),
                                ins=np.asarray(ins, np.int64).reshape(-1),
                                outs=np.asarray(outs, np.int64).reshape(-1),
                                coef=np.asarray(coef, np.float64 if dt == torch.float64 else np.float32),
                                tiles=tl.reshape(-1))
        self.flops = self.gemm1.flops + self.gemm4.flops
        self.count = len(self.structures)

    def _init_sector(self, instances, tune: dict, operators=None) -> None:
        from tenet.engine import RecoupleProgram
        from tenet.network.heff import exit_regions

        if operators is None:
            self.ops = [Heff2(fl, w1, w2, fr, self.precision, layout="sector") for fl, w1, w2, fr, _ in instances]
            self.structures = [inst[4].structure for inst in instances]
        else:
            self.ops = [op for op, _ in operators]
            self.structures = [s for _, s in operators]
        chains = [op.chain(s) for op, s in zip(self.ops, self.structures)]
        for ch in chains:
            if not hasattr(ch, "exit_spec"):
                raise ValueError("BatchedHeff2 needs abelian (sector-path) operators")

        def offsets(sizes):
            o = np.zeros(len(sizes) + 1, dtype=np.int64)
            np.cumsum(sizes, out=o[1:])
            return o
        self.x_off = offsets([st.size for st in self.structures])
        xs_off = offsets([ch.xs_size for ch in chains])
        frs_off = offsets([ch.frs_size for ch in chains])
        fls_off = offsets([ch.fls_size for ch in chains])
        t1_off = offsets([ch.t1_size for ch in chains])
        t3_off = offsets([ch.t3_size for ch in chains])
        p3_off = offsets([len(ch.problems3) for ch in chains])
        self.x_size = int(self.x_off[-1])
        self.xs_size, self.t1_size, self.t3_size = int(xs_off[-1]), int(t1_off[-1]), int(t3_off[-1])
        # stacked environments of every instance, once
        stacks = []
        for op, ch, st in zip(self.ops, chains, self.structures):
            if st not in op._prepared:
                op._prepared[st] = ch.prepare(op._datas)
            stacks.append(op._prepared[st])
        self.fr = torch.cat([a for a, _ in stacks])
        self.fl = torch.cat([b for _, b in stacks])
        entry, p1, p3 = [], [], []
        p1_rows, rows1 = [], []
        jobs, ins, outs, coef, tl = [], [], [], [], []
        for i, ch in enumerate(chains):
            xo, so = int(self.x_off[i]), int(xs_off[i])
            entry += [(d + so, shp, ds, ss, [(t + xo, c) for t, c in terms])
                      for d, shp, ds, ss, terms in ch.xs_regions]
            p1 += _shift_problems(ch.problems1, int(t1_off[i]), so, int(frs_off[i]))
            # the row-gathered form: every instance's step-1 rows read in place in the arena
            p1_rows += _shift_problems(ch.gemm1_rows.problems, int(t1_off[i]), 0, int(frs_off[i]))
            rows1 += [r + xo for r in ch.gemm1_rows.a_rows]
            p3 += _shift_problems(ch.problems3, 0, int(fls_off[i]), int(t3_off[i]))
            i0, o0, c0 = len(ins), len(outs), len(coef)
            cj, ci, co, cc = ch.mpo_host
            for R, S, ib, ie, ob, oe, cf, _ in cj:
                jobs.append((R, S, ib + i0, ie + i0, ob + o0, oe + o0, cf + c0, 0))
                nt = -(-(R * S) // MPO_TILE)
                tl.append(np.stack([np.full(nt, len(jobs) - 1, np.int64), np.arange(nt, dtype=np.int64)], 1))
            ins += [(off + int(t1_off[i]), rs) for off, rs in ci]
            outs += [(off + int(t3_off[i]), rs) for off, rs in co]
            coef += cc
        self.entry = RecoupleProgram(entry, self.x_size, max(self.xs_size, 1))
        self.gemm1 = GemmProgram(p1, max(self.xs_size, 1), int(frs_off[-1]) or 1, max(self.t1_size, 1),
                                 tile=tune.get("gemm1") or (64, 64))
        t1_tile = tune.get("gemm1") or (64, 64)
        self.gemm1_rows = GemmProgram(p1_rows, max(self.x_size, 1), int(frs_off[-1]) or 1, max(self.t1_size, 1),
                                      tile=t1_tile if t1_tile in GemmProgram.ROWS_TILES else (64, 64),
                                      a_rows=rows1)
        self.gemm4 = GemmProgram(p3, int(fls_off[-1]) or 1, max(self.t3_size, 1), 1, trans_a=True,
                                 tile=tune.get("gemm4") or (64, 64), gather=True)
        ex = []
        for i, ch in enumerate(chains):
            ex += exit_regions(ch.exit_spec, self.gemm4.slices, int(self.x_off[i]), int(p3_off[i]))
        self.exit = RecoupleProgram(ex, max(self.gemm4.ws_size, 1), self.x_size)
        tl = np.concatenate(tl) if tl else np.zeros((0, 2), np.int64)
        self.mpo_njobs, self.mpo_ntiles = len(jobs), len(tl)
        self.mpo = _DeviceCache(jobs=np.asarray(jobs, np.int64).reshape(-1),
                                ins=np.asarray(ins, np.int64).reshape(-1),
                                outs=np.asarray(outs, np.int64).reshape(-1),
                                coef=np.asarray(coef, np.float64 if self.dtype == torch.float64 else np.float32),
                                tiles=tl.reshape(-1))
        self.flops = self.gemm1.flops + self.gemm4.flops
        self.count = len(self.structures)

    def _mpo(self, t1: torch.Tensor, t3: torch.Tensor) -> None:
        tb = self.mpo.get(t1.device)
        _native.check(_native.lib().tn_mpo_apply(
            _native.TN_F64 if t1.dtype == torch.float64 else _native.TN_F32,
            t1.data_ptr(), t3.data_ptr(), tb["jobs"].data_ptr(), self.mpo_njobs,
            tb["ins"].data_ptr(), tb["outs"].data_ptr(), tb["coef"].data_ptr(),
            tb["tiles"].data_ptr(), self.mpo_ntiles, _native.stream_ptr()))

    def pack(self, vectors) -> torch.Tensor:
        """One arena buffer (working dtype) holding every instance's input vector."""
        return torch.cat([v.data.to(self.dtype) for v in vectors])

    def unpack(self, x: torch.Tensor) -> list:
        return [SymmetricTensor(s, x[int(a):int(b)]) for s, a, b in zip(self.structures, self.x_off[:-1], self.x_off[1:])]

    def apply64(self, x: torch.Tensor) -> torch.Tensor:
        """The batched matvec on an FP64 arena, returning an FP64 arena: the Krylov vectors
        stay FP64 while the GEMMs run in the batch's arithmetic.  In the sector layout the
        casts ride on the entry gather and the exit scatter; the block layout casts apart."""
        if x.dtype != torch.float64 or x.numel() != self.x_size:
            raise ValueError("BatchedHeff2.apply64: expects the FP64 arena of the batch")
        if self.dtype == torch.float64:
            return self.apply(x)
        if self.layout == "sector":
            return self._apply_sector(x, torch.float64)
        from tenet import ops

        return ops.cast_data(self.apply(ops.cast_data(x, self.dtype)), torch.float64)

    def _apply_sector(self, x: torch.Tensor, out_dtype: torch.dtype, tiles: dict | None = None) -> torch.Tensor:
        t = tiles or {}
        wt = self.dtype
        t1 = torch.empty(max(self.t1_size, 1), dtype=wt, device=x.device)
        tile1 = tuple(t["gemm1"]) if t.get("gemm1") else self.gemm1_rows.tile
        if wt == torch.float32 and tile1 in GemmProgram.ROWS_TILES:
            # FP32 / BF16x3: step 1 reads the arena's blocks in place (FP64 or FP32), no entry pass
            self.gemm1_rows.run(x, self.fr, t1, self.precision, tile=tile1)
        else:
            xs = torch.empty(max(self.xs_size, 1), dtype=wt, device=x.device)
            self.entry.run(x, xs)
            self.gemm1.run(xs, self.fr, t1, self.precision, tile=t.get("gemm1"))
        t3 = torch.empty(max(self.t3_size, 1), dtype=wt, device=x.device)
        self._mpo(t1, t3)
        ws = torch.empty(max(self.gemm4.ws_size, 1), dtype=wt, device=x.device)
        self.gemm4.run(self.fl, t3, ws, self.precision, tile=t.get("gemm4"))
        out = torch.empty(self.x_size, dtype=out_dtype, device=x.device)
        self.exit.run(ws, out)
        return out

    def apply(self, x: torch.Tensor, tiles: dict | None = None) -> torch.Tensor:
        if x.dtype != self.dtype or x.numel() != self.x_size:
            raise ValueError("BatchedHeff2.apply: the arena does not match the batch")
        t = tiles or {}
        if self.layout == "sector":
            return self._apply_sector(x, x.dtype, tiles)
        t1 = torch.empty(self.t1_size, dtype=x.dtype, device=x.device)
        self.gemm1.run(x, self.fr, t1, self.precision, tile=t.get("gemm1"))
        t3 = torch.empty(self.t3_size, dtype=x.dtype, device=x.device)
        self._mpo(t1, t3)
        out = torch.empty(self.x_size, dtype=x.dtype, device=x.device)
        self.gemm4.run(self.fl, t3, out, self.precision, tile=t.get("gemm4"))
        return out
