"""Blockwise factorizations on the device: ``svd``, ``svd_truncated``, ``select_bond``, ``lq``.

Lowering (reference ``ops/linalg.py:72-90``): repartition to ``(left | right)``, then one
``tn_recouple`` launch gathers every block into one dense matrix per coupled sector, all
packed in a single device buffer.  Every sector matrix is factorized in FP64 on the GPU by
one batched one-sided Jacobi SVD (``tn_svd_jacobi_round``/``_finish``: one launch per
block-tournament round for all sectors together, converged when a whole sweep rotates
nothing), the truncation decision runs on the host over the O(D)
singular values with the reference's exact keep rule (global spectrum sorted by
``(-sigma, sector, index)``, cutoff modes, greedy walk under the dense budget
``sum qdim(c)``; reference ``linalg.py:1079-1324``), and one more ``tn_recouple`` launch
per factor scatters the kept columns/rows back into block storage.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np
import torch

from tenet.engine import RecoupleProgram, require_device, upload, wait_upload
from tenet.ops import repartition
from tenet.plans import band_layout
from tenet.space import IN, OUT, GradedSpace, Leg
from tenet.structure import TensorStructure
from tenet.symmetry import StructureChangingError
from tenet.tensor import SymmetricTensor

__all__ = ["svd", "svd_truncated", "select_bond", "BondSelection", "lq", "to_matrices", "from_matrices"]

_CUTOFF_MODES = ("abs", "rel", "sum2", "rsum2", "sum1", "rsum1")


# ----------------------------------------------------------------------------- matrix lowering


class _MatrixLayout:
    """Packed sector matrices of a structure and the two recouple programs to/from them."""

    def __init__(self, s: TensorStructure) -> None:
        pat, _, roff, _, coff, rtot, ctot = band_layout(s)
        self.sectors = pat.sectors
        self.shapes = {c: (int(r), int(k)) for c, r, k in zip(pat.sectors, rtot.tolist(), ctot.tolist())}
        sizes = rtot * ctot
        base = np.zeros(len(sizes), dtype=np.int64)
        if len(sizes) > 1:
            np.cumsum(sizes[:-1], out=base[1:])
        self.offsets = dict(zip(pat.sectors, base.tolist()))
        self.size = int(sizes.sum())
        # one region per block: its matrix offset (row band x column band of its sector)
        blk = pat.blk
        ld = ctot[pat.blk_sector]
        moff = base[pat.blk_sector] + roff[pat.blk_row] * ld + coff[pat.blk_col]
        bl = s.layout
        order = [*s.out_axes, *s.in_axes]
        n_out = len(s.out_axes)
        shape = bl.shapes[blk][:, order] if len(blk) else np.zeros((0, len(order)), np.int64)
        bstr = bl.strides[blk][:, order] if len(blk) else np.zeros((0, len(order)), np.int64)
        # matrix strides: rows (the output axes) of length ld, columns C-contiguous
        mat = np.ones_like(shape)
        for d in range(len(order) - 2, n_out - 1, -1):
            mat[:, d] = mat[:, d + 1] * shape[:, d + 1]
        if n_out:
            mat[:, n_out - 1] = ld
            for d in range(n_out - 2, -1, -1):
                mat[:, d] = mat[:, d + 1] * shape[:, d + 1]
        boff = bl.offsets[blk] if len(blk) else np.zeros(0, np.int64)
        ptr = np.arange(len(blk) + 1, dtype=np.int64)
        ones = np.ones(len(blk))
        self.to_prog = RecoupleProgram.from_arrays(moff, shape, mat, bstr, ptr, boff, ones, s.size, self.size)
        self.from_prog = RecoupleProgram.from_arrays(boff, shape, bstr, mat, ptr, moff, ones, self.size, s.size)


@lru_cache(maxsize=1024)
def _matrix_layout(s: TensorStructure) -> _MatrixLayout:
    return _MatrixLayout(s)


def _packed(t: SymmetricTensor):
    require_device(t)
    ml = _matrix_layout(t.structure)
    buf = torch.empty(ml.size, dtype=t.dtype, device=t.device)
    ml.to_prog.run(t.data, buf)
    return buf, ml


def _views(buf, ml) -> dict:
    return {
        c: buf[ml.offsets[c] : ml.offsets[c] + ml.shapes[c][0] * ml.shapes[c][1]].view(ml.shapes[c])
        for c in ml.sectors
    }


def to_matrices(t: SymmetricTensor) -> dict:
    """``{c: B_c}`` device matrices (views into one packed buffer), rows OUT, columns IN
    (reference ``map_view.py:238-279``)."""
    return _views(*_packed(t))


def from_matrices(structure: TensorStructure, mats: dict) -> SymmetricTensor:
    """Inverse of ``to_matrices`` (reference ``map_view.py:282-335``)."""
    ml = _matrix_layout(structure)
    if set(mats) != set(ml.sectors):
        raise ValueError("from_matrices: sector set does not match the structure")
    ref = next(iter(mats.values()))
    for c in ml.sectors:
        if tuple(mats[c].shape) != tuple(ml.shapes[c]):
            raise ValueError(f"from_matrices: sector {c!r} has shape {tuple(mats[c].shape)}, "
                             f"expected {ml.shapes[c]}")
    # the packed buffer is the sector matrices back to back in sector order: one batched copy
    buf = torch.cat([mats[c].reshape(-1) for c in ml.sectors])
    out = torch.empty(structure.size, dtype=ref.dtype, device=ref.device)
    ml.from_prog.run(buf, out)
    return SymmetricTensor(structure, out)


def _lower(t: SymmetricTensor, axes):
    """``(repartitioned tensor, {c: B_c}, (packed buffer, layout))``."""
    if axes is None:
        axes = (t.structure.out_axes, t.structure.in_axes)
    left, right = axes
    m = repartition(t, left, right)
    buf, ml = _packed(m)
    return m, _views(buf, ml), (buf, ml)


# ----------------------------------------------------------------------------- keep rule


@dataclass(frozen=True)
class BondSelection:
    """The truncation decision (reference ``linalg.py:1159-1271``)."""

    bond: GradedSpace
    dense_dim: float
    reduced_dim: int
    kept: tuple
    discarded: tuple
    discarded_weight: float
    next_dense_cost: float
    max_bond: int | None
    scale: float

    @property
    def next_multiplet(self):
        return self.discarded[0] if self.discarded else None

    @property
    def undershoot(self):
        return None if self.max_bond is None else self.max_bond - self.dense_dim


def _validate(max_bond, cutoff, cutoff_mode, renorm, caller):
    if max_bond is None and cutoff is None:
        raise ValueError(f"{caller} needs at least one of max_bond or cutoff")
    if max_bond is not None and max_bond <= 0:
        raise ValueError(f"max_bond must be a positive dense bond dimension, got {max_bond!r}")
    if cutoff is not None and cutoff < 0:
        raise ValueError(f"cutoff must be non-negative, got {cutoff!r}")
    if not isinstance(cutoff_mode, str) or cutoff_mode not in _CUTOFF_MODES:
        raise ValueError(f"unknown cutoff_mode {cutoff_mode!r}; expected one of {_CUTOFF_MODES}")
    if not isinstance(renorm, bool):
        raise TypeError(f"renorm must be a bool, got {renorm!r}")


def _spectrum(values: dict) -> list:
    """``[(sigma, c, i)]`` sorted by ``(-sigma, c, i)``; ``values`` are host arrays."""
    entries = [(float(s), c, i) for c, arr in values.items() for i, s in enumerate(arr)]
    entries.sort(key=lambda e: (-e[0], e[1], e[2]))
    return entries


def _admissible(spectrum, qdim, cutoff, mode) -> int:
    if cutoff is None:
        return len(spectrum)
    if mode == "abs":
        return sum(1 for s, _, _ in spectrum if s > cutoff)
    if mode == "rel":
        thr = cutoff * spectrum[0][0]
        return sum(1 for s, _, _ in spectrum if s > thr)
    power = 2 if mode in ("sum2", "rsum2") else 1
    weights = [qdim(c) * s**power for s, c, _ in spectrum]
    thr = cutoff * sum(weights) if mode.startswith("r") else cutoff
    kept, dropped = len(spectrum), 0.0
    for w in reversed(weights):
        if dropped + w >= thr:
            break
        dropped += w
        kept -= 1
    return kept


def _decide(spectrum, provider, max_bond, cutoff, cutoff_mode, renorm, caller) -> BondSelection:
    qdim = provider.qdim
    if not spectrum:
        raise ValueError(f"{caller}: this tensor has no singular values at all")
    admissible = _admissible(spectrum, qdim, cutoff, cutoff_mode)
    keep: dict = {}
    kept = []
    kept_w, dense = 0.0, 0.0
    for entry in spectrum[:admissible]:
        s, c, _ = entry
        if max_bond is not None and dense + qdim(c) > max_bond:
            break
        dense += qdim(c)
        keep[c] = keep.get(c, 0) + 1
        kept.append(entry)
        kept_w += qdim(c) * s * s
    if not keep:
        raise ValueError(
            f"{caller}: cutoff={cutoff!r} in cutoff_mode={cutoff_mode!r} with max_bond={max_bond!r} "
            "keeps no singular value at all"
        )
    discarded = tuple(spectrum[len(kept) :])
    total = sum(qdim(c) * s * s for s, c, _ in spectrum)
    return BondSelection(
        bond=GradedSpace.new(provider, keep),
        dense_dim=dense,
        reduced_dim=len(kept),
        kept=tuple(kept),
        discarded=discarded,
        discarded_weight=sum(qdim(c) * s * s for s, c, _ in discarded),
        next_dense_cost=qdim(discarded[0][1]) if discarded else 0.0,
        max_bond=max_bond,
        scale=(total / kept_w) ** 0.5 if renorm else 1.0,
    )


# ----------------------------------------------------------------------------- factorizations


_SVD_LDS_BYTES = None     # per-workgroup LDS the block pair may use (device limit less a margin)
_SVD_MAX_SWEEPS = 60
_SVD_INNER_SWEEPS = 1     # inner sweeps per block-pair visit when a problem spans several blocks
_SVD_SINGLE_SWEEPS = 64   # cap for a problem that fits in one block pair (converges in-launch)
_SVD_LAST_SWEEPS = 0      # sweeps queued by the last call (diagnostics)
_SVD_SWEEP_BATCH = 8      # sweeps queued before the first read of the convergence flags
# per launch-group shape ((w, rounds, single) of every group): the sweeps the last such SVD
# had queued before its final batch, plus one -- the first batch of the next one, so that a
# run of similar bonds (a DMRG sweep) reads the flags about once per SVD
_SVD_SWEEP_HINT: dict = {}
# widest block of a block tournament (vectors): a round's time follows the pairs one
# workgroup rotates, and the sweep's depth (about k inner rounds) does not depend on the
# width, so narrower blocks spread the same rounds over more workgroups (16 measured
# faster than 32 on 105-260-vector sectors; 8 and 4 no faster or slower)
_SVD_MAX_BLOCK = 16


def _lds_budget() -> int:
    global _SVD_LDS_BYTES
    if _SVD_LDS_BYTES is None:
        props = torch.cuda.get_device_properties(torch.cuda.current_device())
        limit = int(getattr(props, "shared_memory_per_block_optin", 0) or props.shared_memory_per_block)
        _SVD_LDS_BYTES = limit - 1024
    return _SVD_LDS_BYTES


def _single_pair(k: int, L: int) -> bool:
    """Whether one workgroup holds the whole problem: its k vectors (k + 1 when k is odd)
    of length L and their Q rows of length k in LDS, at most 2 * SVD_MAX_W vectors.  Such a
    problem converges inside one launch."""
    from tenet._native import SVD_MAX_W

    return k <= 2 * SVD_MAX_W and ((k + (k & 1)) * (L + k) + k + 1) * 8 <= _lds_budget()


def _svd_width(span: int, k: int) -> int:
    """Vectors per block of the block tournament: the largest power of two up to
    ``_SVD_MAX_BLOCK`` whose block pair (2w vectors of length L and their Q rows of length
    k, span = L + k doubles each) fits in LDS, and no wider than the smallest pair that
    holds all k vectors."""
    budget = _lds_budget()
    w = _SVD_MAX_BLOCK
    while w > 1 and (2 * w * span * 8 > budget or w >= k):
        w //= 2
    if 2 * w * span * 8 > budget:
        raise ValueError(f"svd: a sector matrix with rows + cols = {span} exceeds the Jacobi "
                         f"kernel's LDS block pair ({budget // 16} doubles)")
    return w


class _JacobiPlan:
    """Problem tables of the batched Jacobi SVD for one packed matrix layout."""

    def __init__(self, ml: _MatrixLayout) -> None:
        from tenet import _native
        from tenet.engine import _DeviceCache

        regions, rows, widths = [], [], []
        self.slices = {}
        xo = qo = uo = so = vo = 0
        for c in ml.sectors:
            m, n = ml.shapes[c]
            if m == 0 or n == 0:
                self.slices[c] = None
                continue
            cols = m > n
            k, L = (n, m) if cols else (m, n)
            if k > _native.SVD_MAX_K:
                raise ValueError(f"svd: sector {c!r} has {k} singular values, more than the "
                                 f"kernel's {_native.SVD_MAX_K}")
            src = ml.offsets[c]
            if cols:  # x holds A^T: element (j, i) <- A[i, j]
                regions.append((xo, (n, m), (m, 1), (1, n), [(src, 1.0)]))
            else:
                regions.append((xo, (m, n), (n, 1), (n, 1), [(src, 1.0)]))
            if _single_pair(k, L):
                w, nb = 0, 2  # the single-pair launch group (its width is set below)
            else:
                w = _svd_width(L + k, k)
                nb = max(2, -(-k // w))
                nb += nb % 2
            inner = _SVD_SINGLE_SWEEPS if nb == 2 else _SVD_INNER_SWEEPS
            rows.append([k, L, xo, qo, nb, int(cols), uo, so, vo, inner, len(rows)])
            widths.append(w)
            self.slices[c] = (uo, m, k, so, vo, n)
            xo += k * L
            qo += k * k
            uo += m * k
            so += k
            vo += k * n
        self.x_size, self.q_size, self.u_size, self.s_size, self.v_size = xo, qo, uo, so, vo
        self.gather = RecoupleProgram(regions, ml.size, xo) if regions else None
        self.nprobs = len(rows)
        self.probs = _DeviceCache(probs=np.asarray(rows, dtype=np.int64).reshape(-1))
        self.groups = []
        for w in sorted(set(widths), reverse=True):
            sub = [r for r, ww in zip(rows, widths) if ww == w]
            wgs = [(p, slot) for p, r in enumerate(sub) for slot in range(r[4] // 2)]
            span = max(r[0] + r[1] for r in sub)
            single = w == 0
            if w == 0:
                # every problem held by one workgroup, in one launch: blocks of half the
                # largest k (vector t is vector t), LDS for the largest (k + k % 2)(L + k)
                w = max(1, max(-(-r[0] // 2) for r in sub))
                span = -(-max((r[0] + (r[0] & 1)) * (r[0] + r[1]) for r in sub) // (2 * w))
            # rotate while |cos(x_a, x_b)| > tol: 8 eps sqrt(L) sits above the rounding
            # noise of a length-L dot product, so converged pairs stop rotating on noise
            tol = 8 * 2.220446049250313e-16 * max(1.0, max(r[1] for r in sub) ** 0.5)
            self.groups.append(dict(
                w=w, span=span, tol=tol, rounds=max(r[4] for r in sub) - 1, nwgs=len(wgs), single=single,
                tables=_DeviceCache(probs=np.asarray(sub, dtype=np.int64).reshape(-1),
                                    wgs=np.asarray(wgs, dtype=np.int32).reshape(-1))))


_SIDE_STREAMS: dict = {}


def _side_stream(dev: torch.device) -> torch.cuda.Stream:
    """One extra stream per device (default priority) for independent SVD launch groups."""
    key = dev.index if dev.index is not None else torch.cuda.current_device()
    st = _SIDE_STREAMS.get(key)
    if st is None:
        st = _SIDE_STREAMS[key] = torch.cuda.Stream(device=dev)
    return st


def _jacobi_plan(ml: _MatrixLayout) -> _JacobiPlan:
    plan = getattr(ml, "_jacobi", None)
    if plan is None:
        plan = ml._jacobi = _JacobiPlan(ml)
    return plan


_EPS = 2.220446049250313e-16


def _svd_parts(packed, complete: bool = True) -> dict:
    """FP64 thin SVD of every packed sector matrix with the batched Jacobi kernels:
    ``{c: (U, S, Vh)}``, singular values descending, as device views.  ``complete=True``:
    plain one-sided Jacobi to convergence, U and Vh isometries on every column.
    ``complete=False`` (a truncating caller): the sweeps stop once the vectors above the
    numerical-null floor ``k eps |A|_F`` are orthogonal, and the factor rows of the values
    at or below it stay zero (such values are never kept by a cutoff)."""
    return _svd_run(packed, complete)[0]


def _svd_run(packed, complete: bool):
    """``_svd_parts`` and a function that completes the null rows afterwards (in place,
    the returned views see it), for a caller that decides on them later."""
    from tenet import _native

    buf, ml = packed
    if buf.dtype != torch.float64:
        raise TypeError("svd: the Jacobi SVD runs in float64; cast the tensor first")
    plan = _jacobi_plan(ml)
    dev = buf.device
    lib = _native.lib()
    stream = _native.stream_ptr()
    x = torch.empty(plan.x_size, dtype=torch.float64, device=dev)
    q = torch.empty(plan.q_size, dtype=torch.float64, device=dev)
    if plan.gather is not None:
        plan.gather.run(buf, x)
    # a truncating caller stops the sweeps once the vectors above k eps |A|_F are
    # orthogonal; an exact SVD runs plain one-sided Jacobi
    null_scale = 0.0 if complete else _EPS * _EPS
    fro2 = torch.empty(max(plan.nprobs, 1), dtype=torch.float64, device=dev)
    if plan.nprobs:
        _native.check(lib.tn_svd_jacobi_norms(plan.probs.get(dev)["probs"].data_ptr(), plan.nprobs,
                                              x.data_ptr(), fro2.data_ptr(), stream))
    rots = []
    for g in plan.groups:
        g["tb"] = g["tables"].get(dev)
        rots.append(torch.zeros(max(g["rounds"], 1) * g["nwgs"], dtype=torch.int32, device=dev))
    # one convergence flag per launch group, set on the device after a sweep that rotated
    # nothing; sweeps are queued in batches and the host reads the flags once per batch
    stop = torch.zeros(len(plan.groups), dtype=torch.int32, device=dev)
    # the launch groups touch disjoint problems: the one of sectors held by one workgroup
    # (which converges inside its first launch) runs on a second stream beside the block
    # tournaments, so the SVD waits for the longer of the two rather than their sum; the
    # side stream starts after the gather and the zeroed counters and flags
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev) if len(plan.groups) > 1 else None
    if side is not None:
        side.wait_stream(main)
    streams = [side.cuda_stream if (side is not None and g["single"]) else stream for g in plan.groups]
    global _SVD_LAST_SWEEPS
    shape_key = tuple((g["w"], g["rounds"], g["single"]) for g in plan.groups)
    first = max(_SVD_SWEEP_BATCH, _SVD_SWEEP_HINT.get(shape_key, 0))
    done = 0
    while True:
        batch = min(first if done == 0 else _SVD_SWEEP_BATCH // 2, _SVD_MAX_SWEEPS - done)
        for sweep in range(done, done + batch):
            for gi, (g, rot, st) in enumerate(zip(plan.groups, rots, streams)):
                flag = stop[gi:]
                for r in range(g["rounds"]):
                    _native.check(lib.tn_svd_jacobi_round(
                        g["tb"]["probs"].data_ptr(), g["tb"]["wgs"].data_ptr(), g["nwgs"], g["span"],
                        x.data_ptr(), q.data_ptr(), rot[r * g["nwgs"]:].data_ptr(), g["w"], r,
                        int(sweep == 0 and r == 0), g["tol"], flag.data_ptr(), fro2.data_ptr(), null_scale,
                        st))
                _native.check(lib.tn_svd_jacobi_converged(rot.data_ptr(), g["rounds"] * g["nwgs"],
                                                          flag.data_ptr(), st))
        done += batch
        if side is not None:
            main.wait_stream(side)
        if bool(stop.all()):
            break
        if done >= _SVD_MAX_SWEEPS:
            raise RuntimeError(f"svd: the Jacobi sweeps did not converge in {_SVD_MAX_SWEEPS} sweeps")
    _SVD_LAST_SWEEPS = done
    # converged inside the last batch: at least done - batch + 1 sweeps were needed
    _SVD_SWEEP_HINT[shape_key] = done - batch + 1
    U = torch.empty(plan.u_size, dtype=torch.float64, device=dev)
    S = torch.empty(plan.s_size, dtype=torch.float64, device=dev)
    Vh = torch.empty(plan.v_size, dtype=torch.float64, device=dev)

    def finish(full: int) -> None:
        if plan.nprobs:
            _native.check(lib.tn_svd_jacobi_finish(
                plan.probs.get(dev)["probs"].data_ptr(), plan.nprobs, x.data_ptr(), q.data_ptr(),
                U.data_ptr(), S.data_ptr(), Vh.data_ptr(), fro2.data_ptr(), null_scale, full,
                _native.stream_ptr()))

    finish(int(complete))
    parts = {}
    for c in ml.sectors:
        sl = plan.slices[c]
        m, n = ml.shapes[c]
        if sl is None:
            e = torch.empty(0, dtype=torch.float64, device=dev)
            parts[c] = (e.view(m, 0), e, e.view(0, n))
            continue
        uo, m, k, so, vo, n = sl
        parts[c] = (U[uo : uo + m * k].view(m, k), S[so : so + k], Vh[vo : vo + k * n].view(k, n))
    return parts, (lambda: None) if complete else (lambda: finish(1))


def _needs_completion(values: dict, keep: dict) -> bool:
    """Whether a kept singular value may lie at or below its sector's numerical-null floor
    (``k eps |A|_F``, with ``|A|_F^2`` taken as the sum of the sector's squared values and
    a factor two of margin against rounding), so its factor rows must be completed."""
    for c, n in keep.items():
        v = values.get(c)
        if v is None or n <= 0:
            continue
        floor = 2.0 * len(v) * _EPS * float(np.sqrt(np.sum(np.square(v, dtype=np.float64))))
        if float(v[n - 1]) <= max(floor, 1e-290):
            return True
    return False


def _factors(m: SymmetricTensor, parts: dict, keep: dict, scale: float = 1.0):
    provider = m.provider
    bond = GradedSpace.new(provider, keep)
    us = TensorStructure((*m.codomain, Leg(bond, IN)))
    ss = TensorStructure((Leg(bond, OUT), Leg(bond, IN)))
    vs = TensorStructure((Leg(bond, OUT), *m.domain))
    u = from_matrices(us, {c: parts[c][0][:, :k] for c, k in keep.items()})
    s = _diagonal(ss, {c: parts[c][1][:k] for c, k in keep.items()}, scale)
    vh = from_matrices(vs, {c: parts[c][2][:k, :] for c, k in keep.items()})
    return u, s, vh


def _diagonal(ss: TensorStructure, values: dict, scale: float = 1.0) -> SymmetricTensor:
    """The diagonal singular-value tensor of structure ``ss`` (one k x k block per sector)
    from per-sector value vectors: every value scattered onto its diagonal position of the
    packed sector matrices in one launch, instead of one diagonal matrix per sector."""
    ml = _matrix_layout(ss)
    if set(values) != set(ml.sectors):
        raise ValueError("diagonal factor: sector set does not match the structure")
    ref = next(iter(values.values()))
    vals = torch.cat([values[c] for c in ml.sectors]) * scale
    cache = ml.__dict__.setdefault("_diag_index", {})
    key = (ref.device.type, ref.device.index)
    got = cache.get(key)
    if got is None:
        got = cache[key] = upload(np.concatenate([ml.offsets[c] + np.arange(ml.shapes[c][0]) * (ml.shapes[c][0] + 1)
                                                  for c in ml.sectors]).astype(np.int64), ref.device)
    else:
        wait_upload(got[1], got[2])
    idx = got[0]
    if vals.numel() != idx.numel():
        raise ValueError("diagonal factor: value counts do not match the sector sizes")
    buf = torch.zeros(ml.size, dtype=ref.dtype, device=ref.device)
    buf[idx] = vals
    out = torch.empty(ss.size, dtype=ref.dtype, device=ref.device)
    ml.from_prog.run(buf, out)
    return SymmetricTensor(ss, out)


def svd(t: SymmetricTensor, axes=None, *, bond: GradedSpace | None = None):
    """Compact SVD ``T = U S Vh`` per coupled sector, or its projection onto a
    pre-decided ``bond`` (reference ``linalg.py:122-245``)."""
    m, mats, packed = _lower(t, axes)
    parts = _svd_parts(packed)
    full = {c: min(mats[c].shape) for c in mats}
    if bond is None:
        keep = full
    else:
        keep = {}
        for c, k in bond.sectors:
            if k > full.get(c, 0):
                raise ValueError(
                    f"svd: bond asks for {k} singular values in sector {c!r}, but the untruncated "
                    f"bond degeneracy there is {full.get(c, 0)}"
                )
            keep[c] = k
    return _factors(m, parts, keep)


def select_bond(t: SymmetricTensor, axes=None, *, max_bond=None, cutoff=None, cutoff_mode="rsum2",
                renorm=False) -> BondSelection:
    """The truncation decision alone (reference ``linalg.py:1327-1438``)."""
    _validate(max_bond, cutoff, cutoff_mode, renorm, "select_bond")
    m, _, packed = _lower(t, axes)
    parts = _svd_parts(packed, complete=False)  # the decision reads the values only
    return _select(m, parts, max_bond, cutoff, cutoff_mode, renorm, "select_bond")


def _host_values(parts, caller) -> dict:
    cs = list(parts)
    if not cs:
        raise ValueError(f"{caller}: this tensor has no singular values at all")
    flat = torch.cat([parts[c][1] for c in cs]).to("cpu").numpy()
    values, o = {}, 0
    for c in cs:
        n = parts[c][1].numel()
        values[c] = flat[o : o + n]
        o += n
    return values


def _select(m, parts, max_bond, cutoff, cutoff_mode, renorm, caller, values=None):
    values = _host_values(parts, caller) if values is None else values
    return _decide(_spectrum(values), m.provider, max_bond, cutoff, cutoff_mode, renorm, caller)


def _truncated(t, axes, max_bond, cutoff, cutoff_mode, renorm):
    """The truncating SVD: Jacobi without null-row completion, the keep rule on the host
    values, and the completion only when a kept value is numerically null."""
    m, _, packed = _lower(t, axes)
    parts, complete = _svd_run(packed, False)
    values = _host_values(parts, "svd_truncated")
    sel = _select(m, parts, max_bond, cutoff, cutoff_mode, renorm, "svd_truncated", values)
    keep = dict(sel.bond.sectors)
    if _needs_completion(values, keep):
        complete()
    return m, parts, keep, sel


def svd_truncated(t: SymmetricTensor, axes=None, *, max_bond=None, cutoff=None,
                  cutoff_mode: str = "rsum2", renorm: bool = False):
    """Truncating blockwise SVD with a global keep rule (reference ``linalg.py:1441-1612``).

    Returns ``(U, S, Vh)``: ``U`` on ``(*left legs, bond IN)``, ``S`` diagonal on
    ``(bond OUT, bond IN)``, ``Vh`` on ``(bond OUT, *right legs)``.
    """
    _validate(max_bond, cutoff, cutoff_mode, renorm, "svd_truncated")
    m, parts, keep, sel = _truncated(t, axes, max_bond, cutoff, cutoff_mode, renorm)
    return _factors(m, parts, keep, sel.scale)


def svd_truncated_with_selection(t, axes=None, *, max_bond=None, cutoff=None, cutoff_mode="rsum2",
                                 renorm=False):
    """``svd_truncated`` that also returns its ``BondSelection``."""
    _validate(max_bond, cutoff, cutoff_mode, renorm, "svd_truncated")
    m, parts, keep, sel = _truncated(t, axes, max_bond, cutoff, cutoff_mode, renorm)
    return (*_factors(m, parts, keep, sel.scale), sel)


class _LQPlan:
    """Problem table of ``tn_lq`` for one packed matrix layout."""

    def __init__(self, ml: _MatrixLayout) -> None:
        from tenet import _native
        from tenet.engine import _DeviceCache

        rows, self.slices = [], {}
        lo = qo = 0
        lds = 1
        for c in ml.sectors:
            m, n = ml.shapes[c]
            k = min(m, n)
            self.slices[c] = (lo, m, k, qo, n)
            if k:
                rows.append([m, n, ml.offsets[c], lo, qo, 0])
                lds = max(lds, n + 2 * k)
            lo += m * k
            qo += k * n
        if lds * 8 > _lds_budget():
            raise ValueError(f"lq: a sector matrix needs {lds} doubles of LDS (columns + 2 min(rows, cols)); "
                             f"the kernel has {_lds_budget() // 8}")
        self.l_size, self.q_size, self.lds, self.nprobs = lo, qo, lds, len(rows)
        self.probs = _DeviceCache(probs=np.asarray(rows, dtype=np.int64).reshape(-1))
        assert not rows or len(rows[0]) == _native.LQ_PROB_WORDS


def lq(t: SymmetricTensor, axes=None):
    """``T = L Q`` per coupled sector with ``Q`` row-orthonormal, bond ``min(rows, cols)``
    (reference ``linalg.py:549-620``: QR of each sector matrix's transpose; L's diagonal
    carries LAPACK's signs).  All sectors in one ``tn_lq`` launch (FP64 Householder; an
    FP32 tensor is factorised in FP64 and the factors cast back)."""
    from tenet import _native

    m, mats, (buf, ml) = _lower(t, axes)
    provider = m.provider
    keep = {c: min(b.shape) for c, b in mats.items()}
    bond = GradedSpace.new(provider, keep)
    plan = getattr(ml, "_lq", None)
    if plan is None:
        plan = ml._lq = _LQPlan(ml)
    dev = buf.device
    a = buf if buf.dtype == torch.float64 else buf.to(torch.float64)
    Lb = torch.empty(max(plan.l_size, 1), dtype=torch.float64, device=dev)
    Qb = torch.empty(max(plan.q_size, 1), dtype=torch.float64, device=dev)
    if plan.nprobs:
        _native.check(_native.lib().tn_lq(plan.probs.get(dev)["probs"].data_ptr(), plan.nprobs, plan.lds,
                                          a.data_ptr(), Lb.data_ptr(), Qb.data_ptr(), _native.stream_ptr()))
    if buf.dtype != torch.float64:
        Lb, Qb = Lb.to(buf.dtype), Qb.to(buf.dtype)
    ls, qs = {}, {}
    for c in ml.sectors:
        lo, rows, k, qo, cols = plan.slices[c]
        ls[c] = Lb[lo: lo + rows * k].view(rows, k)
        qs[c] = Qb[qo: qo + k * cols].view(k, cols)
    L = from_matrices(TensorStructure((*m.codomain, Leg(bond, IN))), ls)
    Q = from_matrices(TensorStructure((Leg(bond, OUT), *m.domain)), qs)
    return L, Q


_ = StructureChangingError
