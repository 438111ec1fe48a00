"""Compile host plans into device tables and run them through the gfx950 kernels.

Two device programs cover all block movement and arithmetic of a contraction:

* ``RecoupleProgram`` -- one ``tn_recouple`` launch: for every destination region
  ``dst = sum_t coeff_t * permute(src_t)``.  Built from a (perm, terms) plan between two
  flat layouts, or from arbitrary strided regions (sector matrices for the SVD).
* ``GemmProgram`` -- one ``tn_gemm_grouped`` launch: for every output block,
  ``C = sum_s A_s B_s`` over its K segments, operands read in place from flat buffers.

``ContractionExec`` strings them together for ``tensordot``/``einsum``: operand
repartitions that are pure block re-indexings (every U(1) bend, every interleaving
change) cost nothing -- the GEMM reads the source blocks where they are -- and only a
real in-block permutation or a multi-term recoupling launches ``tn_recouple``.  Every
compiled program keeps its tables resident on the device and is cached on the frozen
structures it was built from.
"""

from __future__ import annotations

import math
import os
from collections import defaultdict
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
import torch

from tenet import _native
from tenet._native import JOB_WORDS, MAX_RANK, PROB_WORDS, SEG_WORDS
from tenet.plans import (
    compose_maps,
    contraction_plan,
    permutation_plan,
    repartition_plan,
)
from tenet.structure import FusionBlockKey, TensorStructure, pattern_of

__all__ = [
    "RecoupleProgram",
    "GemmProgram",
    "ContractionExec",
    "gemm_precision",
    "set_gemm_precision",
    "require_device",
]

_LIN_TILE = 2048        # linear tile of mode -1 (256 threads x 8 elements)
_LIN_SMALL_TILE = 512   # linear tile of mode -2 (256 threads x 2 elements)
_LDS_T = 64
_SMALL_PROGRAM_TILES = 2048  # below this many workgroups, LDS jobs use 32 x 32 tiles
_GEMM_PRECISION = {"value": os.environ.get("TENET_GEMM_PRECISION", "fp32")}


def set_gemm_precision(p: str) -> None:
    """Arithmetic of GEMMs on float32 tensors: ``"fp32"`` or ``"bf16x3"``.  float64
    tensors always use FP64 MFMA."""
    if p not in ("fp32", "bf16x3"):
        raise ValueError("gemm precision for float32 tensors is 'fp32' or 'bf16x3'")
    _GEMM_PRECISION["value"] = p


def gemm_precision(dtype: torch.dtype) -> str:
    return "fp64" if dtype == torch.float64 else _GEMM_PRECISION["value"]


def require_device(*tensors) -> None:
    for t in tensors:
        if not t.data.is_cuda:
            raise RuntimeError(
                "tenet's contraction engine runs on the HIP device only; move the tensor "
                "with .to('cuda') (there is no CPU fallback)"
            )
    _native.lib()


def _dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float64:
        return _native.TN_F64
    if dtype == torch.float32:
        return _native.TN_F32
    raise TypeError(f"unsupported dtype {dtype}")


_TORCH_DTYPES = {np.dtype(np.int64): torch.int64, np.dtype(np.int32): torch.int32,
                 np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                 np.dtype(np.uint8): torch.uint8}


def upload(raw: np.ndarray, device: torch.device) -> tuple:
    """Host-to-device copy of a host array without stalling the host: the bytes go
    through a pinned staging buffer (torch's caching host allocator, which keeps the
    buffer until the copy has run) and the copy is queued on the current stream, after
    the work already queued there.  A synchronous copy would wait for that work to finish,
    so plan compilation on the host could never overlap the kernels in flight.  Returns
    ``(device tensor, stream, event)``: the event marks the copy's completion for users on
    other streams."""
    raw = np.ascontiguousarray(raw)
    if torch.device(device).type != "cuda":  # host-side table users (plan checks)
        return torch.from_numpy(raw.copy()), None, None
    pinned = torch.empty(raw.nbytes, dtype=torch.uint8, pin_memory=True)
    pinned.numpy()[:] = raw.view(np.uint8).reshape(-1)
    buf = pinned.to(device, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    return buf.view(_TORCH_DTYPES[raw.dtype]), _native.stream_ptr(), ev


def wait_upload(stream: int, ev) -> None:
    """Order the current stream after an upload queued on ``stream``."""
    if ev is not None and _native.stream_ptr() != stream:
        torch.cuda.current_stream().wait_event(ev)



class _DeviceCache:
    """Lazily uploaded copies of host tables, one per device.  All of a cache's tables
    travel in ONE host-to-device copy (packed at 256-byte aligned offsets into one byte
    buffer) and are handed out as typed views of it.  The copy is queued on the stream
    that first asks for the tables (``upload``); a later user on another stream is
    ordered after it."""

    _ALIGN = 256

    def __init__(self, **arrays) -> None:
        self.host = arrays
        self._dev: dict = {}
        self._ready: dict = {}

    def get(self, device: torch.device) -> dict:
        key = (device.type, device.index)
        d = self._dev.get(key)
        if d is not None:
            st, ev = self._ready[key]
            wait_upload(st, ev)
        else:
            parts, spans, off = [], [], 0
            for k, v in self.host.items():
                a = np.ascontiguousarray(v)
                if a.dtype not in _TORCH_DTYPES:
                    raise TypeError(f"device table {k!r}: unsupported dtype {a.dtype}")
                spans.append((k, off, a.size, a.dtype))
                parts.append((off, a.view(np.uint8).reshape(-1)))
                off += -(-a.nbytes // self._ALIGN) * self._ALIGN
            raw = np.zeros(max(off, self._ALIGN), dtype=np.uint8)
            for o, b in parts:
                raw[o:o + b.size] = b
            buf, st, ev = upload(raw, device)
            d = {}
            for k, o, n, dt in spans:
                d[k] = buf[o:o + n * dt.itemsize].view(_TORCH_DTYPES[dt])
            self._dev[key] = d
            self._ready[key] = (st, ev)
        return d


# ----------------------------------------------------------------------------- recouple


def _pad_regions(regions):
    """``(dst_off, shape, dst_strides, src_strides, [(src_off, coeff), ...])`` regions as
    the tables of ``RecoupleProgram.from_arrays`` (ranks padded with unit dims)."""
    n = len(regions)
    R = max((len(r[1]) for r in regions), default=0)
    doff = np.fromiter((r[0] for r in regions), dtype=np.int64, count=n)
    shp = np.ones((n, R), dtype=np.int64)
    dstr = np.zeros((n, R), dtype=np.int64)
    sstr = np.zeros((n, R), dtype=np.int64)
    for j, (_, shape, ds, ss, _) in enumerate(regions):
        r = len(shape)
        shp[j, :r], dstr[j, :r], sstr[j, :r] = shape, ds, ss
    counts = np.fromiter((len(r[4]) for r in regions), dtype=np.int64, count=n)
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    terms = [t for r in regions for t in r[4]]
    tsrc = np.fromiter((t[0] for t in terms), dtype=np.int64, count=len(terms))
    tco = np.fromiter(((t[1].real if isinstance(t[1], complex) else t[1]) for t in terms),
                      dtype=np.float64, count=len(terms))
    return doff, shp, dstr, sstr, ptr, tsrc, tco


class RecoupleProgram:
    """A compiled ``tn_recouple`` launch.

    ``regions`` is a list of ``(dst_off, shape, dst_strides, src_strides, [(src_off,
    coeff), ...])``; dims are given in destination order.  ``from_arrays`` builds the same
    program from those regions as tables (one row per region, ranks padded with unit dims;
    terms as a CSR list ``term_ptr``/``term_src``/``term_coef``).
    """

    def __init__(self, regions, src_size: int, dst_size: int) -> None:
        self._build(*_pad_regions(regions), src_size, dst_size)

    @classmethod
    def from_arrays(cls, dst_off, shape, dst_strides, src_strides, term_ptr, term_src, term_coef,
                    src_size: int, dst_size: int) -> "RecoupleProgram":
        prog = cls.__new__(cls)
        prog._build(np.asarray(dst_off, dtype=np.int64), np.asarray(shape, dtype=np.int64),
                    np.asarray(dst_strides, dtype=np.int64), np.asarray(src_strides, dtype=np.int64),
                    np.asarray(term_ptr, dtype=np.int64), np.asarray(term_src, dtype=np.int64),
                    np.asarray(term_coef, dtype=np.float64), src_size, dst_size)
        return prog

    def _build(self, doff, shp, dstr, sstr, ptr, tsrc, tco, src_size: int, dst_size: int) -> None:
        n = len(doff)
        c = np.ascontiguousarray
        doff, shp, dstr, sstr, ptr, tsrc = (c(x, dtype=np.int64) for x in (doff, shp, dstr, sstr, ptr, tsrc))
        jobs = np.empty((n, JOB_WORDS), dtype=np.int64)
        counts = np.empty(n, dtype=np.int64)
        stats = np.zeros(3, dtype=np.int64)
        addr = lambda x: x.ctypes.data  # noqa: E731
        # unit dims dropped, contiguous dims merged, mode chosen, every access checked
        # against the buffers, tiles counted (the small-tile rule applied): one host call
        _native.check_host(_native.host_lib().tn_recouple_plan(
            n, shp.shape[1] if shp.ndim == 2 else 0, addr(doff), addr(shp), addr(dstr), addr(sstr), addr(ptr),
            addr(tsrc), src_size, dst_size, _LIN_TILE, _LIN_SMALL_TILE, _LDS_T, _SMALL_PROGRAM_TILES,
            addr(jobs), addr(counts), addr(stats)))
        self.elements, self.moved_terms, total = (int(v) for v in stats)
        self.njobs = n
        # tile table: (job, tile index within the job) for every job's tiles, in job order
        starts = np.cumsum(counts) - counts
        tiles_arr = np.empty((total, 2), dtype=np.int64)
        tiles_arr[:, 0] = np.repeat(np.arange(n, dtype=np.int64), counts)
        tiles_arr[:, 1] = np.arange(total, dtype=np.int64) - np.repeat(starts, counts)
        self.ntiles = total
        self.src_size, self.dst_size = src_size, dst_size
        self._tables = _DeviceCache(
            jobs=jobs.reshape(-1),
            tsrc=tsrc,
            tco=c(tco, dtype=np.float64),
            tiles=tiles_arr.reshape(-1),
        )

    def bytes_moved(self, itemsize: int) -> int:
        """Algorithmic HBM bytes: each term's source read once, each destination written once."""
        return (self.moved_terms + self.elements) * itemsize

    def run(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        """One launch; when ``dst`` has another float type than ``src`` the conversion
        rides along with the permutation."""
        if self.ntiles == 0:
            return
        if src.numel() < self.src_size or dst.numel() < self.dst_size:
            raise ValueError("recouple: buffers smaller than the compiled plan")
        t = self._tables.get(dst.device)
        _native.check(
            _native.lib().tn_recouple_cast(
                _dtype_code(src.dtype),
                _dtype_code(dst.dtype),
                src.data_ptr(),
                dst.data_ptr(),
                t["jobs"].data_ptr(),
                t["tsrc"].data_ptr(),
                t["tco"].data_ptr(),
                t["tiles"].data_ptr(),
                self.ntiles,
                _native.stream_ptr(),
            )
        )


def _plan_arrays(S: TensorStructure, D: TensorStructure, perm, terms):
    """The destination-region tables of a (perm, terms) block map between two flat
    layouts: per target block its offset, shape and strides, the source strides permuted
    into destination order, and its terms (source offsets, coefficients) in plan order."""
    T = len(terms)
    try:
        t = np.array(terms, dtype=np.float64).reshape(T, 3)
    except TypeError:  # complex coefficients: their real parts are the stored ones
        t = np.array(terms, dtype=np.complex128).reshape(T, 3)
    ti, tj, tc = t[:, 0].real.astype(np.int64), t[:, 1].real.astype(np.int64), t[:, 2].real
    counts = np.bincount(tj, minlength=D.num_blocks) if T else np.zeros(D.num_blocks, np.int64)
    if len(counts) != D.num_blocks or int((counts > 0).sum()) != D.num_blocks:
        raise ValueError(
            f"the plan fills {int((counts > 0).sum())} of {D.num_blocks} target blocks; the "
            "provider's coefficients dropped terms"
        )
    order = np.argsort(tj, kind="stable")
    ptr = np.zeros(D.num_blocks + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    sl, dl = S.layout, D.layout
    first = ti[order][ptr[:-1]]
    sstr = sl.strides[first][:, list(perm)] if len(perm) else np.zeros((D.num_blocks, 0), np.int64)
    return dl.offsets, dl.shapes, dl.strides, sstr, ptr, sl.offsets[ti[order]], tc[order]


def regions_for_plan(S: TensorStructure, D: TensorStructure, perm, terms):
    """Destination regions of a (perm, terms) block map between two flat layouts."""
    doff, shp, dstr, sstr, ptr, tsrc, tco = _plan_arrays(S, D, perm, terms)
    d_off, d_shp, d_str, s_str = doff.tolist(), shp.tolist(), dstr.tolist(), sstr.tolist()
    src, co = tsrc.tolist(), tco.tolist()
    return [(d_off[j], tuple(d_shp[j]), tuple(d_str[j]), tuple(s_str[j]),
             list(zip(src[ptr[j]:ptr[j + 1]], co[ptr[j]:ptr[j + 1]])))
            for j in range(D.num_blocks)]


@lru_cache(maxsize=4096)
def recouple_program(S: TensorStructure, D: TensorStructure, perm: tuple, terms: tuple):
    return RecoupleProgram.from_arrays(*_plan_arrays(S, D, perm, terms), S.size, D.size)


def inplace_map(S: TensorStructure, D: TensorStructure, perm, terms):
    """Offsets of D's blocks inside S's buffer when the map is a pure re-indexing
    (identity permutation, one-to-one, coefficient 1); else None."""
    if tuple(perm) != tuple(range(len(perm))):
        return None
    if len(terms) != D.num_blocks or len(terms) != S.num_blocks:
        return None
    offs = np.empty(D.num_blocks, dtype=np.int64)
    seen = np.zeros(D.num_blocks, dtype=bool)
    sl = S.layout.offsets
    for i, j, c in terms:
        if c != 1 or seen[j]:
            return None
        seen[j] = True
        offs[j] = sl[i]
    return offs if seen.all() else None


def scatter_map(S: TensorStructure, D: TensorStructure, perm, terms):
    """Offsets, inside D's buffer, of the target of each of S's blocks when the map is a
    pure re-indexing (the output-side dual of ``inplace_map``); else None."""
    if inplace_map(S, D, perm, terms) is None:
        return None
    offs = np.empty(S.num_blocks, dtype=np.int64)
    dl = D.layout.offsets
    for i, j, _ in terms:
        offs[i] = dl[j]
    return offs


# ----------------------------------------------------------------------------- gemm

_TILES = ((128, 128), (128, 64), (64, 128), (64, 64), (32, 32))
_CUS: dict = {}


def device_cus() -> int:
    """Compute units of the current device (queried once through the C ABI)."""
    if "n" not in _CUS:
        import ctypes

        n = ctypes.c_int(0)
        _native.check(_native.lib().tn_device_cus(ctypes.byref(n)))
        _CUS["n"] = int(n.value)
    return _CUS["n"]


def resident_workgroups(precision: str, tile) -> int:
    """Workgroups of the GEMM kernel resident at once on the device: the LDS footprint of
    the double-buffered K step bounds workgroups per CU (160 KiB of LDS per CU), capped at
    8 (4 waves each, 8 waves per SIMD) and at 2 for the 128x128 tile (register bound)."""
    bm, bn = tile
    per_row = 320 if precision == "bf16x3" else 288
    lds = (bm + bn) * per_row
    per_cu = max(1, min(8, (160 * 1024) // lds))
    if bm * bn >= 128 * 128:
        per_cu = min(per_cu, 2)
    return device_cus() * per_cu


SHORT_K = 64  # output-weighted mean K at or below which a launch is output-write bound


def choose_tile(Ms, Ns, Ks) -> tuple:
    """Pick the workgroup tile from the problem-size histogram: the largest tile whose
    padded work stays within 15% of the best achievable utilisation -- except for short-K
    launches (mean K weighted by output size <= SHORT_K), which gain nothing from a large
    tile's operand reuse and everything from more workgroups in flight: 64 x 64 there
    (measured on the cfg1 einsum, K <= 32: 26.3 us against 31.4 us with 128 x 64, FP64)."""
    Ms, Ns, Ks = (np.asarray(x, dtype=np.float64) for x in (Ms, Ns, Ks))
    useful = float((Ms * Ns * Ks).sum()) or 1.0
    util = {}
    for bm, bn in _TILES:
        padded = float((np.ceil(Ms / bm) * bm * np.ceil(Ns / bn) * bn * Ks).sum()) or 1.0
        util[(bm, bn)] = useful / padded
    best = max(util.values())
    out = float((Ms * Ns).sum()) or 1.0
    if useful / out <= SHORT_K and util[(64, 64)] >= 0.85 * best:
        return (64, 64)
    for cfg in _TILES:
        if util[cfg] >= 0.85 * best:
            return cfg
    return (64, 64)


def _band_layout(tiles: np.ndarray, nprob: int, nxcd: int = 8) -> np.ndarray:
    """Tile table for one workgroup per tile: the tiles in (problem, tile row, tile col)
    order cut into ``nxcd`` contiguous runs of equal work (K plus a per-tile fixed cost),
    XCD x taking run x, each run ordered longest K first (ties in row-major order).  A
    problem's tiles -- which share its operand panels -- land on one or two XCDs, so each
    panel is fetched from HBM by at most two L2s; slices padded with empty-problem tiles."""
    order = np.lexsort((tiles[:, 2], tiles[:, 1], tiles[:, 0]))
    t = tiles[order]
    work = t[:, 3].astype(np.float64) + 64.0
    cum = np.cumsum(work)
    bounds = np.searchsorted(cum, cum[-1] * np.arange(1, nxcd) / nxcd, side="right")
    parts = np.split(np.arange(len(t)), bounds)
    L = max(len(pt) for pt in parts)
    out = np.zeros((nxcd * L, 4), dtype=tiles.dtype)
    out[:, 0] = nprob
    for x, pt in enumerate(parts):
        pt = pt[np.argsort(-t[pt, 3], kind="stable")]
        out[x * L: x * L + len(pt)] = t[pt]
    return out


def _xcd_layout(tiles: np.ndarray, nprob: int, nxcd: int = 8, grid: int = 0) -> np.ndarray:
    """Tile table for a launch of one workgroup per tile on a device whose workgroup b
    runs on XCD ``b % nxcd`` and whose GEMM kernel hands XCD x the table slice
    ``[x n / nxcd, (x + 1) n / nxcd)``.

    Tiles (rows ``(problem, tile row, tile col, K)``) are cut, in row-major order within each
    problem, into units of about 1/48 of the total work -- neighbouring tiles share operand
    panels, so each unit stays on one XCD's L2; units go to the least loaded XCD, largest first
    (LPT); every XCD runs its tiles longest K first.  The slices are padded to equal
    length with tiles of problem ``nprob`` (an empty problem row).

    ``grid`` > 0 (a multiple of ``nxcd``): the layout of a persistent launch of ``grid``
    workgroups, where workgroup i of XCD x walks table positions ``x grid/nxcd + i + k grid``
    (k = 0, 1, ...); each XCD's longest-first list is dealt round-robin over its workgroups."""
    work = tiles[:, 3].astype(np.float64) + 64.0  # K plus a per-tile fixed cost
    cap = max(work.sum() / (nxcd * 6), 1.0)
    order = np.lexsort((tiles[:, 2], tiles[:, 1], tiles[:, 0]))
    # a unit closes once its work reaches cap or its problem ends; every tile of a problem
    # carries the same K, so a unit holds n tiles (n w >= cap, n minimal; integer-valued
    # work, so the products are exact) -- the units follow from the problem boundaries
    # without a walk over the tiles
    ps = tiles[order, 0]
    starts = np.flatnonzero(np.r_[True, ps[1:] != ps[:-1]]) if len(order) else np.zeros(0, np.int64)
    ends = np.r_[starts[1:], len(order)]
    spans, uwork = [], []
    for s0, e0 in zip(starts.tolist(), ends.tolist()):
        w = float(work[order[s0]])
        n = max(1, int(np.ceil(cap / w)))
        while n > 1 and (n - 1) * w >= cap:
            n -= 1
        while n * w < cap:
            n += 1
        for a in range(s0, e0, n):
            b = min(e0, a + n)
            spans.append((a, b))
            uwork.append((b - a) * w)
    loads = np.zeros(nxcd)
    parts = [[] for _ in range(nxcd)]
    for u in sorted(range(len(spans)), key=lambda i: -uwork[i]):
        x = int(np.argmin(loads))
        a, b = spans[u]
        parts[x].extend(order[a:b].tolist())
        loads[x] += uwork[u]
    if grid > 0:
        per = grid // nxcd
        rounds = max(-(-len(pt) // per) for pt in parts)
        out = np.zeros((grid * rounds, 4), dtype=tiles.dtype)
        out[:, 0] = nprob
        for x, pt in enumerate(parts):
            pt = np.asarray(pt, dtype=np.int64)
            pt = pt[np.argsort(-tiles[pt, 3], kind="stable")]
            j = np.arange(len(pt))
            out[x * per + j % per + (j // per) * grid] = tiles[pt]
        return out
    L = max(len(pt) for pt in parts)
    out = np.zeros((nxcd * L, 4), dtype=tiles.dtype)
    out[:, 0] = nprob
    for x, pt in enumerate(parts):
        pt = np.asarray(pt, dtype=np.int64)
        pt = pt[np.argsort(-tiles[pt, 3], kind="stable")]
        out[x * L: x * L + len(pt)] = tiles[pt]
    return out


class GemmProgram:
    """A compiled ``tn_gemm_grouped`` launch over a list of block problems.

    ``problems``: list of ``(M, N, c_off, ldc, [(a_off, b_off, K, lda, ldb), ...])``.

    When the problems give too few tiles to fill the device, long-K problems are split
    into K slices whose partial products go to a workspace; one multi-term recouple launch
    then sums the slices in a fixed order into C (deterministic split-K).  Tiles are
    issued longest-K first so the tail of the launch is short tiles.

    ``gather=True``: every problem writes its slices (one when unsplit) to the workspace,
    row-major with leading dimension N, and no reduction is compiled; ``slices[p]`` lists
    the workspace offsets of problem p so the caller can fold the sum into a pass of its
    own (the problems' ``c_off``/``ldc`` are ignored).
    """

    SPLIT_TARGET_TILES = 1536
    SPLIT_MIN_K = 256
    XCD_LAYOUT = "lpt"  # tile table layout over the 8 XCDs: "lpt", "deal" or "none"
    SKINNY_MAX_W = 8    # the kernel's register line width (kSkinnyMax)
    SKINNY_MAX_K = 64

    ROWS_TILES = ((64, 64), (128, 64), (64, 128), (128, 128))  # tn_gemm_grouped_rows' tiles

    def __init__(self, problems, a_size, b_size, c_size, trans_a=False, trans_b=False,
                 tile=None, split: bool = True, gather: bool = False, a_rows=None) -> None:
        """``a_rows`` (optional, one integer array per problem): row m of problem p's A
        starts ``a_rows[p][m]`` elements after each segment's ``a_off`` (k contiguous), so a
        sector matrix whose rows lie in several blocks is read in place; the launch then
        goes through ``tn_gemm_grouped_rows`` (FP32/BF16x3 arithmetic, A in FP64 or FP32)."""
        self.trans_a, self.trans_b = bool(trans_a), bool(trans_b)
        self.gather = bool(gather)
        self.a_size, self.b_size, self.c_size = a_size, b_size, c_size
        self.problems = problems  # unsplit, for merging into batched launches
        self.flops = 0
        if a_rows is not None:
            if self.trans_a or self.trans_b:
                raise ValueError("a row-gathered A is read k-contiguous against a row-major B")
            if len(a_rows) != len(problems):
                raise ValueError("a_rows needs one row table per problem")
            a_rows = [np.asarray(r, dtype=np.int64).reshape(-1) for r in a_rows]
        self.a_rows = a_rows
        for i, (M, N, c_off, ldc, sl) in enumerate(problems):
            if not gather and (c_off < 0 or c_off + (M - 1) * ldc + N > c_size):
                raise ValueError("gemm problem writes outside the output buffer")
            rows = a_rows[i] if a_rows is not None else None
            if rows is not None and len(rows) != M:
                raise ValueError("a row table needs one entry per row of its problem")
            for a_off, b_off, K, lda, ldb in sl:
                if rows is None:
                    a_lo = a_off
                    a_hi = a_off + ((K - 1) * lda + M if self.trans_a else (M - 1) * lda + K)
                else:
                    a_lo = a_off + (int(rows.min()) if M else 0)
                    a_hi = a_off + (int(rows.max()) if M else 0) + K
                b_hi = b_off + ((N - 1) * ldb + K if self.trans_b else (K - 1) * ldb + N)
                if a_lo < 0 or b_off < 0 or a_hi > a_size or b_hi > b_size:
                    raise ValueError("gemm segment reads outside its operand buffer")
                # the kernel addresses a segment with 32-bit byte offsets from its base and
                # marks rows outside a tile with offsets >= 2^31
                if max(a_hi - a_off, b_hi - b_off) * 8 >= 2**31:
                    raise ValueError("gemm segment spans 2 GiB or more of one operand")
                self.flops += 2 * M * N * K
        Ks0 = [max(sum(sg[2] for sg in sl), 1) for *_, sl in problems]
        self.tile = tuple(tile) if tile else (
            choose_tile([p[0] for p in problems], [p[1] for p in problems], Ks0) if problems else (64, 64))
        # skinny launch: every output at most SKINNY_MAX_W wide in one dimension and K short
        # -- a streaming pass, one thread per output line, instead of mostly-empty MFMA tiles
        self.skinny = None
        if problems and max(Ks0) <= self.SKINNY_MAX_K:
            if all(p[1] <= self.SKINNY_MAX_W for p in problems):
                self.skinny = 1
            elif all(p[0] <= self.SKINNY_MAX_W for p in problems):
                self.skinny = 0
        if self.skinny is not None and (gather or a_rows is not None):
            self.skinny = None  # the skinny kernel writes C directly and reads A by lda
        if self.skinny is not None:
            split = False
        bm, bn = self.tile
        ntiles = sum(-(-p[0] // bm) * -(-p[1] // bn) for p in problems)
        factor = 1
        if split and problems and ntiles < self.SPLIT_TARGET_TILES:
            factor = min(16, -(-self.SPLIT_TARGET_TILES // max(ntiles, 1)))
        rows, segs, red = [], [], []
        self.slices = []
        ws = 0
        # word 7 of a problem row: its first entry in the concatenated row table
        rbase = np.zeros(len(problems), dtype=np.int64)
        if a_rows is not None and len(problems):
            rbase[1:] = np.cumsum([len(r) for r in a_rows])[:-1]
        for pi, ((M, N, c_off, ldc, sl), kt) in enumerate(zip(problems, Ks0)):
            rb = int(rbase[pi])
            nsplit = 1
            if factor > 1 and kt >= self.SPLIT_MIN_K:
                nsplit = max(1, min(factor, kt // (self.SPLIT_MIN_K // 2)))
            if nsplit == 1 and not gather:
                rows.append((M, N, c_off, ldc, len(segs), len(segs) + len(sl), 0, rb, kt))
                segs.extend((*sg, 0) for sg in sl)
                continue
            chunks = self._split_segments(sl, kt, nsplit) if nsplit > 1 else [list(sl)]
            offs = []
            for ch in chunks:
                rows.append((M, N, ws, N, len(segs), len(segs) + len(ch), 1, rb, sum(c[2] for c in ch)))
                segs.extend((*c, 0) for c in ch)
                offs.append(ws)
                ws += M * N
            self.slices.append(offs)
            if not gather:
                red.append((c_off, (M, N), (ldc, 1), (N, 1), [(o, 1.0) for o in offs]))
        self.ws_size = ws
        self.reduce = RecoupleProgram(red, ws, c_size) if red else None
        self._rows = rows
        self.nproblems = len(rows)
        self._probs = np.asarray([r[:8] for r in rows], dtype=np.int64).reshape(-1, PROB_WORDS)
        self._segs = np.asarray(segs, dtype=np.int64).reshape(-1, SEG_WORDS)
        self._arows = (np.concatenate(a_rows) if a_rows else np.zeros(1, np.int64)) \
            if a_rows is not None else None
        if a_rows is not None and self.tile not in self.ROWS_TILES:
            raise ValueError(f"a row-gathered A runs the MFMA tiles {self.ROWS_TILES}, not {self.tile}")
        self.Ms = [r[0] for r in rows]
        self.Ns = [r[1] for r in rows]
        self.Ks = [r[8] for r in rows]
        self._by_tile: dict = {}

    def _split_segments(self, sl, kt, nsplit):
        """Cut the K range into ``nsplit`` slices at multiples of 32, splitting segments."""
        step = -(-kt // nsplit)
        step = -(-step // 32) * 32
        chunks, cur, filled = [], [], 0
        for a_off, b_off, K, lda, ldb in sl:
            k0 = 0
            while k0 < K:
                take = min(K - k0, step - filled)
                ao = a_off + (k0 * lda if self.trans_a else k0)
                bo = b_off + (k0 if self.trans_b else k0 * ldb)
                cur.append((ao, bo, take, lda, ldb))
                filled += take
                k0 += take
                if filled == step:
                    chunks.append(cur)
                    cur, filled = [], 0
        if cur:
            chunks.append(cur)
        return chunks

    def _tables(self, tile, device, grid: int = 0):
        if self.skinny is not None:
            tile = "skinny"
        key = tile if grid <= 0 else (tile, grid)
        d = self._by_tile.get(key)
        if d is None and tile == "skinny":
            lines = np.asarray(self.Ms if self.skinny else self.Ns, dtype=np.int64)
            per = -(-lines // 256)
            total = int(per.sum())
            tiles = np.empty((total, 2), dtype=np.int32)
            tiles[:, 0] = np.repeat(np.arange(len(lines), dtype=np.int32), per)
            tiles[:, 1] = np.arange(total, dtype=np.int64) - np.repeat(np.cumsum(per) - per, per)
            d = _DeviceCache(probs=self._probs.reshape(-1), segs=self._segs.reshape(-1),
                             tiles=tiles.reshape(-1))
            d.ntiles = len(tiles)
            self._by_tile[tile] = d
        if d is None:
            bm, bn = tile
            Ms, Ns = np.asarray(self.Ms, np.int64), np.asarray(self.Ns, np.int64)
            tm, tn = -(-Ms // bm), -(-Ns // bn)
            cnt = tm * tn
            total = int(cnt.sum())
            local = np.arange(total, dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            tn_rep = np.repeat(tn, cnt)
            tiles = np.empty((total, 4), dtype=np.int64)
            tiles[:, 0] = np.repeat(np.arange(len(Ms), dtype=np.int64), cnt)
            tiles[:, 1] = local // np.maximum(tn_rep, 1)
            tiles[:, 2] = local % np.maximum(tn_rep, 1)
            tiles[:, 3] = np.repeat(np.asarray(self.Ks, np.int64), cnt)
            if grid > 0:
                tiles = _xcd_layout(tiles, self.nproblems, grid=grid)
            elif self.XCD_LAYOUT == "lpt" and len(tiles) >= 64:
                tiles = _xcd_layout(tiles, self.nproblems)
            elif self.XCD_LAYOUT == "band" and len(tiles) >= 64:
                tiles = _band_layout(tiles, self.nproblems)
            else:
                # longest K first; ties keep problem order
                tiles = tiles[np.argsort(-tiles[:, 3], kind="stable")]
                if self.XCD_LAYOUT == "deal" and len(tiles) % 8 == 0:
                    tiles = tiles[np.arange(len(tiles)).reshape(-1, 8).T.reshape(-1)]
            # word 3: the problem's first segment (the kernel fetches it beside the problem)
            seg0 = np.append(self._probs[:, 4], 0)
            tiles[:, 3] = seg0[tiles[:, 0]]
            tiles = tiles.astype(np.int32)
            # one all-zero problem row past the real ones: the padding tiles of the XCD
            # layout point at it (M = N = 0, no segments: the workgroup exits at once)
            probs = np.vstack([self._probs, np.zeros((1, PROB_WORDS), np.int64)])
            extra = {} if self._arows is None else {"arows": self._arows}
            d = _DeviceCache(probs=probs.reshape(-1), segs=self._segs.reshape(-1),
                             tiles=tiles.reshape(-1), **extra)
            d.ntiles = len(tiles)
            self._by_tile[key] = d
        return d.get(device), d.ntiles

    def run(self, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, precision: str | None = None,
            tile=None, grid: int = 0, events: list | None = None, reduce_events: list | None = None) -> None:
        """One grouped launch (plus the split-K reduction when the program has one).  With
        ``events`` a list, the GEMM kernel alone is bracketed by two recorded events,
        appended as ``(start, end, flops)``; ``reduce_events`` does the same for the
        reduction with its algorithmic bytes.  A gather program takes its workspace as C."""
        if self.nproblems == 0:
            return
        if self.a_rows is not None:
            return self._run_rows(A, B, C, precision, tile, grid, events)
        if not (A.dtype == B.dtype == C.dtype):
            raise TypeError("gemm: operand dtypes differ")
        c_need = self.ws_size if self.gather else self.c_size
        if A.numel() < self.a_size or B.numel() < self.b_size or C.numel() < c_need:
            raise ValueError("gemm: buffers smaller than the compiled plan")
        prec = precision or gemm_precision(C.dtype)
        tile = tuple(tile) if tile else self.tile
        if grid < 0 and self.skinny is None:
            # persistent: every resident workgroup walks its share of the tile list (dealt
            # per XCD, longest first), prefetching the next tile's first K step while it
            # finishes the current one
            grid = resident_workgroups(prec, tile)
            t, ntiles = self._tables(tile, C.device, grid)
        else:
            grid = max(grid, 0)
            t, ntiles = self._tables(tile, C.device)
        if self.gather:
            W = C
        else:
            W = torch.empty(self.ws_size, dtype=C.dtype, device=C.device) if self.ws_size else C
        if events is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        if self.skinny is not None:
            _native.check(_native.lib().tn_gemm_skinny(
                _dtype_code(C.dtype), int(self.trans_a), int(self.trans_b), self.skinny,
                A.data_ptr(), B.data_ptr(), C.data_ptr(), W.data_ptr(), t["probs"].data_ptr(),
                t["segs"].data_ptr(), t["tiles"].data_ptr(), ntiles, _native.stream_ptr()))
            if events is not None:
                e1.record()
                events.append((e0, e1, self.flops))
            return
        _native.check(
            _native.lib().tn_gemm_grouped(
                _native.PRECISIONS[prec],
                _dtype_code(C.dtype),
                int(self.trans_a),
                int(self.trans_b),
                A.data_ptr(),
                B.data_ptr(),
                C.data_ptr(),
                W.data_ptr(),
                t["probs"].data_ptr(),
                t["segs"].data_ptr(),
                t["tiles"].data_ptr(),
                ntiles,
                tile[0],
                tile[1],
                grid,
                _native.stream_ptr(),
            )
        )
        if events is not None:
            e1.record()
            events.append((e0, e1, self.flops))
        if self.reduce is not None:
            if reduce_events is not None:
                r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                r0.record()
            self.reduce.run(W, C)
            if reduce_events is not None:
                r1.record()
                reduce_events.append((r0, r1, self.reduce.bytes_moved(C.element_size())))


    def _run_rows(self, A, B, C, precision, tile, grid, events) -> None:
        """The launch of a program with a row-gathered A (``tn_gemm_grouped_rows``)."""
        if A.dtype not in (torch.float64, torch.float32) or B.dtype != torch.float32 or C.dtype != torch.float32:
            raise TypeError("gemm with a row table: A FP64 or FP32, B and C FP32")
        prec = precision or "fp32"
        if prec not in ("fp32", "bf16x3"):
            raise ValueError("gemm with a row table: FP32 or BF16x3 arithmetic")
        tile = tuple(tile) if tile else self.tile
        if grid != 0 or tile not in self.ROWS_TILES:
            raise ValueError(f"gemm with a row table: one workgroup per tile, tiles {self.ROWS_TILES}")
        c_need = self.ws_size if self.gather else self.c_size
        if A.numel() < self.a_size or B.numel() < self.b_size or C.numel() < c_need:
            raise ValueError("gemm: buffers smaller than the compiled plan")
        t, ntiles = self._tables(tile, C.device)
        if self.gather:
            W = C
        else:
            W = torch.empty(self.ws_size, dtype=C.dtype, device=C.device) if self.ws_size else C
        if events is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        _native.check(_native.lib().tn_gemm_grouped_rows(
            _native.PRECISIONS[prec], _dtype_code(A.dtype), A.data_ptr(), t["arows"].data_ptr(),
            B.data_ptr(), C.data_ptr(), W.data_ptr(), t["probs"].data_ptr(), t["segs"].data_ptr(),
            t["tiles"].data_ptr(), ntiles, tile[0], tile[1], _native.stream_ptr()))
        if events is not None:
            e1.record()
            events.append((e0, e1, self.flops))
        if self.reduce is not None:
            self.reduce.run(W, C)

    def run_static(self, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, precision: str | None = None,
                   tile=None) -> None:
        """The same product through ``tn_gemm_grouped_static``: the split-K partials go to
        the library's static device scratch and the slice sum follows in the same call, so
        nothing is allocated (for callers that cannot allocate).  Calls that overlap on one
        device share that scratch and must be serialized by the caller."""
        if self.nproblems == 0:
            return
        if self.skinny is not None or self.gather:
            raise ValueError("run_static covers the MFMA grouped launch written to C, not skinny "
                             "or gather-mode programs")
        if not (A.dtype == B.dtype == C.dtype):
            raise TypeError("gemm: operand dtypes differ")
        if A.numel() < self.a_size or B.numel() < self.b_size or C.numel() < self.c_size:
            raise ValueError("gemm: buffers smaller than the compiled plan")
        prec = precision or gemm_precision(C.dtype)
        tile = tuple(tile) if tile else self.tile
        t, ntiles = self._tables(tile, C.device)
        if self.reduce is not None and self.reduce.ntiles:
            r = self.reduce._tables.get(C.device)
            red = (r["jobs"].data_ptr(), r["tsrc"].data_ptr(), r["tco"].data_ptr(), r["tiles"].data_ptr(),
                   self.reduce.ntiles)
        else:
            red = (None, None, None, None, 0)
        _native.check(_native.lib().tn_gemm_grouped_static(
            _native.PRECISIONS[prec], _dtype_code(C.dtype), int(self.trans_a), int(self.trans_b),
            A.data_ptr(), B.data_ptr(), C.data_ptr(), t["probs"].data_ptr(), t["segs"].data_ptr(),
            t["tiles"].data_ptr(), ntiles, tile[0], tile[1], 0, self.ws_size, *red, _native.stream_ptr()))


# ----------------------------------------------------------------------------- contraction


def compose_plans(first, second):
    """(perm, terms) of applying plan ``first`` then plan ``second``."""
    p1, t1 = first
    p2, t2 = second
    return tuple(p1[i] for i in p2), compose_maps(t1, t2)


@dataclass
class _Operand:
    """How a GEMM operand is reached: in place (offsets into the caller's buffer) or via a
    recouple program into a scratch buffer of ``structure``'s layout."""

    structure: TensorStructure
    offsets: np.ndarray
    program: RecoupleProgram | None


def _operand(S: TensorStructure, plan) -> _Operand:
    D = plan.new_structure
    offs = inplace_map(S, D, plan.perm, plan.terms)
    if offs is not None:
        return _Operand(D, offs, None)
    prog = recouple_program(S, D, tuple(plan.perm), tuple(plan.terms))
    return _Operand(D, D.layout.offsets, prog)


class ContractionExec:
    """Compiled ``tensordot(a, b, axes)`` followed by an optional output transpose."""

    def __init__(self, sa: TensorStructure, sb: TensorStructure, axes: tuple, out_axes=None):
        cp = contraction_plan(sa, sb, axes)
        rp_a = repartition_plan(sa, cp.a_outputs, cp.a_inputs)
        rp_b = repartition_plan(sb, cp.b_outputs, cp.b_inputs)
        self.a_plan = (tuple(rp_a.perm), tuple(rp_a.terms))  # sa -> the GEMM's A layout
        self.a = _operand(sa, rp_a)
        self.b = _operand(sb, rp_b)
        Ap, Bp = self.a.structure, self.b.structure
        n_out_a = len(cp.a_outputs)
        n_out_b = len(cp.b_outputs)
        Cp = TensorStructure((*(Ap.legs[i] for i in Ap.out_axes), *(Bp.legs[i] for i in Bp.in_axes)))
        # C' -> final: restore sides, tensordot order, then the caller's transpose
        r = repartition_plan(Cp, cp.restore_outputs, cp.restore_inputs)
        plan = (r.perm, r.terms)
        cur = r.new_structure
        t = permutation_plan(cur, cp.final_transpose)
        plan = compose_plans(plan, (t.axes, t.terms))
        cur = t.new_structure
        if out_axes is not None and tuple(out_axes) != tuple(range(cur.ndim)):
            t = permutation_plan(cur, tuple(out_axes))
            plan = compose_plans(plan, (t.axes, t.terms))
            cur = t.new_structure
        self.out_structure = cur
        self.c_structure = Cp
        self.c_plan = (tuple(plan[0]), tuple(plan[1]))  # the GEMM's C layout -> out
        c_offs = scatter_map(Cp, cur, plan[0], plan[1])
        if c_offs is None:
            self.c_program = recouple_program(Cp, cur, tuple(plan[0]), tuple(plan[1]))
            c_offs = Cp.layout.offsets
            c_size = Cp.size
        else:
            self.c_program = None
            c_size = cur.size
        # ---- block problems of the composition C' = A' o B': which A and B blocks meet in
        # which C block is fixed by the sector pattern (shared by every structure of it);
        # only the extents and offsets are this structure's own
        shapes_a = Ap.shapes
        shapes_c = Cp.shapes
        ptr, seg_a, seg_b = _block_pairing(pattern_of(Ap), pattern_of(Bp), pattern_of(Cp))
        # row / column / inner extents of every block, one vectorised product each
        sc = np.asarray(shapes_c, dtype=np.int64).reshape(len(shapes_c), Cp.ndim)
        sa_ = np.asarray(shapes_a, dtype=np.int64).reshape(len(shapes_a), Ap.ndim)
        Ms = np.prod(sc[:, :n_out_a], axis=1)
        Ns = np.prod(sc[:, n_out_a:], axis=1)
        Ka = np.prod(sa_[:, n_out_a:], axis=1)
        seg_c = np.repeat(np.arange(len(Ms)), np.diff(ptr))
        seg_rows = zip(np.asarray(self.a.offsets, dtype=np.int64)[seg_a].tolist(),
                       np.asarray(self.b.offsets, dtype=np.int64)[seg_b].tolist(),
                       Ka[seg_a].tolist(), Ns[seg_c].tolist())
        segs_all = [(ao, bo, K, K, N) for ao, bo, K, N in seg_rows]
        ptr_l = ptr.tolist()
        problems = [(M, N, c, N, segs_all[lo:hi])
                    for M, N, c, lo, hi in zip(Ms.tolist(), Ns.tolist(), np.asarray(c_offs).tolist(),
                                               ptr_l[:-1], ptr_l[1:])]
        a_size = sa.size if self.a.program is None else Ap.size
        b_size = sb.size if self.b.program is None else Bp.size
        self.gemm = GemmProgram(problems, a_size, b_size, c_size)
        self.flops = self.gemm.flops
        _ = n_out_b

    def prepare_b(self, b_data: torch.Tensor) -> torch.Tensor:
        """``b`` in the layout the GEMM reads (its recouple applied, if it needs one): for an
        operand that stays fixed across many runs (an environment, an MPO tensor), do this
        once and pass the result with ``b_ready=True``."""
        if self.b.program is None:
            return b_data
        bp = torch.empty(self.b.structure.size, dtype=b_data.dtype, device=b_data.device)
        self.b.program.run(b_data, bp)
        return bp

    def run(self, a_data: torch.Tensor, b_data: torch.Tensor, precision=None, tile=None,
            out: torch.Tensor | None = None, events: list | None = None,
            reduce_events: list | None = None, b_ready: bool = False) -> torch.Tensor:
        """Run the contraction; when ``events`` is a list, the GEMM kernel launch is
        bracketed by two recorded events appended as ``(start, end, flops)`` (and a split-K
        reduction, if any, into ``reduce_events``).  ``b_ready``: ``b_data`` already went
        through ``prepare_b``."""
        dtype, dev = a_data.dtype, a_data.device
        if b_data.dtype != dtype:
            raise TypeError("tensordot: operands have different dtypes")
        if self.a.program is not None:
            ap = torch.empty(self.a.structure.size, dtype=dtype, device=dev)
            self.a.program.run(a_data, ap)
        else:
            ap = a_data
        bp = b_data if b_ready else self.prepare_b(b_data)
        if out is None:
            out = torch.empty(self.out_structure.size, dtype=dtype, device=dev)
        target = out if self.c_program is None else torch.empty(
            self.c_structure.size, dtype=dtype, device=dev)
        self.gemm.run(ap, bp, target, precision, tile, events=events, reduce_events=reduce_events)
        if self.c_program is not None:
            self.c_program.run(target, out)
        return out


@lru_cache(maxsize=1024)
def _block_pairing(pa: TensorStructure, pb: TensorStructure, pc: TensorStructure) -> tuple:
    """Segments of the block problems of ``C = A o B`` on sector patterns: for C block i,
    the A/B block pairs ``(seg_a[s], seg_b[s])`` for s in ``[ptr[i], ptr[i+1])``, A's input
    tree matching B's output tree, in A's block order."""
    a_rows = {}
    for j, key in enumerate(pa.block_order):
        a_rows.setdefault((key.coupled, key.output_tree), []).append((key.input_tree, j))
    b_lookup = {(key.output_tree, key.input_tree): j for j, key in enumerate(pb.block_order)}
    ptr, seg_a, seg_b = [0], [], []
    for key in pc.block_order:
        for kt, ja in a_rows.get((key.coupled, key.output_tree), ()):
            jb = b_lookup.get((kt, key.input_tree))
            if jb is not None:
                seg_a.append(ja)
                seg_b.append(jb)
        ptr.append(len(seg_a))
    return (np.asarray(ptr, dtype=np.int64), np.asarray(seg_a, dtype=np.int64),
            np.asarray(seg_b, dtype=np.int64))


@lru_cache(maxsize=2048)
def contraction_exec(sa: TensorStructure, sb: TensorStructure, axes: tuple, out_axes=None):
    return ContractionExec(sa, sb, axes, out_axes)
