"""The Lanczos step of many independent H_eff problems at once, one hipGraph per iteration.

``BatchedLanczos(batch, ncv)`` runs the recurrence of ``tenet.network.lanczos`` (reference
``network/dmrg.py:106-227``: three-term recurrence, happy breakdown below ``tol``, the
ground pair of the small tridiagonal, one recombination) for every problem of a
``BatchedHeff2`` together: the Krylov vectors of all problems live back to back in one
FP64 arena per basis vector, each iteration is ONE batched matvec (one launch per H_eff
stage for the whole batch) plus segmented FP64 vector kernels (``tn_dot_segmented``,
``tn_axpby_segmented``) whose scalars -- one per problem -- stay in device memory, and the
finish is one batched tridiagonal solve (``tn_tridiag_ground_batched``, one workgroup per
problem) plus the segmented recombination.  Each problem's vectors are re-orthogonalized
in FP64 against its own earlier ones, as in the single-problem solver.  Iterations and
the finish are captured as hipGraphs after one eager warm-up pass and replayed; the host
reads one energy per problem after the replays.

``refine`` (a ``BatchedHeff2`` of the same problems at FP64) closes every problem's step
with the residual-refinement Rayleigh-Ritz step of ``residual_refine`` on the span of the
Ritz vector and its FP64 residual, batched the same way.
"""

from __future__ import annotations

import numpy as np
import torch

from tenet import _native
from tenet.engine import _DeviceCache
from tenet.tensor import SymmetricTensor

__all__ = ["BatchedLanczos", "segment_chunks"]

CHUNK = 16384  # elements per workgroup of the segmented vector kernels


def segment_chunks(offsets, chunk: int = CHUNK):
    """``(chunks, seg_chunks)``: ``(segment, begin, end)`` triples covering every segment
    ``[offsets[s], offsets[s+1])`` in pieces of at most ``chunk`` elements, and the index
    range of each segment's chunks."""
    rows, starts = [], [0]
    for s in range(len(offsets) - 1):
        lo, hi = int(offsets[s]), int(offsets[s + 1])
        for b in range(lo, hi, chunk):
            rows.append((s, b, min(hi, b + chunk)))
        starts.append(len(rows))
    return np.asarray(rows, dtype=np.int64).reshape(-1, 3), np.asarray(starts, dtype=np.int64)


class _Segments:
    """Segmented FP64 vector operations over one arena layout."""

    def __init__(self, offsets, device) -> None:
        chunks, seg_chunks = segment_chunks(offsets)
        self.nseg = len(offsets) - 1
        self.nchunks = len(chunks)
        self.tables = _DeviceCache(chunks=chunks.reshape(-1), seg=seg_chunks).get(device)
        self.partial = torch.empty(max(self.nchunks, 1), dtype=torch.float64, device=device)

    def dot(self, x: torch.Tensor, y: torch.Tensor, out: torch.Tensor, stride: int = 1) -> None:
        """``out[s * stride] = <x_s, y_s>`` for every segment (FP64 arenas)."""
        _native.check(_native.lib().tn_dot_segmented(
            _native.TN_F64, x.data_ptr(), y.data_ptr(), None, self.tables["chunks"].data_ptr(), self.nchunks,
            self.partial.data_ptr(), self.tables["seg"].data_ptr(), self.nseg, out.data_ptr(), stride,
            _native.stream_ptr()))

    def axpby(self, a: float, a_dev, a_stride: int, a_inv: bool, x: torch.Tensor, b: float, b_dev,
              b_stride: int, y: torch.Tensor) -> None:
        """``y_s = (a * A_s) x_s + (b * B_s) y_s`` with ``A_s = a_dev[s * a_stride]``
        (inverted when ``a_inv``), ``B_s`` likewise (1 when ``b_dev`` is None)."""
        _native.check(_native.lib().tn_axpby_segmented(
            _native.TN_F64, self.tables["chunks"].data_ptr(), self.nchunks, a,
            None if a_dev is None else a_dev.data_ptr(), a_stride, int(a_inv), x.data_ptr(), b,
            None if b_dev is None else b_dev.data_ptr(), b_stride, 0, y.data_ptr(), _native.stream_ptr()))

    def normalize(self, x: torch.Tensor, tmp: torch.Tensor) -> None:
        """``x_s /= |x_s|`` in place (``tmp``: nseg doubles)."""
        self.dot(x, x, tmp)
        torch.sqrt(tmp, out=tmp)
        self.axpby(1.0, tmp, 1, True, x, 0.0, None, 1, x)


def _tridiag_batched(alphas, betas, n: int, ld: int, nprob: int, tol: float, values, coefs) -> None:
    _native.check(_native.lib().tn_tridiag_ground_batched(
        alphas.data_ptr(), betas.data_ptr(), n, ld, nprob, tol, values.data_ptr(), coefs.data_ptr(),
        _native.stream_ptr()))


class BatchedLanczos:
    """One Lanczos step for every problem of ``batch`` (a ``BatchedHeff2``): ``ncv``
    captured iteration graphs and one captured finish."""

    def __init__(self, batch, ncv: int = 3, tol: float = 1e-13, reorthogonalize: bool = True,
                 refine=None, graphs: bool = True) -> None:
        from tenet._native import TRIDIAG_MAX

        if not 1 <= ncv <= TRIDIAG_MAX:
            raise ValueError(f"ncv must lie in [1, {TRIDIAG_MAX}]")
        if refine is not None and (refine.x_size != batch.x_size or refine.count != batch.count):
            raise ValueError("refine must be a batch of the same problems")
        self.batch, self.ncv, self.tol, self.reorth, self.refine = batch, ncv, tol, reorthogonalize, refine
        dev = batch.fl.device if hasattr(batch, "fl") else batch.fr.device
        self.device = dev
        self.nseg = batch.count
        n = batch.x_size
        f64 = dict(dtype=torch.float64, device=dev)
        self.seg = _Segments(batch.x_off, dev)
        self.V = torch.zeros(ncv + 1, n, **f64)
        self.alphas = torch.zeros(self.nseg, ncv, **f64)   # problem-major: one tridiagonal per row
        self.betas = torch.zeros(self.nseg, ncv, **f64)
        self.g = torch.zeros(self.nseg, ncv, **f64)
        self.values = torch.zeros(self.nseg, **f64)
        self.coef = torch.zeros(self.nseg, **f64)
        self.tmp = torch.zeros(self.nseg, **f64)
        self.out = torch.zeros(n, **f64)
        if refine is not None:
            self.r_a = torch.zeros(self.nseg, 2, **f64)
            self.r_b = torch.zeros(self.nseg, 2, **f64)
            self.r_g = torch.zeros(self.nseg, 2, **f64)
            self.r_u = torch.zeros(n, **f64)
        self.graphs = None
        if graphs:
            # eager warm-up on random start vectors: compiles and uploads every table, so
            # the captured work holds no host-to-device copy
            self.V[0].normal_()
            self.start_(self.V[0])
            for j in range(ncv):
                self._iteration(j)
            self._finish()
            torch.cuda.synchronize(dev)
            self.graphs = []
            pool = torch.cuda.graph_pool_handle()
            for step in [*range(ncv), "finish"]:
                gr = torch.cuda.CUDAGraph()
                with torch.cuda.graph(gr, pool=pool):
                    self._finish() if step == "finish" else self._iteration(step)
                self.graphs.append(gr)

    def _matvec(self, b, x: torch.Tensor) -> torch.Tensor:
        return b.apply64(x)

    def start_(self, v: torch.Tensor) -> None:
        """``V[0] = v / |v|`` per problem (``v`` may be ``V[0]`` itself)."""
        if v.data_ptr() != self.V[0].data_ptr():
            self.V[0].copy_(v)
        self.seg.normalize(self.V[0], self.tmp)

    def _iteration(self, j: int) -> None:
        S, v = self.seg, self.V[j]
        w = self._matvec(self.batch, v)
        S.dot(v, w, self.alphas[:, j:], stride=self.ncv)
        S.axpby(-1.0, self.alphas[:, j:], self.ncv, False, v, 1.0, None, 1, w)
        if j:
            S.axpby(-1.0, self.betas[:, j - 1:], self.ncv, False, self.V[j - 1], 1.0, None, 1, w)
        if self.reorth:
            for k in range(j + 1):
                S.dot(self.V[k], w, self.coef)
                S.axpby(-1.0, self.coef, 1, False, self.V[k], 1.0, None, 1, w)
        S.dot(w, w, self.tmp)
        self.betas[:, j].copy_(torch.sqrt(self.tmp))
        S.axpby(1.0, self.betas[:, j:], self.ncv, True, w, 0.0, None, 1, self.V[j + 1])

    def _finish(self) -> None:
        S = self.seg
        _tridiag_batched(self.alphas, self.betas, self.ncv, self.ncv, self.nseg, self.tol, self.values, self.g)
        for k in range(self.ncv):
            S.axpby(1.0, self.g[:, k:], self.ncv, False, self.V[k], 1.0 if k else 0.0, None, 1, self.out)
        S.normalize(self.out, self.tmp)
        if self.refine is not None:
            self._refine()

    def _refine(self) -> None:
        """Per problem: ``w = H psi`` (FP64 batch), ``E = <psi|w>``, ``r = w - E psi``, the
        2x2 Rayleigh-Ritz problem on ``span{psi, r/|r|}``, its ground pair replacing
        ``(E, psi)``."""
        S, psi, u = self.seg, self.out, self.r_u
        w = self._matvec(self.refine, psi)
        S.dot(psi, w, self.r_a[:, 0:], stride=2)
        u.copy_(w)
        S.axpby(-1.0, self.r_a[:, 0:], 2, False, psi, 1.0, None, 1, u)
        S.dot(u, u, self.tmp)
        self.r_b[:, 0].copy_(torch.sqrt(self.tmp))
        # r == 0 leaves u == 0 and the ground pair (E, psi)
        self.tmp.copy_(torch.clamp(self.r_b[:, 0], min=1e-300))
        S.axpby(1.0, self.tmp, 1, True, u, 0.0, None, 1, u)
        hu = self._matvec(self.refine, u)
        S.dot(u, hu, self.r_a[:, 1:], stride=2)
        _tridiag_batched(self.r_a, self.r_b, 2, 2, self.nseg, 1e-300, self.values, self.r_g)
        S.axpby(1.0, self.r_g[:, 1:], 2, False, u, 0.0, None, 1, w)
        S.axpby(1.0, self.r_g[:, 0:], 2, False, psi, 1.0, None, 1, w)
        psi.copy_(w)
        S.normalize(psi, self.tmp)

    def run(self) -> None:
        if self.graphs is None:
            for j in range(self.ncv):
                self._iteration(j)
            self._finish()
            return
        for gr in self.graphs:
            gr.replay()

    def solve(self, vectors) -> tuple[list, list]:
        """The Lanczos step of every problem from start vectors ``vectors`` (one
        ``SymmetricTensor`` per problem): returns ``(energies, ground vectors)``; the host
        reads the energies only."""
        packed = torch.cat([v.data.to(torch.float64) for v in vectors])
        if packed.numel() != self.V.shape[1]:
            raise ValueError("BatchedLanczos: the vectors do not match the batch")
        self.start_(packed)
        self.run()
        out = self.out.clone()
        vals = self.values.cpu().tolist()
        offs = self.batch.x_off
        return vals, [SymmetricTensor(s, out[int(offs[i]):int(offs[i + 1])])
                      for i, s in enumerate(self.batch.structures)]
