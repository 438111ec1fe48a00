"""The Krylov loop of one compiled H_eff captured as hipGraphs, one graph per Lanczos
iteration.

Iteration ``j`` of the recurrence of ``tenet.network.lanczos`` (reference
``network/dmrg.py:106-227``) is a fixed launch sequence on fixed buffers once the
operator and the vector's structure are fixed:

    w = H v_j                                   (the H_eff launches)
    c = V^T w;  w -= V c;  d = V^T w;  w -= V d   (V = v_0..v_j, two classical Gram-Schmidt
                                                 passes in one tn_krylov_orth call)
    alpha_j = c_j;  beta_j = |w|;  v_{j+1} = w / beta_j

so it is captured once as a hipGraph and replayed.  Every Krylov scalar lives in device
memory; the host reads ``alpha``/``beta`` once after the last replay, solves the small
tridiagonal problem and recombines the Ritz vector with one FP64 GEMV.  Captured work
holds no host synchronisation, no allocation outside the graph's private pool and no
memset node.  Plan tables are uploaded and chains compiled by one eager warm-up pass
before capture.

``KrylovGraph`` is valid while the operator's static operands (environments, MPO) keep
their buffers: it belongs to one compiled ``Heff2`` and one vector structure.
"""

from __future__ import annotations

import ctypes

import torch

from tenet import _native, ops
from tenet.structure import TensorStructure
from tenet.tensor import SymmetricTensor

__all__ = ["KrylovGraph", "GraphedMatvec"]


def _dot(x: torch.Tensor, y: torch.Tensor, w, work: torch.Tensor, out: torch.Tensor) -> None:
    _native.check(_native.lib().tn_dot(
        _native.TN_F64, x.data_ptr(), y.data_ptr(), 0 if w is None else w.data_ptr(), x.numel(),
        work.data_ptr(), out.data_ptr(), _native.stream_ptr()))


def _axpby(a: float, a_dev, a_inv: int, x: torch.Tensor, b: float, y: torch.Tensor) -> None:
    """``y = (a * s) x + b y`` with ``s`` the device scalar ``a_dev`` (or its reciprocal)."""
    _native.check(_native.lib().tn_axpby(
        _native.TN_F64, x.numel(), a, None if a_dev is None else a_dev.data_ptr(), a_inv,
        x.data_ptr(), b, None, 0, y.data_ptr(), _native.stream_ptr()))


class GraphedMatvec:
    """``y = H x`` on static FP64 buffers ``x``/``y``, replayed from one captured graph."""

    def __init__(self, op, structure: TensorStructure) -> None:
        dev = op.fl.device
        self.structure = structure
        self.x = torch.randn(structure.size, dtype=torch.float64, device=dev)
        op.apply_data(structure, self.x)  # warm-up: compile, upload tables
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            # the matvec's own output buffer (from the graph's private pool) is the static
            # output: every replay rewrites it in place
            self.y = op.apply_data(structure, self.x)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        self.x.copy_(x)
        self.graph.replay()
        return self.y


class KrylovGraph:
    """One Lanczos step of ``op`` on vectors of ``structure``: ``ncv`` iteration graphs and
    one finishing graph (device tridiagonal solve, recombination, normalisation)."""

    def __init__(self, op, structure: TensorStructure, ncv: int = 3, reorthogonalize: bool = True,
                 tol: float = 1e-13) -> None:
        from tenet._native import TRIDIAG_MAX

        if not 1 <= ncv <= TRIDIAG_MAX:
            raise ValueError(f"ncv must lie in [1, {TRIDIAG_MAX}]")
        dev = op.fl.device
        self.op, self.structure, self.ncv, self.tol = op, structure, ncv, tol
        self.reorth = reorthogonalize
        n = structure.size
        f64 = dict(dtype=torch.float64, device=dev)
        self.V = torch.zeros(ncv + 1, n, **f64)
        self.alphas = torch.zeros(ncv, **f64)
        self.betas = torch.zeros(ncv, **f64)
        self.coef = torch.zeros(max(ncv + 1, ops.ORTH_MAX), **f64)
        self.owork = torch.empty(int(_native.lib().tn_krylov_orth_work()), **f64)
        self.g = torch.zeros(ncv, **f64)
        self.value = torch.zeros((), **f64)
        self.out = torch.zeros(n, **f64)
        self.tmp = torch.zeros((), **f64)
        self.work = torch.empty(1024, **f64)
        self.weights = ops._weights(structure, dev)
        # warm-up pass on a random start vector: compiles the H_eff chain and uploads
        # every table, so the captured work holds no host-to-device copy
        self.V[0].normal_()
        self.start_(self.V[0])
        for j in range(ncv):
            self._iteration(j)
        self._finish()
        torch.cuda.synchronize(dev)
        self.graphs = []
        pool = torch.cuda.graph_pool_handle()
        for step in [*range(ncv), "finish"]:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, pool=pool):
                self._finish() if step == "finish" else self._iteration(step)
            self.graphs.append(g)

    def start_(self, v: torch.Tensor) -> None:
        """``V[0] = v / |v|`` (qdim-weighted norm), eagerly."""
        _dot(v, v, self.weights, self.work, self.tmp)
        torch.sqrt(self.tmp, out=self.tmp)
        _axpby(1.0, self.tmp, 1, v, 0.0, self.V[0])  # elementwise, so v may be V[0] itself

    def _iteration(self, j: int) -> None:
        v = self.V[j]
        w = self.op.apply_data(self.structure, v)
        if w.dtype != torch.float64:
            raise TypeError("KrylovGraph expects an FP64 matvec result")
        if self.reorth and j + 1 <= ops.ORTH_MAX:
            # two classical Gram-Schmidt passes against V[0..j] in one native call: the
            # first pass's coefficient of V[j] is alpha_j, the norm after the second beta_j
            ptrs = (ctypes.c_void_p * (j + 1))(*[self.V[k].data_ptr() for k in range(j + 1)])
            _native.check(_native.lib().tn_krylov_orth(
                _native.TN_F64, w.numel(), j + 1, ptrs, w.data_ptr(),
                None if self.weights is None else self.weights.data_ptr(), 2, self.coef.data_ptr(),
                self.alphas[j].data_ptr(), self.betas[j].data_ptr(), self.owork.data_ptr(),
                _native.stream_ptr()))
        else:
            _dot(v, w, self.weights, self.work, self.alphas[j])
            _axpby(-1.0, self.alphas[j], 0, v, 1.0, w)
            if j:
                _axpby(-1.0, self.betas[j - 1], 0, self.V[j - 1], 1.0, w)
            if self.reorth:
                for k in range(j + 1):
                    _dot(self.V[k], w, self.weights, self.work, self.coef[k])
                    _axpby(-1.0, self.coef[k], 0, self.V[k], 1.0, w)
            _dot(w, w, self.weights, self.work, self.tmp)
            torch.sqrt(self.tmp, out=self.betas[j])
        _axpby(1.0, self.betas[j], 1, w, 0.0, self.V[j + 1])

    def _finish(self) -> None:
        from tenet.network.dmrg import ritz

        ritz(self.alphas, self.betas, [self.V[k] for k in range(self.ncv)], self.tol,
             coef=self.g, value=self.value, out=self.out)
        _dot(self.out, self.out, self.weights, self.work, self.tmp)
        torch.sqrt(self.tmp, out=self.tmp)
        _axpby(1.0, self.tmp, 1, self.out, 0.0, self.out)

    def replay(self) -> None:
        for g in self.graphs:
            g.replay()

    def solve(self, v: SymmetricTensor, refine=None) -> tuple[float, SymmetricTensor]:
        """The Lanczos step of ``tenet.network.lanczos`` from start vector ``v``: the
        replayed graphs leave the Ritz pair on the device; ``refine`` (an FP64 matvec)
        closes it with ``residual_refine``.  The host reads the energy only."""
        from tenet.network.dmrg import residual_refine

        if v.structure != self.structure:
            raise ValueError("KrylovGraph: the vector's structure differs from the captured one")
        x = v.data if v.dtype == torch.float64 else ops.astype(v, torch.float64).data
        self.start_(x)
        self.replay()
        out = SymmetricTensor(self.structure, self.out.clone())
        value = self.value
        if refine is not None:
            value, out, _ = residual_refine(refine, out)
        return float(value), out
