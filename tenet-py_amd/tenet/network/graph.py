"""The Krylov loop of one compiled H_eff captured as hipGraphs, one graph per Lanczos
iteration.

Iteration ``j`` of the recurrence of ``tenet.network.lanczos`` (reference
``network/dmrg.py:106-227``) is a fixed launch sequence on fixed buffers once the
operator and the vector's structure are fixed:

    w = H v_j                                   (the H_eff launches)
    alpha_j = <v_j|w>;  w -= alpha_j v_j + beta_{j-1} v_{j-1}
    w -= sum_k <v_k|w> v_k   (k <= j, FP64 re-orthogonalisation)
    beta_j = |w|;  v_{j+1} = w / beta_j

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

import numpy as np
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
        self.y = torch.empty_like(self.x)
        self.y.copy_(op.apply_data(structure, self.x))  # warm-up: compile, upload tables
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.y.copy_(op.apply_data(structure, self.x))

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        self.x.copy_(x)
        self.graph.replay()
        return self.y


class KrylovGraph:
    """``ncv`` Lanczos iterations of ``op`` on vectors of ``structure``, one graph each."""

    def __init__(self, op, structure: TensorStructure, ncv: int = 3, reorthogonalize: bool = True) -> None:
        if ncv < 1:
            raise ValueError("ncv must be at least 1")
        dev = op.fl.device
        self.op, self.structure, self.ncv = op, structure, ncv
        self.reorth = reorthogonalize
        n = structure.size
        f64 = dict(dtype=torch.float64, device=dev)
        self.V = torch.zeros(ncv + 1, n, **f64)
        self.alphas = torch.zeros(ncv, **f64)
        self.betas = torch.zeros(ncv, **f64)
        self.coef = torch.zeros(ncv + 1, **f64)
        self.tmp = torch.zeros((), **f64)
        self.work = torch.empty(1024, **f64)
        self.weights = ops._weights(structure, dev)
        # warm-up pass on a random start vector: compiles the H_eff chain and uploads
        # every table, so the captured work holds no host-to-device copy
        self.V[0].normal_()
        self.start_(self.V[0])
        for j in range(ncv):
            self._iteration(j)
        torch.cuda.synchronize(dev)
        self.graphs = []
        pool = torch.cuda.graph_pool_handle()
        for j in range(ncv):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, pool=pool):
                self._iteration(j)
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

    def replay(self) -> None:
        for g in self.graphs:
            g.replay()

    def solve(self, v: SymmetricTensor, tol: float = 1e-13, refine=None) -> tuple[float, SymmetricTensor]:
        """The Lanczos step of ``tenet.network.lanczos`` from start vector ``v``."""
        if v.structure != self.structure:
            raise ValueError("KrylovGraph: the vector's structure differs from the captured one")
        x = v.data if v.dtype == torch.float64 else ops.astype(v, torch.float64).data
        self.start_(x)
        self.replay()
        a_h = self.alphas.cpu().numpy()
        b_h = self.betas.cpu().numpy()
        m = self.ncv
        for j in range(self.ncv - 1):
            if b_h[j] < tol:
                m = j + 1
                break
        tri = np.diag(a_h[:m]) + np.diag(b_h[: m - 1], 1) + np.diag(b_h[: m - 1], -1)
        values, states = np.linalg.eigh(tri)
        g = torch.from_numpy(np.ascontiguousarray(states[:, 0])).to(self.V.device)
        out = SymmetricTensor(self.structure, g @ self.V[:m])
        out = ops.divide(out, ops.norm(out))
        value = float(values[0])
        if refine is not None:
            value = float(ops.inner(out, refine(out)))
        return value, out
