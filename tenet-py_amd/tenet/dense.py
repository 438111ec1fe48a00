"""The dense boundary: ``to_dense`` and ``from_dense`` (reference ``ops/dense.py``).

``T = sum_tau A^(tau) (x) C^(tau)``: every block ``A`` (degeneracy indices) is tensored
with its fusion channel's Clebsch-Gordan tensor ``C`` (irrep indices), built along the
left-associated spine of the output tree and contracted with the conjugate of the input
tree's (reference ``ops/dense.py:173-283``).  A dual leg runs over ``V_a`` in the dense
array, so the provider's ``Z_a`` is contracted onto that axis.  Axis ``i`` of the dense
array runs over ``legs[i]``'s sectors in canonical order, index ``alpha * d_a + m``
inside sector ``a``'s slab.

``from_dense`` inverts it cell by cell with the pseudo-inverse of the cell's stacked CG
tensors and refuses input that is not symmetric (the residual includes the whole mass of
symmetry-forbidden cells).  Both are host-side, for small tensors: building operators
(MPO site tensors) and checking results; nothing on the contraction path calls them.
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np
import torch

from tenet.structure import TensorStructure

__all__ = ["to_dense", "from_dense", "dense_cells"]


def _tree_cgt(provider, tree, duals) -> np.ndarray:
    """A tree's CG tensor, shape ``(d_u0, ..., d_u{N-1}, d_coupled)``."""
    if tree.rank == 0:
        return np.ones((provider.irrep_dim(tree.coupled),))
    x = np.eye(provider.irrep_dim(tree.uncoupled[0]))
    spine = tree.lines()
    for k in range(1, tree.rank):
        x = np.tensordot(x, provider.cgc(spine[k - 1], tree.uncoupled[k], spine[k]), axes=([-1], [0]))
    for i, d in enumerate(duals):
        if d:
            z = provider.z_matrix(provider.dual(tree.uncoupled[i]))
            x = np.moveaxis(np.tensordot(x, z, axes=([i], [1])), -1, i)
    return x


@lru_cache(maxsize=256)
def dense_cells(structure: TensorStructure):
    """``{axis sectors: (block indices, stacked CG matrix (K, prod d), dims)}``."""
    prov = structure.provider
    legs = structure.legs
    outs, ins = structure.out_axes, structure.in_axes
    order = (*outs, *ins)
    to_public = sorted(range(len(legs)), key=order.__getitem__)
    cells: dict = {}
    for i, key in enumerate(structure.block_order):
        xo = _tree_cgt(prov, key.output_tree, [legs[a].dual for a in outs])
        xi = _tree_cgt(prov, key.input_tree, [legs[a].dual for a in ins])
        cgt = np.transpose(np.tensordot(xo, xi.conj(), axes=([-1], [-1])), to_public)
        cells.setdefault(structure.axis_sectors(key), []).append((i, cgt))
    out = {}
    for secs, entries in cells.items():
        dims = entries[0][1].shape
        mat = np.stack([c.reshape(-1) for _, c in entries])
        out[secs] = (tuple(i for i, _ in entries), mat, dims)
    return out


def _slab_offsets(leg) -> dict:
    o, acc = {}, 0
    for a, m in leg.space.sectors:
        o[a] = acc
        acc += m * leg.provider.irrep_dim(a)
    return o, acc


def to_dense(t) -> np.ndarray:
    """Host float64 dense array of a (small) symmetric tensor."""
    s = t.structure
    n = s.ndim
    offs = [_slab_offsets(leg) for leg in s.legs]
    out = np.zeros(tuple(acc for _, acc in offs))
    blocks = t.numpy_blocks()
    sub = [x for i in range(n) for x in (i, n + i)]
    for secs, (idx, mat, dims) in dense_cells(s).items():
        degs = tuple(leg.degeneracy(a) for leg, a in zip(s.legs, secs))
        sl = tuple(slice(o[a], o[a] + m * d) for (o, _), a, m, d in zip(offs, secs, degs, dims))
        acc = np.zeros(tuple(m * d for m, d in zip(degs, dims)))
        for k, i in enumerate(idx):
            full = np.einsum(np.asarray(blocks[i], dtype=np.float64), list(range(n)),
                             mat[k].reshape(dims), list(range(n, 2 * n)), sub)
            acc += full.reshape(acc.shape)
        out[sl] = acc
    return out


def from_dense(array, legs, *, atol: float | None = 1e-10, dtype=torch.float64, device=None):
    """Project a dense array onto the symmetric subspace of ``legs``.

    Raises ``ValueError`` when the part the projection cannot reproduce exceeds
    ``atol * |array|`` (``atol=None`` projects without checking)."""
    from tenet.tensor import SymmetricTensor, default_device

    s = TensorStructure(tuple(legs))
    a = np.asarray(array, dtype=np.float64)
    offs = [_slab_offsets(leg) for leg in s.legs]
    shape = tuple(acc for _, acc in offs)
    if a.shape != shape:
        raise ValueError(f"from_dense: array of shape {a.shape} does not match the legs' dense shape {shape}")
    n = s.ndim
    flat = np.zeros(s.size)
    lay = s.layout
    forbidden = a.copy()
    resid_sq = 0.0
    for secs, (idx, mat, dims) in dense_cells(s).items():
        degs = tuple(leg.degeneracy(x) for leg, x in zip(s.legs, secs))
        sl = tuple(slice(o[x], o[x] + m * d) for (o, _), x, m, d in zip(offs, secs, degs, dims))
        piece = a[sl].reshape(tuple(v for p in zip(degs, dims) for v in p))
        piece = piece.transpose(tuple(range(0, 2 * n, 2)) + tuple(range(1, 2 * n, 2)))
        m2 = piece.reshape(math.prod(degs), math.prod(dims))
        coeffs = m2 @ np.linalg.pinv(mat)
        resid_sq += float(np.sum((m2 - coeffs @ mat) ** 2))
        forbidden[sl] = 0.0
        for k, i in enumerate(idx):
            o = int(lay.offsets[i])
            flat[o : o + int(lay.sizes[i])] = coeffs[:, k]
    total = float(np.sum(a * a))
    resid = math.sqrt(resid_sq + float(np.sum(forbidden * forbidden)))
    if atol is not None and resid > atol * max(1.0, math.sqrt(total)):
        raise ValueError(f"from_dense: the array is not symmetric on these legs (residual {resid:.3g})")
    dev = torch.device(device) if device is not None else default_device()
    return SymmetricTensor(s, torch.from_numpy(flat).to(device=dev, dtype=dtype))
