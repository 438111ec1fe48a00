"""The plain FP64 per-block GPU path: a comparison arm, not the engine.

This is what running the reference unchanged with torch-ROCm blocks amounts to: every
``ar.do`` of the reference's lowering becomes one torch call per block or per sector --
block transposes for repartition (reference ``ops/repartition.py:458-468``),
``reshape``+``concatenate`` per band for ``to_matrices`` (``map_view.py:274-278``), one
``matmul`` per coupled sector (``ops/map.py:160-161``), slices and transposes for
``from_matrices`` (``map_view.py:331-334``).  It shares only host plans with the engine
and is used by the benchmark to price the engine against the reference's execution
model on the same GPU.
"""

from __future__ import annotations

import torch

from tenet.plans import (
    contraction_plan,
    map_layout,
    permutation_plan,
    repartition_plan,
)
from tenet.space import OUT
from tenet.structure import TensorStructure

__all__ = ["BlockTensor", "einsum", "heff2"]


class BlockTensor:
    """A structure plus one torch tensor per block, exactly as the reference stores blocks."""

    def __init__(self, structure: TensorStructure, blocks: list) -> None:
        self.structure = structure
        self.blocks = blocks

    @classmethod
    def from_symmetric(cls, t) -> "BlockTensor":
        return cls(t.structure, [b.clone() for b in t.blocks])

    @property
    def legs(self):
        return self.structure.legs


def _apply(t: BlockTensor, new: TensorStructure, perm, terms) -> BlockTensor:
    moved = {s: t.blocks[s].permute(perm) for s in {s for s, _, _ in terms}}
    out: dict = {}
    for s, d, c in terms:
        x = moved[s] if c == 1 else moved[s] * float(c)
        out[d] = x if d not in out else out[d] + x
    return BlockTensor(new, [out[i] for i in range(new.num_blocks)])


def transpose(t: BlockTensor, axes) -> BlockTensor:
    p = permutation_plan(t.structure, tuple(axes))
    return _apply(t, p.new_structure, p.axes, p.terms)


def repartition(t: BlockTensor, outs, ins) -> BlockTensor:
    p = repartition_plan(t.structure, tuple(outs), tuple(ins))
    return _apply(t, p.new_structure, p.perm, p.terms)


def _to_matrices(t: BlockTensor) -> dict:
    lay = map_layout(t.structure)
    order = (*t.structure.out_axes, *t.structure.in_axes)
    ident = order == tuple(range(len(order)))
    mats = {}
    for c, cells in lay.grid:
        cb = lay.col_bands(c)
        rows = []
        for (_, _, dr), idx in zip(lay.row_bands(c), cells):
            parts = []
            for i, (_, _, dc) in zip(idx, cb):
                b = t.blocks[i] if ident else t.blocks[i].permute(order)
                parts.append(b.reshape(dr, dc))
            rows.append(parts[0] if len(parts) == 1 else torch.cat(parts, 1))
        mats[c] = rows[0] if len(rows) == 1 else torch.cat(rows, 0)
    return mats


def _from_matrices(s: TensorStructure, mats: dict) -> BlockTensor:
    lay = map_layout(s)
    order = (*s.out_axes, *s.in_axes)
    inverse = tuple(sorted(range(len(order)), key=order.__getitem__))
    blocks = [None] * s.num_blocks
    shapes = s.shapes
    for c, cells in lay.grid:
        m = mats[c]
        cb = lay.col_bands(c)
        for (_, ro, dr), idx in zip(lay.row_bands(c), cells):
            for i, (_, co, dc) in zip(idx, cb):
                piece = m[ro : ro + dr, co : co + dc].reshape(tuple(shapes[i][a] for a in order))
                blocks[i] = piece.permute(inverse)
    return BlockTensor(s, blocks)


def compose(a: BlockTensor, b: BlockTensor) -> BlockTensor:
    s = TensorStructure((*(a.legs[i] for i in a.structure.out_axes), *(b.legs[i] for i in b.structure.in_axes)))
    ma, mb = _to_matrices(a), _to_matrices(b)
    lay = map_layout(s)
    ref = a.blocks[0]
    mats = {}
    for c in lay.sectors:
        if c in ma and c in mb:
            mats[c] = torch.matmul(ma[c], mb[c])
        else:
            mats[c] = torch.zeros(lay.shape(c), dtype=ref.dtype, device=ref.device)
    return _from_matrices(s, mats)


def tensordot(a: BlockTensor, b: BlockTensor, axes) -> BlockTensor:
    cp = contraction_plan(a.structure, b.structure, axes)
    c = compose(repartition(a, cp.a_outputs, cp.a_inputs), repartition(b, cp.b_outputs, cp.b_inputs))
    return transpose(repartition(c, cp.restore_outputs, cp.restore_inputs), cp.final_transpose)


def einsum(eq: str, a: BlockTensor, b: BlockTensor) -> BlockTensor:
    lhs, out = eq.split("->")
    t0, t1 = lhs.split(",")
    shared = [x for x in t0 if x in t1]
    axes = (tuple(t0.index(x) for x in shared), tuple(t1.index(x) for x in shared))
    free = [x for x in t0 if x not in shared] + [x for x in t1 if x not in shared]
    return transpose(tensordot(a, b, axes), tuple(free.index(x) for x in out))


def composed(eq: str, a: BlockTensor, b: BlockTensor, bend: str = "") -> BlockTensor:
    if bend:
        lhs, out = eq.split("->")
        ta, tb = lhs.split(",")

        def bent(t, term):
            flip = set(bend)
            outs = tuple(i for i, c in enumerate(term) if (t.legs[i].side is OUT) != (c in flip))
            ins = tuple(i for i in range(len(term)) if i not in outs)
            return repartition(t, outs, ins), "".join(term[i] for i in (*outs, *ins))

        a, ta = bent(a, ta)
        b, tb = bent(b, tb)
        eq = f"{ta},{tb}->{out}"
    return einsum(eq, a, b)


def heff2(fl: BlockTensor, w1: BlockTensor, w2: BlockTensor, fr: BlockTensor, aa: BlockTensor) -> BlockTensor:
    t = einsum("apqr,rys->apqys", aa, fr)
    t = composed("apqys,mQqy->apQms", t, w2, bend="q")
    t = composed("apQms,xPpm->aPQxs", t, w1, bend="p")
    return composed("aPQxs,axB->BPQs", t, fl, bend="a")
