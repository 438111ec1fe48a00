"""Tensor structure: ordered legs, the sorted block set, shapes and the flat HBM layout.

``TensorStructure(legs)`` derives everything categorical about a tensor (reference
``structure.py:98-405``):

* ``block_order`` -- the sorted tuple of allowed ``FusionBlockKey(output_tree,
  input_tree)``.  It is the storage contract: block ``i`` of a tensor belongs to key
  ``i``.  Sector matching is the cross product of leg sectors, filtered by coupled-sector
  agreement of the two trees (reference ``structure.py:349-371``).
* ``block_shape`` -- degeneracies in *public axis order*.

On top of that the device layout is fixed: every block is C-contiguous and the
blocks follow each other in ``block_order`` in ONE flat buffer, so ``offsets[i]`` (an
int64 table, computed once per structure) locates block ``i``.

Everything that depends only on sectors (not degeneracies) is cached on the structure's
*pattern* -- the same legs with every degeneracy 1 -- so a DMRG bond whose degeneracies
move does not re-enumerate its blocks (reference ``structure.py:311-338``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from dataclasses import replace as _replace
from functools import cache
from itertools import product

import numpy as np

from tenet.fusion_tree import FusionTree, _all_trees
from tenet.space import OUT, Leg

__all__ = ["FusionBlockKey", "TensorStructure"]


@dataclass(frozen=True, slots=True, order=True)
class FusionBlockKey:
    """An (output tree, input tree) pair sharing one coupled sector; field order sorts."""

    output_tree: FusionTree
    input_tree: FusionTree
    _h: int = field(init=False, repr=False, compare=False)  # the generated hash, computed once

    def __post_init__(self) -> None:
        object.__setattr__(self, "_h", hash((self.output_tree, self.input_tree)))

    def __hash__(self) -> int:
        return self._h

    @property
    def coupled(self):
        return self.output_tree.coupled


@dataclass(frozen=True, eq=False)
class TensorStructure:
    """Ordered legs plus everything derivable from them.  Immutable and hashable."""

    legs: tuple

    def __post_init__(self) -> None:
        if not self.legs:
            raise ValueError("TensorStructure needs at least one leg")
        object.__setattr__(self, "legs", tuple(self.legs))
        object.__setattr__(self, "_hash", hash(self.legs))

    def __hash__(self) -> int:
        return self._hash

    def __eq__(self, other) -> bool:
        if self is other:
            return True
        if not isinstance(other, TensorStructure):
            return NotImplemented
        return self._hash == other._hash and self.legs == other.legs

    @property
    def provider(self):
        return self.legs[0].provider

    @property
    def ndim(self) -> int:
        return len(self.legs)

    @property
    def out_axes(self) -> tuple:
        return _axes(self)[0]

    @property
    def in_axes(self) -> tuple:
        return _axes(self)[1]

    @property
    def block_order(self) -> tuple:
        return _block_order(self)

    @property
    def num_blocks(self) -> int:
        return len(_block_order(self))

    def index_of(self, key: FusionBlockKey) -> int:
        return _index_map(self)[key]

    def axis_sectors(self, key: FusionBlockKey) -> tuple:
        return _axis_sectors(self)[self.index_of(key)]

    def block_shape(self, key: FusionBlockKey) -> tuple:
        return _shapes(self)[self.index_of(key)]

    @property
    def shapes(self) -> tuple:
        """Block shapes aligned with ``block_order``."""
        return _shapes(self)

    @property
    def layout(self) -> "FlatLayout":
        """The flat HBM layout (offsets, sizes, C strides) of this structure."""
        return _layout(self)

    @property
    def size(self) -> int:
        """Total number of stored elements."""
        return _layout(self).size


@cache
def _flat_space(space):
    return space.flat()


@cache
def _pattern(s: TensorStructure) -> TensorStructure:
    legs = tuple(_replace(leg, space=_flat_space(leg.space)) for leg in s.legs)
    return s if legs == s.legs else TensorStructure(legs)


@cache
def _axes(s: TensorStructure):
    return (
        tuple(i for i, leg in enumerate(s.legs) if leg.side is OUT),
        tuple(i for i, leg in enumerate(s.legs) if leg.side is not OUT),
    )


def _side_trees(s: TensorStructure, axes) -> dict:
    """Every fusion tree one side of ``s`` can carry, grouped by coupled sector."""
    provider = s.provider
    legs = [s.legs[i] for i in axes]
    groups: dict = {}
    for assignment in product(*(leg.sectors for leg in legs)):
        fused = tuple(leg.fused_sector(a) for leg, a in zip(legs, assignment))
        for t in _all_trees(provider, fused):
            groups.setdefault(t.coupled, []).append(t)
    return groups


@cache
def _block_order(s: TensorStructure) -> tuple:
    """Sector matching: every (output tree, input tree) pair that meets in one coupled
    sector, sorted.  Each side is enumerated once and the two are joined on the coupled
    sector, so the cost is the two sides' sector products plus the number of blocks, not
    the product over all legs."""
    p = _pattern(s)
    if p is not s:
        return _block_order(p)
    out_axes, in_axes = _axes(s)
    outs = _side_trees(s, out_axes)
    ins = _side_trees(s, in_axes)
    keys = [FusionBlockKey(ot, it) for c, ots in outs.items() for ot in ots for it in ins.get(c, ())]
    return tuple(sorted(keys))


@cache
def _index_map(s: TensorStructure) -> dict:
    p = _pattern(s)
    if p is not s:
        return _index_map(p)
    return {k: i for i, k in enumerate(_block_order(s))}


@cache
def _axis_sectors(s: TensorStructure) -> tuple:
    """Space (de-dualized) sector per public axis for every key, aligned with block order."""
    p = _pattern(s)
    if p is not s:
        return _axis_sectors(p)
    out_axes, in_axes = _axes(s)
    rows = []
    for key in _block_order(s):
        by_axis = dict(zip(out_axes, key.output_tree.uncoupled))
        by_axis.update(zip(in_axes, key.input_tree.uncoupled))
        rows.append(tuple(leg.space_sector(by_axis[ax]) for ax, leg in enumerate(s.legs)))
    return tuple(rows)


@cache
def _axis_index(s: TensorStructure) -> np.ndarray:
    """Per block and axis, the position of the block's sector in that leg's space (the
    spaces of a pattern and of every structure sharing it list the same sectors in the
    same order), as an int64 (blocks x ndim) table."""
    p = _pattern(s)
    if p is not s:
        return _axis_index(p)
    pos = [{a: j for j, (a, _) in enumerate(leg.space.sectors)} for leg in s.legs]
    rows = _axis_sectors(s)
    return np.array([[pm[a] for pm, a in zip(pos, secs)] for secs in rows],
                    dtype=np.int64).reshape(len(rows), s.ndim)


def _block_shapes(s: TensorStructure) -> np.ndarray:
    """Block shapes as an int64 (blocks x ndim) table: each leg's degeneracies gathered
    at the blocks' sector positions."""
    idx = _axis_index(s)
    out = np.empty(idx.shape, dtype=np.int64)
    for d, leg in enumerate(s.legs):
        degs = np.fromiter((m for _, m in leg.space.sectors), dtype=np.int64, count=len(leg.space.sectors))
        out[:, d] = degs[idx[:, d]] if len(idx) else 0
    return out


@cache
def _shapes(s: TensorStructure) -> tuple:
    return tuple(map(tuple, _layout(s).shapes.tolist()))


@dataclass(frozen=True)
class FlatLayout:
    """Where each block of a structure lives in the one flat device buffer.

    ``offsets[i]``/``sizes[i]`` are element offsets/counts of block ``i``;
    ``shapes[i, d]`` and ``strides[i, d]`` (C order) are int64 tables with one row per
    block; ``size`` is the buffer length.
    """

    offsets: np.ndarray
    sizes: np.ndarray
    shapes: np.ndarray
    strides: np.ndarray
    size: int


@cache
def _layout(s: TensorStructure) -> FlatLayout:
    shapes = _block_shapes(s)
    sizes = np.prod(shapes, axis=1, dtype=np.int64) if len(shapes) else np.zeros(0, np.int64)
    offsets = np.zeros(len(sizes), dtype=np.int64)
    if len(sizes) > 1:
        np.cumsum(sizes[:-1], out=offsets[1:])
    strides = np.ones_like(shapes)
    for d in range(s.ndim - 2, -1, -1):
        strides[:, d] = strides[:, d + 1] * shapes[:, d + 1]
    return FlatLayout(offsets, sizes, shapes, strides, int(sizes.sum()))


def pattern_of(s: TensorStructure) -> TensorStructure:
    """The degeneracy-free structure sharing ``s``'s block set."""
    return _pattern(s)


def make_leg(space, side, dual=False, name=None) -> Leg:
    return Leg(space, side, dual, name)
