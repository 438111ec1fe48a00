"""Host-side categorical plans: where every source block goes, with which coefficient, when a
tensor's legs are permuted, bent across the map, repartitioned or contracted.

A *block map* is a tuple of ``(src_block, dst_block, coeff)`` triples -- a sparse linear
map between the block index spaces of two structures.  Abelian symmetries give unit-weight
bijections; SU(2) expands a source into several targets and sums several sources into
one; fZ2 weights are +-1.  Plans are pure functions of frozen structures and are cached;
the tree algebra underneath is evaluated on the degeneracy-free *pattern* of a structure
(every degeneracy set to 1), so a bond whose degeneracies move reuses every coefficient
(the caching scheme of reference ``ops/permutation.py:137-197`` and
``ops/repartition.py:117-385``; the plans themselves follow the reference's semantics:
``permutation_plan`` = ``permutation.py:160-197``, ``bend_plan`` = ``repartition.py:
130-167``, ``repartition_plan`` = ``repartition.py:342-385``, ``contraction_plan`` =
``contraction.py:248-306``, ``map_layout`` = ``map_view.py:155-235``).

Design: each elementary move (a permutation of the public axes, or a bend of the last
public axis) is a block map built from the provider's tree operations over every block
key; a repartition is a *move list* derived from the legs' sides alone (every leg that
changes side is carried to the last public position and bent, in ascending axis order,
then one permutation puts the axes in the requested order) folded into a single map.
Composition of maps is a join on the intermediate block index done with array
arithmetic; unit-weight bijections compose by index lookup.
"""

from __future__ import annotations

import operator
from dataclasses import dataclass, replace
from functools import cache

import numpy as np

from tenet.space import IN, OUT, Leg
from tenet.structure import FusionBlockKey, TensorStructure, pattern_of
from tenet.symmetry import CapabilityError

__all__ = [
    "PermutationPlan",
    "RepartitionPlan",
    "ContractionPlan",
    "MapLayout",
    "permutation_plan",
    "bend_plan",
    "repartition_plan",
    "contraction_plan",
    "map_layout",
    "compose_maps",
    "validated_axes",
]


# ----------------------------------------------------------------------------- block maps


@cache
def identity_map(n: int) -> tuple:
    return tuple((i, i, 1.0) for i in range(n))


def _unit_bijection(terms) -> bool:
    """Every block goes to exactly one block with weight 1 (and no two to the same)."""
    if any(c != 1 for _, _, c in terms):
        return False
    srcs = {s for s, _, _ in terms}
    dsts = {d for _, d, _ in terms}
    return len(srcs) == len(dsts) == len(terms)


_COMPOSED: dict = {}  # (id(first), id(second)) -> (first, second, composition)
_COMPOSED_MAX = 8192


def compose_maps(first, second) -> tuple:
    """``_compose_maps`` remembered by the identity of its operands: the block maps come
    from pattern-level plans (one object per sector pattern, shared by every structure of
    that pattern), so the structures of consecutive bond steps, whose degeneracies change
    but whose sector patterns do not, compose the same two objects again.  The entry keeps
    both operands alive, so an id is never reused while its entry exists."""
    if type(first) is not tuple or type(second) is not tuple:
        return _compose_maps(first, second)
    key = (id(first), id(second))
    hit = _COMPOSED.get(key)
    if hit is not None and hit[0] is first and hit[1] is second:
        return hit[2]
    out = _compose_maps(first, second)
    if len(_COMPOSED) >= _COMPOSED_MAX:
        _COMPOSED.clear()
    _COMPOSED[key] = (first, second, out)
    return out


def _compose_maps(first, second) -> tuple:
    """The block map of applying ``first`` and then ``second``: every path
    ``src -> mid -> dst`` contributes ``c_first * c_second``, contributions to one
    ``(src, dst)`` pair summed in the order the paths are met (first's terms in order,
    each through second's terms in order).  A pair whose paths cancel keeps a zero
    weight, so every target block stays present in the map."""
    if not first or not second:
        return ()
    if _unit_bijection(first) and _unit_bijection(second):
        to = {m: d for m, d, _ in second}
        return tuple((s, to[m], 1.0) for s, m, _ in first if m in to)
    f = np.asarray(first, dtype=object)
    s1 = np.asarray(f[:, 0], dtype=np.int64)
    m1 = np.asarray(f[:, 1], dtype=np.int64)
    c1 = np.asarray(f[:, 2], dtype=np.float64)
    g = np.asarray(second, dtype=object)
    m2 = np.asarray(g[:, 0], dtype=np.int64)
    d2 = np.asarray(g[:, 1], dtype=np.int64)
    c2 = np.asarray(g[:, 2], dtype=np.float64)
    order = np.argsort(m2, kind="stable")
    m2s = m2[order]
    lo = np.searchsorted(m2s, m1, side="left")
    hi = np.searchsorted(m2s, m1, side="right")
    fan = hi - lo
    if not fan.any():
        return ()
    # expand: term k of `first` pairs with second's terms order[lo[k] : hi[k]]
    rep = np.repeat(np.arange(len(m1)), fan)
    within = np.arange(int(fan.sum())) - np.repeat(np.cumsum(fan) - fan, fan)
    j = order[lo[rep] + within]
    src, dst, w = s1[rep], d2[j], c1[rep] * c2[j]
    # merge duplicate (src, dst) pairs, keeping the order of their first appearance
    width = int(dst.max()) + 1
    key = src * width + dst
    uniq, first_at, inverse = np.unique(key, return_index=True, return_inverse=True)
    total = np.zeros(len(uniq))
    for pos in range(len(key)):  # sequential, so the summation order is the path order
        total[inverse[pos]] += w[pos]
    seq = np.argsort(first_at, kind="stable")
    return tuple((int(uniq[i] // width), int(uniq[i] % width), float(total[i])) for i in seq)


def _tree_map(structure: TensorStructure, new: TensorStructure, expand) -> tuple:
    """Block map whose terms come from ``expand(key) -> ((new_key, coeff), ...)``; exact
    zero coefficients are not recorded."""
    where = new.index_of
    return tuple((src, where(nk), c) for src, key in enumerate(structure.block_order)
                 for nk, c in expand(key) if c != 0)


# ----------------------------------------------------------------------------- permutation


@dataclass(frozen=True)
class PermutationPlan:
    axes: tuple
    new_structure: TensorStructure
    terms: tuple


def validated_axes(ndim: int, axes) -> tuple:
    """A permutation of ``range(ndim)`` as plain ints, or a ``ValueError`` naming why not."""
    try:
        axes = tuple(operator.index(a) for a in axes)
    except TypeError as exc:
        raise ValueError(f"transpose: axes {tuple(axes)} must all be integers") from exc
    if len(axes) != ndim:
        raise ValueError(f"transpose: axes {axes} has length {len(axes)}, expected {ndim}")
    seen = set()
    for a in axes:
        if not 0 <= a < ndim:
            raise ValueError(f"transpose: axis {a} is out of range for a {ndim}-dimensional tensor")
        if a in seen:
            raise ValueError(f"transpose: axis {a} is repeated in {axes}")
        seen.add(a)
    return axes


def _within_side(axes, old_axes, new_axes) -> tuple:
    """The reordering a public permutation induces on one side: entry j is the old side
    position of the leg that ends up at side position j."""
    rank_on_side = dict(zip(old_axes, range(len(old_axes))))
    return tuple(rank_on_side[axes[a]] for a in new_axes)


def _with_legs(plan, structure: TensorStructure, legs: tuple):
    """A pattern-level plan re-targeted at the real legs (same block order, real shapes)."""
    if legs == plan.new_structure.legs:
        return plan
    return replace(plan, new_structure=TensorStructure(legs))


@cache
def permutation_plan(structure: TensorStructure, axes: tuple) -> PermutationPlan:
    """Public axis ``j`` of the result is axis ``axes[j]`` of ``structure``."""
    legs = tuple(structure.legs[a] for a in axes)
    return _with_legs(_permutation_on_pattern(pattern_of(structure), axes), structure, legs)


@cache
def _permutation_on_pattern(pattern: TensorStructure, axes: tuple) -> PermutationPlan:
    new = TensorStructure(tuple(pattern.legs[a] for a in axes))
    on_out = _within_side(axes, pattern.out_axes, new.out_axes)
    on_in = _within_side(axes, pattern.in_axes, new.in_axes)
    if on_out == tuple(sorted(on_out)) and on_in == tuple(sorted(on_in)):
        # only the interleaving of OUT and IN legs changed: the trees, hence the keys and
        # their sorted order, are untouched
        return PermutationPlan(axes, new, identity_map(pattern.num_blocks))
    provider = pattern.provider
    if not hasattr(provider, "permute_tree"):
        raise CapabilityError(f"{provider.name} does not implement permutation coefficients")
    # each distinct tree of a side is expanded once
    out_exp = {t: provider.permute_tree(t, on_out) for t in {k.output_tree for k in pattern.block_order}}
    in_exp = {t: provider.permute_tree(t, on_in) for t in {k.input_tree for k in pattern.block_order}}

    def expand(key):
        return [(FusionBlockKey(ot, it), co * ci)
                for ot, co in out_exp[key.output_tree] for it, ci in in_exp[key.input_tree]]

    return PermutationPlan(axes, new, _tree_map(pattern, new, expand))


# ----------------------------------------------------------------------------- bend


def _flipped(leg: Leg) -> Leg:
    """The same wire seen from the other side of the map: side swapped, dual toggled."""
    return replace(leg, side=IN if leg.side is OUT else OUT, dual=not leg.dual)


def _bent_legs(legs: tuple, axis: int) -> tuple:
    return (*legs[:axis], *legs[axis + 1:], _flipped(legs[axis]))


@cache
def bend_plan(structure: TensorStructure, axis: int) -> PermutationPlan:
    """Bend ``axis`` -- the last leg of its side -- to the end of the other side."""
    return _with_legs(_bend_on_pattern(pattern_of(structure), axis), structure,
                      _bent_legs(structure.legs, axis))


@cache
def _bend_on_pattern(pattern: TensorStructure, axis: int) -> PermutationPlan:
    leg = pattern.legs[axis]
    provider = pattern.provider
    move = provider.bend_right if leg.side is OUT else provider.bend_left
    new = TensorStructure(_bent_legs(pattern.legs, axis))
    terms = _tree_map(pattern, new, lambda key: move(key, dual=leg.dual))
    order = tuple(a for a in range(pattern.ndim) if a != axis) + (axis,)
    return PermutationPlan(order, new, terms)


# ----------------------------------------------------------------------------- repartition


@dataclass(frozen=True)
class RepartitionPlan:
    """One composed axis permutation ``perm`` plus the composed block map."""

    new_structure: TensorStructure
    perm: tuple
    terms: tuple

    @property
    def trivial(self) -> bool:
        """True when every block maps to itself with weight 1 and no axis moves."""
        return self.perm == tuple(range(len(self.perm))) and all(s == d and c == 1 for s, d, c in self.terms)


def validated_partition(ndim: int, outputs, inputs) -> tuple:
    parts = []
    for name, axes in (("outputs", outputs), ("inputs", inputs)):
        try:
            parts.append(tuple(operator.index(a) for a in axes))
        except TypeError as exc:
            raise ValueError(f"repartition: {name} {tuple(axes)} must all be integers") from exc
    seen = set()
    for name, axes in zip(("outputs", "inputs"), parts):
        for a in axes:
            if not 0 <= a < ndim:
                raise ValueError(
                    f"repartition: axis {a} in {name} is out of range for a {ndim}-dimensional tensor"
                )
            if a in seen:
                raise ValueError(f"repartition: axis {a} appears twice in outputs/inputs")
            seen.add(a)
    missing = sorted(set(range(ndim)) - seen)
    if missing:
        raise ValueError(
            f"repartition: axis {missing[0]} is in neither outputs {parts[0]} nor inputs "
            f"{parts[1]}; together they must be a permutation of range({ndim})"
        )
    return tuple(parts)


def _target_sides(outputs: tuple, inputs: tuple) -> dict:
    return {**{a: OUT for a in outputs}, **{a: IN for a in inputs}}


@cache
def repartition_plan(structure: TensorStructure, outputs: tuple, inputs: tuple) -> RepartitionPlan:
    """``outputs`` become OUT and ``inputs`` IN, public order ``(*outputs, *inputs)``."""
    side = _target_sides(outputs, inputs)
    if all(structure.legs[a].side is side[a] for a in range(structure.ndim)):
        p = permutation_plan(structure, (*outputs, *inputs))
        return RepartitionPlan(p.new_structure, p.axes, p.terms)
    legs = tuple(leg if leg.side is side[a] else _flipped(leg)
                 for a, leg in ((a, structure.legs[a]) for a in (*outputs, *inputs)))
    return _with_legs(_repartition_on_pattern(pattern_of(structure), outputs, inputs), structure, legs)


def _moves(sides: tuple, outputs: tuple, inputs: tuple) -> list:
    """The elementary moves of a repartition, from the legs' current sides alone:
    ``("permute", axes)`` and ``("bend", position)``.  Every axis that changes side, in
    ascending order, is carried to the last public position and bent there; a final
    permutation lists the axes as ``(*outputs, *inputs)``."""
    want = _target_sides(outputs, inputs)
    n = len(sides)
    at = list(range(n))  # at[p] = original axis now at public position p
    moves = []
    for a in range(n):
        if sides[a] is want[a]:
            continue
        p = at.index(a)
        moves.append(("permute", (*range(p), *range(p + 1, n), p)))
        at.append(at.pop(p))
        moves.append(("bend", n - 1))
    moves.append(("permute", tuple(at.index(a) for a in (*outputs, *inputs))))
    return moves


@cache
def _repartition_on_pattern(pattern: TensorStructure, outputs: tuple, inputs: tuple) -> RepartitionPlan:
    cur, perm, terms = pattern, tuple(range(pattern.ndim)), identity_map(pattern.num_blocks)
    for kind, arg in _moves(tuple(l.side for l in pattern.legs), outputs, inputs):
        step = permutation_plan(cur, arg) if kind == "permute" else bend_plan(cur, arg)
        perm = tuple(perm[i] for i in step.axes)
        terms = compose_maps(terms, step.terms)
        cur = step.new_structure
    return RepartitionPlan(cur, perm, terms)


# ----------------------------------------------------------------------------- map layout


@dataclass(frozen=True)
class MapLayout:
    """Per coupled sector: row bands over OUT trees, column bands over IN trees, and the
    grid of block indices (reference ``map_view.py:69-235``).

    ``rows``/``cols`` hold ``(sector, tree, offset, extent)``; ``grid`` holds, per
    sector, a tuple of rows of block indices.
    """

    structure: TensorStructure
    sectors: tuple
    rows: tuple
    cols: tuple
    grid: tuple

    def row_bands(self, c) -> tuple:
        return tuple(b[1:] for b in self.rows if b[0] == c)

    def col_bands(self, c) -> tuple:
        return tuple(b[1:] for b in self.cols if b[0] == c)

    def shape(self, c) -> tuple:
        return (
            sum(e for _, _, e in self.row_bands(c)),
            sum(e for _, _, e in self.col_bands(c)),
        )


@dataclass(frozen=True)
class _BandPattern:
    """The degeneracy-free part of a map layout, shared by every structure of one block
    pattern: the sectors; per row (column) band its sector's position, its tree and one
    block lying in it; per block (in sector, row band, column band order) its block index,
    sector position and band numbers; and the grid of block indices."""

    sectors: tuple
    row_sector: np.ndarray
    row_tree: tuple
    row_rep: np.ndarray
    col_sector: np.ndarray
    col_tree: tuple
    col_rep: np.ndarray
    blk: np.ndarray
    blk_sector: np.ndarray
    blk_row: np.ndarray
    blk_col: np.ndarray
    grid: tuple


@cache
def _band_pattern(pattern: TensorStructure) -> _BandPattern:
    """Blocks of one coupled sector tile that sector's matrix: OUT trees index row bands,
    IN trees column bands, both in order of first appearance in the block order (which is
    sorted, so the bands are too).  Every (row tree, column tree) pair of a sector must be
    a block, or the structure is not a per-sector cross product."""
    per_sector: dict = {}  # c -> ({out tree: band}, {in tree: band}, {(band, band): block})
    for i, key in enumerate(pattern.block_order):
        r, k, cells = per_sector.setdefault(key.coupled, ({}, {}, {}))
        cells[(r.setdefault(key.output_tree, len(r)), k.setdefault(key.input_tree, len(k)))] = i
    sectors = tuple(sorted(per_sector))
    row_sector, row_tree, row_rep, col_sector, col_tree, col_rep = [], [], [], [], [], []
    blk, blk_sector, blk_row, blk_col, grid = [], [], [], [], []
    for ci, c in enumerate(sectors):
        r, k, cells = per_sector[c]
        if len(cells) != len(r) * len(k):
            raise RuntimeError("map_layout: block order is not a per-sector cross product")
        r0, k0 = len(row_tree), len(col_tree)
        rows = tuple(tuple(cells[(a, b)] for b in range(len(k))) for a in range(len(r)))
        row_tree.extend(r)
        row_sector.extend([ci] * len(r))
        row_rep.extend(row[0] for row in rows)
        col_tree.extend(k)
        col_sector.extend([ci] * len(k))
        col_rep.extend(rows[0])
        for a, row in enumerate(rows):
            blk.extend(row)
            blk_sector.extend([ci] * len(row))
            blk_row.extend([r0 + a] * len(row))
            blk_col.extend(range(k0, k0 + len(row)))
        grid.append((c, rows))
    ix = lambda v: np.asarray(v, dtype=np.int64)  # noqa: E731
    return _BandPattern(sectors, ix(row_sector), tuple(row_tree), ix(row_rep), ix(col_sector),
                        tuple(col_tree), ix(col_rep), ix(blk), ix(blk_sector), ix(blk_row),
                        ix(blk_col), tuple(grid))


def _band_extents(shapes: np.ndarray, rep: np.ndarray, axes: tuple, sector: np.ndarray, nsec: int):
    """Extents and in-sector offsets of bands: a band's extent is the product of the
    degeneracies on that side's legs of any block in it; offsets run within each sector."""
    if not len(rep):
        z = np.zeros(0, dtype=np.int64)
        return z, z, np.zeros(nsec, dtype=np.int64)
    ext = np.prod(shapes[rep][:, list(axes)], axis=1, dtype=np.int64) if axes else np.ones(len(rep), np.int64)
    total = np.zeros(nsec, dtype=np.int64)
    np.add.at(total, sector, ext)
    start = np.zeros(nsec, dtype=np.int64)
    np.cumsum(total[:-1], out=start[1:])
    off = np.cumsum(ext) - ext - start[sector]
    return ext, off, total


@cache
def band_layout(structure: TensorStructure) -> tuple:
    """``(pattern, row extents, row offsets, column extents, column offsets, sector rows,
    sector columns)`` of a structure's map layout as int64 arrays (bands in the pattern's
    order, offsets within their sector)."""
    pat = _band_pattern(pattern_of(structure))
    shapes = structure.layout.shapes
    nsec = len(pat.sectors)
    rext, roff, rtot = _band_extents(shapes, pat.row_rep, structure.out_axes, pat.row_sector, nsec)
    cext, coff, ctot = _band_extents(shapes, pat.col_rep, structure.in_axes, pat.col_sector, nsec)
    return pat, rext, roff, cext, coff, rtot, ctot


@cache
def map_layout(structure: TensorStructure) -> MapLayout:
    """Row bands over OUT trees, column bands over IN trees and the block grid of every
    coupled sector (the band pattern is shared by all structures of one block pattern;
    only the extents are read from this structure's degeneracies)."""
    pat, rext, roff, cext, coff, _, _ = band_layout(structure)
    secs = pat.sectors
    rows = tuple((secs[c], t, o, e) for c, t, o, e in zip(pat.row_sector.tolist(), pat.row_tree,
                                                          roff.tolist(), rext.tolist()))
    cols = tuple((secs[c], t, o, e) for c, t, o, e in zip(pat.col_sector.tolist(), pat.col_tree,
                                                          coff.tolist(), cext.tolist()))
    return MapLayout(structure, secs, rows, cols, pat.grid)


# ----------------------------------------------------------------------------- contraction


def outward_dual(leg: Leg) -> bool:
    """Whether the wire, read as leaving the tensor, carries the dual space: exactly when
    the leg sits on the domain side xor is marked dual."""
    return (leg.side is IN) != leg.dual


def contractible(x: Leg, y: Leg) -> bool:
    """Two legs join into one wire when they carry the same space and point opposite ways."""
    return x.space == y.space and outward_dual(x) is not outward_dual(y)


@dataclass(frozen=True)
class ContractionPlan:
    """Which repartitions a pairwise contraction runs, and the output structure
    (reference ``contraction.py:211-306``)."""

    a_outputs: tuple
    a_inputs: tuple
    b_outputs: tuple
    b_inputs: tuple
    restore_outputs: tuple
    restore_inputs: tuple
    final_transpose: tuple
    new_structure: TensorStructure


def validated_contraction_axes(a: TensorStructure, b: TensorStructure, axes) -> tuple:
    if hasattr(axes, "__index__"):
        n = operator.index(axes)
        if n < 0:
            raise ValueError(f"tensordot: axes={n} is negative; the integer form counts axes")
        if n > a.ndim or n > b.ndim:
            raise ValueError(f"tensordot: axes={n} exceeds the operands' ranks")
        return tuple(range(a.ndim - n, a.ndim)), tuple(range(n))
    try:
        raw_a, raw_b = axes
    except (TypeError, ValueError) as exc:
        raise ValueError(f"tensordot: axes {axes!r} must be an integer or a pair") from exc
    parts = []
    for name, t, raw in (("a", a, raw_a), ("b", b, raw_b)):
        raw = (raw,) if hasattr(raw, "__index__") else raw
        try:
            got = tuple(operator.index(x) for x in raw)
        except TypeError as exc:
            raise ValueError(f"tensordot: axes for {name} must all be integers") from exc
        seen = set()
        for x in got:
            if not 0 <= x < t.ndim:
                raise ValueError(
                    f"tensordot: axis {x} of {name} is out of range for a {t.ndim}-dimensional tensor"
                )
            if x in seen:
                raise ValueError(f"tensordot: axis {x} appears twice in {name}'s axes {got}")
            seen.add(x)
        parts.append(got)
    if len(parts[0]) != len(parts[1]):
        raise ValueError(
            f"tensordot: axes {parts[0]} of a and {parts[1]} of b have different lengths"
        )
    return tuple(parts)


def _refusal(k: int, i: int, j: int, x: Leg, y: Leg) -> ValueError:
    why = ("their spaces differ" if x.space != y.space
           else "both ends of the wire point the same way (equal outward duals)")
    return ValueError(f"tensordot: pair {k}, axis {i} of a and axis {j} of b do not contract: {why}")


@cache
def contraction_plan(a: TensorStructure, b: TensorStructure, axes: tuple) -> ContractionPlan:
    """Lower ``tensordot(a, b, axes)`` to ``a: (free | contracted)``, ``b: (contracted |
    free)``, one composition, and a restoring repartition of the composed legs: each free
    leg returns to its own side, public order = a's free legs then b's."""
    ca, cb = axes
    if a.provider != b.provider:
        raise ValueError(f"tensordot: a has provider {a.provider.name} and b has {b.provider.name}")
    bad = next(((k, i, j) for k, (i, j) in enumerate(zip(ca, cb)) if not contractible(a.legs[i], b.legs[j])), None)
    if bad is not None:
        k, i, j = bad
        raise _refusal(k, i, j, a.legs[i], b.legs[j])
    free_a = tuple(sorted(set(range(a.ndim)) - set(ca)))
    free_b = tuple(sorted(set(range(b.ndim)) - set(cb)))
    if not free_a and not free_b:
        raise ValueError("tensordot: this contraction leaves no free leg")
    legs = tuple(a.legs[i] for i in free_a) + tuple(b.legs[j] for j in free_b)
    sides = np.array([leg.side is OUT for leg in legs], dtype=bool)
    outs = tuple(int(k) for k in np.flatnonzero(sides))
    ins = tuple(int(k) for k in np.flatnonzero(~sides))
    # the composed tensor lists (*outs, *ins); the last transpose puts each leg back
    where = np.empty(len(legs), dtype=np.int64)
    where[list(outs + ins)] = np.arange(len(legs))
    return ContractionPlan(
        a_outputs=free_a, a_inputs=ca, b_outputs=cb, b_inputs=free_b,
        restore_outputs=outs, restore_inputs=ins,
        final_transpose=tuple(int(x) for x in where),
        new_structure=TensorStructure(legs),
    )
