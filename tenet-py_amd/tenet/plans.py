"""Host-side categorical plans: which source block lands in which target block, with what
coefficient, under a permutation, a bend, a repartition or a contraction.

All plans are array-free and cached; the bodies that enumerate terms are cached on the
degeneracy-free *pattern* of a structure, so moving bond degeneracies never re-derives
coefficients (reference ``ops/permutation.py:137-197``, ``ops/repartition.py:117-385``).

A term list is ``((src_block, dst_block, coeff), ...)``.  For U(1)/Trivial every term has
coefficient exactly 1 and the map is one-to-one; SU(2) expands a source into several
targets and several sources sum into one target.
"""

from __future__ import annotations

import operator
from dataclasses import dataclass, replace
from functools import cache

from tenet.space import IN, OUT, Leg, Side
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
    "validated_axes",
]


# ----------------------------------------------------------------------------- transpose


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


def _side_perm(axes, old_side, new_side) -> tuple:
    """Permutation induced on one side's legs: entry j = old in-side position of the leg
    that becomes in-side position j."""
    pos = {ax: k for k, ax in enumerate(old_side)}
    return tuple(pos[axes[new_ax]] for new_ax in new_side)


@cache
def permutation_plan(structure: TensorStructure, axes: tuple) -> PermutationPlan:
    plan = _pattern_permutation(pattern_of(structure), axes)
    legs = tuple(structure.legs[i] for i in axes)
    if legs == plan.new_structure.legs:
        return plan
    return replace(plan, new_structure=TensorStructure(legs))


@cache
def _pattern_permutation(structure: TensorStructure, axes: tuple) -> PermutationPlan:
    new = TensorStructure(tuple(structure.legs[i] for i in axes))
    out_perm = _side_perm(axes, structure.out_axes, new.out_axes)
    in_perm = _side_perm(axes, structure.in_axes, new.in_axes)
    if out_perm == tuple(range(len(out_perm))) and in_perm == tuple(range(len(in_perm))):
        # Only the OUT/IN interleaving moved: every key survives unchanged and the
        # sorted block order is identical (reference permutation.py:166-170).
        return PermutationPlan(axes, new, tuple((i, i, 1.0) for i in range(structure.num_blocks)))
    provider = structure.provider
    if not hasattr(provider, "permute_tree"):
        raise CapabilityError(f"{provider.name} does not implement permutation coefficients")
    terms = []
    for src, key in enumerate(structure.block_order):
        for ot, c_out in provider.permute_tree(key.output_tree, out_perm):
            for it, c_in in provider.permute_tree(key.input_tree, in_perm):
                coeff = c_out * c_in
                if coeff == 0:
                    continue
                terms.append((src, new.index_of(FusionBlockKey(ot, it)), coeff))
    return PermutationPlan(axes, new, tuple(terms))


# ----------------------------------------------------------------------------- bend


def _on_side(leg: Leg, side: Side) -> Leg:
    return leg if leg.side is side else replace(leg, side=side, dual=not leg.dual)


@cache
def bend_plan(structure: TensorStructure, axis: int) -> PermutationPlan:
    """Bend ``axis`` (the last leg of its side) to the end of the other side."""
    plan = _pattern_bend(pattern_of(structure), axis)
    leg = structure.legs[axis]
    moved = _on_side(leg, IN if leg.side is OUT else OUT)
    legs = (*(o for i, o in enumerate(structure.legs) if i != axis), moved)
    if legs == plan.new_structure.legs:
        return plan
    return replace(plan, new_structure=TensorStructure(legs))


@cache
def _pattern_bend(structure: TensorStructure, axis: int) -> PermutationPlan:
    provider = structure.provider
    leg = structure.legs[axis]
    right = leg.side is OUT
    bend = provider.bend_right if right else provider.bend_left
    moved = replace(leg, side=IN if right else OUT, dual=not leg.dual)
    new = TensorStructure((*(o for i, o in enumerate(structure.legs) if i != axis), moved))
    terms = []
    for src, key in enumerate(structure.block_order):
        for new_key, coeff in bend(key, dual=leg.dual):
            if coeff == 0:
                continue
            terms.append((src, new.index_of(new_key), coeff))
    perm = (*(i for i in range(structure.ndim) if i != axis), axis)
    return PermutationPlan(perm, new, tuple(terms))


# ----------------------------------------------------------------------------- repartition


@dataclass(frozen=True)
class RepartitionPlan:
    """One composed block permutation ``perm`` plus the composed sparse block map."""

    new_structure: TensorStructure
    perm: tuple
    terms: tuple

    @property
    def trivial(self) -> bool:
        """True when every block maps one-to-one with coefficient 1 and no axis moves."""
        return _is_identity_map(self.terms) and self.perm == tuple(range(len(self.perm)))


def _is_identity_map(terms) -> bool:
    return all(s == d and c == 1 for s, d, c in terms)


def _compose_terms(first, second) -> tuple:
    """Sparse product of two block maps, duplicate (src, dst) pairs summed in first-seen
    order (reference ``repartition.py:298-316``)."""
    by_src: dict = {}
    for s, d, c in second:
        by_src.setdefault(s, []).append((d, c))
    merged: dict = {}
    for s, mid, c in first:
        for d, c2 in by_src.get(mid, ()):
            k = (s, d)
            merged[k] = c * c2 if k not in merged else merged[k] + c * c2
    return tuple((s, d, c) for (s, d), c in merged.items())


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


@cache
def repartition_plan(structure: TensorStructure, outputs: tuple, inputs: tuple) -> RepartitionPlan:
    """Plan ``repartition``: outputs become OUT, inputs IN, public order ``(*outputs, *inputs)``."""
    want = {ax: OUT for ax in outputs} | {ax: IN for ax in inputs}
    if not any(structure.legs[ax].side is not want[ax] for ax in range(structure.ndim)):
        p = permutation_plan(structure, (*outputs, *inputs))
        return RepartitionPlan(p.new_structure, p.axes, p.terms)
    plan = _pattern_repartition(pattern_of(structure), outputs, inputs)
    legs = tuple(_on_side(structure.legs[ax], want[ax]) for ax in (*outputs, *inputs))
    if legs == plan.new_structure.legs:
        return plan
    return replace(plan, new_structure=TensorStructure(legs))


@cache
def _pattern_repartition(structure: TensorStructure, outputs: tuple, inputs: tuple):
    """Walk transpose -> bend per crossing leg -> final transpose, composing the block maps
    and the axis permutations (reference ``repartition.py:341-385``)."""
    ndim = structure.ndim
    want = {ax: OUT for ax in outputs} | {ax: IN for ax in inputs}
    crossing = [ax for ax in range(ndim) if structure.legs[ax].side is not want[ax]]
    perm = tuple(range(ndim))
    terms = tuple((i, i, 1.0) for i in range(structure.num_blocks))
    labels = list(range(ndim))
    for ax in crossing:
        p = labels.index(ax)
        axes = tuple(i for i in range(ndim) if i != p) + (p,)
        tp = permutation_plan(structure, axes)
        perm = tuple(perm[i] for i in axes)
        terms = _compose_terms(terms, tp.terms)
        structure = tp.new_structure
        labels.append(labels.pop(p))
        bp = bend_plan(structure, ndim - 1)
        terms = _compose_terms(terms, bp.terms)
        structure = bp.new_structure
    position = {a: p for p, a in enumerate(labels)}
    axes = tuple(position[ax] for ax in (*outputs, *inputs))
    tp = permutation_plan(structure, axes)
    return RepartitionPlan(
        tp.new_structure, tuple(perm[i] for i in axes), _compose_terms(terms, tp.terms)
    )


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


@cache
def map_layout(structure: TensorStructure) -> MapLayout:
    out_axes, in_axes = structure.out_axes, structure.in_axes
    row_trees: dict = {}
    col_trees: dict = {}
    for key in structure.block_order:
        row_trees.setdefault(key.coupled, {})[key.output_tree] = None
        col_trees.setdefault(key.coupled, {})[key.input_tree] = None
    sectors = tuple(sorted(row_trees))
    rows, cols, grid = [], [], []
    for c in sectors:
        rb = _bands(structure, out_axes, c, row_trees[c])
        cb = _bands(structure, in_axes, c, col_trees[c])
        rows.extend(rb)
        cols.extend(cb)
        grid.append(
            (
                c,
                tuple(
                    tuple(structure.index_of(FusionBlockKey(ot, it)) for _, it, _, _ in cb)
                    for _, ot, _, _ in rb
                ),
            )
        )
    covered = sum(len(cells) * len(cells[0]) for _, cells in grid)
    if covered != structure.num_blocks:
        raise RuntimeError("map_layout: block order is not a per-sector cross product")
    return MapLayout(structure, sectors, tuple(rows), tuple(cols), tuple(grid))


def _bands(structure, axes, c, trees) -> tuple:
    bands, offset = [], 0
    for tree in trees:
        extent = 1
        for ax, u in zip(axes, tree.uncoupled):
            leg = structure.legs[ax]
            extent *= leg.degeneracy(leg.space_sector(u))
        bands.append((c, tree, offset, extent))
        offset += extent
    return tuple(bands)


# ----------------------------------------------------------------------------- contraction


def outward_dual(leg: Leg) -> bool:
    """The all-outgoing object of a leg: dualized exactly when ``side is IN`` xor ``dual``."""
    return leg.dual != (leg.side is IN)


def contractible(x: Leg, y: Leg) -> bool:
    return x.space == y.space and outward_dual(x) != outward_dual(y)


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


@cache
def contraction_plan(a: TensorStructure, b: TensorStructure, axes: tuple) -> ContractionPlan:
    ca, cb = axes
    if a.provider != b.provider:
        raise ValueError(
            f"tensordot: a has provider {a.provider.name} and b has {b.provider.name}"
        )
    for k, (i, j) in enumerate(zip(ca, cb)):
        x, y = a.legs[i], b.legs[j]
        if not contractible(x, y):
            reason = (
                "their spaces differ"
                if x.space != y.space
                else "both ends of the wire point the same way (equal outward duals)"
            )
            raise ValueError(
                f"tensordot: pair {k}, axis {i} of a and axis {j} of b do not contract: {reason}"
            )
    sca, scb = set(ca), set(cb)
    fa = tuple(i for i in range(a.ndim) if i not in sca)
    fb = tuple(j for j in range(b.ndim) if j not in scb)
    if not fa and not fb:
        raise ValueError("tensordot: this contraction leaves no free leg")
    free = (*(a.legs[i] for i in fa), *(b.legs[j] for j in fb))
    outputs = tuple(k for k, leg in enumerate(free) if leg.side is OUT)
    inputs = tuple(k for k, leg in enumerate(free) if leg.side is IN)
    position = {k: p for p, k in enumerate((*outputs, *inputs))}
    return ContractionPlan(
        a_outputs=fa,
        a_inputs=ca,
        b_outputs=cb,
        b_inputs=fb,
        restore_outputs=outputs,
        restore_inputs=inputs,
        final_transpose=tuple(position[k] for k in range(len(free))),
        new_structure=TensorStructure(free),
    )
