"""Left-associated fusion trees, the basis label of a list of sectors.

``FusionTree(uncoupled, inner, multiplicities, coupled)`` with ``e_0 = u_0``,
``e_{k+1} in e_k x u_{k+1}``, ``inner = (e_1..e_{N-2})`` and ``coupled = e_{N-1}``.  The
field order defines the canonical sort that the storage order of blocks rests on
(reference ``fusion_tree.py:33-267``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from functools import cache

__all__ = ["FusionTree", "fusion_trees", "coupled_sectors"]


@dataclass(frozen=True, slots=True, order=True)
class FusionTree:
    """A left-associated tree: frozen, hashable, totally ordered by its fields."""

    uncoupled: tuple
    inner: tuple
    multiplicities: tuple
    coupled: object
    # trees key every block lookup: the hash (the one the generated __hash__ would compute
    # from the four fields) is computed once
    _h: int = field(init=False, repr=False, compare=False)

    def __post_init__(self) -> None:
        object.__setattr__(self, "_h", hash((self.uncoupled, self.inner, self.multiplicities, self.coupled)))

    def __hash__(self) -> int:
        return self._h

    @property
    def rank(self) -> int:
        return len(self.uncoupled)

    def lines(self) -> tuple:
        """The spine ``(e_0, ..., e_{N-1})``."""
        if not self.uncoupled:
            return ()
        if len(self.uncoupled) == 1:
            return (self.coupled,)
        return (self.uncoupled[0], *self.inner, self.coupled)


def fusion_trees(provider, uncoupled, coupled) -> tuple:
    """Every valid tree for ``uncoupled -> coupled``, sorted."""
    return _trees_to(provider, tuple(uncoupled), coupled)


def coupled_sectors(provider, uncoupled) -> tuple:
    """Every sector reachable from ``uncoupled``, sorted."""
    return _coupled(provider, tuple(uncoupled))


@cache
def _trees_to(provider, uncoupled, coupled):
    return tuple(t for t in _all_trees(provider, uncoupled) if t.coupled == coupled)


@cache
def _coupled(provider, uncoupled):
    return tuple(sorted({t.coupled for t in _all_trees(provider, uncoupled)}))


@cache
def _all_trees(provider, uncoupled):
    """All trees over ``uncoupled`` to any coupled sector, grown one line at a time."""
    if not uncoupled:
        return (FusionTree((), (), (), provider.unit),)
    if len(uncoupled) == 1:
        return (FusionTree(uncoupled, (), (), uncoupled[0]),)
    u = uncoupled[-1]
    grown = []
    for tree in _all_trees(provider, uncoupled[:-1]):
        e = tree.coupled
        inner = (*tree.inner, e) if tree.rank >= 2 else ()
        for c in provider.fusion(e, u):
            for mu in range(provider.n_symbol(e, u, c)):
                grown.append(FusionTree(uncoupled, inner, (*tree.multiplicities, mu), c))
    return tuple(sorted(grown))
