"""Graded spaces ``V = (+)_a C^{m_a} (x) V_a`` and the per-axis ``Leg``.

``GradedSpace`` is a provider plus a canonical (sorted) sector -> degeneracy tuple
(reference ``space.py:25-249``).  ``Leg`` adds ``side`` (IN/OUT: domain or codomain) and
``dual`` (V or V*) as independent per-axis flags (reference ``leg.py:33-187``).  Both
are frozen and hashable so they can key plan caches.
"""

from __future__ import annotations

import enum
from collections.abc import Hashable, Iterable, Mapping
from dataclasses import dataclass, field, replace

__all__ = ["GradedSpace", "Leg", "Side", "IN", "OUT"]


@dataclass(frozen=True, slots=True)
class GradedSpace:
    """Immutable graded space; ``sectors`` is sorted by sector, every ``m >= 1``.

    Build one with ``GradedSpace.new``; the raw constructor trusts its input.
    """

    provider: object
    sectors: tuple
    # derived once: plan caches hash spaces constantly, and degeneracy lookups are hot
    _hash: int = field(default=0, init=False, repr=False, compare=False)
    _map: dict = field(default=None, init=False, repr=False, compare=False)

    def __post_init__(self) -> None:
        object.__setattr__(self, "_hash", hash((self.provider, self.sectors)))
        object.__setattr__(self, "_map", dict(self.sectors))

    def __hash__(self) -> int:
        return self._hash

    @classmethod
    def new(cls, provider, sectors: Mapping | Iterable) -> "GradedSpace":
        """Sort, reject duplicates, non-positive degeneracies and foreign sector types."""
        pairs = list(sectors.items() if isinstance(sectors, Mapping) else sectors)
        expected = type(provider.unit)
        seen = set()
        for a, m in pairs:
            if type(a) is not expected:
                raise TypeError(
                    f"sector {a!r} of type {type(a).__name__} does not belong to provider "
                    f"{provider.name} (expects {expected.__name__})"
                )
            if m <= 0:
                raise ValueError(f"degeneracy of {a!r} must be positive, got {m}")
            if a in seen:
                raise ValueError(f"duplicate sector {a!r}")
            seen.add(a)
        return cls(provider, tuple(sorted(((a, int(m)) for a, m in pairs), key=lambda p: p[0])))

    def degeneracy(self, a) -> int:
        """``m_a``, or 0 when ``a`` is absent."""
        return self._map.get(a, 0)

    def _lookup(self) -> dict:
        return self._map

    def __contains__(self, a) -> bool:
        return self.degeneracy(a) > 0

    def __iter__(self):
        return (a for a, _ in self.sectors)

    def __len__(self) -> int:
        return len(self.sectors)

    @property
    def reduced_dim(self) -> int:
        """``sum_a m_a``."""
        return sum(m for _, m in self.sectors)

    @property
    def dim(self) -> int:
        """``sum_a m_a * qdim(a)`` as an integer dense dimension."""
        return int(sum(m * self.provider.qdim(a) for a, m in self.sectors))

    def flat(self) -> "GradedSpace":
        """Every degeneracy set to 1: the part of a space a block set depends on."""
        return GradedSpace(self.provider, tuple((a, 1) for a, _ in self.sectors))


class Side(enum.Enum):
    """Which half of ``Hom(domain, codomain)`` a leg belongs to."""

    OUT = "out"
    IN = "in"


OUT = Side.OUT
IN = Side.IN


@dataclass(frozen=True, slots=True)
class Leg:
    """One tensor axis: ``space``, ``side``, ``dual`` and an optional ``name``."""

    space: GradedSpace
    side: Side
    dual: bool = False
    name: Hashable | None = None

    @property
    def provider(self):
        return self.space.provider

    @property
    def sectors(self) -> tuple:
        return tuple(self.space)

    def degeneracy(self, a) -> int:
        return self.space.degeneracy(a)

    def fused_sector(self, a):
        """The label this leg contributes to a fusion tree: ``dual(a)`` on a dual leg."""
        return self.provider.dual(a) if self.dual else a

    def space_sector(self, u):
        """Inverse of ``fused_sector``."""
        return self.provider.dual(u) if self.dual else u

    def dualized(self) -> "Leg":
        return replace(self, dual=not self.dual)

    def renamed(self, name) -> "Leg":
        return replace(self, name=name)
