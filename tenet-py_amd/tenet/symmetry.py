"""Symmetry providers: sector labels, fusion rules and recoupling coefficients.

Sectors are frozen, ordered labels; everything a symmetry *does* lives on its provider.
Three providers cover the contraction engine's workloads:

* ``Trivial``  -- one sector, every coefficient 1 (reference ``symmetry/base.py:1193-1262``).
* ``U1``       -- integer charges, ``dual(q) = -q``, F = R = B = FS = 1
  (reference ``symmetry/u1.py:42-126``).
* ``SU2``      -- labels are ``2j``; triangle fusion, ``qdim = 2j+1``, F from the Racah
  6j formula in the TensorKit gauge, ``R = (-1)^(ja+jb-jc)``, ``B`` derived from F and
  ``chi = (-1)^(2j)`` (reference ``symmetry/su2.py:150-252``, ``_sun_coeff.py:75-117``).
* ``fZ2``      -- fermion parity: XOR fusion, F = B = FS = 1, and the whole fermionic
  content in the braid: a permutation of a tree's lines costs the Koszul sign
  ``(-1)^(number of odd-odd inversions)`` (reference ``symmetry/fz2.py:85-130,185-200``).

Every provider also carries its dense-basis data (``irrep_dim``, ``cgc``, ``z_matrix``)
for ``to_dense``/``from_dense``.  ``bosonic`` says that every braid and bend coefficient
is exactly 1, which is what lets the two-site H_eff run as a pure index computation.

Coefficients are host-side plan data: they enter only when a plan is compiled, never per
block element on the device.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import cache
from math import sqrt

import numpy as np
from typing import TYPE_CHECKING

from tenet import _su2coeff

if TYPE_CHECKING:
    from tenet.fusion_tree import FusionTree
    from tenet.structure import FusionBlockKey

__all__ = [
    "CapabilityError",
    "StructureChangingError",
    "Sector",
    "TrivialSector",
    "U1Sector",
    "SU2Sector",
    "FZ2Sector",
    "Trivial",
    "U1",
    "SU2",
    "fZ2",
    "koszul_sign",
]


class CapabilityError(TypeError):
    """A provider lacks a capability an operation requires (reference ``base.py:700``)."""


class StructureChangingError(TypeError):
    """An output structure depends on block values inside a traced region (``base.py:704``)."""


@dataclass(frozen=True, slots=True, order=True)
class Sector:
    """Marker base for sector labels; subclasses are frozen, slotted and ordered."""


@dataclass(frozen=True, slots=True, order=True)
class TrivialSector(Sector):
    """The single sector of the trivial symmetry."""


@dataclass(frozen=True, slots=True, order=True)
class U1Sector(Sector):
    """A U(1) irrep labelled by its integer charge; negation is duality."""

    charge: int


@dataclass(frozen=True, slots=True, order=True)
class SU2Sector(Sector):
    """An SU(2) irrep labelled by ``two_j`` (so labels stay integers)."""

    two_j: int

    def __post_init__(self) -> None:
        if not isinstance(self.two_j, int) or isinstance(self.two_j, bool):
            raise TypeError(f"two_j must be an int, got {self.two_j!r}")
        if self.two_j < 0:
            raise ValueError(f"two_j must be non-negative, got {self.two_j}")


@dataclass(frozen=True, slots=True, order=True)
class FZ2Sector(Sector):
    """A fermion-parity sector: ``parity`` 0 (even, the unit) or 1 (odd)."""

    parity: int

    def __post_init__(self) -> None:
        if not isinstance(self.parity, int) or isinstance(self.parity, bool):
            raise TypeError(f"parity must be an int, got {self.parity!r}")
        if self.parity not in (0, 1):
            raise ValueError(f"parity must be 0 or 1, got {self.parity}")


# --------------------------------------------------------------------------------------
# Unique-fusion helpers (Trivial, U1): one term, coefficient exactly 1.
# Reference: ``permute_unique_tree`` and ``bend_unique`` in ``symmetry/base.py:778-1074``.
# The spine of the new tree is recomputed from the new uncoupled labels.
# --------------------------------------------------------------------------------------


def _only_tree(provider, uncoupled, coupled) -> "FusionTree":
    from tenet.fusion_tree import fusion_trees

    trees = fusion_trees(provider, uncoupled, coupled)
    if len(trees) != 1:
        raise CapabilityError(
            f"{provider.name}: {uncoupled} -> {coupled!r} has {len(trees)} trees; a "
            "unique-fusion recoupling needs exactly one"
        )
    return trees[0]


def _permute_unique(provider, tree, perm, coeff: float = 1.0):
    uncoupled = tuple(tree.uncoupled[i] for i in perm)
    return ((_only_tree(provider, uncoupled, tree.coupled), coeff),)


def koszul_sign(uncoupled, perm) -> float:
    """``(-1)^(number of inversions of perm among odd lines)``; ``perm[j]`` is the old
    position that becomes position ``j`` (reference ``symmetry/fz2.py:85-130``)."""
    odd = [i for i in perm if uncoupled[i].parity]
    inv = sum(1 for j in range(len(odd)) for k in range(j + 1, len(odd)) if odd[j] > odd[k])
    return -1.0 if inv % 2 else 1.0


_ONE_CGC = np.ones((1, 1, 1))
_ONE_CGC.flags.writeable = False
_ONE_Z = np.ones((1, 1))
_ONE_Z.flags.writeable = False


def _bent_key(provider, key, right: bool):
    """The block key after bending the last uncoupled line of one tree onto the end of the
    other: ``right`` takes it from the output tree, else from the input tree.  The source
    vertex is ``a x b -> c`` (``a`` the unit for a rank-1 source tree); the shortened source
    tree keeps its spine up to ``a``, and the other tree grows by ``dual(b)`` fusing to
    ``a``.  Returns ``(new key, a, b, c)``."""
    from tenet.fusion_tree import FusionTree
    from tenet.structure import FusionBlockKey

    src, dst = (key.output_tree, key.input_tree) if right else (key.input_tree, key.output_tree)
    if src.rank == 0:
        raise ValueError(f"{provider.name}: cannot bend from an empty tree of {key}")
    b, c = src.uncoupled[-1], src.coupled
    spine = src.lines()
    a = spine[-2] if src.rank >= 2 else provider.unit
    shorter = FusionTree(src.uncoupled[:-1], src.inner[:-1], src.multiplicities[:-1], a)
    longer = FusionTree((*dst.uncoupled, provider.dual(b)), dst.lines()[1:],
                        (*dst.multiplicities, 0) if dst.rank else (), a)
    new_key = FusionBlockKey(shorter, longer) if right else FusionBlockKey(longer, shorter)
    return new_key, a, b, c


def _bend_unique(provider, key, right: bool):
    """Unique fusion (Trivial, U(1), fZ2): the bent key with coefficient exactly 1
    (reference ``bend_unique``, ``symmetry/base.py:999-1074``)."""
    return ((_bent_key(provider, key, right)[0], 1.0),)


@dataclass(frozen=True, slots=True)
class TrivialProvider:
    """The trivial symmetry: one sector, every coefficient 1."""

    name: str = "Trivial"

    @property
    def unit(self) -> TrivialSector:
        return TrivialSector()

    def dual(self, a):
        return a

    def fusion(self, a, b):
        return (TrivialSector(),)

    def n_symbol(self, a, b, c) -> int:
        return 1

    def qdim(self, a) -> float:
        return 1.0

    def frobenius_schur(self, a) -> float:
        return 1.0

    def irrep_dim(self, a) -> int:
        return 1

    def cgc(self, a, b, c) -> np.ndarray:
        return _ONE_CGC

    def z_matrix(self, a) -> np.ndarray:
        return _ONE_Z

    @property
    def bosonic(self) -> bool:
        return True

    def permute_tree(self, tree, perm):
        return _permute_unique(self, tree, perm)

    def bend_right(self, key, *, dual: bool):
        return _bend_unique(self, key, True)

    def bend_left(self, key, *, dual: bool):
        return _bend_unique(self, key, False)

    @property
    def unique_fusion(self) -> bool:
        return True


@dataclass(frozen=True, slots=True)
class U1Provider:
    """U(1): abelian, multiplicity free, every recoupling coefficient exactly 1.

    Reference ``symmetry/u1.py:42-126``.
    """

    name: str = "U1"

    @property
    def unit(self) -> U1Sector:
        return U1Sector(0)

    def dual(self, a: U1Sector) -> U1Sector:
        return U1Sector(-a.charge)

    def fusion(self, a: U1Sector, b: U1Sector) -> tuple[U1Sector, ...]:
        return (U1Sector(a.charge + b.charge),)

    def n_symbol(self, a, b, c) -> int:
        return int(c.charge == a.charge + b.charge)

    def qdim(self, a) -> float:
        return 1.0

    def frobenius_schur(self, a) -> float:
        return 1.0

    def irrep_dim(self, a) -> int:
        return 1

    def cgc(self, a, b, c) -> np.ndarray:
        return _ONE_CGC

    def z_matrix(self, a) -> np.ndarray:
        return _ONE_Z

    @property
    def bosonic(self) -> bool:
        return True

    def permute_tree(self, tree, perm):
        return _permute_unique(self, tree, perm)

    def bend_right(self, key, *, dual: bool):
        return _bend_unique(self, key, True)

    def bend_left(self, key, *, dual: bool):
        return _bend_unique(self, key, False)

    @property
    def unique_fusion(self) -> bool:
        return True


def _triangle(a: int, b: int, c: int) -> bool:
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


@dataclass(frozen=True, slots=True)
class SU2Provider:
    """SU(2) with sectors labelled by ``2j``.

    Fusion ``a x b = |a-b|, |a-b|+2, ..., a+b``; ``qdim = 2j+1``; self-dual labels.
    F-symbols come from the host Racah/6j routine in the TensorKit gauge
    (``tenet._su2coeff``), R from the exchange phase, B and the Frobenius-Schur
    indicator derived from F exactly as reference ``_sun_coeff.py:90-106`` defines them.
    Braids follow reference ``permute_braided_tree``/``_artin_braid``
    (``symmetry/base.py:828-996``) and bends ``bend_braided`` (``base.py:1077-1190``).
    """

    name: str = "SU2"

    @property
    def unit(self) -> SU2Sector:
        return SU2Sector(0)

    def dual(self, a: SU2Sector) -> SU2Sector:
        return a

    def fusion(self, a: SU2Sector, b: SU2Sector) -> tuple[SU2Sector, ...]:
        return tuple(SU2Sector(t) for t in range(abs(a.two_j - b.two_j), a.two_j + b.two_j + 1, 2))

    def n_symbol(self, a, b, c) -> int:
        return int(_triangle(a.two_j, b.two_j, c.two_j))

    def qdim(self, a: SU2Sector) -> float:
        return float(a.two_j + 1)

    def f_symbol(self, a, b, c, d, e, f) -> float:
        return _su2coeff.f_symbol(a.two_j, b.two_j, c.two_j, d.two_j, e.two_j, f.two_j)

    def r_symbol(self, a, b, c) -> float:
        return _su2coeff.r_symbol(a.two_j, b.two_j, c.two_j)

    def b_symbol(self, a, b, c) -> float:
        return _su2coeff.b_symbol(a.two_j, b.two_j, c.two_j)

    def frobenius_schur(self, a: SU2Sector) -> float:
        return _su2coeff.frobenius_schur(a.two_j)

    def irrep_dim(self, a: SU2Sector) -> int:
        return a.two_j + 1

    def cgc(self, a: SU2Sector, b: SU2Sector, c: SU2Sector) -> np.ndarray:
        """``(d_a, d_b, d_c)`` Clebsch-Gordan tensor, magnetic indices descending."""
        return _su2coeff.cg(a.two_j, b.two_j, c.two_j)

    def z_matrix(self, a: SU2Sector) -> np.ndarray:
        return _su2coeff.z_matrix(a.two_j)

    @property
    def bosonic(self) -> bool:
        return False

    def permute_tree(self, tree, perm):
        return _permute_braided(self, tree, tuple(perm))

    def bend_right(self, key, *, dual: bool):
        return _bend_braided(self, key, True, dual)

    def bend_left(self, key, *, dual: bool):
        return _bend_braided(self, key, False, dual)

    @property
    def unique_fusion(self) -> bool:
        return False


@dataclass(frozen=True, slots=True)
class FZ2Provider:
    """Fermion parity: Z2 fusion with fermionic braiding.  Abelian, ``d = qdim = 1``;
    ``F = B = FS = 1``, so a bend costs nothing and a permutation costs the Koszul sign
    of the odd lines it moves past each other (reference ``symmetry/fz2.py:133-240``)."""

    name: str = "fZ2"

    @property
    def unit(self) -> FZ2Sector:
        return FZ2Sector(0)

    def dual(self, a: FZ2Sector) -> FZ2Sector:
        return a

    def fusion(self, a: FZ2Sector, b: FZ2Sector) -> tuple[FZ2Sector, ...]:
        return (FZ2Sector(a.parity ^ b.parity),)

    def n_symbol(self, a, b, c) -> int:
        return int(c.parity == a.parity ^ b.parity)

    def qdim(self, a) -> float:
        return 1.0

    def frobenius_schur(self, a) -> float:
        return 1.0

    def twist(self, a: FZ2Sector) -> float:
        return -1.0 if a.parity else 1.0

    def irrep_dim(self, a) -> int:
        return 1

    def cgc(self, a, b, c) -> np.ndarray:
        return _ONE_CGC

    def z_matrix(self, a) -> np.ndarray:
        return _ONE_Z

    def permute_tree(self, tree, perm):
        return _permute_unique(self, tree, perm, koszul_sign(tree.uncoupled, perm))

    def bend_right(self, key, *, dual: bool):
        return _bend_unique(self, key, True)

    def bend_left(self, key, *, dual: bool):
        return _bend_unique(self, key, False)

    @property
    def unique_fusion(self) -> bool:
        return True

    @property
    def bosonic(self) -> bool:
        return False


# --------------------------------------------------------------------------------------
# Braided recoupling for multiplicity-free providers with real F and R (SU(2)).
# --------------------------------------------------------------------------------------


def _adjacent_swaps(perm) -> list:
    """A reduced word for ``perm`` (``perm[j]`` = old position that ends at ``j``) as the
    positions of adjacent exchanges, applied left to right.  Positions are filled from the
    right: the line that belongs last is carried rightwards to its place, then the next."""
    cur = list(range(len(perm)))
    word = []
    for j in range(len(perm) - 1, -1, -1):
        p = cur.index(perm[j])
        for i in range(p, j):
            word.append(i)
            cur[i], cur[i + 1] = cur[i + 1], cur[i]
    return word


@cache
def _permute_braided(provider, tree, perm):
    """``tree`` with its uncoupled lines permuted by ``perm`` (``perm[j]`` = old position of
    the line at new position ``j``), expanded in the tree basis: the product of one braid
    generator per exchange of a reduced word for ``perm``.  For a symmetric braiding such
    as SU(2)'s (``R^2 = 1``) every reduced word gives the same map, so the word need not be
    the reference's (``permute_braided_tree``, ``symmetry/base.py:829-895``).  Returns
    ``((tree, coeff), ...)``."""
    state = {tree: 1.0}
    for i in _adjacent_swaps(perm):
        nxt: dict = {}
        for t, w in state.items():
            for t2, g in _braid_generator(provider, t, i):
                nxt[t2] = nxt.get(t2, 0.0) + w * g
        state = nxt
    return tuple(state.items())


@cache
def _braid_generator(provider, tree, i: int):
    """The braid generator exchanging uncoupled lines ``i`` and ``i + 1`` of a
    left-associated tree, as ``((tree, coeff), ...)``.

    ``i == 0``: the two lines meet at the first vertex; one R-symbol
    ``R^{u0 u1}_{e1}``.  ``i >= 1``: the local subtree is ``((a x b)_e x c)_f`` with
    ``a = e_{i-1}``, ``b = u_i``, ``c = u_{i+1}``, ``f = e_{i+1}``.  An F-move reassociates
    it to ``(a x (b x c)_d)_f`` (row ``e`` of ``F^{abc}_f``), the R-symbol ``R^{bc}_d``
    exchanges ``b`` and ``c`` inside ``d``, and the inverse F-move of the exchanged triple
    (``F^{acb}_f`` transposed, F being real orthogonal in this gauge) returns to the
    left-associated basis with a new inner line ``e'``.
    """
    from tenet.fusion_tree import FusionTree

    u = tree.uncoupled
    spine = tree.lines()
    exchanged = (*u[:i], u[i + 1], u[i], *u[i + 2:])
    if i == 0:
        return ((FusionTree(exchanged, tree.inner, tree.multiplicities, tree.coupled),
                 provider.r_symbol(u[0], u[1], spine[1])),)
    a, b, c, e, f = spine[i - 1], u[i], u[i + 1], spine[i], spine[i + 1]
    mids = [d for d in provider.fusion(b, c) if provider.n_symbol(a, d, f)]
    # F^{abc}_f[e, d] R^{bc}_d: the reassociated, exchanged amplitudes over the middle line d
    amp = {d: provider.f_symbol(a, b, c, f, e, d) * provider.r_symbol(b, c, d) for d in mids}
    terms = []
    for e2 in provider.fusion(a, c):
        if not provider.n_symbol(e2, b, f):
            continue
        w = 0.0
        for d in mids:
            w += amp[d] * provider.f_symbol(a, c, b, f, e2, d)
        inner = (*tree.inner[: i - 1], e2, *tree.inner[i:])
        terms.append((FusionTree(exchanged, inner, tree.multiplicities, tree.coupled), w))
    return tuple(terms)


def _bend_braided(provider, key, right: bool, dual: bool):
    """Bend for SU(2): the bent key of ``_bent_key`` with coefficient
    ``sqrt(qdim(c) / qdim(a)) * B(a, b, c)``, times the Frobenius-Schur indicator of
    ``dual(b)`` when the moved leg was already dual (the input-tree reading is the
    conjugate, equal in this real gauge).  Reference ``bend_braided``,
    ``symmetry/base.py:1077-1190``."""
    new_key, a, b, c = _bent_key(provider, key, right)
    w = sqrt(provider.qdim(c) / provider.qdim(a)) * provider.b_symbol(a, b, c)
    if dual:
        w *= provider.frobenius_schur(provider.dual(b))
    return ((new_key, w),)


Trivial = TrivialProvider()
U1 = U1Provider()
SU2 = SU2Provider()
fZ2 = FZ2Provider()
