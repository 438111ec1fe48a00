"""ORACLE -- test infrastructure only.  A NumPy float64 dense restatement of SU(2) (and
fermion-parity) symmetric tensors, used by ``tests/`` as the checker of block values that
go through multi-term fusion-tree recoupling.  Nothing in the product imports it.

Independent of the product package on purpose: legs are plain tuples, trees are plain
tuples, and every coefficient comes from the Clebsch-Gordan closed form below.

* ``cg(j1, j2, j3)`` -- Clebsch-Gordan tensor ``<j1 m1; j2 m2 | j3 m3>`` of shape
  ``(d1, d2, d3)``, Condon-Shortley phase, magnetic indices descending (index ``k``
  carries ``m = j - k``), exactly the layout of reference ``symmetry/su2.py:162-171``
  (``cgc``, trailing multiplicity axis dropped).  Racah's single-sum formula, evaluated
  in exact rational arithmetic before one final square root.
* ``z_matrix(j)`` -- ``Z[i, d-1-i] = (-1)^i`` (reference ``su2.py:258-270``,
  ``_sun_coeff.py:109-117``: ``sqrt(d) * cg(j, j, 0)``).
* ``block_keys(legs)`` -- the reference's canonical block order (``structure.py:349-371``,
  ``fusion_tree.py:33-267``): per side the left-associated trees over the fused labels,
  sorted as ``(uncoupled, inner, multiplicities, coupled)`` with the output tree first.
* ``to_dense``/``from_dense`` -- ``T = sum_tau A^(tau) (x) C^(tau)`` with the tree's CG
  tensor contracted along the spine, ``Z`` applied on dual legs and the input tree
  conjugated (reference ``ops/dense.py:173-283``), and its pseudo-inverse.
* ``f_symbol_from_cg`` -- ``[F^{abc}_d]_{e,f}`` recomputed from CG tensors: the overlap of
  the ``((ab)_e c)_d`` and ``(a (bc)_f)_d`` bases divided by ``d_d``.  Comparing it with the
  reference fixture ``su2_f.txt`` pins this module's CG gauge to the reference's F gauge.
* ``supersign`` -- the dense Koszul sign of a graded transpose (reference
  ``tests/helpers.py:72-93``), for fermion-parity legs.

Pinned by: sympy's exact Clebsch-Gordan values, CG orthonormality, F-from-CG against
every fixture row tested, and ``to_dense`` norm identity ``|T| = |dense|`` with qdim
weights (reference ``basic.py:340-346``) -- see ``tests/test_oracle_su2.py``.
"""

from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from itertools import product
from math import factorial, sqrt

import numpy as np


# ------------------------------------------------------------------ Clebsch-Gordan


def _tri(a: int, b: int, c: int) -> bool:
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def _cg_value(j1: int, m1: int, j2: int, m2: int, j3: int, m3: int) -> float:
    """``<j1 m1; j2 m2 | j3 m3>`` with every argument doubled (Racah's formula)."""
    if m1 + m2 != m3 or not _tri(j1, j2, j3):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = lambda x2: factorial(x2 // 2)  # noqa: E731  argument is a doubled integer
    pre = Fraction((j3 + 1) * f(j3 + j1 - j2) * f(j3 - j1 + j2) * f(j1 + j2 - j3), f(j1 + j2 + j3 + 2))
    pre *= f(j3 + m3) * f(j3 - m3) * f(j1 - m1) * f(j1 + m1) * f(j2 - m2) * f(j2 + m2)
    s = Fraction(0)
    for k in range(0, j1 + j2 + 1):
        args = (j1 + j2 - j3 - 2 * k, j1 - m1 - 2 * k, j2 + m2 - 2 * k,
                j3 - j2 + m1 + 2 * k, j3 - j1 - m2 + 2 * k)
        if min(args) < 0:
            continue
        den = factorial(k)
        for a in args:
            den *= f(a)
        s += Fraction((-1) ** k, den)
    if s == 0:
        return 0.0
    sign = 1.0 if s > 0 else -1.0
    val = s * s * pre  # the square of the coefficient, exact
    return sign * sqrt(val.numerator) / sqrt(val.denominator)


@lru_cache(maxsize=None)
def cg(j1: int, j2: int, j3: int) -> np.ndarray:
    """``(d1, d2, d3)`` Clebsch-Gordan tensor; doubled spins; descending magnetic index."""
    if not _tri(j1, j2, j3):
        raise ValueError(f"{j3} is not in {j1} x {j2}")
    out = np.zeros((j1 + 1, j2 + 1, j3 + 1))
    for a in range(j1 + 1):
        for b in range(j2 + 1):
            m1, m2 = j1 - 2 * a, j2 - 2 * b
            m3 = m1 + m2
            if abs(m3) <= j3:
                out[a, b, (j3 - m3) // 2] = _cg_value(j1, m1, j2, m2, j3, m3)
    out.flags.writeable = False
    return out


def z_matrix(j: int) -> np.ndarray:
    """``Z_j: V_j -> V_j^*`` in the descending basis."""
    return sqrt(j + 1) * cg(j, j, 0)[:, :, 0]


def f_symbol_from_cg(a: int, b: int, c: int, d: int, e: int, f: int) -> float:
    """``[F^{abc}_d]_{e,f}`` from CG tensors (zero off the four triangles)."""
    if not (_tri(a, b, e) and _tri(e, c, d) and _tri(b, c, f) and _tri(a, f, d)):
        return 0.0
    left = np.einsum("xyE,Ezw->xyzw", cg(a, b, e), cg(e, c, d))
    right = np.einsum("yzF,xFw->xyzw", cg(b, c, f), cg(a, f, d))
    return float(np.sum(left * right) / (d + 1))


# ------------------------------------------------------------------ legs, trees, keys


class SLeg:
    """``sectors``: ((label, degeneracy), ...) sorted by label; ``out`` side; ``dual``."""

    def __init__(self, sectors, out: bool, dual: bool = False) -> None:
        self.sectors = tuple(sorted((int(a), int(m)) for a, m in sectors))
        self.out = bool(out)
        self.dual = bool(dual)

    def deg(self, a: int) -> int:
        return dict(self.sectors).get(a, 0)

    def offset(self, a: int, irrep_dim) -> int:
        o = 0
        for s, m in self.sectors:
            if s == a:
                return o
            o += m * irrep_dim(s)
        raise KeyError(a)

    def dim(self, irrep_dim) -> int:
        return sum(m * irrep_dim(s) for s, m in self.sectors)


class SU2Rules:
    """Doubled-spin labels: triangle fusion, self-dual, ``d = 2j+1``."""

    unit = 0

    @staticmethod
    def fuse(a: int, b: int):
        return tuple(range(abs(a - b), a + b + 1, 2))

    @staticmethod
    def dual(a: int) -> int:
        return a

    @staticmethod
    def irrep_dim(a: int) -> int:
        return a + 1

    qdim = irrep_dim


class FZ2Rules:
    """Parity labels 0/1: XOR fusion, self-dual, ``d = 1``."""

    unit = 0

    @staticmethod
    def fuse(a: int, b: int):
        return (a ^ b,)

    @staticmethod
    def dual(a: int) -> int:
        return a

    @staticmethod
    def irrep_dim(a: int) -> int:
        return 1

    qdim = irrep_dim


def trees(rules, uncoupled: tuple) -> list:
    """Every left-associated tree ``(uncoupled, inner, coupled)`` over ``uncoupled``."""
    if not uncoupled:
        return [((), (), rules.unit)]
    if len(uncoupled) == 1:
        return [(uncoupled, (), uncoupled[0])]
    out = []
    for unc, inner, e in trees(rules, uncoupled[:-1]):
        new_inner = (*inner, e) if len(unc) >= 2 else ()
        for c in rules.fuse(e, uncoupled[-1]):
            out.append((uncoupled, new_inner, c))
    return out


def block_keys(rules, legs) -> list:
    """``[(out_tree, in_tree, axis_labels)]`` in canonical block order."""
    outs = [i for i, l in enumerate(legs) if l.out]
    ins = [i for i, l in enumerate(legs) if not l.out]
    keys = []
    for combo in product(*([a for a, _ in l.sectors] for l in legs)):
        uo = tuple(rules.dual(combo[i]) if legs[i].dual else combo[i] for i in outs)
        ui = tuple(rules.dual(combo[i]) if legs[i].dual else combo[i] for i in ins)
        tin = trees(rules, ui)
        for to in trees(rules, uo):
            for ti in tin:
                if ti[2] == to[2]:
                    keys.append((to, ti, combo))
    # multiplicity labels are all zero for these multiplicity-free rules, so the
    # (uncoupled, inner, coupled) tuple order is the reference's field order
    keys.sort(key=lambda k: (k[0], k[1]))
    return keys


def block_shape(legs, key) -> tuple:
    return tuple(l.deg(a) for l, a in zip(legs, key[2]))


def random_blocks(rules, legs, seed: int) -> list:
    """``SymmetricTensor.random``: one ``default_rng(seed)`` normal draw per block in block
    order (reference ``tensor.py:252-262``)."""
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(block_shape(legs, k)) for k in block_keys(rules, legs)]


# ------------------------------------------------------------------ dense expansion


def _tree_cgt(tree, duals) -> np.ndarray:
    unc, inner, coupled = tree
    if not unc:
        return np.ones((1,))
    x = np.eye(unc[0] + 1)
    spine = (unc[0], *inner, coupled) if len(unc) >= 2 else (coupled,)
    for k in range(1, len(unc)):
        x = np.tensordot(x, cg(spine[k - 1], unc[k], spine[k]), axes=([-1], [0]))
    for i, d in enumerate(duals):
        if d:
            x = np.moveaxis(np.tensordot(x, z_matrix(unc[i]), axes=([i], [1])), -1, i)
    return x


def _cell_cgt(legs, key) -> np.ndarray:
    outs = [i for i, l in enumerate(legs) if l.out]
    ins = [i for i, l in enumerate(legs) if not l.out]
    xo = _tree_cgt(key[0], [legs[i].dual for i in outs])
    xi = _tree_cgt(key[1], [legs[i].dual for i in ins])
    c = np.tensordot(xo, xi.conj(), axes=([-1], [-1]))
    order = (*outs, *ins)
    return np.transpose(c, sorted(range(len(legs)), key=order.__getitem__))


def _slabs(rules, legs, labels):
    return tuple(slice(l.offset(a, rules.irrep_dim), l.offset(a, rules.irrep_dim) + l.deg(a) * rules.irrep_dim(a))
                 for l, a in zip(legs, labels))


def to_dense(rules, legs, blocks) -> np.ndarray:
    """Dense array: axis ``i`` runs over ``legs[i]``'s sectors in order, index
    ``alpha * d_a + m`` inside sector ``a``'s slab."""
    n = len(legs)
    out = np.zeros(tuple(l.dim(rules.irrep_dim) for l in legs))
    for key, blk in zip(block_keys(rules, legs), blocks):
        cgt = _cell_cgt(legs, key) if rules is SU2Rules else np.ones((1,) * n)
        full = np.einsum(blk, list(range(n)), cgt, list(range(n, 2 * n)),
                         [x for i in range(n) for x in (i, n + i)])
        out[_slabs(rules, legs, key[2])] += full.reshape(tuple(m * d for m, d in zip(blk.shape, cgt.shape)))
    return out


def from_dense(rules, legs, dense) -> tuple[list, float]:
    """Blocks (in block order) of the projection of ``dense`` onto the symmetric subspace,
    and the residual norm of what the projection leaves out."""
    n = len(legs)
    keys = block_keys(rules, legs)
    cells: dict = {}
    for i, k in enumerate(keys):
        cells.setdefault(k[2], []).append(i)
    blocks = [None] * len(keys)
    forbidden = np.array(dense, dtype=np.float64)
    resid_sq = 0.0
    for labels, idx in cells.items():
        degs = tuple(l.deg(a) for l, a in zip(legs, labels))
        dims = tuple(rules.irrep_dim(a) for a in labels)
        piece = dense[_slabs(rules, legs, labels)].reshape(tuple(x for p in zip(degs, dims) for x in p))
        piece = piece.transpose(tuple(range(0, 2 * n, 2)) + tuple(range(1, 2 * n, 2)))
        mat = piece.reshape(int(np.prod(degs)), int(np.prod(dims)))
        cmat = np.stack([(_cell_cgt(legs, keys[i]) if rules is SU2Rules else np.ones(dims)).reshape(-1) for i in idx])
        coeffs = mat @ np.linalg.pinv(cmat)
        for c, i in enumerate(idx):
            blocks[i] = coeffs[:, c].reshape(degs)
        resid_sq += float(np.sum((mat - coeffs @ cmat) ** 2))
        forbidden[_slabs(rules, legs, labels)] = 0.0
    return blocks, sqrt(resid_sq + float(np.sum(forbidden ** 2)))


def norm(rules, legs, blocks) -> float:
    """``sqrt(sum_tau qdim(c_tau) |A_tau|^2)`` (reference ``basic.py:361-367``)."""
    keys = block_keys(rules, legs)
    return sqrt(sum(rules.qdim(k[0][2]) * float(np.sum(b * b)) for k, b in zip(keys, blocks)))


# ------------------------------------------------------------------ fermion parity


def supersign(legs, p, per_side: bool = True) -> np.ndarray:
    """Koszul sign array shaped like ``np.transpose(dense, p)`` for parity-graded legs
    (labels are parities); ``per_side`` counts only inversions within one side."""
    pars = [np.concatenate([np.full(m, a) for a, m in legs[ax].sectors]) for ax in p]
    sides = [legs[ax].out for ax in p]
    n = len(p)
    sign = np.ones(tuple(len(v) for v in pars))
    for j in range(n):
        for k in range(j + 1, n):
            if p[j] <= p[k] or (per_side and sides[j] != sides[k]):
                continue
            sj = [1] * n
            sj[j] = len(pars[j])
            sk = [1] * n
            sk[k] = len(pars[k])
            sign = sign * (-1.0) ** (pars[j].reshape(sj) * pars[k].reshape(sk))
    return sign
