"""ORACLE -- test infrastructure only.  A NumPy float64 restatement of the reference's
U(1) block-sparse lowering, used by ``tests/`` (and ``bench.py``'s cpu_baseline leg) as the
checker.  Nothing in the product imports it.

Independent of the product package on purpose: tensors are dicts keyed by the tuple of
*space* charges per public axis, legs are plain tuples, and every step is written from the
reference's own algorithm:

* block order: keys sorted by (fused OUT charges, fused IN charges) -- the U(1) reading
  of ``structure.py:349-371`` + ``fusion_tree.py:251-267`` (left-associated trees of a
  unique-fusion symmetry are determined by their uncoupled labels);
* repartition: ``ops/repartition.py:388-476`` -- for U(1) every bend/braid coefficient is
  1 (``u1.py:90-126``), so a repartition is the block transpose to ``(*outputs,
  *inputs)`` with crossing legs flipped in side and dual;
* compose: ``ops/map.py:105-172`` with ``to_matrices``/``from_matrices``
  (``map_view.py:238-335``): one ``np.matmul`` per coupled sector over bands ordered as
  block order restricted;
* tensordot/einsum: ``ops/contraction.py:309-376,695-794``;
* inner/norm: ``contraction.py:584-589``, ``basic.py:361-367``.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import product

import numpy as np


@dataclass(frozen=True)
class OLeg:
    """(charge, degeneracy) pairs sorted by charge; ``out`` side flag; ``dual`` flag."""

    sectors: tuple
    out: bool
    dual: bool = False

    def deg(self, q: int) -> int:
        return dict(self.sectors).get(q, 0)

    def fused(self, q: int) -> int:
        return -q if self.dual else q


def space(d: dict) -> tuple:
    return tuple(sorted((int(q), int(m)) for q, m in d.items()))


def block_keys(legs) -> list:
    """Allowed keys (space charges per axis) in the reference's canonical block order."""
    outs = [i for i, l in enumerate(legs) if l.out]
    ins = [i for i, l in enumerate(legs) if not l.out]
    keys = []
    for combo in product(*([q for q, _ in l.sectors] for l in legs)):
        fo = tuple(legs[i].fused(combo[i]) for i in outs)
        fi = tuple(legs[i].fused(combo[i]) for i in ins)
        if sum(fo) == sum(fi):
            keys.append(((fo, fi), combo))
    keys.sort(key=lambda x: x[0])
    return [k for _, k in keys]


def block_shape(legs, key) -> tuple:
    return tuple(l.deg(q) for l, q in zip(legs, key))


class OTensor:
    """legs + {key: float64 block}."""

    def __init__(self, legs, blocks: dict) -> None:
        self.legs = tuple(legs)
        self.blocks = blocks

    @classmethod
    def random(cls, legs, seed):
        """``SymmetricTensor.random``: one ``default_rng(seed)`` draw per block in block
        order (reference ``tensor.py:252-262``)."""
        rng = np.random.default_rng(seed)
        return cls(legs, {k: rng.standard_normal(block_shape(legs, k)) for k in block_keys(legs)})

    @classmethod
    def ones(cls, legs):
        return cls(legs, {k: np.ones(block_shape(legs, k)) for k in block_keys(legs)})

    @classmethod
    def from_flat(cls, legs, flat: np.ndarray):
        out, off = {}, 0
        for k in block_keys(legs):
            shp = block_shape(legs, k)
            n = int(np.prod(shp))
            out[k] = np.asarray(flat[off : off + n], dtype=np.float64).reshape(shp)
            off += n
        if off != len(flat):
            raise ValueError("flat buffer length does not match the block set")
        return cls(legs, out)

    def flat(self) -> np.ndarray:
        ks = block_keys(self.legs)
        return np.concatenate([self.blocks[k].reshape(-1) for k in ks]) if ks else np.zeros(0)

    def keys(self) -> list:
        return block_keys(self.legs)


# ------------------------------------------------------------------ permutations / bends


def transpose(t: OTensor, axes) -> OTensor:
    """U(1) transpose: legs and block axes permuted, coefficient 1."""
    legs = tuple(t.legs[a] for a in axes)
    return OTensor(legs, {tuple(k[a] for a in axes): np.transpose(b, axes) for k, b in t.blocks.items()})


def repartition(t: OTensor, outputs, inputs) -> OTensor:
    """``outputs`` become OUT, ``inputs`` IN; a crossing leg flips side and dual."""
    order = (*outputs, *inputs)
    want = {a: True for a in outputs} | {a: False for a in inputs}
    legs = []
    for a in order:
        l = t.legs[a]
        legs.append(l if l.out == want[a] else OLeg(l.sectors, want[a], not l.dual))
    return OTensor(tuple(legs), {tuple(k[a] for a in order): np.transpose(b, order) for k, b in t.blocks.items()})


def adjoint(t: OTensor) -> OTensor:
    """Every leg flips side; for real blocks keyed by space charges nothing else moves."""
    return OTensor(tuple(OLeg(l.sectors, not l.out, l.dual) for l in t.legs), dict(t.blocks))


# ------------------------------------------------------------------ map lowering


def _layout(legs):
    """Per coupled sector: ordered row bands (OUT charge tuple -> extent) and column bands."""
    outs = [i for i, l in enumerate(legs) if l.out]
    ins = [i for i, l in enumerate(legs) if not l.out]
    rows, cols = {}, {}
    for k in block_keys(legs):
        c = sum(legs[i].fused(k[i]) for i in outs)
        ro = tuple(k[i] for i in outs)
        co = tuple(k[i] for i in ins)
        rows.setdefault(c, {}).setdefault(ro, int(np.prod([legs[i].deg(k[i]) for i in outs])))
        cols.setdefault(c, {}).setdefault(co, int(np.prod([legs[i].deg(k[i]) for i in ins])))
    return outs, ins, rows, cols


def to_matrices(t: OTensor) -> dict:
    outs, ins, rows, cols = _layout(t.legs)
    order = (*outs, *ins)
    mats = {}
    for c in sorted(rows):
        rparts = []
        for ro, dr in rows[c].items():
            parts = []
            for co, dc in cols[c].items():
                key = [0] * len(t.legs)
                for i, q in zip(outs, ro):
                    key[i] = q
                for i, q in zip(ins, co):
                    key[i] = q
                parts.append(np.transpose(t.blocks[tuple(key)], order).reshape(dr, dc))
            rparts.append(np.concatenate(parts, axis=1))
        mats[c] = np.concatenate(rparts, axis=0)
    return mats


def from_matrices(legs, mats: dict) -> OTensor:
    outs, ins, rows, cols = _layout(legs)
    order = (*outs, *ins)
    inverse = tuple(np.argsort(order))
    blocks = {}
    for c in rows:
        m = mats[c]
        r0 = 0
        for ro, dr in rows[c].items():
            c0 = 0
            for co, dc in cols[c].items():
                key = [0] * len(legs)
                for i, q in zip(outs, ro):
                    key[i] = q
                for i, q in zip(ins, co):
                    key[i] = q
                key = tuple(key)
                shp = block_shape(legs, key)
                piece = m[r0 : r0 + dr, c0 : c0 + dc].reshape(tuple(shp[a] for a in order))
                blocks[key] = np.transpose(piece, inverse)
                c0 += dc
            r0 += dr
    return OTensor(legs, blocks)


def compose(a: OTensor, b: OTensor) -> OTensor:
    """One np.matmul per coupled sector of the output layout; a sector carried by one
    operand only is a zero matrix (reference ``ops/map.py:154-171``)."""
    legs = tuple(a.legs[i] for i, l in enumerate(a.legs) if l.out) + tuple(
        b.legs[i] for i, l in enumerate(b.legs) if not l.out
    )
    ma, mb = to_matrices(a), to_matrices(b)
    _, _, rows, cols = _layout(legs)
    mats = {}
    for c in rows:
        if c in ma and c in mb:
            mats[c] = np.matmul(ma[c], mb[c])
        else:
            mats[c] = np.zeros((sum(rows[c].values()), sum(cols[c].values())))
    return from_matrices(legs, mats)


def tensordot(a: OTensor, b: OTensor, axes) -> OTensor:
    """repartition -> compose -> restore sides -> tensordot order."""
    ca, cb = (tuple(x) for x in axes)
    for i, j in zip(ca, cb):
        x, y = a.legs[i], b.legs[j]
        # contractible: same space, opposite all-outgoing duality (contraction.py:99-114)
        if x.sectors != y.sectors or (x.dual != (not x.out)) == (y.dual != (not y.out)):
            raise ValueError(f"axis {i} of a and axis {j} of b do not contract")
    fa = tuple(i for i in range(len(a.legs)) if i not in ca)
    fb = tuple(j for j in range(len(b.legs)) if j not in cb)
    free = [a.legs[i] for i in fa] + [b.legs[j] for j in fb]
    c = compose(repartition(a, fa, ca), repartition(b, cb, fb))
    outputs = tuple(k for k, l in enumerate(free) if l.out)
    inputs = tuple(k for k, l in enumerate(free) if not l.out)
    c = repartition(c, outputs, inputs)
    position = {k: p for p, k in enumerate((*outputs, *inputs))}
    return transpose(c, tuple(position[k] for k in range(len(free))))


def einsum(eq: str, a: OTensor, b: OTensor) -> OTensor:
    lhs, out = eq.replace(" ", "").split("->")
    t0, t1 = lhs.split(",")
    shared = [x for x in t0 if x in t1]
    axes = (tuple(t0.index(x) for x in shared), tuple(t1.index(x) for x in shared))
    free = [x for x in t0 if x not in shared] + [x for x in t1 if x not in shared]
    return transpose(tensordot(a, b, axes), tuple(free.index(x) for x in out))


def composed(eq: str, a: OTensor, b: OTensor, bend: str = "") -> OTensor:
    """Reference ``network/env.py:60-90``: bend the named wires on both operands, then einsum."""
    if bend:
        lhs, out = eq.split("->")
        ta, tb = lhs.split(",")

        def bent(t, term):
            flip = set(bend)
            outs = tuple(i for i, ch in enumerate(term) if t.legs[i].out != (ch in flip))
            ins = tuple(i for i in range(len(term)) if i not in outs)
            return repartition(t, outs, ins), "".join(term[i] for i in (*outs, *ins))

        a, ta = bent(a, ta)
        b, tb = bent(b, tb)
        eq = f"{ta},{tb}->{out}"
    return einsum(eq, a, b)


# ------------------------------------------------------------------ vector space


def inner(a: OTensor, b: OTensor) -> float:
    return float(sum(np.sum(a.blocks[k] * b.blocks[k]) for k in a.blocks))


def norm(t: OTensor) -> float:
    return float(np.sqrt(inner(t, t)))


def scale(t: OTensor, s: float) -> OTensor:
    return OTensor(t.legs, {k: v * s for k, v in t.blocks.items()})


def axpy(a: float, x: OTensor, y: OTensor) -> OTensor:
    return OTensor(y.legs, {k: a * x.blocks[k] + y.blocks[k] for k in y.blocks})


# ------------------------------------------------------------------ dense pin


def densify(t: OTensor) -> np.ndarray:
    """Place every block at its sectors' offsets (U(1): no Clebsch-Gordan factor)."""
    offs = []
    for l in t.legs:
        o, acc = {}, 0
        for q, m in l.sectors:
            o[q] = acc
            acc += m
        offs.append((o, acc))
    out = np.zeros(tuple(n for _, n in offs))
    for k, b in t.blocks.items():
        sl = tuple(slice(o[q], o[q] + l.deg(q)) for (o, _), l, q in zip(offs, t.legs, k))
        out[sl] = b
    return out
