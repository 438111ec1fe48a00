"""MPS and MPO containers on device-resident tensors.

``MPS`` keeps every site on the partition ``(left bond OUT, phys OUT | right bond IN)``
through a write barrier (reference ``network/mps.py:55-352,677-697``).  ``MPO.from_w``
builds the site-tensor operator ``(wl IN, p OUT, p IN, wr OUT)`` from one dense ``W``
(reference ``mps.py:2780-2845``); it carries no edge description, so ``Env.heff2`` takes
the site-tensor path.
"""

from __future__ import annotations

from collections.abc import Iterable, Sequence

import numpy as np
import torch

from tenet import linalg, ops
from tenet.space import IN, OUT, GradedSpace, Leg
from tenet.structure import TensorStructure
from tenet.tensor import SymmetricTensor, default_device

__all__ = ["MPS", "MPO", "from_dense", "overlap"]


def _as_site(t: SymmetricTensor) -> SymmetricTensor:
    if t.ndim != 3:
        raise ValueError(f"an MPS site tensor is rank 3, got rank {t.ndim}")
    return ops.repartition(t, (0, 1), (2,))


class MPS:
    """A finite MPS: ``sites[n]`` on ``(bond_n OUT, phys OUT | bond_{n+1} IN)``."""

    def __init__(self, sites: Iterable[SymmetricTensor], center: int | None = None) -> None:
        self.sites = [_as_site(t) for t in sites]
        self.center = center

    def __len__(self) -> int:
        return len(self.sites)

    def __getitem__(self, n: int) -> SymmetricTensor:
        return self.sites[n]

    def __setitem__(self, n: int, t: SymmetricTensor) -> None:
        if not isinstance(n, int):
            raise TypeError(f"MPS index must be an int, got {type(n).__name__}")
        self.sites[n] = _as_site(t)

    def __iter__(self):
        return iter(self.sites)

    def copy(self) -> "MPS":
        return MPS(self.sites, self.center)

    @classmethod
    def random(cls, phys: GradedSpace, bonds: Sequence[GradedSpace], *, seed: int = 0,
               dtype=None, device=None) -> "MPS":
        """Site ``i`` drawn with seed ``seed + i`` (reference ``mps.py:162-190``)."""
        return cls(
            SymmetricTensor.random(
                (Leg(bonds[i], OUT), Leg(phys, OUT), Leg(bonds[i + 1], IN)),
                seed=seed + i, dtype=dtype, device=device,
            )
            for i in range(len(bonds) - 1)
        )

    @classmethod
    def product(cls, phys: GradedSpace, states: Sequence, *, dtype=None, device=None) -> "MPS":
        """A norm-1 product state, site ``n`` in the basis vector of sector ``states[n]``
        (reference ``mps.py:193-290``).  The bonds follow from the charges, right to left:
        the last bond is the unit sector and bond ``n`` is bond ``n + 1`` fused with the
        dual of ``states[n]``, so bond 0 carries the total charge; its leg shows the total
        itself and is dual when that is not the unit.  Refuses a sector absent from
        ``phys`` or of degeneracy above 1, and a fusion with more than one channel."""
        prov = phys.provider
        states = list(states)
        for n, a in enumerate(states):
            if phys.degeneracy(a) != 1:
                raise ValueError(f"MPS.product: sector {a!r} at site {n} must appear in phys with "
                                 f"degeneracy 1, it has {phys.degeneracy(a)}")
        bonds = [prov.unit]
        for n in range(len(states) - 1, -1, -1):
            out = prov.fusion(bonds[0], prov.dual(states[n]))
            if len(out) != 1:
                raise ValueError(f"MPS.product: bond {n + 1} x dual({states[n]!r}) has {len(out)} "
                                 "fusion channels; a product state needs one")
            bonds.insert(0, out[0])
        total = prov.dual(bonds[0])
        first = Leg(GradedSpace.new(prov, {total: 1}), OUT, total != prov.unit)
        offsets, pos = {}, 0
        for a, m in phys.sectors:
            offsets[a] = pos
            pos += int(m * prov.qdim(a))
        sites = []
        for n, a in enumerate(states):
            arr = np.zeros((1, phys.dim, 1))
            arr[0, offsets[a], 0] = 1.0
            left = first if n == 0 else Leg(GradedSpace.new(prov, {bonds[n]: 1}), OUT)
            legs = (left, Leg(phys, OUT), Leg(GradedSpace.new(prov, {bonds[n + 1]: 1}), IN))
            t = from_dense(arr, legs, device=device)
            sites.append(t if dtype is None else ops.astype(t, dtype))
        return cls(sites)

    def norm(self) -> float:
        """``sqrt(<psi|psi>)``: an LQ pass from the right moves the whole state's weight
        into site 0 through right isometries, whose norm is then the state's (the state
        itself is not changed)."""
        if not self.sites:
            return 0.0
        head = self.sites[-1]
        for n in range(len(self) - 1, 0, -1):
            left, _ = linalg.lq(head, ((0,), (1, 2)))
            head = ops.einsum("apx,xy->apy", self.sites[n - 1], left)
        return float(ops.norm(head))

    def canonize_(self, to: int = 0) -> "MPS":
        """Right-canonical form by an LQ sweep from the right (reference ``mps.py:312-352``)."""
        if to != 0:
            raise NotImplementedError("only to=0 (right-canonical) is implemented")
        for n in range(len(self) - 1, 0, -1):
            left, q = linalg.lq(self[n], ((0,), (1, 2)))
            self[n] = q
            self[n - 1] = ops.einsum("apx,xy->apy", self[n - 1], left)
        self[0] = ops.divide(self[0], ops.norm(self[0]))
        self.center = to
        return self


def overlap(bra: MPS, ket: MPS) -> float:
    """``<bra|ket>``, undivided by either norm (reference ``mps.py:714-760``): the two-state
    environment pass over ``MPO.identity``."""
    from tenet.network.env import Env

    if len(bra) != len(ket):
        raise ValueError(f"overlap: chains of {len(bra)} and {len(ket)} sites")
    ident = MPO.identity(len(ket), ket[0].legs[1].space, device=ket[0].device)
    return Env(ket, ident, bra=bra).measure()


def from_dense(array, legs: Sequence[Leg], *, atol: float = 1e-10, device=None) -> SymmetricTensor:
    """Blocks of a dense array whose every axis runs over its leg's sectors in canonical
    order (``tenet.dense.from_dense``: Clebsch-Gordan projection for SU(2), block
    placement for abelian providers).  Raises when the array is not symmetric."""
    from tenet.dense import from_dense as _from_dense

    return _from_dense(array, legs, atol=atol, device=device)


class MPO:
    """A finite MPO of site tensors ``(wl IN, p OUT, p IN, wr OUT)``.

    ``edges`` is the channel table of an MPO built by ``from_terms`` (which open term
    occupies which graded channel at which cut) and ``None`` otherwise; ``materialize``
    returns the same site tensors without it (reference ``network/mps.py:2551-2690``).
    The two-site H_eff always runs on the site tensors."""

    def __init__(self, sites: Iterable[SymmetricTensor], *, edges=None) -> None:
        self.sites = list(sites)
        self.edges = edges

    def __len__(self) -> int:
        return len(self.sites)

    def __getitem__(self, n: int) -> SymmetricTensor:
        return self.sites[n]

    def __iter__(self):
        return iter(self.sites)

    def materialize(self) -> "MPO":
        """The same operator as plain site tensors, with the channel table dropped."""
        return MPO(list(self.sites))

    @classmethod
    def from_terms(cls, n_sites: int, terms, *, cutoff: float | None = 1e-13, device=None) -> "MPO":
        """The MPO of a sum of terms ``(coef, [(op, site), ...])`` -- rank-3 charge-leg
        factors from ``local_op`` on increasing sites, or one invariant two-site operator on
        ``(i, i + 1)`` (reference ``mps.py:3135-3342``; ``tenet.network.terms``).
        ``cutoff`` drops split channels whose weight is below it relative to the operator."""
        from tenet.network.terms import build_mpo

        sites, table = build_mpo(n_sites, list(terms), tol=1e-13 if cutoff is None else cutoff, device=device)
        return cls(sites, edges=table)

    @classmethod
    def identity(cls, n_sites: int, phys: GradedSpace, *, device=None) -> "MPO":
        """The identity operator: unit-sector ``D = 1`` bonds and ``eye`` on every site
        (reference ``mps.py:2731-2778``).  A two-state ``Env`` over it gives overlaps and
        the per-bond projection vectors of an excited-state sweep."""
        unit = GradedSpace.new(phys.provider, {phys.provider.unit: 1})
        legs = (Leg(unit, IN), Leg(phys, OUT), Leg(phys, IN), Leg(unit, OUT))
        eye = from_dense(np.eye(phys.dim).reshape(1, phys.dim, phys.dim, 1), legs, device=device)
        return cls([eye] * n_sites)

    @classmethod
    def from_w(cls, w, n_sites: int, *, phys: GradedSpace, bond: GradedSpace,
               boundary: GradedSpace, start: int, end: int, device=None) -> "MPO":
        """One bulk ``W[wl, p_out, p_in, wr]`` with boundary rows/columns cut at ``start``
        and ``end`` (reference ``mps.py:2780-2845``)."""
        w = np.asarray(w, dtype=np.float64)

        def make(arr, left, right):
            legs = (Leg(left, IN), Leg(phys, OUT), Leg(phys, IN), Leg(right, OUT))
            return from_dense(arr, legs, device=device)

        first = make(w[start : start + 1], boundary, bond)
        bulk = make(w, bond, bond)
        last = make(w[:, :, :, end : end + 1], bond, boundary)
        return cls([first, *[bulk] * (n_sites - 2), last])
