"""Environments and the two-site effective Hamiltonian on the device.

``Env`` is the directed-bond cache of reference ``network/env.py:506-770``:
``F[(n, n+1)]`` is ``(ket IN, mpo OUT, bra OUT)`` built from sites ``<= n`` and
``F[(n, n-1)]`` is ``(ket OUT, mpo IN, bra IN)`` built from sites ``>= n``.  ``clear_``
pops both directed bonds of a changed site, so a missed update is a ``KeyError``, never a
stale number.

``heff2`` is the site-tensor path ``_heff2_full`` (reference ``env.py:416-432``): right
environment, then ``W_{n+1}``, then ``W_n``, then the left environment, each wire that
turns around in the planar diagram bent explicitly (``_composed``, reference
``env.py:60-90``).  ``Env.operator(n, precision)`` compiles the same four contractions
once into a fused device program (``tenet.network.heff.Heff2``) for the Krylov loop.
"""

from __future__ import annotations

import torch

from tenet import ops
from tenet.space import IN, OUT, Leg
from tenet.structure import TensorStructure
from tenet.tensor import SymmetricTensor

__all__ = ["Env", "ones", "composed", "heff2_full", "measure_mpo"]


def ones(legs, *, dtype=torch.float64, device=None) -> SymmetricTensor:
    s = TensorStructure(tuple(legs))
    dev = device if device is not None else ("cuda" if torch.cuda.is_available() else "cpu")
    return SymmetricTensor(s, torch.ones(s.size, dtype=dtype, device=dev))


def composed(equation: str, a: SymmetricTensor, b: SymmetricTensor, bend: str = "") -> SymmetricTensor:
    """A two-operand einsum with the wires named in ``bend`` bent first on both operands."""
    if bend:
        lhs, out = equation.split("->")
        ta, tb = lhs.split(",")

        def bent(t, term):
            flip = set(bend)
            outs = tuple(i for i, c in enumerate(term) if (t.legs[i].side is OUT) != (c in flip))
            ins = tuple(i for i in range(len(term)) if i not in outs)
            return ops.repartition(t, outs, ins), "".join(term[i] for i in (*outs, *ins))

        a, ta = bent(a, ta)
        b, tb = bent(b, tb)
        equation = f"{ta},{tb}->{out}"
    return ops.einsum(equation, a, b)


def heff2_full(h, fl: SymmetricTensor, fr: SymmetricTensor, n: int, aa: SymmetricTensor) -> SymmetricTensor:
    """``<bra env| W_n W_{n+1} |aa, ket env>`` by the four contractions."""
    t = ops.einsum("apqr,rys->apqys", aa, fr)
    t = composed("apqys,mQqy->apQms", t, h[n + 1], bend="q")
    t = composed("apQms,xPpm->aPQxs", t, h[n], bend="p")
    return composed("aPQxs,axB->BPQs", t, fl, bend="a")


class Env:
    """``<bra|H|psi>`` partial contractions for one ``(psi, h)`` pair."""

    def __init__(self, psi, h, *, bra=None, compile=None) -> None:
        self.psi = psi
        self.bra = psi if bra is None else bra
        self.h = h
        self.compile = compile
        n = len(psi)
        kl, kr = psi[0].legs[0], psi[n - 1].legs[2]
        bl, br = self.bra[0].legs[0], self.bra[len(self.bra) - 1].legs[2]
        wl, wr = h[0].legs[0], h[n - 1].legs[3]
        dev = psi[0].device
        dt = psi[0].dtype
        self.F = {
            (-1, 0): ones((Leg(kl.space, IN, kl.dual), Leg(wl.space, OUT, wl.dual),
                           Leg(bl.space, OUT, bl.dual)), dtype=dt, device=dev),
            (n, n - 1): ones((Leg(kr.space, OUT, kr.dual), Leg(wr.space, IN, wr.dual),
                              Leg(br.space, IN, br.dual)), dtype=dt, device=dev),
        }
        self._ops: dict = {}

    def setup_(self, to: int = 0) -> "Env":
        if to != 0:
            raise NotImplementedError("only to=0 is implemented; canonize_ has the same note")
        for n in range(len(self.psi) - 1, 0, -1):
            self.update_(n, to="first")
        return self

    def update_(self, n: int, *, to: str) -> None:
        """Write one directed-bond entry from its neighbour (reference ``env.py:593-633``)."""
        a, bra = self.psi[n], ops.adjoint(self.bra[n])
        if to == "last":
            t = ops.einsum("axB,apr->xBpr", self.F[n - 1, n], a)
            t = ops.einsum("xPpm,xBpr->BrPm", self.h[n], t)
            self.F[n, n + 1] = ops.einsum("BPs,BrPm->rms", bra, t)
        else:
            t = ops.einsum("apr,rys->apys", a, self.F[n + 1, n])
            t = composed("apys,xPpy->axPs", t, self.h[n], bend="p")
            self.F[n, n - 1] = composed("axPs,BPs->axB", t, bra, bend="P")

    def clear_(self, *sites: int) -> None:
        for n in sites:
            self.F.pop((n, n - 1), None)
            self.F.pop((n, n + 1), None)
        # drop compiled operators whose environments are gone
        live = {id(t) for t in self.F.values()}
        self._ops = {k: v for k, v in self._ops.items() if id(v.fl) in live and id(v.fr) in live}

    def heff2(self, n: int, aa: SymmetricTensor) -> SymmetricTensor:
        """``H_eff @ aa`` on bond ``(n, n+1)``; the result has ``aa``'s structure."""
        return heff2_full(self.h, self.F[n - 1, n], self.F[n + 2, n + 1], n, aa)

    def project2(self, n: int, aa: SymmetricTensor) -> SymmetricTensor:
        """``<bra env| W_n W_{n+1} |aa, ket env>`` on a two-state ``Env`` (reference
        ``env.py:843-906``): ``aa`` is the ket chain's two-site tensor at bond ``(n, n+1)``
        and the result lives on the bra chain's two-site space.  Over ``MPO.identity`` it
        is the ket state read in the bra's environment gauge, the vector an excited-state
        sweep projects out of the bra's variational space."""
        return heff2_full(self.h, self.F[n - 1, n], self.F[n + 2, n + 1], n, aa)

    def measure(self) -> float:
        """``<bra|H|psi>``, not divided by either norm, by a left-to-right pass in a fresh
        ``Env`` (the sweep's cache is not touched; reference ``env.py:948-977``).  The pass
        ends on the right boundary, whose three legs carry one sector each: the closing
        contraction with that boundary is the qdim-weighted sum of its entries."""
        n = len(self.psi)
        env = Env(self.psi, self.h, bra=self.bra)
        for site in range(n):
            env.update_(site, to="last")
        last = env.F[n - 1, n]
        for leg in last.legs:
            if len(leg.space.sectors) != 1 or leg.space.dim != leg.space.provider.qdim(leg.space.sectors[0][0]):
                raise ValueError("measure() closes on boundary legs of one degeneracy-1 sector each")
        return float(ops.inner(ones(last.legs, dtype=last.dtype, device=last.device), last))

    def heff2_families(self, n: int, aa: SymmetricTensor) -> tuple:
        """``heff2``'s term families applied separately (reference ``env.py:772-841``).  The
        two-site H_eff here always runs on the MPO's site tensors, where the operator has
        no families to resolve: the single vector ``(heff2(n, aa),)``, as the reference's
        site-tensor path returns."""
        return (self.heff2(n, aa),)

    def operator(self, n: int, precision: str = "fp64"):
        """The compiled H_eff of bond ``(n, n+1)`` at ``precision`` (cached until the
        bond's environments are cleared)."""
        from tenet.network.heff import Heff2

        key = (n, precision)
        fl, fr = self.F[n - 1, n], self.F[n + 2, n + 1]
        op = self._ops.get(key)
        if op is None or op.fl is not fl or op.fr is not fr:
            op = Heff2(fl, self.h[n], self.h[n + 1], fr, precision)
            self._ops[key] = op
        return op


def measure_mpo(bra, h, ket) -> float:
    """``<bra|H|ket>``, undivided by either norm (reference ``env.py:996-1040``): the closing
    contraction of a fresh two-state ``Env`` pass."""
    return Env(ket, h, bra=bra).measure()
