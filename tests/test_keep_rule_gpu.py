"""``svd_truncated``/``select_bond``/``svd(bond=)`` on the device against the keep rule's
independent prefix-sum formulation over host LAPACK singular values.

U(1): the sector matrices come from the oracle (``oracle.u1.to_matrices`` of the same seeded
tensor).  SU(2): they are read back from the product's own lowering, and pinned first by
the dense oracle -- the singular values of the dense matrix are exactly the sector values,
each repeated ``2j + 1`` times.  Every decision (six cutoff modes, ``max_bond`` alone and
with a cutoff, budgets where a ``qdim = 3`` value stops the walk) must equal the host
selection; the factors must be isometries on the kept bond, reconstruct the tensor up to
exactly the discarded weight, and, for SU(2), reach the dense Eckart-Young error.
"""

import numpy as np
import pytest
import torch

from oracle import su2 as OS
from oracle import u1 as O
from tests.helpers import leg, to_oracle
from tests.test_keep_rule import prefix_selection

pytestmark = pytest.mark.gpu

CASES = [
    (None, 0.3, "abs"), (None, 0.2, "rel"), (None, 0.05, "sum2"), (None, 0.02, "rsum2"),
    (None, 0.5, "sum1"), (None, 0.05, "rsum1"), (7, None, "rsum2"), (12, None, "rsum2"),
    (23, None, "rsum2"), (16, 0.02, "rsum2"), (9, 0.2, "rel"),
]


def _entries(values):
    e = [(float(s), c, i) for c, arr in values.items() for i, s in enumerate(arr)]
    e.sort(key=lambda x: (-x[0], x[1], x[2]))
    return e


def _kept(bond):
    return {(c, i) for c, k in bond.sectors for i in range(k)}


def _u1_tensor():
    import tenet

    d = {-1: 4, 0: 6, 1: 5, 2: 2}
    legs = (leg(d, 1), leg(d, 1), leg(d, 1), leg(d, 0))
    return tenet.SymmetricTensor.random(legs, seed=21, device="cuda")


def _su2_tensor():
    import tenet
    from tenet import IN, OUT, GradedSpace, Leg
    from tenet.symmetry import SU2, SU2Sector

    V = GradedSpace.new(SU2, {SU2Sector(0): 3, SU2Sector(1): 2, SU2Sector(2): 2})
    W = GradedSpace.new(SU2, {SU2Sector(1): 2, SU2Sector(2): 1, SU2Sector(3): 1})
    legs = (Leg(V, OUT), Leg(W, OUT), Leg(V, IN), Leg(W, IN))
    return tenet.SymmetricTensor.random(legs, seed=22, device="cuda")


def _host_values(t, kind):
    """Per-sector singular values from LAPACK on host sector matrices."""
    from tenet import linalg
    from tenet.symmetry import U1Sector

    if kind == "u1":
        mats = O.to_matrices(O.repartition(to_oracle(t), (0, 1), (2, 3)))
        return {U1Sector(q): np.linalg.svd(b, compute_uv=False) for q, b in mats.items()}
    import tenet

    mats = {c: b.cpu().numpy() for c, b in linalg.to_matrices(tenet.repartition(t, (0, 1), (2, 3))).items()}
    vals = {c: np.linalg.svd(b, compute_uv=False) for c, b in mats.items() if b.size}
    # the dense oracle pins these matrices' spectra: dense sigma = sector sigma x (2j+1)
    rules = OS.SU2Rules
    ol = [OS.SLeg([(a.two_j, m) for a, m in l.space.sectors], l.side is tenet.OUT, l.dual) for l in t.legs]
    dense = OS.to_dense(rules, ol, [np.asarray(b, dtype=np.float64) for b in t.numpy_blocks()])
    d = dense.reshape(dense.shape[0] * dense.shape[1], -1)
    want = np.sort(np.linalg.svd(d, compute_uv=False))[::-1]
    got = np.sort(np.concatenate([np.repeat(v, c.two_j + 1) for c, v in vals.items()]))[::-1]
    assert np.abs(got - want[: len(got)]).max() <= 1e-12 * want[0]
    assert np.abs(want[len(got):]).max(initial=0.0) <= 1e-12 * want[0]
    return vals, d


@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_truncation_decisions_match_the_host_keep_rule(kind):
    import tenet
    from tenet import linalg

    t = _u1_tensor() if kind == "u1" else _su2_tensor()
    hv = _host_values(t, kind)
    values = hv if kind == "u1" else hv[0]
    entries = _entries(values)
    qdim = t.provider.qdim
    total = sum(qdim(c) * s * s for s, c, _ in entries)
    for max_bond, cutoff, mode in CASES:
        if mode in ("sum2", "sum1") and cutoff is not None:
            cutoff = cutoff * (total if mode == "sum2" else sum(qdim(c) * s for s, c, _ in entries))
        want = prefix_selection(entries, qdim, max_bond, cutoff, mode)
        u, s, vh = linalg.svd_truncated(t, ((0, 1), (2, 3)), max_bond=max_bond, cutoff=cutoff,
                                        cutoff_mode=mode)
        bond = u.legs[-1].space
        assert _kept(bond) == set(want), (max_bond, cutoff, mode)
        sel = linalg.select_bond(t, ((0, 1), (2, 3)), max_bond=max_bond, cutoff=cutoff, cutoff_mode=mode)
        assert sel.bond == bond
        assert [(c, i) for _, c, i in sel.kept] == want
        dropped = sum(qdim(c) * s_ * s_ for s_, c, i in entries if (c, i) not in set(want))
        assert sel.discarded_weight == pytest.approx(dropped, rel=1e-10, abs=1e-12 * total)
        # Pythagoras: |T|^2 - |U S Vh|^2 is the discarded weight
        rec = tenet.compose(tenet.compose(u, s), vh)
        err = float(tenet.norm(t)) ** 2 - float(tenet.norm(rec)) ** 2
        assert err == pytest.approx(dropped, abs=1e-10 * total)
        # the factors are isometries on the kept bond
        eye = tenet.compose(tenet.adjoint(u), u)
        for blk in eye.numpy_blocks():
            assert np.abs(blk - np.eye(blk.shape[0])).max() < 1e-12
        # svd(bond=) projects onto the same bond and gives the same factors
        u2, s2, vh2 = linalg.svd(t, ((0, 1), (2, 3)), bond=bond)
        assert u2.legs == u.legs and vh2.legs == vh.legs
        assert torch.allclose(s2.data, s.data, rtol=0, atol=1e-13)
        assert torch.allclose(tenet.compose(tenet.compose(u2, s2), vh2).data, rec.data, rtol=0, atol=1e-12)
        if kind == "su2":
            dense_sigma = np.sort(np.linalg.svd(hv[1], compute_uv=False))[::-1]
            assert dropped == pytest.approx(float(np.sum(dense_sigma[int(bond.dim):] ** 2)),
                                            abs=1e-10 * total)


def test_su2_qdim_3_value_stops_the_walk():
    """A budget at which the next value is a spin-1 multiplet (3 dense dimensions) that does
    not fit while a later spin-0 value would: the device decision stops there."""
    from tenet import linalg

    t = _su2_tensor()
    entries = _entries(_host_values(t, "su2")[0])
    qdim = t.provider.qdim
    found = False
    for max_bond in range(2, 60):
        bond = linalg.svd_truncated(t, ((0, 1), (2, 3)), max_bond=max_bond)[0].legs[-1].space
        kept = _kept(bond)
        rest = [e for e in entries if (e[1], e[2]) not in kept]
        if not rest or qdim(rest[0][1]) != 3:
            continue
        room = max_bond - bond.dim
        cheap = [e for e in rest[1:] if qdim(e[1]) <= room]
        if room >= 1 and cheap:
            assert (cheap[0][1], cheap[0][2]) not in kept
            assert bond.dim + 3 > max_bond
            found = True
            break
    assert found, "no budget exercised the overflow stop"


def test_renorm_rescales_only_s():
    import tenet
    from tenet import linalg

    t = _su2_tensor()
    plain = linalg.svd_truncated(t, ((0, 1), (2, 3)), max_bond=9)
    scaled = linalg.svd_truncated(t, ((0, 1), (2, 3)), max_bond=9, renorm=True)
    assert plain[0].legs == scaled[0].legs
    assert torch.equal(plain[0].data, scaled[0].data) and torch.equal(plain[2].data, scaled[2].data)
    rec = tenet.compose(tenet.compose(*scaled[:2]), scaled[2])
    assert float(tenet.norm(rec)) == pytest.approx(float(tenet.norm(t)), rel=1e-12)
