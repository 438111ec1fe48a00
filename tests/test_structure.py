"""Sector matching and storage contract (CPU): the product's block order, shapes and flat
offsets are bit-identical to the oracle's enumeration; pattern caching; index maps."""

import itertools

import numpy as np
import pytest

from oracle import u1 as O
from tests.helpers import leg, oleg

D1 = {-2: 1, -1: 2, 0: 3, 1: 2}
D2 = {-1: 3, 1: 1, 3: 2}


@pytest.mark.parametrize("sides,duals", [
    ((1, 1, 0), (0, 0, 0)), ((1, 0, 0), (0, 1, 0)), ((0, 0, 0), (1, 0, 1)),
    ((1, 1, 1), (0, 1, 0)), ((1, 1, 0, 0), (0, 0, 1, 0)), ((0, 1, 0, 1), (1, 1, 0, 0)),
])
def test_block_order_matches_oracle(sides, duals):
    from tenet import TensorStructure

    ds = [D1, D2, D1, D2][: len(sides)]
    s = TensorStructure(tuple(leg(d, o, u) for d, o, u in zip(ds, sides, duals)))
    olegs = tuple(oleg(d, o, u) for d, o, u in zip(ds, sides, duals))
    keys = O.block_keys(olegs)
    got = [tuple(a.charge for a in s.axis_sectors(k)) for k in s.block_order]
    assert got == keys
    assert [tuple(x) for x in s.shapes] == [O.block_shape(olegs, k) for k in keys]
    sizes = [int(np.prod(O.block_shape(olegs, k))) for k in keys]
    assert list(s.layout.offsets) == list(np.concatenate([[0], np.cumsum(sizes)[:-1]]))
    assert all(s.index_of(k) == i for i, k in enumerate(s.block_order))


def test_pattern_shares_block_order_across_degeneracies():
    from tenet import TensorStructure
    from tenet.structure import pattern_of

    a = TensorStructure((leg(D1, 1), leg(D2, 0)))
    b = TensorStructure((leg({q: m + 3 for q, m in D1.items()}, 1), leg(D2, 0)))
    assert a.block_order == b.block_order
    assert pattern_of(a) == pattern_of(b)
    assert a.shapes != b.shapes


def test_su2_block_order_counts_trees():
    from tenet import GradedSpace, IN, OUT, Leg, TensorStructure
    from tenet.symmetry import SU2, SU2Sector

    half = GradedSpace.new(SU2, {SU2Sector(1): 1})
    s = TensorStructure((Leg(half, OUT), Leg(half, OUT), Leg(half, OUT), Leg(half, IN)))
    # (1/2)^3 -> 1/2 has two trees (inner 0 or 1); 3/2 cannot meet a single 1/2 input
    assert s.num_blocks == 2
