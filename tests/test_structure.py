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


@pytest.mark.parametrize("sides,duals", [
    ((1, 1, 0), (0, 0, 0)), ((1, 0, 0), (0, 1, 0)), ((1, 1, 0, 0), (0, 0, 1, 0)),
    ((0, 1, 0, 1), (1, 1, 0, 0)), ((1, 1, 1, 0), (0, 1, 0, 1)),
])
def test_matrix_layout_matches_oracle_lowering(sides, duals):
    """Per coupled sector, the map layout's row and column bands (order, offsets, extents),
    the block at every (row band, column band) and the packed matrix shapes and offsets are
    those of the oracle's ``to_matrices`` (reference ``map_view.py:155-279``)."""
    from tenet import TensorStructure
    from tenet.linalg import _MatrixLayout
    from tenet.plans import map_layout

    ds = [D1, D2, D1, D2][: len(sides)]
    s = TensorStructure(tuple(leg(d, o, u) for d, o, u in zip(ds, sides, duals)))
    olegs = tuple(oleg(d, o, u) for d, o, u in zip(ds, sides, duals))
    outs, ins, rows, cols = O._layout(olegs)
    keys = O.block_keys(olegs)
    lay = map_layout(s)
    assert [c.charge for c in lay.sectors] == sorted(rows)
    ml = _MatrixLayout(s)
    base = 0
    for c, (cs, grid) in zip(sorted(rows), lay.grid):
        rb = [(o, e) for sc, _, o, e in lay.rows if sc == cs]
        cb = [(o, e) for sc, _, o, e in lay.cols if sc == cs]
        assert [e for _, e in rb] == list(rows[c].values())
        assert [e for _, e in cb] == list(cols[c].values())
        assert [o for o, _ in rb] == list(np.cumsum([0, *rows[c].values()])[:-1])
        assert [o for o, _ in cb] == list(np.cumsum([0, *cols[c].values()])[:-1])
        for ro, row in zip(rows[c], grid):
            for co, i in zip(cols[c], row):
                key = [0] * len(olegs)
                for ax, q in zip(outs, ro):
                    key[ax] = q
                for ax, q in zip(ins, co):
                    key[ax] = q
                assert i == keys.index(tuple(key))
        shape = (sum(rows[c].values()), sum(cols[c].values()))
        assert ml.shapes[cs] == shape and ml.offsets[cs] == base
        base += shape[0] * shape[1]
    assert ml.size == base
