"""The pair schedule of the batched Jacobi SVD (``csrc/svd.hip``, ``jacobi_round``), restated
on the host: a sweep is nb - 1 launches of a circle-method tournament over nb blocks of w
vectors; round 0 rotates every pair of each block pair's 2w vectors, later rounds only the
w x w pairs across the two blocks.  The checks are the properties one-sided Jacobi needs:
every pair of the problem is visited exactly once per sweep, and the pairs of one inner
round are disjoint (their 16-lane groups rotate side by side)."""

import itertools

import pytest


def pair_of(n, r, p):
    """Circle method over n players: player 0 stays, the others turn one seat per round."""
    ia, ib = (p - 1 + r) % (n - 1), (n - 2 - p + r) % (n - 1)
    return (0 if p == 0 else ia + 1), ib + 1


def sweep_pairs(nb, w):
    """Inner rounds of one sweep, each a list of (i, j) problem-vector pairs."""
    rounds = []
    for rnd in range(nb - 1):
        for slot in range(nb // 2):
            bi, bj = pair_of(nb, rnd, slot)

            def vec(t):
                return bi * w + t if t < w else bj * w + (t - w)

            if rnd == 0 or nb == 2:
                inner = [[pair_of(2 * w, r, p) for p in range(w)] for r in range(2 * w - 1)]
            else:
                inner = [[(p, w + (p + r) % w) for p in range(w)] for r in range(w)]
            rounds += [[(vec(a), vec(b)) for a, b in rr] for rr in inner]
    return rounds


@pytest.mark.parametrize("nb,w", [(2, 1), (2, 8), (4, 1), (4, 16), (6, 4), (10, 8), (18, 16), (66, 8)])
def test_every_pair_once_per_sweep(nb, w):
    seen = {}
    for rr in sweep_pairs(nb, w):
        slots = [v for pair in rr for v in pair]
        assert len(slots) == len(set(slots)), "an inner round uses a vector twice"
        for a, b in rr:
            key = (min(a, b), max(a, b))
            assert a != b
            seen[key] = seen.get(key, 0) + 1
    n = nb * w
    assert set(seen) == set(itertools.combinations(range(n), 2))
    assert set(seen.values()) == {1}


def test_cross_rounds_halve_the_inner_rounds():
    # nb = 18 blocks of 16 (a 288-vector problem), summed over the 9 workgroups of a round:
    # round 0 runs 31 inner rounds, the 16 later rounds 16 each (17 x 31 if every round
    # re-ran the pairs inside its blocks)
    assert len(sweep_pairs(18, 16)) == 9 * (31 + 16 * 16)
