"""The truncation keep rule, pinned without a GPU.

Two implementations are checked: the oracle's restatement (``oracle.u1_network.keep_rule``)
and the product's host decision (``tenet.linalg._decide``, which ``svd_truncated`` and
``select_bond`` run on the device's singular values).  Both are held to

* an independent formulation written here with prefix sums (admitted prefix = the
  shortest prefix whose discarded tail weighs less than the threshold; kept prefix = the
  longest admitted prefix whose cumulative ``qdim`` fits the budget), and
* the reference's property checks (``tests/ops/test_svd_truncated.py:139-355``): bare
  cutoffs keep exactly the values above the threshold, cumulative cutoffs are admissible
  and maximal, relative modes are scale-invariant, ``max_bond`` with a cutoff keeps the
  intersection, the greedy walk stops at the first ``qdim`` overflow (a ``qdim = 3`` value
  blocks a cheaper later one), the dense budget undershoots by less than the largest
  ``qdim``, kept sets nest across budgets, and the discarded weight equals the dense
  Eckart-Young error.

Spectra come from ``numpy.linalg.svd`` of random sector matrices labelled with U(1)
charges (``qdim = 1``) or SU(2) spins (``qdim = 2j + 1``).  Planted mistakes (``>=`` for
``>``, the wrong power, an unweighted sum, a walk that scans past an overflow) are each
caught by these checks.
"""

import numpy as np
import pytest

from oracle import u1_network as N

MODES = ("abs", "rel", "sum2", "rsum2", "sum1", "rsum1")


# --- spectra ------------------------------------------------------------------------


def _su2():
    from tenet.symmetry import SU2, SU2Sector

    return SU2, SU2Sector


def _u1():
    from tenet.symmetry import U1, U1Sector

    return U1, U1Sector


def spectrum(kind: str, seed: int):
    """``(provider, [(sigma, c, i)] sorted by (-sigma, c, i))`` from random sector matrices."""
    rng = np.random.default_rng(seed)
    if kind == "su2":
        provider, mk = _su2()
        shapes = {0: (9, 7), 1: (6, 8), 2: (7, 7), 3: (4, 6), 4: (3, 3)}
    else:
        provider, mk = _u1()
        shapes = {-2: (5, 3), -1: (8, 6), 0: (9, 9), 1: (7, 8), 2: (4, 4)}
    values = {}
    for lab, (m, n) in shapes.items():
        scale = np.exp(-0.35 * abs(lab))
        values[mk(lab)] = np.linalg.svd(scale * rng.standard_normal((m, n)), compute_uv=False)
    entries = [(float(s), c, i) for c, arr in values.items() for i, s in enumerate(arr)]
    entries.sort(key=lambda e: (-e[0], e[1], e[2]))
    return provider, entries


# --- the two implementations under test, as kept (c, i) sets ------------------------


def oracle_kept(entries, provider, max_bond=None, cutoff=None, mode="rsum2"):
    kept, _, _ = N.keep_rule(entries, max_bond, cutoff, mode, provider.qdim)
    return [(c, i) for _, c, i in kept]


def product_kept(entries, provider, max_bond=None, cutoff=None, mode="rsum2"):
    from tenet.linalg import _decide

    sel = _decide(entries, provider, max_bond, cutoff, mode, False, "test")
    return [(c, i) for _, c, i in sel.kept]


IMPLS = {"oracle": oracle_kept, "product": product_kept}


# --- an independent formulation with prefix sums -------------------------------------


def prefix_selection(entries, qdim, max_bond=None, cutoff=None, mode="rsum2"):
    sig = np.array([s for s, _, _ in entries])
    q = np.array([qdim(c) for _, c, _ in entries], dtype=np.float64)
    n = len(entries)
    if cutoff is None:
        adm = n
    elif mode == "abs":
        adm = int(np.count_nonzero(sig > cutoff))
    elif mode == "rel":
        adm = int(np.count_nonzero(sig > cutoff * sig[0]))
    else:
        w = q * sig ** (2 if mode.endswith("2") else 1)
        thr = cutoff * w.sum() if mode.startswith("r") else cutoff
        # tail[k] = weight discarded when the first k entries are kept
        tail = np.concatenate([np.cumsum(w[::-1])[::-1], [0.0]])
        ok = np.nonzero(tail < thr)[0]
        adm = int(ok[0]) if len(ok) else n
    keep = adm
    if max_bond is not None:
        keep = int(np.searchsorted(np.cumsum(q[:adm]), max_bond, side="right"))
    return [(c, i) for _, c, i in entries[:keep]]


# --- property checks -----------------------------------------------------------------


def check_bare(select, entries, provider, mode):
    if mode == "rel":  # largest value exactly 1, so the relative threshold is exact below
        entries = [(s / entries[0][0], c, i) for s, c, i in entries]
    cutoff = 0.4 if mode == "rel" else 0.4 * entries[0][0]
    thr = cutoff * entries[0][0] if mode == "rel" else cutoff
    kept = set(select(entries, provider, cutoff=cutoff, mode=mode))
    assert kept == {(c, i) for s, c, i in entries if s > thr}
    assert 0 < len(kept) < len(entries)
    # a value exactly at the threshold is discarded (strict >)
    at = entries[len(kept)][0]
    cut = at / entries[0][0] if mode == "rel" else at
    assert (entries[len(kept)][1], entries[len(kept)][2]) not in set(
        select(entries, provider, cutoff=cut, mode=mode))


def check_cumulative(select, entries, provider, mode):
    qdim = provider.qdim
    power = 2 if mode.endswith("2") else 1
    weights = {(c, i): qdim(c) * s**power for s, c, i in entries}
    total = sum(weights.values())
    cutoff, thr = (0.05, 0.05 * total) if mode.startswith("r") else (0.05 * total, 0.05 * total)
    kept = set(select(entries, provider, cutoff=cutoff, mode=mode))
    dropped = sum(w for k, w in weights.items() if k not in kept)
    assert dropped < thr  # admissible
    smallest = min((s, c, i) for s, c, i in entries if (c, i) in kept)
    assert dropped + weights[(smallest[1], smallest[2])] >= thr  # maximal
    assert 0 < len(kept) < len(entries)


def check_scale_invariant(select, entries, provider, mode, cutoff):
    one = select(entries, provider, cutoff=cutoff, mode=mode)
    two = select([(2.0 * s, c, i) for s, c, i in entries], provider, cutoff=cutoff, mode=mode)
    assert one == two
    assert len(one) < len(entries)


def check_intersection(select, entries, provider, max_bond, cutoff):
    both = set(select(entries, provider, max_bond=max_bond, cutoff=cutoff, mode="rsum2"))
    by_bond = set(select(entries, provider, max_bond=max_bond))
    by_cut = set(select(entries, provider, cutoff=cutoff, mode="rsum2"))
    assert both == by_bond & by_cut


def check_budget_walk(select, entries, provider):
    """Dense budget, stop on the first overflow, nesting; returns the budget at which a
    qdim-3 value blocked a cheaper qdim-1 value that would have fitted."""
    qdim = provider.qdim
    largest = max(qdim(c) for _, c, _ in entries)
    total_dense = sum(qdim(c) for _, c, _ in entries)
    blocked_at, prev = None, set()
    for max_bond in range(1, int(total_dense) + 2):
        kept = select(entries, provider, max_bond=max_bond)
        dense = sum(qdim(c) for c, _ in kept)
        assert dense <= max_bond
        assert kept == [(c, i) for _, c, i in entries[: len(kept)]]  # a prefix
        assert set(kept) >= prev  # nesting
        prev = set(kept)
        if len(kept) < len(entries):
            blocker = entries[len(kept)]
            assert dense + qdim(blocker[1]) > max_bond  # stopped at an overflow
            assert max_bond - dense < largest
            room = max_bond - dense
            later = [e for e in entries[len(kept) + 1:] if qdim(e[1]) <= room]
            if qdim(blocker[1]) == 3 and later and blocked_at is None:
                assert (later[0][1], later[0][2]) not in set(kept)
                assert later[0][0] <= blocker[0]
                blocked_at = max_bond
    return blocked_at


def check_eckart_young(select, entries, provider):
    """The SU(2) dense matrix of a block-diagonal map has each sector value with
    multiplicity qdim; keeping a bare-descending prefix of dense dimension k discards
    exactly the best rank-k error of the dense matrix."""
    qdim = provider.qdim
    dense = np.sort(np.concatenate([[s] * int(qdim(c)) for s, c, _ in entries]))[::-1]
    for max_bond in (5, 11, 23, 40):
        kept = set(select(entries, provider, max_bond=max_bond))
        k = int(sum(qdim(c) for c, _ in kept))
        dropped = sum(qdim(c) * s * s for s, c, i in entries if (c, i) not in kept)
        assert dropped == pytest.approx(float(np.sum(dense[k:] ** 2)), rel=1e-12, abs=1e-14)


# --- tests ---------------------------------------------------------------------------


@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_matches_prefix_formulation(impl, kind):
    provider, entries = spectrum(kind, 1)
    select = IMPLS[impl]
    total2 = sum(provider.qdim(c) * s * s for s, c, _ in entries)
    cases = [(None, c, m) for m in MODES for c in (1e-3, 0.05, 0.3)]
    cases[-1:] = [(None, 0.3, "rsum1")]
    cases += [(b, None, "rsum2") for b in (1, 3, 7, 13, 29, 61)]
    cases += [(b, c, m) for b in (6, 20) for c, m in ((0.02, "rsum2"), (0.1 * total2, "sum2"), (0.3, "rel"))]
    cases += [(None, 0.0, "sum2"), (None, 0.0, "abs")]
    for max_bond, cutoff, mode in cases:
        got = select(entries, provider, max_bond=max_bond, cutoff=cutoff, mode=mode)
        want = prefix_selection(entries, provider.qdim, max_bond, cutoff, mode)
        assert got == want, (max_bond, cutoff, mode)


@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("mode", ["abs", "rel"])
def test_bare_cutoff_keeps_exactly_the_values_above_the_threshold(impl, mode):
    provider, entries = spectrum("su2", 6)
    check_bare(IMPLS[impl], entries, provider, mode)


@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("kind", ["u1", "su2"])
@pytest.mark.parametrize("mode", ["sum2", "rsum2", "sum1", "rsum1"])
def test_cumulative_cutoff_is_admissible_and_maximal(impl, kind, mode):
    provider, entries = spectrum(kind, 6)
    check_cumulative(IMPLS[impl], entries, provider, mode)


@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize(("mode", "cutoff"), [("rsum2", 0.02), ("rel", 0.4), ("rsum1", 0.05)])
def test_relative_modes_are_scale_invariant(impl, mode, cutoff):
    provider, entries = spectrum("su2", 6)
    check_scale_invariant(IMPLS[impl], entries, provider, mode, cutoff)


@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("max_bond", [4, 9, 16, 40])
@pytest.mark.parametrize("cutoff", [1e-8, 0.02, 0.2])
def test_max_bond_and_cutoff_take_the_intersection(impl, max_bond, cutoff):
    provider, entries = spectrum("su2", 6)
    check_intersection(IMPLS[impl], entries, provider, max_bond, cutoff)


@pytest.mark.parametrize("impl", IMPLS)
def test_stop_on_overflow_with_qdim_3(impl):
    provider, entries = spectrum("su2", 3)
    assert check_budget_walk(IMPLS[impl], entries, provider) is not None, \
        "no budget on this spectrum exercises the overflow stop"


@pytest.mark.parametrize("impl", IMPLS)
def test_u1_budget_is_a_count(impl):
    provider, entries = spectrum("u1", 3)
    for max_bond in (1, 5, 17):
        assert len(IMPLS[impl](entries, provider, max_bond=max_bond)) == max_bond
    check_budget_walk(IMPLS[impl], entries, provider)


@pytest.mark.parametrize("impl", IMPLS)
def test_discarded_weight_is_the_dense_eckart_young_error(impl):
    provider, entries = spectrum("su2", 4)
    check_eckart_young(IMPLS[impl], entries, provider)


def test_oracle_discarded_weight_and_dense_dimension():
    provider, entries = spectrum("su2", 4)
    kept, dense, dw = N.keep_rule(entries, 17, None, "rsum2", provider.qdim)
    ks = {(c, i) for _, c, i in kept}
    assert dense == sum(provider.qdim(c) for _, c, _ in kept) <= 17
    assert dw == pytest.approx(sum(provider.qdim(c) * s * s for s, c, i in entries if (c, i) not in ks))


def test_product_selection_record():
    from tenet.linalg import _decide

    provider, entries = spectrum("su2", 4)
    qdim = provider.qdim
    sel = _decide(entries, provider, 17, None, "rsum2", True, "test")
    kept = {(c, i) for _, c, i in sel.kept}
    total = sum(qdim(c) * s * s for s, c, _ in entries)
    kw = sum(qdim(c) * s * s for s, c, i in entries if (c, i) in kept)
    assert sel.dense_dim == sum(qdim(c) for c, _ in kept) <= 17
    assert sel.reduced_dim == len(kept) < sel.dense_dim
    assert sel.discarded_weight == pytest.approx(total - kw, rel=1e-12)
    assert sel.next_dense_cost == qdim(sel.discarded[0][1])
    assert sel.undershoot == 17 - sel.dense_dim
    assert sel.scale == pytest.approx((total / kw) ** 0.5, rel=1e-14)
    assert dict(sel.bond.sectors) == {c: sum(1 for cc, _ in kept if cc == c) for c in {c for c, _ in kept}}
    assert _decide(entries, provider, 17, None, "rsum2", False, "test").scale == 1.0


def test_product_refusals():
    from tenet.linalg import _decide, _validate

    provider, entries = spectrum("u1", 2)
    with pytest.raises(ValueError, match="keeps no singular value"):
        _decide(entries, provider, None, 10 * entries[0][0], "abs", False, "test")
    with pytest.raises(ValueError, match="no singular values"):
        _decide([], provider, 3, None, "rsum2", False, "test")
    with pytest.raises(ValueError, match="at least one"):
        _validate(None, None, "rsum2", False, "test")
    with pytest.raises(ValueError, match="positive"):
        _validate(0, None, "rsum2", False, "test")
    with pytest.raises(ValueError, match="non-negative"):
        _validate(None, -1.0, "rsum2", False, "test")
    with pytest.raises(ValueError, match="cutoff_mode"):
        _validate(4, None, "sum3", False, "test")
    with pytest.raises(TypeError, match="bool"):
        _validate(4, None, "rsum2", 2, "test")


# --- the checks are discriminating: planted mistakes in copies of the oracle's rule ----


def _mutant(kind):
    def select(entries, provider, max_bond=None, cutoff=None, mode="rsum2"):
        qdim = provider.qdim if kind != "unweighted" else (lambda c: 1.0)
        if cutoff is None:
            adm = len(entries)
        elif mode in ("abs", "rel"):
            thr = cutoff * entries[0][0] if mode == "rel" else cutoff
            adm = sum(1 for s, _, _ in entries if (s >= thr if kind == "ge" else s > thr))
        else:
            p = 2 if mode.endswith("2") else 1
            if kind == "power":
                p = 2
            w = [qdim(c) * s**p for s, c, _ in entries]
            thr = cutoff * sum(w) if mode.startswith("r") else cutoff
            adm, dropped = len(entries), 0.0
            for x in reversed(w):
                if dropped + x >= thr:
                    break
                dropped += x
                adm -= 1
        kept, dense = [], 0.0
        for s, c, i in entries[:adm]:
            q = provider.qdim(c)
            if max_bond is not None and dense + q > max_bond:
                if kind == "scan":
                    continue
                break
            dense += q
            kept.append((c, i))
        return kept

    return select


@pytest.mark.parametrize("kind", ["ge", "power", "unweighted", "scan"])
def test_planted_mistakes_are_caught(kind):
    select = _mutant(kind)
    caught = []
    provider, entries = spectrum("su2", 6)
    for name, fn in (
        ("bare", lambda: [check_bare(select, entries, provider, m) for m in ("abs", "rel")]),
        ("cumulative", lambda: [check_cumulative(select, entries, provider, m) for m in ("sum1", "rsum1")]),
        ("walk", lambda: check_budget_walk(select, *reversed(spectrum("su2", 3)))),
        ("prefix", lambda: [assert_prefix(select, provider, entries, m) for m in MODES]),
    ):
        try:
            fn()
        except AssertionError:
            caught.append(name)
    assert caught, f"planted mistake {kind!r} passed every check"


def assert_prefix(select, provider, entries, mode):
    for cutoff in (0.02, 0.3):
        for max_bond in (None, 9):
            got = select(entries, provider, max_bond=max_bond, cutoff=cutoff, mode=mode)
            assert got == prefix_selection(entries, provider.qdim, max_bond, cutoff, mode)
