"""SU(2) coefficients (CPU), pinned against the reference's fixture tables (every row of
su2_f.txt / su2_r.txt, generated by TensorKitSectors), an independent exact 6j routine
(sympy), and the pentagon/unitarity identities."""

import itertools
from pathlib import Path

import numpy as np
import pytest

from tenet import _su2coeff as C

FIX = Path(__file__).resolve().parent / "golden"


def _rows(name):
    path = FIX / name
    if not path.exists():
        pytest.skip(f"{name} fixture missing")
    return np.loadtxt(path, comments="#")


def test_f_symbols_match_reference_fixture_table():
    rows = _rows("su2_f.txt")
    assert len(rows) == 109900
    worst = max(abs(C.f_symbol(*map(int, r[:6])) - r[6]) for r in rows)
    assert worst < 1e-13


def test_r_symbols_match_reference_fixture_table():
    rows = _rows("su2_r.txt")
    for a, b, c, v in rows:
        assert C.r_symbol(int(a), int(b), int(c)) == v


def test_six_j_against_sympy_exact():
    from sympy import Rational
    from sympy.physics.wigner import wigner_6j

    rng = np.random.default_rng(1)
    done = 0
    while done < 40:
        js = tuple(int(x) for x in rng.integers(0, 15, 6))
        j1, j2, j3, j4, j5, j6 = js
        if not (C._tri(j1, j2, j3) and C._tri(j1, j5, j6) and C._tri(j4, j2, j6) and C._tri(j4, j5, j3)):
            continue
        ref = float(wigner_6j(*(Rational(j, 2) for j in js)))
        assert abs(C.six_j(*js) - ref) < 1e-14
        done += 1


def test_f_move_unitarity_and_pentagon_beyond_fixture_range():
    # unitarity: sum_f F[e,f] F[e',f] = delta(e,e') at spins past the fixture's dj <= 12
    a, b, c, d = 13, 12, 13, 14
    es = [e for e in range(0, 30) if C._tri(a, b, e) and C._tri(e, c, d)]
    fs = [f for f in range(0, 30) if C._tri(b, c, f) and C._tri(a, f, d)]
    F = np.array([[C.f_symbol(a, b, c, d, e, f) for f in fs] for e in es])
    assert np.allclose(F @ F.T, np.eye(len(es)), atol=1e-13)
    # pentagon on a sample: F^{fcd}_e[g,l] F^{abl}_e[f,k] = sum_h F^{abc}_g[f,h] F^{ahd}_e[g,k] F^{bcd}_k[h,l]
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(400):
        a, b, c, d = (int(x) for x in rng.integers(0, 5, 4))
        for e, f, g, k, l in itertools.product(range(0, 9), repeat=5):
            if not (C._tri(a, b, f) and C._tri(f, c, g) and C._tri(g, d, e) and C._tri(c, d, l)
                    and C._tri(b, l, k) and C._tri(a, k, e)):
                continue
            lhs = C.f_symbol(f, c, d, e, g, l) * C.f_symbol(a, b, l, e, f, k)
            rhs = sum(C.f_symbol(a, b, c, g, f, h) * C.f_symbol(a, h, d, e, g, k) * C.f_symbol(b, c, d, k, h, l)
                      for h in range(0, 9))
            assert abs(lhs - rhs) < 1e-12
            checked += 1
        if checked > 300:
            break
    assert checked > 50


def test_six_j_against_sympy_exact_high_spins():
    """Spins above the reference fixture's 2j <= 12 (the SU(2) workload's odd bonds reach
    2j = 13 and their inner lines more): sympy's exact 6j symbols."""
    from sympy import Rational
    from sympy.physics.wigner import wigner_6j

    rng = np.random.default_rng(7)
    done = 0
    while done < 25:
        js = tuple(int(x) for x in rng.integers(12, 27, 6))
        j1, j2, j3, j4, j5, j6 = js
        if not (C._tri(j1, j2, j3) and C._tri(j1, j5, j6) and C._tri(j4, j2, j6) and C._tri(j4, j5, j3)):
            continue
        ref = float(wigner_6j(*(Rational(j, 2) for j in js)))
        assert abs(C.six_j(*js) - ref) < 1e-13
        done += 1
