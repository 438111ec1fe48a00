"""SU(2) recoupling coefficients from the Racah formula, exact until the last rounding.

Labels are ``two_j`` integers.  The 6j symbol is evaluated with the Racah single-sum
formula (G. Racah, Phys. Rev. 62, 438 (1942); Edmonds, *Angular Momentum in Quantum
Mechanics*, eq. 6.3.7) in exact integer/rational arithmetic: the product of the four
triangle coefficients is a rational ``R`` and the alternating sum a rational ``S``, so
``{6j} = sign(S) * sqrt(R * S^2)`` and the only floating-point steps are one correctly
rounded ``Fraction -> float`` and one ``sqrt``.

Gauge (TensorKit ``SU2Irrep``, the gauge of the reference's racah backend):

* ``F^{abc}_d[e, f] = (-1)^{ja+jb+jc+jd} sqrt((2je+1)(2jf+1)) {ja jb je; jc jd jf}``
* ``R^{ab}_c = (-1)^{ja+jb-jc}``
* ``B(a, b, c) = sqrt(d_a d_b / d_c) F^{a b b}_a[c, 0]`` (reference ``_sun_coeff.py:90-100``)
* ``chi_a = sign F^{a a a}_a[0, 0] = (-1)^{2ja}`` (reference ``_sun_coeff.py:103-106``)

These replace the racah-py 0.1.1 Rust crate the reference calls
(``symmetry/_sun_coeff.py:45-117``); they are pinned against the reference's fixture
tables ``tests/fixtures/su2_f.txt`` / ``su2_r.txt`` by the test suite.
"""

from __future__ import annotations

from fractions import Fraction
from functools import cache
from math import factorial, sqrt

import numpy as np

__all__ = ["six_j", "f_symbol", "r_symbol", "b_symbol", "frobenius_schur", "cg", "z_matrix"]


def _tri(a: int, b: int, c: int) -> bool:
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def _delta_sq(a: int, b: int, c: int) -> Fraction:
    """Squared triangle coefficient for doubled spins (all sums even by ``_tri``)."""
    return Fraction(
        factorial((a + b - c) // 2) * factorial((a - b + c) // 2) * factorial((-a + b + c) // 2),
        factorial((a + b + c) // 2 + 1),
    )


@cache
def _six_j_exact(j1: int, j2: int, j3: int, j4: int, j5: int, j6: int) -> tuple[int, Fraction]:
    """``(sign, square)`` of ``{j1 j2 j3; j4 j5 j6}`` for doubled spins; ``(0, 0)`` if absent."""
    if not (_tri(j1, j2, j3) and _tri(j1, j5, j6) and _tri(j4, j2, j6) and _tri(j4, j5, j3)):
        return 0, Fraction(0)
    t1 = (j1 + j2 + j3) // 2
    t2 = (j1 + j5 + j6) // 2
    t3 = (j4 + j2 + j6) // 2
    t4 = (j4 + j5 + j3) // 2
    p1 = (j1 + j2 + j4 + j5) // 2
    p2 = (j2 + j3 + j5 + j6) // 2
    p3 = (j3 + j1 + j6 + j4) // 2
    total = Fraction(0)
    for z in range(max(t1, t2, t3, t4), min(p1, p2, p3) + 1):
        den = (
            factorial(z - t1)
            * factorial(z - t2)
            * factorial(z - t3)
            * factorial(z - t4)
            * factorial(p1 - z)
            * factorial(p2 - z)
            * factorial(p3 - z)
        )
        term = Fraction(factorial(z + 1), den)
        total += -term if z % 2 else term
    if total == 0:
        return 0, Fraction(0)
    radicand = _delta_sq(j1, j2, j3) * _delta_sq(j1, j5, j6) * _delta_sq(j4, j2, j6)
    radicand *= _delta_sq(j4, j5, j3)
    return (1 if total > 0 else -1), radicand * total * total


def six_j(j1: int, j2: int, j3: int, j4: int, j5: int, j6: int) -> float:
    """Wigner 6j symbol ``{j1 j2 j3; j4 j5 j6}`` for doubled spins."""
    s, sq = _six_j_exact(j1, j2, j3, j4, j5, j6)
    return s * sqrt(float(sq)) if s else 0.0


@cache
def f_symbol(a: int, b: int, c: int, d: int, e: int, f: int) -> float:
    """``[F^{abc}_d]_{e,f}`` for doubled spins; exactly ``0.0`` off the fusion rules."""
    if not (_tri(a, b, e) and _tri(e, c, d) and _tri(b, c, f) and _tri(a, f, d)):
        return 0.0
    s, sq = _six_j_exact(a, b, e, c, d, f)
    if not s:
        return 0.0
    phase = -1 if ((a + b + c + d) // 2) % 2 else 1
    return phase * s * sqrt(float(sq * (e + 1) * (f + 1)))


def r_symbol(a: int, b: int, c: int) -> float:
    """``R^{ab}_c = (-1)^{ja+jb-jc}`` (``0`` off the fusion rules)."""
    if not _tri(a, b, c):
        return 0.0
    return -1.0 if ((a + b - c) // 2) % 2 else 1.0


@cache
def b_symbol(a: int, b: int, c: int) -> float:
    """``B(a,b,c) = sqrt(d_a d_b / d_c) F^{a b b}_a[c, 0]``."""
    if not _tri(a, b, c):
        return 0.0
    return sqrt((a + 1) * (b + 1) / (c + 1)) * f_symbol(a, b, b, a, c, 0)


def frobenius_schur(a: int) -> float:
    """``chi_a = (-1)^{2ja}``: ``-1`` on half-integer spin."""
    return -1.0 if a % 2 else 1.0


def _ladder(two_j: int, two_m: int, up: bool) -> float:
    """``<j, m +- 1| J_+- |j, m>`` for doubled ``j`` and ``m`` (Condon-Shortley: positive)."""
    j, m = two_j / 2.0, two_m / 2.0
    return sqrt(max(j * (j + 1) - m * (m + 1 if up else m - 1), 0.0))


@cache
def _cg_cached(a: int, b: int, c: int) -> np.ndarray:
    """Highest-weight construction of the ``(d_a, d_b, d_c)`` Clebsch-Gordan tensor.

    The top state ``|c, c>`` is the normalised kernel of ``J_+ = J_+^a + J_+^b`` inside the
    ``m = c`` subspace (a two-term recursion over ``m_a``), with the Condon-Shortley phase
    ``<a, a; b, c-a | c, c> > 0``; every lower ``|c, m>`` follows by ``J_-``.  Magnetic
    indices run descending (index ``k`` carries ``m = j - k``), the dense basis of
    reference ``symmetry/su2.py:162-171``.
    """
    if not _tri(a, b, c):
        raise ValueError(f"SU2: {c} is not in {a} x {b}")
    da, db, dc = a + 1, b + 1, c + 1
    out = np.zeros((da, db, dc))
    # top state: amplitudes x[m_a] on |m_a, c - m_a> (doubled m), m_a descending from a
    mas = [ma for ma in range(a, -a - 1, -2) if abs(c - ma) <= b]
    amp = {mas[0]: 1.0}
    for hi, lo in zip(mas, mas[1:]):
        # J_+ kills the top state: coefficient of |hi, c - lo> must vanish
        amp[lo] = -amp[hi] * _ladder(b, c - hi, True) / _ladder(a, lo, True)
    nrm = sqrt(sum(v * v for v in amp.values()))
    state = np.zeros((da, db))
    for ma, v in amp.items():
        state[(a - ma) // 2, (b - (c - ma)) // 2] = v / nrm
    out[:, :, 0] = state
    for k in range(1, dc):
        m = c - 2 * (k - 1)  # lower |c, m> to |c, m - 1>
        nxt = np.zeros((da, db))
        for ia in range(da):
            for ib in range(db):
                v = state[ia, ib]
                if v == 0.0:
                    continue
                ma, mb = a - 2 * ia, b - 2 * ib
                if ia + 1 < da:
                    nxt[ia + 1, ib] += v * _ladder(a, ma, False)
                if ib + 1 < db:
                    nxt[ia, ib + 1] += v * _ladder(b, mb, False)
        state = nxt / _ladder(c, m, False)
        out[:, :, k] = state
    out.flags.writeable = False
    return out


def cg(a: int, b: int, c: int) -> np.ndarray:
    """``<a m_a; b m_b | c m_c>`` as a read-only ``(d_a, d_b, d_c)`` array (doubled spins)."""
    return _cg_cached(a, b, c)


@cache
def z_matrix(a: int) -> np.ndarray:
    """``Z_a: V_a -> V_a^*`` in the descending basis: ``Z[i, d-1-i] = (-1)^i``."""
    d = a + 1
    z = np.zeros((d, d))
    for i in range(d):
        z[i, d - 1 - i] = -1.0 if i % 2 else 1.0
    z.flags.writeable = False
    return z
