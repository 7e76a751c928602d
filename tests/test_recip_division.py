"""The FP64 field-map lookup divides by a bin width with a reciprocal and
two fused multiply-add corrections (divRecip in acts_amd/csrc/kf_common.hpp)
instead of a division.  The claim is that this
returns exactly the correctly rounded quotient, so that lookups, and with
them every step of the fit, are bit-identical to the division the reference
does (InterpolatedBFieldMap / Axis::getBin, Grid.hpp:470-494).

Checked here with exact rational arithmetic (every operation rounded once,
to nearest, from its exact value, as IEEE 754 and the hardware's FMA do):
the same operation sequence equals the correctly rounded quotient on the
field-map axes of the workloads (bin edges, points next to them, random
points) and on random operands, in binary64 (and, as the theorem is
independent of the precision, in binary32); and the plain reciprocal
product, or the sequence with the residual's sign flipped, is caught.
"""
from fractions import Fraction

import numpy as np
import pytest


def rn64(x: Fraction) -> float:
    return float(x)  # int / int true division: correctly rounded


def rn32(x: Fraction) -> float:
    """x rounded to the nearest binary32 value (ties to even), as a float."""
    if x == 0:
        return 0.0
    s = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    ulp = Fraction(2) ** (max(e, -126) - 23)
    n = x / ulp
    q, r = divmod(n.numerator, n.denominator)
    rem = Fraction(r, n.denominator)
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and q % 2 == 1):
        q += 1
    return s * float(q * ulp)


def fma(a, b, c, rn):
    return rn(Fraction(a) * Fraction(b) + Fraction(c))


def div_recip(a, b, y, rn, corrections=2, sign=1):
    q = rn(Fraction(a) * Fraction(y))
    for _ in range(corrections):
        q = fma(fma(-q, b, a, rn) * sign, y, q, rn)
    return q


def quotient(a, b, rn):
    return rn(Fraction(a) / Fraction(b))


# (min, point spacing, points) per axis of the field maps the workloads use,
# the width as the library computes it: (max + step - min) / n
AXES = [(-2000.0, 20.0, 201), (-4000.0, 20.0, 401), (-1234.5, 7.3, 57)]


def _axis_width(lo, step, n):
    hi = lo + (n - 1) * step
    return ((hi + (hi - lo) / (n - 1)) - lo) / n


def _axis_cases(rng, lo, w, n, count):
    xs = []
    for b in rng.integers(0, n, size=count):
        edge = lo + float(b) * w
        xs += [edge, np.nextafter(edge, np.inf), np.nextafter(edge, -np.inf),
               edge + rng.random() * w]
    return [x for x in xs if x >= lo]


@pytest.mark.parametrize("lo,step,n", AXES)
def test_binary64_axis_divisions_are_exact(lo, step, n):
    rng = np.random.default_rng(5)
    w = _axis_width(lo, step, n)
    yw = rn64(1 / Fraction(w))
    for x in _axis_cases(rng, lo, w, n, 150):
        a = x - lo
        assert div_recip(a, w, yw, rn64) == quotient(a, w, rn64), (x, w)
        b = min(max(int(np.floor(a / w)) + 1, 1), n)
        ll = lo + float(b - 1) * w
        ur = lo + float(b) * w
        den = ur - ll
        yd = rn64(1 / Fraction(den))
        assert div_recip(x - ll, den, yd, rn64) == quotient(x - ll, den, rn64), (x, den)


@pytest.mark.parametrize("lo,step,n", AXES)
def test_binary32_axis_divisions_are_exact(lo, step, n):
    rng = np.random.default_rng(6)
    w = rn32(Fraction(_axis_width(lo, step, n)))
    lo32 = rn32(Fraction(lo))
    yw = rn32(1 / Fraction(w))
    for x in _axis_cases(rng, lo, w, n, 100):
        x = rn32(Fraction(x))
        if x < lo32:
            continue
        a = rn32(Fraction(x) - Fraction(lo32))
        assert div_recip(a, w, yw, rn32) == quotient(a, w, rn32), (x, w)


def test_random_operands_are_exact():
    rng = np.random.default_rng(7)
    for _ in range(1500):
        a = float(rng.uniform(-1e4, 1e4) * 2.0 ** rng.integers(-20, 20))
        b = float(rng.uniform(0.5, 1.0) * 2.0 ** rng.integers(-30, 30))
        y = rn64(1 / Fraction(b))
        assert div_recip(a, b, y, rn64) == quotient(a, b, rn64), (a, b)
        a32, b32 = rn32(Fraction(a)), rn32(Fraction(b))
        y32 = rn32(1 / Fraction(b32))
        assert div_recip(a32, b32, y32, rn32) == quotient(a32, b32, rn32), (a32, b32)


def test_planted_mistakes_are_caught():
    """A plain reciprocal product, one correction only, or a residual with
    the wrong sign gives quotients that differ from the division."""
    rng = np.random.default_rng(8)
    lo, w = -2000.0, 20.0
    yw = rn64(1 / Fraction(w))
    wrong = {"product": 0, "one_correction": 0, "flipped": 0}
    xs = [lo + rng.random() * 4000.0 for _ in range(3000)]
    for x in xs:
        a = x - lo
        exact = quotient(a, w, rn64)
        wrong["product"] += rn64(Fraction(a) * Fraction(yw)) != exact
        wrong["one_correction"] += div_recip(a, w, yw, rn64, corrections=1) != exact
        wrong["flipped"] += div_recip(a, w, yw, rn64, sign=-1) != exact
    assert wrong["product"] > 0 and wrong["flipped"] > 0, wrong
    # the single correction is the faithful step: it may or may not round
    # correctly, the second one always does; count it for information only
    print(wrong)
