"""Derive the numeric constants of the deterministic math helpers.

Used by the oracle (oracle/dgswe_oracle.c) and the GPU library
(dg-rain_amd/csrc/dgr_math.h); both embed the printed values.

  * cubic seed for t^(-1/3) on [1, 8): least-squares fit on Chebyshev nodes
    (only a starting point for the Newton iterations; accuracy ~1e-2 is enough);
  * three-part split of pi/2 (33 + 33 + 53 significant bits) for Cody-Waite
    range reduction;
  * Taylor coefficients 1/n! rounded to nearest double;
  * ln 2, 4/3 - fl(4/3).
Run: python oracle/gen_constants.py
"""
from decimal import Decimal, getcontext
from fractions import Fraction
import math

import numpy as np

getcontext().prec = 80

# pi to 70 digits (Machin formula, exact rational arithmetic)
def _arctan_inv(n, terms=120):
    s = Fraction(0)
    for k in range(terms):
        s += Fraction((-1) ** k, (2 * k + 1) * n ** (2 * k + 1))
    return s

PI = 16 * _arctan_inv(5) - 4 * _arctan_inv(239)
PIO2 = PI / 2


def round_bits(fr, bits):
    """Round a positive Fraction to `bits` significant bits."""
    e = math.floor(math.log2(float(fr)))
    scale = Fraction(2) ** (bits - 1 - e)
    return Fraction(round(fr * scale)) / scale


def main():
    c1 = round_bits(PIO2, 33)
    c2 = round_bits(PIO2 - c1, 33)
    c3 = PIO2 - c1 - c2
    print("PIO2_1 = %.17g" % float(c1))
    print("PIO2_2 = %.17g" % float(c2))
    print("PIO2_3 = %.17g" % float(c3))
    print("TWO_OVER_PI = %.17g" % float(1 / PIO2))
    for n in range(2, 17):
        print("1/%d! = %.17g" % (n, float(Fraction(1, math.factorial(n)))))
    ln2 = Decimal(2).ln()
    print("LN2 = %.17g" % float(ln2))
    y = 4.0 / 3.0
    delta = Fraction(4, 3) - Fraction(y)
    print("DELTA43 = %.17g  (2^-52/3 = %.17g)" % (float(delta), 2.0 ** -52 / 3))
    # seed for t^(-1/3) on [1,8)
    k = np.arange(64)
    t = 4.5 + 3.5 * np.cos(np.pi * (k + 0.5) / 64)
    A = np.vander(t, 4, increasing=True)
    coef, *_ = np.linalg.lstsq(A, t ** (-1.0 / 3), rcond=None)
    tt = np.linspace(1, 8, 100001)
    err = np.max(np.abs(np.polyval(coef[::-1], tt) * tt ** (1.0 / 3) - 1))
    for i, c in enumerate(coef):
        print("P43_C%d = %.17g" % (i, c))
    print("seed max rel err = %.3g" % err)


if __name__ == "__main__":
    main()
