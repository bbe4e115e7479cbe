#!/usr/bin/env python3
"""Coefficients of the exp polynomial used by exp_neg in pair_core.h.

e^f = 1 + f + f^2 q(f) on [-ln2/2, ln2/2]; q is the degree-9 Chebyshev
interpolant of (e^f - 1 - f) / f^2 (near-minimax), computed at 60 digits and
rounded to double.  Prints the coefficients (highest power first) and the
worst relative error of 1 + f + f^2 q(f) against e^f on a fine grid, with the
rounded coefficients evaluated exactly.  Run by hand; needs mpmath.
"""
import mpmath as mp

mp.mp.dps = 60
HALF_LN2 = mp.log(2) / 2
DEGREE_Q = 9


def g(f):
    if abs(f) < mp.mpf("1e-30"):
        return mp.mpf(1) / 2
    return (mp.exp(f) - 1 - f) / f ** 2


def main():
    poly = mp.chebyfit(g, [-HALF_LN2, HALF_LN2], DEGREE_Q + 1)
    coefs = [float(c) for c in poly]
    worst = mp.mpf(0)
    for k in range(4001):
        f = -HALF_LN2 + 2 * HALF_LN2 * k / 4000
        q = mp.mpf(0)
        for c in coefs:
            q = q * f + mp.mpf(c)
        worst = max(worst, abs(1 + f + f * f * q - mp.exp(f)) / mp.exp(f))
    print("coefficients:", ", ".join(repr(c) for c in coefs))
    print("worst relative error:", mp.nstr(worst, 4), "=", mp.nstr(worst / mp.mpf(2) ** -52, 3), "ulp")


if __name__ == "__main__":
    main()
