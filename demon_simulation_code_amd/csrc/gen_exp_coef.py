#!/usr/bin/env python3
"""Coefficients of the exp polynomials of pair_core.h: exp_neg's, and
exp_neg_tab's (after its table e^{-k/64}: the degree-4 Chebyshev interpolant
of e^-d on |d| <= 1/128, printed second).

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


# --- exp_neg_tab: e^-d on |d| <= 1/128 after the table e^{-k/64} --------------
TAB_HALF_STEP = mp.mpf(1) / 128
DEGREE_TAB = 4


def tab_main():
    poly = mp.chebyfit(lambda d: mp.exp(-d), [-TAB_HALF_STEP, TAB_HALF_STEP], DEGREE_TAB + 1)
    # the kernel adds the constant term as the literal 1.0
    coefs = [float(c) for c in poly[:-1]] + [1.0]
    worst = mp.mpf(0)
    for k in range(4001):
        d = -TAB_HALF_STEP + 2 * TAB_HALF_STEP * k / 4000
        p = mp.mpf(0)
        for c in coefs:
            p = p * d + mp.mpf(c)
        worst = max(worst, abs(p - mp.exp(-d)) / mp.exp(-d))
    print("table-exp coefficients (highest power first):", ", ".join(repr(c) for c in coefs[:-1]), "and 1.0")
    print("worst relative error:", mp.nstr(worst, 4), "=", mp.nstr(worst / mp.mpf(2) ** -52, 3), "ulp")


if __name__ == "__main__":
    main()
    tab_main()
