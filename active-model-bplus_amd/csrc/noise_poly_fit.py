"""Coefficients of the FP64 noise polynomials in amb_math.h (ln_unit, sincos_turn).

Each is a Chebyshev fit (mpmath.chebyfit, near-minimax) at 60 digits, rounded to double:
  g(t) = (atanh(sqrt t) / sqrt t - 1) / t      on t in [0, (3 - 2 sqrt 2)^2], 7 coefficients
         (ln m = 2 s + 2 s^3 g(s^2), s = (m - 1) / (m + 1), m in [sqrt(1/2), sqrt(2)))
  S(t) = (sin(sqrt t) - sqrt t) / t^(3/2)      on t in [0, (pi/4)^2], 6 coefficients
  C(t) = (cos(sqrt t) - 1 + t/2) / t^2         on t in [0, (pi/4)^2], 6 coefficients
and prints them, highest degree first, with the worst relative error of the double-rounded
polynomial over 4001 points of its interval.  Run: python3 noise_poly_fit.py (needs mpmath).
"""
import mpmath as mp

mp.mp.dps = 60


def g(t):
    if t == 0:
        return mp.mpf(1) / 3
    s = mp.sqrt(t)
    return (mp.atanh(s) / s - 1) / t


def S(t):
    if t == 0:
        return -mp.mpf(1) / 6
    x = mp.sqrt(t)
    return (mp.sin(x) - x) / (t * x)


def C(t):
    if t == 0:
        return mp.mpf(1) / 24
    x = mp.sqrt(t)
    return (mp.cos(x) - 1 + t / 2) / (t * t)


FITS = [("kLnG", g, (3 - 2 * mp.sqrt(2)) ** 2, 7), ("kSinS", S, (mp.pi / 4) ** 2, 6),
        ("kCosC", C, (mp.pi / 4) ** 2, 6)]

if __name__ == "__main__":
    for name, f, tmax, n in FITS:
        coefs = [float(c) for c in mp.chebyfit(f, [0, tmax], n)]
        worst = mp.mpf(0)
        for k in range(4001):
            t = tmax * k / 4000
            p = mp.mpf(0)
            for c in coefs:
                p = p * t + c
            worst = max(worst, abs(p / f(t) - 1))
        body = ", ".join(c.hex() for c in coefs)
        print(f"constexpr double {name}[{n}] = {{{body}}};  // rel. error {mp.nstr(worst, 3)}")
