"""Coefficients of the scaled complementary error function used by the GKS flux.

The half-space moments of a Maxwellian need 0.5 erfc(x) and exp(-x^2) at the
same x (csrc/gkm_flux.hpp).  With E = exp(-x^2) already formed,

    erfc(t) = E erfcx(t)            (t = |x| >= 0)
    erfc(-t) = 2 - erfc(t)

and erfcx on [0, inf) is evaluated as

    erfcx(t) = 2 / (1 + 2t) * P(y),   y = (t - K) / (t + K) in [-1, 1),

with P a polynomial of degree N in y (the map y(t) and the 2/(1 + 2t) factor
follow the rational-transform idea of J. A. C. Weideman, "Computation of the
complex error function", SIAM J. Numer. Anal. 31 (1994) 1497-1518).  P is the
least-squares Chebyshev fit, on Chebyshev points in y, of 40-digit values of
erfcx(t) (1 + 2t) / 2 (mpmath), converted to monomial coefficients for Horner
evaluation.  Printed in the form pasted into csrc/gkm_math.hpp (kErfcxP).

    python -m dns_gkm_lpt_amd.erfcx_fit [--check]

--check evaluates 0.5 erfc(x) with the printed coefficients in double
precision, as the device code does, on 40001 points of [-8, 8] against mpmath
(about a minute): measured 3.3e-16 absolute.
"""
import sys

import numpy as np

K = 3.0
N = 24
M = 6000  # fit points


def coefficients():
    import mpmath as mp

    mp.mp.dps = 40
    j = np.arange(M)
    ys = np.cos(np.pi * (j + 0.5) / M)
    F = []
    for y in ys:
        ym = mp.mpf(float(y))
        t = K * (1 + ym) / (1 - ym)
        F.append(float(mp.exp(t * t) * mp.erfc(t) * (1 + 2 * t) / 2))
    c = np.polynomial.chebyshev.chebfit(ys, np.array(F), N)
    return np.polynomial.chebyshev.cheb2poly(c)


def half_erfc(x, p):
    """0.5 erfc(x) in double precision, the device code's operation order."""
    t = abs(x)
    E = np.exp(-x * x)
    a, b = t + K, 1.0 + 2.0 * t
    r = 1.0 / (a * b)
    y = (t - K) * b * r
    acc = p[-1]
    for c in p[-2::-1]:
        acc = acc * y + c
    v = E * acc * a * r
    return v if x >= 0 else 1.0 - v


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    p = coefficients()
    print(f"constexpr double kErfcxP[{N + 1}] = {{")
    for a in p:
        print(f"    {float(a)!r},")
    print("};")
    if "--check" in argv:
        import mpmath as mp

        worst = 0.0
        for x in np.linspace(-8.0, 8.0, 40001):
            ref = float(mp.erfc(mp.mpf(float(x))) / 2)
            worst = max(worst, abs(half_erfc(float(x), p) - ref))
        print(f"# largest absolute error on [-8, 8]: {worst:.2e}")


if __name__ == "__main__":
    main()
