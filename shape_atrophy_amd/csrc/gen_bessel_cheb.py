"""Writes bessel_cheb.hpp, the coefficient tables of the 2D Matern kernels'
modified Bessel functions K0 and K1:

* x <= 2: power series in y = x^2 / 4 (exact rationals, 14 terms; the
  remainder is below 1e-19 of the sum):
      K0 = -(ln(x/2) + gamma) I0 + S0,   I0 = sum y^k/(k!)^2,  S0 = sum H_k y^k/(k!)^2
      K1 = 1/x + ln(x/2) (x/2) I1s - (x/4) S1,
      I1s = sum y^k/(k!(k+1)!),  S1 = sum (2(H_k - gamma) + 1/(k+1)) y^k/(k!(k+1)!)
* x > 2: Chebyshev series of g_nu(u) = sqrt(x) e^x K_nu(x), x = 4/(u+1),
  interpolated at 32 Chebyshev points (discrete cosine transform) from
  values computed in long double by Temme's continued fraction for K0 and
  K1 (Numerical Recipes, bessik, x >= 2), truncated below 1e-18 of the
  leading coefficient; max relative error 2e-16 on x in [2, 700].

Checked against scipy.special.k0e / k1e.  Run from this directory:
    python gen_bessel_cheb.py > bessel_cheb.hpp
"""
from fractions import Fraction
import math

import numpy as np
from scipy.special import k0e, k1e

L = np.longdouble
PI = L("3.14159265358979323846264338327950288")
EULER = Fraction(57721566490153286060651209, 10 ** 26)


def scaled_k01(x):
    x = L(x)
    b = 2 * (1 + x)
    d = 1 / b
    h = d
    delh = d
    q1, q2, q, c, a = L(0), L(1), L(0.25), L(0.25), L(-0.25)
    s = 1 + q * delh
    for i in range(2, 5000):
        a -= 2 * (i - 1)
        c = -a * c / i
        qn = (q1 - b * q2) / a
        q1, q2 = q2, qn
        q += c * qn
        b += 2
        d = 1 / (b + a * d)
        delh = (b * d - 1) * delh
        h += delh
        dels = q * delh
        s += dels
        if abs(dels / s) < L(1e-21):
            break
    g0 = np.sqrt(PI / 2) / s
    return g0, g0 * (x + L(0.5) - h / 4) / x


def cheb(j, N=32):
    th = PI * (np.arange(N).astype(L) + L(0.5)) / N
    G = np.array([scaled_k01(4 / (u + 1))[j] for u in np.cos(th)])
    c = np.array([2 * np.sum(G * np.cos(m * th)) / N for m in range(N)])
    c[0] /= 2
    c = c.astype(np.float64)
    keep = np.nonzero(np.abs(c) > 1e-18 * abs(c[0]))[0].max() + 1
    return c[:keep]


def clenshaw(c, u):
    b1 = b2 = 0.0
    for cj in c[:0:-1]:
        b1, b2 = 2 * u * b1 - b2 + cj, b1
    return u * b1 - b2 + c[0]


def series(n=14):
    I0, S0, I1, S1 = [], [], [], []
    H = Fraction(0)
    for k in range(n):
        if k:
            H += Fraction(1, k)
        f0 = Fraction(1, math.factorial(k) ** 2)
        f1 = Fraction(1, math.factorial(k) * math.factorial(k + 1))
        I0.append(f0)
        S0.append(H * f0)
        I1.append(f1)
        S1.append((2 * (H - EULER) + Fraction(1, k + 1)) * f1)
    return [[float(v) for v in s] for s in (I0, S0, I1, S1)]


c0, c1 = cheb(0), cheb(1)
for c, f in ((c0, k0e), (c1, k1e)):
    for x in np.geomspace(2.0, 700.0, 2000):
        v = clenshaw(c, 4.0 / x - 1.0)
        assert abs(v / (np.sqrt(x) * f(x)) - 1.0) < 1.5e-15, (x, v)

print("// generated by gen_bessel_cheb.py (coefficients of the 2D Matern kernels' K0, K1)")
print("#pragma once")
names = ("kSerI0", "kSerS0", "kSerI1", "kSerS1")
for name, s in zip(names, series()):
    print(f"__device__ constexpr double {name}[{len(s)}] = {{")
    for i in range(0, len(s), 3):
        print("    " + ", ".join(f"{v:.17e}" for v in s[i:i + 3]) + ",")
    print("};")
for name, c in (("kCheb0", c0), ("kCheb1", c1)):
    print(f"constexpr int {name}N = {len(c)};")
    print(f"__device__ constexpr double {name}[{len(c)}] = {{")
    for i in range(0, len(c), 3):
        print("    " + ", ".join(f"{v:.17e}" for v in c[i:i + 3]) + ",")
    print("};")
