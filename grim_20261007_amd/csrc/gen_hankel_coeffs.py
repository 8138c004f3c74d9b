"""Generate hankel_coeffs.h: polynomial coefficients of the device Hankel kernels.

The device evaluates (j/4) H0^(2)(x) and (j/4) H1^(2)(x) for real x > 0 in
FP64 from three representations (DLMF 10.8.1, 10.17.3):

* x < XS, with y = x^2/4 and L = ln(x/2):
      J0 = A0(y)                 J1 = (x/2) A1(y)
      Y0 = (2/pi) L J0 + B0(y)   Y1 = (2/pi) L J1 - 2/(pi x) + (x/2) B1(y)
  A0, A1, B0, B1 are entire in y; their exact Taylor coefficients (to
  degree DEG_SMALL) are used with Horner on the device.
* x >= XS, the modulus / phase form
      H_nu^(2)(x) = sqrt(2/(pi x)) (P_nu(x) - j Q_nu(x)) exp(-j (x - (2 nu + 1) pi/4))
  with P_nu, Q_nu from two Chebyshev fits converted to monomials:
  - XS <= x < XM: w = (XS/x)^2 in (W_LO, 1], W_LO = (XS/XM)^2, mapped to
    u = MID_SCALE w + MID_SHIFT in (-1, 1]; P_nu = PnM(u), Q_nu = (XS/x) QnM(u),
    degree DEG_MID;
  - x >= XM: w = (XM/x)^2 in (0, 1]; P_nu = Pn(w), Q_nu = (XM/x) Qn(w),
    degree DEG_LARGE.  Far from the origin the asymptotic expansion
    converges fast, so this range needs few terms; it is the common case of
    the far-pair assembly.

The header also carries the Gauss-Legendre rules (8, 16 and 20 points,
mapped to [0, 1] as t = (x + 1)/2, w/2) of the element-pair quadratures.

Reference values come from mpmath at 40 digits.  Run once; the header is
committed:  python gen_hankel_coeffs.py > hankel_coeffs.h
"""
from __future__ import annotations

import sys

import mpmath as mp
import numpy as np
from numpy.polynomial import chebyshev as C
from numpy.polynomial import polynomial as P

mp.mp.dps = 40
XS = 6.0
XM = 20.0
DEG_SMALL = 18
DEG_MID = 11
DEG_LARGE = 5
W_LO = (XS / XM) ** 2                  # 0.09
MID_SCALE = 2.0 / (1.0 - W_LO)
MID_SHIFT = -(1.0 + W_LO) / (1.0 - W_LO)
NODES = 600


def _cheb_nodes(a, b, n):
    j = np.arange(n)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos(np.pi * (j + 0.5) / n)


def small_parts(y):
    """A0, A1, B0, B1 at y (mpmath)."""
    y = mp.mpf(y)
    x = 2 * mp.sqrt(y)
    if y == 0:
        a0 = mp.mpf(1)
        a1 = mp.mpf(1)
        b0 = 2 / mp.pi * mp.euler
        b1 = -(1 / mp.pi) * (mp.digamma(1) + mp.digamma(2))
        return a0, a1, b0, b1
    L = mp.log(x / 2)
    J0, J1, Y0, Y1 = mp.besselj(0, x), mp.besselj(1, x), mp.bessely(0, x), mp.bessely(1, x)
    a0 = J0
    a1 = J1 / (x / 2)
    b0 = Y0 - 2 / mp.pi * L * J0
    b1 = (Y1 - 2 / mp.pi * L * J1 + 2 / (mp.pi * x)) / (x / 2)
    return a0, a1, b0, b1


def modulus_parts(w, x0):
    """P0, Q0n, P1, Q1n at w = (x0/x)^2 (mpmath); Qn = Q * x / x0."""
    w = mp.mpf(w)
    out = []
    for nu in (0, 1):
        if w == 0:
            out += [mp.mpf(1), mp.mpf(4 * nu * nu - 1) / 8 / x0]
            continue
        x = x0 / mp.sqrt(w)
        h = mp.hankel2(nu, x)
        v = h * mp.exp(1j * (x - (2 * nu + 1) * mp.pi / 4)) * mp.sqrt(mp.pi * x / 2)
        out += [mp.re(v), -mp.im(v) * x / x0]
    return out


def mid_parts(u):
    """P0, Q0n, P1, Q1n of XS <= x < XM at u = MID_SCALE w + MID_SHIFT, w = (XS/x)^2."""
    lo = (mp.mpf(XS) / XM) ** 2
    return modulus_parts(((1 - lo) * mp.mpf(u) + 1 + lo) / 2, XS)


def large_parts(w):
    """P0, Q0n, P1, Q1n of x >= XM at w = (XM/x)^2."""
    return modulus_parts(w, XM)


def taylor_small(deg):
    """Exact Taylor coefficients of A0, A1, B0, B1 in y (DLMF 10.8.1)."""
    a0, a1, b0, b1 = [], [], [], []
    for m in range(deg + 1):
        sign = (-1) ** m
        f0 = mp.factorial(m) ** 2
        f1 = mp.factorial(m) * mp.factorial(m + 1)
        harmonic = mp.fsum(mp.mpf(1) / j for j in range(1, m + 1))
        a0.append(sign / f0)
        a1.append(sign / f1)
        # Y0 - (2/pi) ln(x/2) J0 = (2/pi) sum (gamma - H_m) ... = -(2/pi) sum psi(m+1) (-y)^m/(m!)^2
        b0.append(-(2 / mp.pi) * mp.digamma(m + 1) * sign / f0)
        b1.append(-(1 / mp.pi) * (mp.digamma(m + 1) + mp.digamma(m + 2)) * sign / f1)
    return [np.array([float(v) for v in arr]) for arr in (a0, a1, b0, b1)]


def fit(fun, a, b, deg):
    t = _cheb_nodes(a, b, NODES)
    vals = np.array([[float(v) for v in fun(ti)] for ti in t])
    u = (2 * t - (a + b)) / (b - a)
    coeffs = []
    for col in range(vals.shape[1]):
        c = C.chebfit(u, vals[:, col], deg)
        mono_u = C.cheb2poly(c)                     # in u
        # u = (2 t - (a+b)) / (b-a) = s * t + o
        s, o = 2.0 / (b - a), -(a + b) / (b - a)
        mono_t = P.polyzero
        power = np.array([1.0])
        lin = np.array([o, s])
        for m, cm in enumerate(mono_u):
            mono_t = P.polyadd(mono_t, cm * power)
            power = P.polymul(power, lin)
        coeffs.append(np.asarray(mono_t, float))
    return coeffs


def main(out=sys.stdout):
    small = taylor_small(DEG_SMALL)
    mid = fit(mid_parts, -1.0, 1.0, DEG_MID)
    large = fit(large_parts, 0.0, 1.0, DEG_LARGE)
    names = ("A0", "A1", "B0", "B1"), ("P0M", "Q0M", "P1M", "Q1M"), ("P0", "Q0", "P1", "Q1")
    print("// Generated by gen_hankel_coeffs.py -- do not edit.", file=out)
    print("#pragma once", file=out)
    print(f"#define GRIM_HANKEL_XS {XS!r}", file=out)
    print(f"#define GRIM_HANKEL_XM {XM!r}", file=out)
    print(f"#define GRIM_HANKEL_DEG_SMALL {DEG_SMALL}", file=out)
    print(f"#define GRIM_HANKEL_DEG_MID {DEG_MID}", file=out)
    print(f"#define GRIM_HANKEL_MID_SCALE {MID_SCALE!r}", file=out)
    print(f"#define GRIM_HANKEL_MID_SHIFT {MID_SHIFT!r}", file=out)
    print(f"#define GRIM_HANKEL_DEG_LARGE {DEG_LARGE}", file=out)
    for group, arrays in zip(names, (small, mid, large)):
        for name, arr in zip(group, arrays):
            body = ", ".join(repr(float(v)) for v in arr)
            print(f"static constexpr double GRIM_HK_{name}[{len(arr)}] = {{{body}}};", file=out)
    # Gauss-Legendre rules on [0, 1] used by the element-pair quadratures.
    for order in (8, 16, 20):
        x, w = np.polynomial.legendre.leggauss(order)
        t, wt = 0.5 * (x + 1.0), 0.5 * w
        print(f"static constexpr double GRIM_GL{order}_T[{order}] = {{{', '.join(repr(float(v)) for v in t)}}};", file=out)
        print(f"static constexpr double GRIM_GL{order}_W[{order}] = {{{', '.join(repr(float(v)) for v in wt)}}};", file=out)


if __name__ == "__main__":
    main()
