"""The device Hankel kernels' coefficient tables (grim_20261007_amd/csrc/hankel_coeffs.h),
evaluated in FP64 with the same three representations as hankel_quarter_v, against SciPy's
Hankel functions: a wrong, truncated or swapped table fails here before any GPU run."""
import re
from pathlib import Path

import numpy as np
import pytest
from scipy import special

HEADER = Path(__file__).resolve().parent.parent / "grim_20261007_amd" / "csrc" / "hankel_coeffs.h"


def _tables():
    src = HEADER.read_text()
    co = {m.group(1): np.array([float(v) for v in m.group(2).split(",")])
          for m in re.finditer(r"GRIM_HK_(\w+)\[\d+\] = \{([^}]*)\}", src)}
    xs = float(re.search(r"GRIM_HANKEL_XS ([\d.]+)", src).group(1))
    xm = float(re.search(r"GRIM_HANKEL_XM ([\d.]+)", src).group(1))
    scale = float(re.search(r"GRIM_HANKEL_MID_SCALE (\S+)", src).group(1))
    shift = float(re.search(r"GRIM_HANKEL_MID_SHIFT (\S+)", src).group(1))
    return co, xs, xm, scale, shift


def _horner(c, t):
    r = np.full_like(t, c[-1])
    for v in c[-2::-1]:
        r = r * t + v
    return r


def _device_form(x):
    """(j/4) H0^(2)(x), (j/4) H1^(2)(x) by the device's three ranges."""
    co, xs, xm, scale, shift = _tables()
    x = np.asarray(x, float)
    g = np.empty(x.shape, complex)
    h = np.empty(x.shape, complex)
    s = x < xs
    y = 0.25 * x[s] ** 2
    lg = np.log(0.5 * x[s])
    j0 = _horner(co["A0"], y)
    j1 = 0.5 * x[s] * _horner(co["A1"], y)
    y0 = (2 / np.pi) * lg * j0 + _horner(co["B0"], y)
    y1 = (2 / np.pi) * lg * j1 - 2 / (np.pi * x[s]) + 0.5 * x[s] * _horner(co["B1"], y)
    g[s], h[s] = 0.25 * y0 + 0.25j * j0, 0.25 * y1 + 0.25j * j1
    for sel, x0, keys, var in (((x >= xs) & (x < xm), xs, ("P0M", "Q0M", "P1M", "Q1M"),
                                lambda w: w * scale + shift),
                               (x >= xm, xm, ("P0", "Q0", "P1", "Q1"), lambda w: w)):
        xl = x[sel]
        r = x0 / xl
        t = var(r * r)
        amp = np.sqrt(2 / (np.pi * xl))
        for nu, out in ((0, g), (1, h)):
            p = _horner(co[keys[2 * nu]], t)
            q = r * _horner(co[keys[2 * nu + 1]], t)
            out[sel] = 0.25j * amp * (p - 1j * q) * np.exp(-1j * (xl - (2 * nu + 1) * np.pi / 4))
    return g, h


@pytest.mark.parametrize("lo,hi,tol", [(1e-6, 1.0, 2e-14), (1.0, 6.0, 2e-14), (6.0, 20.0, 2e-14),
                                       (20.0, 60.0, 2e-14), (60.0, 2000.0, 1e-12)])
def test_device_hankel_tables_match_scipy(lo, hi, tol):
    x = np.geomspace(lo, hi, 4001)
    g, h = _device_form(x)
    ref_g = 0.25j * special.hankel2(0, x)
    ref_h = 0.25j * special.hankel2(1, x)
    # beyond x = 60 both sides carry the eps * x conditioning of the phase
    assert np.max(np.abs(g - ref_g) / np.abs(ref_g)) < tol
    assert np.max(np.abs(h - ref_h) / np.abs(ref_h)) < tol


def test_range_boundaries_are_continuous():
    _, xs, xm, _, _ = _tables()
    for b in (xs, xm):
        x = np.array([np.nextafter(b, 0.0), b])
        g, h = _device_form(x)
        assert abs(g[0] - g[1]) < 1e-14 and abs(h[0] - h[1]) < 1e-14


def _sincos_block():
    src = (HEADER.parent / "common.h").read_text()
    block = src[src.index("if (x < 1.0e9) {") + len("if (x < 1.0e9) {"):]
    block = block[:block.index("} else {")]
    nums = [float(v) for v in re.findall(r"(?<![\w.])-?\d+\.\d+(?:e-?\d+)?", block)]
    return nums


def test_device_phase_reduction_matches_sin_cos():
    """cos x + sin x and cos x - sin x as hankel_quarter_v forms them (Cody-Waite reduction
    by pi/2 with FMA, Taylor polynomials of sin and cos on [-pi/4, pi/4]), with the constants
    read from common.h and exact FMA rounding."""
    from fractions import Fraction

    def fma(a, b, c):
        return float(Fraction(a) * Fraction(b) + Fraction(c))

    nums = _sincos_block()
    two_over_pi, pio2_hi, pio2_lo = nums[0], nums[1], nums[2]
    ps_c, pc_c = nums[3:10], nums[10:17]
    assert nums[17:] == [-0.5, 1.0]
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(6.0, 60.0, 400), rng.uniform(60.0, 1e5, 400),
                         np.arange(4, 400) * (np.pi / 4)])
    worst = 0.0
    for x in xs:
        x = float(x)
        n = float(np.rint(x * two_over_pi))
        t = fma(-n, pio2_lo, fma(-n, pio2_hi, x))
        t2 = t * t
        ps = ps_c[0]
        for c in ps_c[1:]:
            ps = fma(ps, t2, c)
        sn = fma(ps * t2, t, t)
        pc = pc_c[0]
        for c in pc_c[1:]:
            pc = fma(pc, t2, c)
        cs = fma(fma(pc, t2, -0.5), t2, 1.0)
        q = int(n) & 3
        a, b = cs + sn, cs - sn
        cps, cms = (b, -a) if q & 1 else (a, b)
        if q & 2:
            cps, cms = -cps, -cms
        worst = max(worst, abs(cps - (np.cos(x) + np.sin(x))), abs(cms - (np.cos(x) - np.sin(x))))
    # numpy's sin/cos are correctly rounded to about an ulp; x up to 1e5 adds nothing here
    assert worst < 1e-15
