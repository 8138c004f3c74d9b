"""The complex-Hankel quadrature of the lossy operators on the CPU: the generated 32-point
generalized Gauss-Laguerre rule (weight u^(-1/2) e^(-u)) against its closed-form moments, and the
header against its generator."""
import io
import math
import os
import runpy

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, os.pardir, "grim_20261007_amd", "csrc")


def _rule():
    text = open(os.path.join(CSRC, "lossy_rules.h")).read()
    vals = {}
    for line in text.splitlines():
        if line.startswith("static constexpr double"):
            name = line.split()[3].split("[")[0]
            vals[name] = np.array([float(v) for v in line.split("{")[1].split("}")[0].split(",")])
    return vals["GRIM_LAG32_X"], vals["GRIM_LAG32_W"]


def test_laguerre_rule_moments():
    """sum w_i x_i^m = Gamma(m + 1/2) for m = 0 .. 63 (a 32-point Gauss rule is exact to degree 63)."""
    x, w = _rule()
    assert len(x) == 32 and np.all(np.diff(x) > 0) and np.all(w > 0)
    for m in range(0, 64, 3):
        want = math.gamma(m + 0.5)
        got = float(np.sum(w * x ** m))
        assert abs(got - want) <= 1e-12 * want, (m, got, want)


def test_laguerre_header_matches_generator():
    gen = runpy.run_path(os.path.join(CSRC, "gen_lossy_rules.py"))
    buf = io.StringIO()
    gen["main"](buf)
    assert buf.getvalue() == open(os.path.join(CSRC, "lossy_rules.h")).read()


def test_asymptotic_coefficients():
    """a_k(nu) = prod (4 nu^2 - (2j - 1)^2) / (8 j) from exact rational arithmetic."""
    from fractions import Fraction
    text = open(os.path.join(CSRC, "lossy_rules.h")).read()
    for nu in (0, 1):
        line = [l for l in text.splitlines() if f"GRIM_HASYM{nu}[" in l][0]
        got = [float(v) for v in line.split("{")[1].split("}")[0].split(",")]
        a = Fraction(1)
        for k in range(20):
            if k:
                a *= Fraction(4 * nu * nu - (2 * k - 1) ** 2, 8 * k)
            assert abs(got[k] - float(a)) <= 1e-15 * abs(float(a)), (nu, k)
