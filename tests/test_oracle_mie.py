"""Pins of the Mie series oracle (oracle/mie.py) without running the reference's code.

The series is checked against (1) small-ka closed forms derived from the Bessel
function limits, (2) losslessness of each PEC coefficient, (3) the reference's
own cylinder gates (tests/test_rcs_physics_regression.py:648-669, 671-719)
re-run with the oracle's BEM solve, and (4) the reference-written golden
amplitudes of two circles.  The last test shows that two plausible mistakes in
the series (the other Hankel kind, the forward-scatter sum) fail these pins.
"""
import math

import numpy as np
import pytest
from scipy import special as sp

from oracle import bem2d, mie
from tests import golden_cases

EULER_GAMMA = 0.5772156649015329
J0_FIRST_ZERO = 2.404825557695773


@pytest.mark.parametrize("x", [1e-4, 1e-3])
def test_small_ka_limits(x):
    """TM: a_0 = -J_0/H_0^(2) -> -1/(1 - (2j/pi)(ln(x/2) + gamma)), the other terms O(x^2);
    TE: a_0 -> j pi x^2/4, a_{+-1} -> -j pi x^2/4, so B = -4j (a_0 - a_1 - a_-1) -> 3 pi x^2."""
    k = 1.0
    a0 = -1.0 / (1.0 - 2j / math.pi * (math.log(x / 2.0) + EULER_GAMMA))
    tm = mie.pec_backscatter_amplitude(x, k, "TM")
    assert abs(tm / (-4j * a0) - 1.0) < 20.0 * x * x * abs(math.log(x))
    te = mie.pec_backscatter_amplitude(x, k, "TE")
    assert abs(te / (3.0 * math.pi * x * x) - 1.0) < 2.0 * x * x
    # sigma of the TE limit: |3 pi (ka)^2|^2 / (4k) = (9 pi^2 / 4) k^3 a^4
    assert abs(mie.pec_sigma(x, k, "TE") / (9.0 * math.pi ** 2 / 4.0 * x ** 4) - 1.0) < 4.0 * x * x


@pytest.mark.parametrize("pol", ["TM", "TE"])
@pytest.mark.parametrize("ka", [0.5, 1.0, 10.0, 37.0])
def test_pec_coefficients_are_lossless(pol, ka):
    """1 + 2 a_n = -H_n^(1)/H_n^(2) (or the derivative ratio) has modulus 1 for a PEC body."""
    n, a = mie.pec_coefficients(ka, pol)
    assert np.max(np.abs(np.abs(1.0 + 2.0 * a) - 1.0)) < 1e-12
    # the truncation keeps every term above the rounding level
    assert np.max(np.abs(a[[0, -1]])) < 1e-16 * np.max(np.abs(a))


def _oracle_backscatter(radius, chords, ka, pol):
    from grim_20261007_amd import workloads
    f = workloads.frequency_for_ka(ka, radius)
    mesh = bem2d.build_mesh(workloads.circle(radius, chords), "meters", f)
    res = bem2d.solve_monostatic(mesh, f, [0.0], pols=(pol,))
    return complex(res[pol]["amplitude"][0]), bem2d.frequency_k(f)


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_reference_gate_ka1_96_chords(pol):
    """tests/test_rcs_physics_regression.py:648-669: 96-chord PEC circle, ka = 1, |dB| < 0.01."""
    amp, k = _oracle_backscatter(0.1, 96, 1.0, pol)
    want = mie.pec_sigma(0.1, k, pol)
    got = abs(amp) ** 2 / (4.0 * k)
    assert abs(10.0 * math.log10(got / want)) < 0.01


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_reference_gate_first_j0_zero(pol):
    """tests/test_rcs_physics_regression.py:671-719: 64-chord circle at the first zero of J_0,
    complex amplitude within 5e-3 of the series (the interior resonance must not leak)."""
    amp, k = _oracle_backscatter(0.1, 64, J0_FIRST_ZERO, pol)
    want = mie.pec_backscatter_amplitude(0.1, k, pol)
    assert abs(amp - want) / abs(want) < 5e-3
    if pol == "TM":      # a_0 vanishes there: the series is carried by |n| >= 1 alone
        n, a = mie.pec_coefficients(k * 0.1, pol)
        assert abs(a[n == 0][0]) < 1e-14


@pytest.mark.parametrize("name,radius,tol", [("circle64_ka1", 0.1, 2.5e-3), ("c1_circle400_ka10", 1.0, 6e-4)])
def test_reference_golden_amplitudes(name, radius, tol):
    """The reference solver's own amplitudes (tests/golden, every angle) agree with the series
    to the discretisation error of the polygon: 1.7e-3 at 64 chords (ka = 1), 4.3e-4 at 400
    chords (ka = 10)."""
    case = golden_cases.load(name)
    k = bem2d.frequency_k(float(case["frequency_ghz"]))
    for pol in ("TE", "TM"):
        want = mie.pec_backscatter_amplitude(radius, k, pol)
        amp = case[f"sample_amp_{pol}"]
        assert np.max(np.abs(amp - want)) / abs(want) < tol, pol


def _wrong_series(ka, pol, kind):
    n = np.arange(-mie.terms_for_ka(ka), mie.terms_for_ka(ka) + 1)
    if kind == "hankel1":
        a = (-sp.jv(n, ka) / sp.hankel1(n, ka) if pol == "TM"
             else -sp.jvp(n, ka, 1) / sp.h1vp(n, ka, 1))
        return complex(-4.0j * np.sum(a * (-1.0) ** n))
    _, a = mie.pec_coefficients(ka, pol)
    return complex(-4.0j * np.sum(a))          # forward scatter instead of back-scatter


@pytest.mark.parametrize("kind", ["hankel1", "forward"])
def test_pins_reject_plausible_mistakes(kind):
    """A series with the other time convention or the forward-scatter sum fails the complex
    pins above by orders of magnitude more than their tolerances."""
    case = golden_cases.load("c1_circle400_ka10")
    k = bem2d.frequency_k(float(case["frequency_ghz"]))
    for pol in ("TE", "TM"):
        wrong = _wrong_series(k * 1.0, pol, kind)
        amp = case[f"sample_amp_{pol}"]
        assert np.max(np.abs(amp - wrong)) / abs(wrong) > 0.1, (kind, pol)
