"""Independent pins of the oracle's building blocks (closed forms, invariants, brute force)."""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy import integrate

from oracle import bem2d, mie


def test_green_kernels_match_mpmath_and_wronskian():
    k = 7.3
    r = np.array([1e-6, 0.01, 0.3, 1.0, 5.0, 40.0])
    g, h = bem2d.green_kernels(k, r)
    for ri, gi, hi in zip(r, g, h):
        x = k * ri
        want_g = complex(0.25j * mp.hankel2(0, x))
        want_h = complex(0.25j * k * mp.hankel2(1, x))
        tol = max(1e-14, 4 * np.finfo(float).eps * x)   # argument conditioning eps * k r
        assert abs(gi - want_g) <= tol * abs(want_g)
        assert abs(hi - want_h) <= tol * abs(want_h)
        # Wronskian J1 Y0 - J0 Y1 = 2 / (pi x): from g = (Y0 + j J0)/4, h = k (Y1 + j J1)/4
        j0, y0 = 4 * gi.imag, 4 * gi.real
        j1, y1 = 4 * hi.imag / k, 4 * hi.real / k
        assert abs((j1 * y0 - j0 * y1) - 2 / (math.pi * x)) <= max(1e-13, 8 * tol) * 2 / (math.pi * x)


@pytest.mark.parametrize("order", [8, 16, 20])
def test_gauss_rules_exact_for_polynomials(order):
    t, w = bem2d.gauss(order)
    for p in range(2 * order):
        assert abs(np.sum(w * t**p) - 1.0 / (p + 1)) < 1e-14


def test_self_block_matches_brute_force():
    """Exact-series self S block against adaptive 2-D quadrature of the log-singular integrand."""
    length, k = 0.2, 9.0
    s = bem2d.self_single_layer(length, k)

    def entry(a, b):
        def f(u, t, part):
            r = abs(t - u) * length
            val = complex(0.25j * mp.hankel2(0, k * max(r, 1e-300)))
            pa = (1 - t) if a == 0 else t
            pb = (1 - u) if b == 0 else u
            v = val * pa * pb
            return v.real if part == 0 else v.imag
        out = []
        for part in (0, 1):
            lo = integrate.dblquad(lambda u, t: f(u, t, part), 0, 1, 0, lambda t: t, epsabs=1e-13, epsrel=1e-12)[0]
            hi = integrate.dblquad(lambda u, t: f(u, t, part), 0, 1, lambda t: t, 1, epsabs=1e-13, epsrel=1e-12)[0]
            out.append(lo + hi)
        return complex(out[0], out[1]) * length * length

    for a, b in ((0, 0), (0, 1)):
        want = entry(a, b)
        assert abs(s[a, b] - want) <= 1e-9 * abs(want)


def test_touching_rule_collinear_against_adaptive_quadrature():
    """Two collinear touching elements: the t^4 rule against scipy's adaptive quadrature."""
    pts = np.array([[0.0, 0.0], [0.1, 0.0], [0.2, 0.0]])
    p0 = np.array([pts[0], pts[1]])
    p1 = np.array([pts[1], pts[2]])
    seg = p1 - p0
    mesh = bem2d.Mesh(pts, np.array([[0, 1], [1, 2]]), p0, p1, 0.5 * (p0 + p1), np.linalg.norm(seg, axis=1),
                      np.array([[0.0, 1.0], [0.0, 1.0]]))
    k = 11.0
    os_, ss_ = bem2d.shared_endpoint(mesh, 0, 1)
    s, _ = bem2d.touching_blocks(mesh, k, 0, 1, os_, ss_)

    def f(u, t, part):
        x = 0.1 * t
        y = 0.1 + 0.1 * u
        val = complex(0.25j * mp.hankel2(0, k * (y - x))) * t * (1 - u)   # phi_1(obs) phi_0(src)
        return val.real if part == 0 else val.imag
    want = complex(*[integrate.dblquad(lambda u, t: f(u, t, p), 0, 1, 0, 1, epsabs=1e-14, epsrel=1e-12)[0]
                     for p in (0, 1)]) * 0.01
    assert abs(s[1, 0] - want) <= 1e-10 * abs(want)


def test_plane_wave_loads_against_gauss_quadrature():
    pts = np.array([[0.3, -0.2], [0.5, 0.1], [0.1, 0.4]])[::-1]
    mesh = bem2d.mesh_from_points(pts)
    k = 17.0
    angles = [0.0, 33.0, 250.0]
    bu, bdn = bem2d.incident_loads(mesh, k, angles)
    t, w = bem2d.gauss(40)
    d = bem2d.directions(angles)
    want_u = np.zeros_like(bu)
    want_dn = np.zeros_like(bdn)
    for e in range(len(mesh.length)):
        xs = mesh.p0[e][None, :] + t[:, None] * mesh.seg[e][None, :]
        ph = np.exp(1j * k * (xs @ d.T)) * (w * mesh.length[e])[:, None]
        dn = 1j * k * (d @ mesh.normal[e])
        for a, phi in ((0, 1 - t), (1, t)):
            want_u[mesh.elements[e, a]] += phi @ ph
            want_dn[mesh.elements[e, a]] += (phi @ ph) * dn
    assert np.max(np.abs(bu - want_u)) < 1e-13
    assert np.max(np.abs(bdn - want_dn)) < 1e-12


def test_far_field_against_brute_force_projection():
    rng = np.random.default_rng(3)
    mesh = bem2d.mesh_from_points(np.array([[1.0, 0.0], [0.0, -1.0], [-1.0, 0.2], [0.1, 1.1]]))
    k = 5.0
    angles = [10.0, 100.0]
    x = rng.standard_normal((4, 2)) + 1j * rng.standard_normal((4, 2))
    got = bem2d.farfield(mesh, k, angles, x)
    t, w = bem2d.gauss(40)
    d = bem2d.directions(angles)
    gamma = 1 / (1j * k)
    want = np.zeros(2, complex)
    for a in range(2):
        for e in range(4):
            xs = mesh.p0[e][None, :] + t[:, None] * mesh.seg[e][None, :]
            rho = (1 - t) * x[mesh.elements[e, 0], a] + t * x[mesh.elements[e, 1], a]
            ph = np.exp(1j * k * (xs @ d[a])) * w * mesh.length[e]
            want[a] += np.sum(rho * ph) + gamma * np.sum(rho * ph) * 1j * k * (d[a] @ mesh.normal[e])
    assert np.max(np.abs(got - want)) < 1e-13


def test_mie_rayleigh_limit_and_symmetry():
    # TM small-ka limit of the PEC cylinder: B ~ -4j * (-J0/H0) -> pi / (2 ln(ka))-type; check against the
    # n = 0 term alone, which dominates as ka -> 0.
    a, k = 0.01, 1.0
    full = mie.pec_backscatter_amplitude(a, k, "TM")
    from scipy import special as sp
    n0 = -4j * (-sp.jv(0, k * a) / sp.hankel2(0, k * a))
    assert abs(full - n0) / abs(n0) < 1e-3
    assert mie.terms_for_ka(10.0) >= 10 + 10


def test_refined_solve_matches_double_solve():
    rng = np.random.default_rng(1)
    n = 120
    z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 20 * np.eye(n)
    b = rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))
    x, steps = bem2d.refined_solve(z, b)
    want = np.linalg.solve(z, b)
    assert np.max(np.abs(x - want)) / np.max(np.abs(want)) < 1e-10
    assert steps >= 1
