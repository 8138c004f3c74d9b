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


def _golden(name):
    import json
    z = np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False)
    return json.loads(str(z["snapshot_json"])), float(z["frequency_ghz"]), z["angles_deg"], z


def test_oracle_lossy_dielectric_matches_reference():
    """The oracle's two-density system at complex k1 = k0 n against the reference's lossy solves."""
    import pytest
    from oracle import bem2d
    for name, eps, mu in (("lossy_diel_circle48", 4.0 - 1.0j, 1.0 + 0j),
                          ("lossy_diel_magnetic64", 2.5 - 0.3j, 1.6 - 0.5j)):
        snap, f, angles, z = _golden(name)
        om = bem2d.mesh_from_nodes(z["nodes"], z["elements"])
        k0 = bem2d.frequency_k(f)
        n = len(om.nodes)
        z_tm, z_te = bem2d.assemble(om, k0)
        k1 = complex(k0 * np.sqrt(eps * mu))
        assert k1.imag < 0
        s1, kp1 = bem2d.plain_operators(om, k1)
        for pol, factor in (("TE", 1.0 / eps), ("TM", 1.0 / mu)):
            a, b = bem2d.dielectric_system(om, k0, z_tm, z_te, s1, kp1, factor, angles)
            if f"Z_{pol}" in z:
                assert np.max(np.abs(a - z[f"Z_{pol}"])) < 1e-14 * np.max(np.abs(z[f"Z_{pol}"]))
            x = bem2d.solve(a, b)
            assert np.max(np.abs(x[:n] - z[f"density_{pol}"])) < 1e-12 * np.max(np.abs(x[:n]))
            amp = bem2d.farfield(om, k0, angles, 1j * k0 * x[:n])
            assert np.max(np.abs(amp - z[f"amp_{pol}"]) / np.abs(z[f"amp_{pol}"])) < 1e-12


def test_oracle_lossy_regions_match_reference():
    """The oracle's regional system with lossy media (coating, two layers, airfoil.geo's medium)
    against the reference's solves."""
    from grim_20261007_amd.mesh import build_mesh
    from oracle import bem2d
    for name in ("lossy_region_coated_pec", "lossy_region_two_layer", "lossy_airfoil_coating"):
        snap, f, angles, z = _golden(name)
        mesh = build_mesh(snap, "meters", f)
        om = bem2d.mesh_from_nodes(z["nodes"], z["elements"])
        layout = bem2d.region_layout(om, [tuple(r) for r in mesh.regions], mesh.region_media)
        k0 = float(z["k0"])
        for pol in ("TE", "TM"):
            a, b, exterior = bem2d.region_system(om, layout, k0, pol, angles)
            if f"Z_{pol}" in z:
                assert np.max(np.abs(a - z[f"Z_{pol}"])) < 1e-14 * np.max(np.abs(z[f"Z_{pol}"]))
            assert np.max(np.abs(b - z[f"rhs_{pol}"])) < 1e-15 * np.max(np.abs(b))
            rho = bem2d.exterior_density(om, exterior, bem2d.solve(a, b))
            assert np.max(np.abs(rho - z[f"density_{pol}"])) < 1e-12 * np.max(np.abs(rho))
            amp = bem2d.farfield(om, k0, angles, rho)
            assert np.max(np.abs(amp - z[f"amp_{pol}"]) / np.abs(z[f"amp_{pol}"])) < 1e-12, name


def test_impedance_mie_pins():
    """The impedance cylinder series tends to the PEC series as Zs -> 0; a coating of eps = mu = 1
    over an impedance core is the bare impedance cylinder of the core's radius."""
    from oracle import mie
    k = 20.0
    for pol in ("TM", "TE"):
        pec = mie.pec_backscatter_amplitude(0.1, k, pol)
        near = mie.impedance_backscatter_amplitude(0.1, k, 1e-9 + 1e-9j, pol)
        assert abs(near - pec) <= 1e-8 * abs(pec), pol
        for zs in (50 + 20j, 120 - 60j, 5j):
            bare = mie.impedance_backscatter_amplitude(0.07, k, zs, pol)
            coated = mie.coated_impedance_backscatter_amplitude(0.07, 0.1, k, 1.0, 1.0, zs, pol)
            assert abs(coated - bare) <= 1e-12 * abs(bare), (pol, zs)
