"""Lossy media on the GPU against the reference's solves of the same snapshots
(tests/golden/lossy_*.npz, written by make_golden.py --lossy): the device complex Hankel functions
against SciPy's hankel2 (the reference's own evaluation), the dielectric and regional systems at a
complex interior wavenumber entry by entry, densities and amplitudes through the public entry in
both precision forms, and the lossy dielectric cylinder against the Mie series as the mesh refines."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["lossy_diel_circle48", "lossy_diel_magnetic64", "lossy_region_coated_pec", "lossy_region_two_layer",
         "lossy_airfoil_coating"]


def _case(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    return json.loads(str(z["snapshot_json"])), float(z["frequency_ghz"]), z["angles_deg"].tolist(), z


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_complex_hankel_matches_scipy():
    """(j/4) H0^(2)(z) and (j/4) H1^(2)(z) over the fourth quadrant, |z| from 1e-6 to 300, both
    forms (power series below |z| = 2, Gauss-Laguerre integral above), against SciPy."""
    _need_gpu()
    from scipy.special import hankel2
    from grim_20261007_amd import native
    native.ensure_device(0)
    rng = np.random.default_rng(7)
    mag = np.exp(rng.uniform(np.log(1e-6), np.log(300.0), 20000))
    ang = rng.uniform(-0.4999 * np.pi, 0.0, 20000)
    z = mag * np.exp(1j * ang)
    z = np.concatenate((z, [2.0 - 1e-13, 2.0 + 1e-13, 1.9999999 - 0.5j, 1.5 - 1.3229j, 20.0, 1e-6, 0.25 - 0.25j]))
    dz = torch.as_tensor(z, device="cuda")
    out = torch.empty((len(z), 4), dtype=torch.float64, device="cuda")
    assert native.lib().grim_hankel2_probe(len(z), dz.data_ptr(), out.data_ptr(),
                                            torch.cuda.current_stream().cuda_stream) == 0
    got = out.cpu().numpy()
    g = got[:, 0] + 1j * got[:, 1]
    h = got[:, 2] + 1j * got[:, 3]
    for name, val, ref in (("H0", g, 0.25j * hankel2(0, z)), ("H1", h, 0.25j * hankel2(1, z))):
        keep = np.abs(ref) > 1e-280                       # past underflow both are zero
        err = np.abs(val[keep] - ref[keep]) / np.abs(ref[keep])
        print(f"{name}: max rel err {err.max():.2e} over {keep.sum()} points", flush=True)
        assert err.max() < 5e-14
        assert np.all(np.abs(val[~keep]) < 1e-250)


@pytest.mark.parametrize("name", ["lossy_diel_circle48", "lossy_region_coated_pec"])
def test_lossy_systems_match_reference(name):
    _need_gpu()
    from grim_20261007_amd import engine
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import wavenumber
    snap, f, angles, z = _case(name)
    mesh = build_mesh(snap, "meters", f)
    dev = torch.device("cuda:0")
    k = wavenumber(f)
    contour = engine.DeviceContour(mesh, dev)
    ws = engine.assembly_workspace(mesh.n, dev)
    if mesh.regions is not None:
        plan = engine.region_plan(mesh)
        pair = engine.region_systems(contour, k, ws, plan)
    else:
        pair = engine.dielectric_systems(contour, k, ws)
    torch.cuda.synchronize()
    systems = pair.transpose(1, 2).cpu().numpy()
    for i, pol in enumerate(("TE", "TM")):
        ref = z[f"Z_{pol}"]
        err = np.max(np.abs(systems[i] - ref)) / np.max(np.abs(ref))
        print(f"{name} {pol}: system max rel err {err:.2e}", flush=True)
        assert err < 1e-11


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("name", CASES)
def test_lossy_public_entry_matches_reference(name, precision):
    _need_gpu()
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import solve_frequency, solve_monostatic_rcs_2d
    snap, f, angles, z = _case(name)
    res = solve_monostatic_rcs_2d(snap, [f], angles, geometry_units="meters", max_panels=100000,
                                  solver_method="direct", precision=precision)
    tol = 1e-10 if precision == "double" else 1e-8
    for export, pol in (("VV", "TE"), ("HH", "TM")):
        rows = res["co_solved_samples"][export]
        amp = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in rows])
        ref = z[f"sample_amp_{pol}"]
        err = np.max(np.abs(amp - ref)) / np.max(np.abs(ref))
        print(f"{name} {precision} {export}: amplitude max rel err {err:.2e}", flush=True)
        assert err < tol
    out = solve_frequency(build_mesh(snap, "meters", f), f, angles, ("TE", "TM"), precision, return_density=True)
    for pol in ("TE", "TM"):
        want = z[f"density_{pol}"]
        err = np.max(np.abs(out[pol]["density"] - want)) / np.max(np.abs(want))
        print(f"{name} {precision} {pol}: density max rel err {err:.2e}", flush=True)
        assert err < tol


def test_lossy_dielectric_cylinder_converges_to_mie():
    """eps = 4 - 1j cylinder, ka = 1: the solve's backscatter amplitude approaches the lossy
    dielectric Mie series (the reference's validation/cylinder.py:133-191) at second order."""
    _need_gpu()
    from grim_20261007_amd import workloads
    from grim_20261007_amd.solver import solve_monostatic_rcs_2d, wavenumber
    from oracle import mie
    r = 0.1
    f = workloads.frequency_for_ka(1.0, r)
    errs = {}
    for pol, export in (("TM", "HH"), ("TE", "VV")):
        want = mie.dielectric_backscatter_amplitude(r, wavenumber(f), 4.0 - 1.0j, 1.0 + 0j, pol)
        errs[pol] = []
        for count in (48, 96, 192):
            snap = workloads._snapshot(workloads.circle_points(r, count), "c")
            seg = snap["segments"][0]
            seg["seg_type"] = 3
            seg["properties"] = ["3", "1", "0", "1", "0"]
            snap["dielectrics"] = [["1", "4", "-1", "1", "0"]]
            res = solve_monostatic_rcs_2d(snap, [f], [0.0], geometry_units="meters", max_panels=100000)
            row = res["co_solved_samples"][export][0]
            got = complex(row["rcs_amp_real"], row["rcs_amp_imag"])
            errs[pol].append(abs(got - want) / abs(want))
        print(pol, ["%.2e" % e for e in errs[pol]], flush=True)
        assert errs[pol][2] < errs[pol][1] < errs[pol][0]
        assert errs[pol][1] / errs[pol][2] > 3.0            # second order: about 4 per halving
        assert errs[pol][2] < 5e-3


def test_lossy_served_and_bistatic():
    """A lossy coated cylinder served (alone, then from the factor cache with the same angles) gives
    the direct solve's amplitudes bit for bit; the bistatic grid's diagonal is the monostatic
    amplitude."""
    _need_gpu()
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.serving import Request, Server
    from grim_20261007_amd.solver import solve_bistatic_rcs_2d, solve_frequency
    snap, f, angles, z = _case("lossy_region_coated_pec")
    mesh = build_mesh(snap, "meters", f)
    direct = solve_frequency(mesh, f, angles, ("TE", "TM"), "double")
    srv = Server(workers=2, batch=4)
    try:
        first = srv.submit(Request(build_mesh(snap, "meters", f), f, np.asarray(angles), precision="double",
                                   key="lossy")).result()
        again = srv.submit(Request(build_mesh(snap, "meters", f), f, np.asarray(angles), precision="double",
                                   key="lossy")).result()
    finally:
        srv.close()
    assert again.cached and not first.cached
    for res in (first, again):
        for pol in ("TE", "TM"):
            assert np.array_equal(res.amplitude[pol], direct[pol]["amplitude"])
    bi = solve_bistatic_rcs_2d(snap, [f], angles, angles, geometry_units="meters", max_panels=100000)
    for export, pol in (("VV", "TE"), ("HH", "TM")):
        rows = [r for r in bi["samples"] if r["polarization"] == export and r["theta_inc_deg"] == r["theta_scat_deg"]]
        diag = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in rows])
        err = np.max(np.abs(diag - z[f"sample_amp_{pol}"])) / np.max(np.abs(z[f"sample_amp_{pol}"]))
        print(f"bistatic diagonal {pol}: {err:.2e}", flush=True)
        assert err < 1e-10
