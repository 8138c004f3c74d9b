"""Sheets (TYPE 1 open contours, alone and beside a PEC body) on the GPU against the reference's own
solves (tests/golden/sheet_*.npz, written by make_golden.py --sheets): each polarization's system
matrix (pinned strip-end rows included) and right-hand sides through the C ABI, the public entries'
amplitudes and densities in both precision forms, the bistatic diagonal, the transparent-sheet
limit of the reference's acceptance test, and the refusal of the boundary-density export."""
import json

import numpy as np
import pytest

from tests import golden_cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _layout(case):
    from grim_20261007_amd.mesh import build_mesh
    return build_mesh(json.loads(str(case["snapshot_json"])), "meters", float(case["frequency_ghz"]))


@pytest.mark.parametrize("name", golden_cases.SHEETS)
@pytest.mark.parametrize("pol", ["TE", "TM"])
def test_sheet_system_and_loads_match_reference(name, pol):
    from grim_20261007_amd import engine
    dev = _dev()
    case = golden_cases.load(name)
    mesh = _layout(case).mesh(pol)
    k = float(case["k0"])
    print(f"{name} {pol}: {mesh.ne} panels, {mesh.n} nodes, ends {mesh.endpoints.tolist()}", flush=True)
    sheet = engine.DeviceSheet(mesh, dev)
    z = engine.sheet_system(sheet, k, pol)
    got = z.cpu().numpy().T
    ref = case[f"Z_{pol}"]
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"{name} {pol}: matrix max rel err {err:.2e}", flush=True)
    assert err < 1e-12
    angles = torch.tensor(case["angles_deg"], dtype=torch.float64, device=dev)
    bu, bdn = engine.sheet_loads(sheet, k, angles)
    rhs = engine.sheet_rhs(sheet, pol, bu, bdn).cpu().numpy().T
    err = np.max(np.abs(rhs - case[f"rhs_{pol}"])) / np.max(np.abs(case[f"rhs_{pol}"]))
    print(f"{name} {pol}: rhs max rel err {err:.2e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("name", golden_cases.SHEETS)
def test_sheet_public_entry_matches_reference(name, precision):
    from grim_20261007_amd.solver import solve_frequency, solve_monostatic_rcs_2d
    _dev()
    case = golden_cases.load(name)
    snap = json.loads(str(case["snapshot_json"]))
    f, angles = float(case["frequency_ghz"]), [float(a) for a in case["angles_deg"]]
    result = solve_monostatic_rcs_2d(snap, [f], angles, geometry_units="meters", strict_quality_gate=False,
                                     precision=precision)
    tol = 1e-10 if precision == "double" else 1e-8
    for export, pol in (("VV", "TE"), ("HH", "TM")):
        rows = result["co_solved_samples"][export]
        amp = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in rows])
        ref = case[f"sample_amp_{pol}"]
        err = np.max(np.abs(amp - ref)) / np.max(np.abs(ref))
        db = np.max(np.abs(np.array([r["rcs_db"] for r in rows]) - case[f"sample_db_{pol}"]))
        print(f"{name} {pol} {precision}: amplitude max rel err {err:.2e}, dB {db:.2e}", flush=True)
        assert err < tol
    res = solve_frequency(_layout(case), f, angles, precision=precision, return_density=True)
    for pol in ("TE", "TM"):
        ref = case[f"density_{pol}"]
        err = np.max(np.abs(res[pol]["density"] - ref)) / np.max(np.abs(ref))
        print(f"{name} {pol} {precision}: density max rel err {err:.2e}", flush=True)
        assert err < tol
        assert res[pol]["formulation"].startswith("2D sheet BIE" if "pec" not in name else "2D mixed sheet+PEC")


def test_sheet_single_polarization_and_bistatic():
    from grim_20261007_amd.solver import solve_bistatic_rcs_2d, solve_monostatic_rcs_2d_single_polarization
    _dev()
    case = golden_cases.load("sheet_beside_pec")
    snap = json.loads(str(case["snapshot_json"]))
    f, angles = float(case["frequency_ghz"]), [float(a) for a in case["angles_deg"]]
    for pol in ("TE", "TM"):
        out = solve_monostatic_rcs_2d_single_polarization(snap, [f], angles, pol, geometry_units="meters",
                                                          strict_quality_gate=False)
        amp = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in out["samples"]])
        assert np.max(np.abs(amp - case[f"sample_amp_{pol}"])) < 1e-10 * np.max(np.abs(amp))
        assert out["metadata"]["formulation"].startswith("2D mixed sheet+PEC BIE (" + pol)
    bi = solve_bistatic_rcs_2d(snap, [f], angles, angles, geometry_units="meters", strict_quality_gate=False)
    for export, pol in (("VV", "TE"), ("HH", "TM")):
        rows = [r for r in bi["samples"] if r["polarization"] == export]
        diag = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in rows
                         if r["theta_inc_deg"] == r["theta_scat_deg"]])
        ref = case[f"sample_amp_{pol}"]
        err = np.max(np.abs(diag - ref)) / np.max(np.abs(ref))
        print(f"bistatic diagonal {pol}: max rel err {err:.2e}", flush=True)
        assert err < 1e-10


def test_sheet_transparency_limit():
    """The reference's acceptance check (tests/test_2d_capability_acceptance.py:303-322): a 48-panel
    strip of 1e10 ohm scatters under 1e-5 of what the 75 ohm strip does, at 0.8 GHz."""
    from grim_20261007_amd.solver import solve_monostatic_rcs_2d
    _dev()
    case = golden_cases.load("sheet_strip48")
    snap = json.loads(str(case["snapshot_json"]))
    angles = [-35.0, 0.0, 40.0]
    ordinary = solve_monostatic_rcs_2d(snap, [0.8], angles, geometry_units="meters", strict_quality_gate=False)
    snap["ibcs"] = [["1", "constant", "1e10", "0", "0", "0"]]
    transparent = solve_monostatic_rcs_2d(snap, [0.8], angles, geometry_units="meters", strict_quality_gate=False)
    for export in ("VV", "HH"):
        ref = np.array([abs(complex(r["rcs_amp_real"], r["rcs_amp_imag"]))
                        for r in ordinary["co_solved_samples"][export]])
        lim = np.array([abs(complex(r["rcs_amp_real"], r["rcs_amp_imag"]))
                        for r in transparent["co_solved_samples"][export]])
        print(f"{export}: transparent / ordinary {lim.max() / ref.max():.2e}", flush=True)
        assert lim.max() < 1e-5 * ref.max()


def test_sheet_boundary_densities_refused():
    from grim_20261007_amd.solver import compute_boundary_densities
    _dev()
    case = golden_cases.load("sheet_strip48")
    with pytest.raises(ValueError, match="not supported by compute_boundary_densities"):
        compute_boundary_densities(json.loads(str(case["snapshot_json"])), 0.8, 0.0, "TM", geometry_units="meters")


def test_sheet_served_request():
    """A sheet request in the serving loop runs on its own (no batching, no factor cache) and gives
    the direct solve's amplitudes, beside a plain PEC request in the same server."""
    from grim_20261007_amd import workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.serving import Request, Server
    from grim_20261007_amd.solver import solve_frequency
    _dev()
    case = golden_cases.load("sheet_two_named")
    f = float(case["frequency_ghz"])
    angles = np.linspace(0.0, 350.0, 36)
    direct = solve_frequency(_layout(case), f, angles.tolist(), ("TE", "TM"))
    pec = build_mesh(workloads._snapshot(workloads.circle_points(0.1, 64), "c"), "meters", 1.0)
    srv = Server(workers=2, batch=4)
    try:
        futs = [srv.submit(Request(_layout(case), f, angles, precision="double", rid=1)),
                srv.submit(Request(pec, 1.0, angles, precision="double", rid=2))]
        out = [fu.result() for fu in futs]
    finally:
        srv.close()
    assert not out[0].cached
    for pol in ("TE", "TM"):
        assert np.array_equal(out[0].amplitude[pol], direct[pol]["amplitude"])
        assert np.all(np.isfinite(out[1].amplitude[pol]))


@pytest.mark.parametrize("name", ["sheet_strip48", "lossy_diel_circle48"])
def test_certified_pair_on_sheet_and_lossy_layouts(name):
    """The certified base/fine pair runs on a sheet layout and on a lossy dielectric body: the gate
    passes on these resolved meshes and the published samples are the fine solve's, bit for bit."""
    from grim_20261007_amd.certify import certification_fine_snapshot, solve_monostatic_rcs_2d_certified
    from grim_20261007_amd.solver import solve_monostatic_rcs_2d
    _dev()
    case = golden_cases.load(name)
    snap = json.loads(str(case["snapshot_json"]))
    f, angles = float(case["frequency_ghz"]), [float(a) for a in case["angles_deg"]]
    cert = solve_monostatic_rcs_2d_certified(snap, [f], angles, geometry_units="meters", max_panels=100000)
    for export in ("VV", "HH"):
        gate = cert["metadata"]["mesh_convergence"][export]
        print(name, export, {k: gate[k] for k in gate if "max" in k}, flush=True)
        assert gate["passed"]
    fine = solve_monostatic_rcs_2d(certification_fine_snapshot(snap, 1.5), [f], angles, geometry_units="meters",
                                   max_panels=100000, compute_condition_number=True)
    for a, b in zip(cert["samples"], fine["samples"]):
        assert a["rcs_amp_real"] == b["rcs_amp_real"] and a["rcs_amp_imag"] == b["rcs_amp_imag"]
