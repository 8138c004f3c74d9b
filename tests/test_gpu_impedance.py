"""Contours with a constant surface impedance on the GPU: the public entry against the reference's
solves of the same snapshots (tests/golden/ibc_*.npz) in both precision forms, and the serving
loop against the public entry."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["ibc_circle64_ka1", "ibc_circle64_resistive", "ibc_pec_and_ibc", "ibc_naca80"]


def _case(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    return json.loads(str(z["snapshot_json"])), float(z["frequency_ghz"]), z["angles_deg"].tolist(), z


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("name", CASES)
def test_impedance_matches_reference(name, precision):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd.solver import solve_monostatic_rcs_2d
    snap, f, angles, z = _case(name)
    print(f"{name} {precision}", flush=True)
    res = solve_monostatic_rcs_2d(snap, [f], angles, geometry_units="meters", max_panels=100000,
                                  solver_method="direct", precision=precision)
    for export in ("VV", "HH"):
        rows = res["co_solved_samples"][export]
        amp = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in rows])
        ref = z[f"amp_{export}"]
        err = np.max(np.abs(amp - ref) / np.abs(ref))
        db = np.max(np.abs(np.array([r["rcs_db"] for r in rows]) - z[f"db_{export}"]))
        print(f"{name} {precision} {export}: amplitude {err:.2e}, dB {db:.2e}", flush=True)
        assert err < 1e-10 and db < 1e-8
        assert max(r["linear_residual"] for r in rows) < 1e-10


def test_serving_impedance_request_matches_public_entry():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.serving import Request, Server
    snap, f, angles, z = _case("ibc_pec_and_ibc")
    mesh = build_mesh(snap, "meters", f)
    srv = Server(workers=2)
    try:
        got = srv.submit(Request(mesh, f, np.asarray(angles), precision="double")).result()
        again = srv.submit(Request(mesh, f, np.asarray(angles[:2]), precision="double")).result()
    finally:
        srv.close()
    assert again.cached
    for export, internal in (("VV", "TE"), ("HH", "TM")):
        ref = z[f"amp_{export}"]
        assert np.max(np.abs(got.amplitude[internal] - ref) / np.abs(ref)) < 1e-10
        assert np.max(np.abs(again.amplitude[internal] - ref[:2]) / np.abs(ref[:2])) < 1e-10


def test_impedance_bistatic_diagonal_is_monostatic():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd.solver import solve_bistatic_rcs_2d
    snap, f, angles, z = _case("ibc_pec_and_ibc")
    res = solve_bistatic_rcs_2d(snap, [f], angles, angles, geometry_units="meters", max_panels=100000)
    n = len(angles)
    for export in ("VV", "HH"):
        rows = res["co_solved_samples"][export]
        grid = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in rows]).reshape(n, n)
        ref = z[f"amp_{export}"]
        err = np.max(np.abs(np.diagonal(grid) - ref) / np.abs(ref))
        print(f"bistatic diagonal {export}: {err:.2e}", flush=True)
        assert err < 1e-10
