"""Bistatic grids against the reference's bistatic solve (tests/golden/bistatic_*.npz), and, at
the NACA 0012 N = 2000 size, the diagonal of a bistatic grid against the monostatic solve."""
import numpy as np
import pytest

from tests import golden_cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("name", ["bistatic_circle64_ka1", "bistatic_naca80"])
def test_bistatic_matches_reference(name):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd.solver import solve_bistatic_rcs_2d
    case = golden_cases.load(name)
    inc, obs = case["incidence_deg"].tolist(), case["observation_deg"].tolist()
    print(f"{name}: {len(inc)} incidences x {len(obs)} observations", flush=True)
    res = solve_bistatic_rcs_2d(golden_cases.snapshot(case), [float(case["frequency_ghz"])], inc, obs,
                                geometry_units="meters", max_panels=100000)
    assert sorted(res) == sorted(case["result_keys"].tolist())
    assert sorted(res["samples"][0]) == sorted(case["sample_keys"].tolist())
    assert res["scattering_mode"] == str(case["scattering_mode"])
    for export in ("VV", "HH"):
        rows = res["co_solved_samples"][export]
        assert [r["theta_inc_deg"] for r in rows] == case[f"inc_{export}"].tolist()
        assert [r["theta_scat_deg"] for r in rows] == case[f"obs_{export}"].tolist()
        amp = np.array([complex(r["rcs_amp_real"], r["rcs_amp_imag"]) for r in rows]).reshape(len(inc), len(obs))
        ref = case[f"amp_{export}"]
        err = np.max(np.abs(amp - ref)) / np.max(np.abs(ref))
        print(f"{name} {export}: max |B - B_ref| / max |B_ref| = {err:.2e}", flush=True)
        assert err < 1e-10
    flat = res["samples"]
    assert [r["polarization"] for r in flat[:4]] == ["VV", "HH", "VV", "HH"]


def test_bistatic_diagonal_is_monostatic_naca2000():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd import workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import solve_bistatic_frequency, solve_frequency
    pts = workloads.naca0012_points(1000)
    f = workloads.frequency_for_segments_per_wavelength(pts, 10.0)
    mesh = build_mesh(workloads._snapshot(pts, "naca"), "meters", f)
    angles = np.linspace(0.0, 357.0, 64).tolist()
    print(f"NACA N={mesh.n}: bistatic 64 x 64 against monostatic", flush=True)
    bi = solve_bistatic_frequency(mesh, f, angles, angles)
    mono = solve_frequency(mesh, f, angles)
    for pol in ("TE", "TM"):
        diag = np.diagonal(bi[pol]["amplitude"])
        ref = mono[pol]["amplitude"]
        err = np.max(np.abs(diag - ref) / np.abs(ref))
        print(f"{pol}: diagonal vs monostatic {err:.2e}", flush=True)
        assert err < 1e-12
        # every column of the grid is one solve: the residual gate holds per incidence
        assert np.max(bi[pol]["residual"]) < 1e-10
