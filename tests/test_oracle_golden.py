"""The CPU oracle against the reference solver's own outputs (golden vectors)."""
import numpy as np
import pytest

from oracle import bem2d
from tests import golden_cases


@pytest.mark.parametrize("name", golden_cases.ALL)
def test_oracle_reproduces_reference(name):
    case = golden_cases.load(name)
    mesh = bem2d.mesh_from_points(case["points"])
    assert np.array_equal(mesh.elements, case["elements"])
    assert np.array_equal(mesh.nodes, case["nodes"])
    k = bem2d.frequency_k(float(case["frequency_ghz"]))
    assert abs(k - case["k0"]) <= 1e-15 * k
    z_tm, z_te = bem2d.assemble(mesh, k)
    for pol, z in (("TM", z_tm), ("TE", z_te)):
        if f"Z_{pol}" in case:
            ref = case[f"Z_{pol}"]
            assert np.max(np.abs(z - ref)) / np.max(np.abs(ref)) < 1e-13
        else:
            rows = case["matrix_rows"]
            ref = case[f"Zrows_{pol}"]
            assert np.max(np.abs(z[rows] - ref)) / np.max(np.abs(ref)) < 1e-13
        angles = case["angles_deg"]
        b = bem2d.rhs(mesh, k, angles, pol)
        cols = case["density_columns"]
        if f"rhs_{pol}" in case:
            assert np.max(np.abs(b[:, cols] - case[f"rhs_{pol}"])) / np.max(np.abs(b)) < 1e-13
        x = bem2d.solve(z, b)
        xr = case[f"density_{pol}"]
        assert np.max(np.abs(x[:, cols] - xr)) / np.max(np.abs(xr)) < 1e-11
        amp = bem2d.farfield(mesh, k, angles, x)
        ar = case[f"sample_amp_{pol}"]
        assert np.max(np.abs(amp - ar) / np.abs(ar)) < 1e-11
        db = bem2d.db_from_sigma(bem2d.sigma_from_amplitude(amp, k))
        assert np.max(np.abs(db - case[f"sample_db_{pol}"])) < 1e-9


def test_near_pair_census_covers_every_rule():
    """The fixtures exercise self, touching, fixed and adaptive near pairs."""
    seen = set()
    for name in golden_cases.SMALL:
        case = golden_cases.load(name)
        census = bem2d.near_pair_census(bem2d.mesh_from_points(case["points"]))
        seen |= {k for k, v in census.items() if v}
    assert seen == {"self", "touching", "fixed", "adaptive"}
