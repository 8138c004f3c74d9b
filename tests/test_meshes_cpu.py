"""Mesh builder against the reference's panel builder: base meshes (explicit N, panels per
wavelength, automatic density) and the fine meshes of certification pairs, plus the fine-count
rule on synthetic chains (fixtures from tests/golden/make_golden.py --meshes)."""
import json
from pathlib import Path

import numpy as np
import pytest

from grim_20261007_amd.certify import certification_fine_snapshot
from grim_20261007_amd.mesh import build_mesh, refined_chain_counts

GOLDEN = np.load(Path(__file__).parent / "golden" / "meshes.npz")
CASES = [str(c) for c in GOLDEN["mesh_cases"]]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tag", ["base", "fine"])
def test_mesh_matches_reference_panels(case, tag):
    snap = json.loads(str(GOLDEN[f"{case}_snapshot_json"]))
    freq, factor = GOLDEN[f"{case}_freq_factor"]
    if tag == "fine":
        snap = certification_fine_snapshot(snap, float(factor))
    mesh = build_mesh(snap, "meters", float(freq), max_panels=100000)
    p0, p1 = GOLDEN[f"{case}_{tag}_p0"], GOLDEN[f"{case}_{tag}_p1"]
    assert mesh.n == len(p0), (case, tag, mesh.n, len(p0))
    np.testing.assert_allclose(mesh.p0, p0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(mesh.p1, p1, rtol=0, atol=1e-14)


def test_fine_count_rule_matches_reference_chains():
    chains = json.loads(str(GOLDEN["chains_json"]))
    for chain in chains:
        got = refined_chain_counts(chain["base"], chain["lengths"], chain["factor"])
        assert got == chain["fine"], chain


def test_fine_count_rule_invariants():
    rng = np.random.default_rng(5)
    for _ in range(200):
        m = int(rng.integers(1, 30))
        base = rng.integers(1, 8, m).tolist()
        lengths = rng.uniform(0.01, 2.0, m).tolist()
        factor = float(rng.uniform(1.01, 3.0))
        fine = refined_chain_counts(base, lengths, factor)
        target = max(sum(base) + 1, int(np.ceil(sum(base) * factor)))
        # never coarser than the base on any primitive, at least the target in total, and only
        # a whole equal-length group (at most 1 % of the target, two panels minimum) beyond it
        assert all(f >= b for f, b in zip(fine, base))
        assert target <= sum(fine) <= target + max(2, int(np.ceil(0.01 * target)))
    # one primitive takes the whole target; a symmetric chain stays symmetric
    assert refined_chain_counts([4], [1.0], 1.5) == [6]
    sym = refined_chain_counts([1] * 8, [1.0, 2.0, 3.0, 4.0, 4.0, 3.0, 2.0, 1.0], 1.5)
    assert sym == sym[::-1]
    # 64 equal chords at factor 1.5: the 32 first in bit-reversed order (the even indices) split
    circle = refined_chain_counts([1] * 64, [1.0] * 64, 1.5)
    assert circle[0::2] == [2] * 32 and circle[1::2] == [1] * 32
