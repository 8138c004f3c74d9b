"""Loaders of the committed golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = ["circle64_ka1", "naca80", "star96", "star128_aspect4"]
ALL = SMALL + ["c1_circle400_ka10"]
# degree-2 and degree-3 Galerkin bases (execution option basis_order)
POLY = [f"poly{d}_{case}" for d in (2, 3) for case in ("circle32_ka1", "naca48", "star40")]


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False))


def snapshot(case):
    from grim_20261007_amd.workloads import _snapshot
    return _snapshot(case["points"], "contour")

# TYPE 1 sheets alone and beside a PEC body (make_golden.py --sheets)
SHEETS = ["sheet_bent192", "sheet_strip48", "sheet_two_named", "sheet_beside_pec", "sheet_taper_auto", "sheet_ring"]
