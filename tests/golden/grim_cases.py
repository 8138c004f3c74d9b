"""Result dicts of the 2-D monostatic solve used as inputs for the `.grim` archive fixtures
(tests/golden/make_golden.py --grim writes the reference exporter's archives for them;
tests/test_grim_export_cpu.py writes this build's and compares)."""
from __future__ import annotations

import math

import numpy as np

C0 = 299_792_458.0


def _rows(rng, freqs, angles, pol, internal):
    rows = []
    for f in freqs:
        k = 2.0 * math.pi * f * 1e9 / C0
        for a in angles:
            b = complex(rng.standard_normal(), rng.standard_normal()) * 0.3
            if a == angles[-1] and f == freqs[0]:
                b = 0j                                    # a zero amplitude: phase 0
            lin = (b.real * b.real + b.imag * b.imag) / (4.0 * k)
            rows.append({"frequency_ghz": float(f), "theta_inc_deg": float(a), "theta_scat_deg": float(a),
                         "rcs_linear": lin, "rcs_db": 10.0 * math.log10(max(lin, 1e-12)),
                         "rcs_amp_real": b.real, "rcs_amp_imag": b.imag,
                         "rcs_amp_phase_deg": math.degrees(math.atan2(b.imag, b.real)),
                         "linear_residual": float(abs(rng.standard_normal()) * 1e-13),
                         "polarization": pol, "polarization_internal": internal})
    return rows


def cases():
    rng = np.random.default_rng(20261204)
    freqs = [1.5, 0.75]
    angles = [0.0, 90.0, 45.0, 180.0, 270.0, 359.5]
    vv = _rows(rng, freqs, angles, "VV", "TE")
    hh = _rows(rng, freqs, angles, "HH", "TM")
    meta = {"frequency_metadata": [{"frequency_ghz": 1.5, "panel_count": 400, "condition": float("inf")},
                                   {"frequency_ghz": 0.75, "panel_count": 200, "note": float("nan")}],
            "solver_method": "direct", "precision": "double", "tags": {"b", "a"},
            "complex_probe": complex(1.0, -2.0), "array": np.arange(3.0), "numpy_scalar": np.float32(0.5)}
    co = {"solver": "ghost-2d-bem", "scattering_mode": "monostatic",
          "amplitude_convention": "A_physical_asymptotic = +j * B_stored", "amplitude_version": 2,
          "rcs_log_unit": "dBke", "rcs_linear_quantity": "sigma_2d", "polarizations": ["VV", "HH"],
          "polarization_mapping": {"VV": "TE", "HH": "TM"}, "samples": vv + hh,
          "co_solved_samples": {"VV": vv, "HH": hh}, "metadata": meta}
    tm = _rows(rng, [2.0], [10.0, 20.0, 30.0], "TM", "TM")
    single = {"solver": "ghost-2d-bem", "scattering_mode": "monostatic", "rcs_log_unit": "dBke",
              "rcs_linear_quantity": "sigma_2d", "polarization": "TM", "samples": tm, "metadata": {"run": 1}}
    return {"co_polarized": (co, {"source_path": "geometry.json", "history": "case co"}),
            "single_tm": (single, {"history": "case single"})}


def error_cases():
    """Results the exporter must refuse, each with the reference's message."""
    base, _ = cases()["single_tm"]

    def variant(**change):
        out = dict(base)
        out.update(change)
        return out

    rows = base["samples"]
    bad_power = [dict(rows[0], rcs_linear=rows[0]["rcs_linear"] * 1.001)] + rows[1:]
    co, _ = cases()["co_polarized"]
    return {
        "duplicate_sample": variant(samples=rows + [dict(rows[1])]),
        "inconsistent_power": variant(samples=bad_power),
        "incomplete_grid": variant(samples=rows + [dict(
            rows[0], frequency_ghz=3.0,
            rcs_linear=(rows[0]["rcs_amp_real"] ** 2 + rows[0]["rcs_amp_imag"] ** 2)
            / (4.0 * 2.0 * math.pi * 3.0e9 / C0))]),
        "nonfinite_amplitude": variant(samples=[dict(rows[0], rcs_amp_real=float("nan"))] + rows[1:]),
        "no_samples": variant(samples=[]),
        "co_without_mapping": dict(co, polarization_mapping={}),
        "co_single_channel": dict(co, polarizations=["VV"]),
    }
