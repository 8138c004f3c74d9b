"""Certified base/fine pair: two co-polarized solves, the second on a mesh refined by
``fine_factor``, compared point by point in magnitude, complex field and phase; only a
converged fine result is published.

Mirrors the reference's certified production entry
(``tools/GHOST/ghost_backend/twod/solver.py:3663-3878`` and ``:3979-4005``) with its
mesh-convergence gate (``ghost_backend/runs/quality.py:23-35, 166-520``).  The two
solves of a pair run on two HIP streams at the same time: the base and fine systems have
different N, so they cannot share launches, but their kernels overlap on the device.
"""
from __future__ import annotations

import copy
import math
import threading
from typing import Any, Callable, Dict, List, Optional

import numpy as np

from .solver import CO_POLARIZED_CHANNELS, MAX_PANELS_DEFAULT, solve_monostatic_rcs_2d

MESH_CONVERGENCE_LEVEL_FLOOR = 3.0e-2          # runs/quality.py:23
DEFAULT_PANELS_PER_WAVELENGTH = 20              # twod/solver.py (auto density)
PRODUCTION_MESH_CONVERGENCE_DEFAULTS = {        # runs/quality.py:25-35
    "fine_factor": 1.5,
    "rms_limit_db": 1.0,
    "max_abs_limit_db": 3.0,
    "complex_rms_limit": 0.02,
    "complex_max_limit": 0.05,
    "phase_rms_limit_deg": 5.0,
    "phase_max_limit_deg": 15.0,
    "phase_floor_relative": MESH_CONVERGENCE_LEVEL_FLOOR,
    "db_floor_relative": MESH_CONVERGENCE_LEVEL_FLOOR,
}
_NONNEGATIVE = ("rms_limit_db", "max_abs_limit_db", "complex_rms_limit", "complex_max_limit",
                "phase_rms_limit_deg", "phase_max_limit_deg")
_FRACTIONS = ("phase_floor_relative", "db_floor_relative")
_TINY = np.finfo(float).tiny


def scale_snapshot_panel_density(snapshot: Dict[str, Any], fine_factor: float) -> Dict[str, Any]:
    """Deep copy with the per-segment discretization ``properties[1]`` refined by ``fine_factor``
    (runs/quality.py:166-216): explicit N -> max(N + 1, ceil(N f)); panels per wavelength -p ->
    -max(p + 1, ceil(p f)); automatic (0) -> -max(2, ceil(20 f))."""
    factor = float(fine_factor)
    if factor <= 1.0:
        raise ValueError("fine_factor must be > 1.0.")
    out = copy.deepcopy(snapshot)
    segments = list(out.get("segments", []) or [])
    for seg in segments:
        props = [str(p) for p in list(seg.get("properties", []) or [])]
        props += [""] * (6 - len(props))
        try:
            base_n = int(round(float(str(props[1]).strip() or 0)))
        except Exception:
            base_n = 0
        if base_n == 0:
            props[1] = str(-max(2, int(math.ceil(DEFAULT_PANELS_PER_WAVELENGTH * factor))))
        elif base_n > 0:
            props[1] = str(max(base_n + 1, int(math.ceil(base_n * factor))))
        else:
            ppw = abs(base_n)
            props[1] = str(-max(ppw + 1, int(math.ceil(ppw * factor))))
        seg["properties"] = props
        seg["seg_type"] = props[0] if props[0] else seg.get("seg_type")
    out["segments"] = segments
    return out


def certification_fine_snapshot(snapshot: Dict[str, Any], fine_factor: float) -> Dict[str, Any]:
    """The fine snapshot of a certification pair (twod/solver.py:3747-3759): the density-scaled
    copy, tagged with the factor and each segment's base N so that the mesh builder refines
    every chain from its base panel counts."""
    fine = scale_snapshot_panel_density(snapshot, fine_factor)
    fine["_2d_certification_refinement_factor"] = float(fine_factor)
    fine["_2d_certification_base_segment_n"] = [
        (list(seg.get("properties", []) or []) + [0, 0])[1] for seg in snapshot.get("segments", []) or []]
    return fine


def validate_mesh_convergence_policy(policy: Optional[Dict[str, Any]] = None) -> Dict[str, float]:
    """Complete, finite, fail-closed policy (runs/quality.py:219-262)."""
    supplied = dict(policy or {})
    unknown = sorted(set(supplied) - set(PRODUCTION_MESH_CONVERGENCE_DEFAULTS))
    if unknown:
        raise ValueError("Unknown mesh convergence policy field(s): " + ", ".join(unknown))
    merged = dict(PRODUCTION_MESH_CONVERGENCE_DEFAULTS)
    merged.update(supplied)
    out = {}
    for name, raw in merged.items():
        try:
            value = float(raw)
        except (TypeError, ValueError, OverflowError) as exc:
            raise ValueError(f"{name} must be a finite number.") from exc
        if not math.isfinite(value):
            raise ValueError(f"{name} must be a finite number.")
        out[name] = value
    if out["fine_factor"] <= 1.0:
        raise ValueError("fine_factor must be > 1.0.")
    for name in _NONNEGATIVE:
        if out[name] < 0.0:
            raise ValueError(f"{name} must be nonnegative.")
    for name in _FRACTIONS:
        if not 0.0 <= out[name] < 1.0:
            raise ValueError(f"{name} must be in [0, 1).")
    return out


def _coordinates(row):
    return (float(row.get("frequency_ghz", 0.0)), float(row.get("theta_inc_deg", 0.0)),
            float(row.get("theta_scat_deg", 0.0)))


def _key(row):
    return tuple(round(v, 9) for v in _coordinates(row))


def _finite(row, name):
    if name not in row:
        raise ValueError("Mesh convergence requires authoritative complex amplitudes; "
                         f"sample {_key(row)} is missing {name!r}.")
    value = float(row[name])
    if not math.isfinite(value):
        raise ValueError(f"Mesh convergence sample {_key(row)} has non-finite {name}={row[name]!r}.")
    return value


def _sample_db(row):
    if "rcs_db" in row:
        value = float(row["rcs_db"])
    else:
        value = 10.0 * math.log10(max(_finite(row, "rcs_linear"), 1e-12))
    if not math.isfinite(value):
        raise ValueError(f"Mesh convergence sample {_key(row)} has non-finite rcs_db={value!r}.")
    return value


def _matched(base_samples, fine_samples):
    """Pairs of samples matched on (frequency, incidence, scattering), 1e-9 rounding."""
    def group(samples, label):
        grouped, seen = {}, set()
        for row in samples:
            c = _coordinates(row)
            if c in seen:
                raise ValueError(f"Mesh convergence {label} contains duplicate exact sample coordinates {c}.")
            seen.add(c)
            grouped.setdefault(_key(row), []).append(row)
        for rows in grouped.values():
            rows.sort(key=_coordinates)
        return grouped

    base, fine = group(base_samples, "base_result"), group(fine_samples, "fine_result")
    keys = set(base) | set(fine)
    missing = sum(max(0, len(base.get(k, ())) - len(fine.get(k, ()))) for k in keys)
    extra = sum(max(0, len(fine.get(k, ())) - len(base.get(k, ()))) for k in keys)
    if missing or extra:
        raise ValueError(f"Mesh convergence sample grids differ: {missing} missing and {extra} extra "
                         "fine-result sample point(s).")
    pairs = []
    for k in sorted(base):
        pairs.extend(zip(base[k], fine[k]))
    if not pairs:
        raise ValueError("Mesh convergence comparison produced no overlapping samples.")
    return pairs


def evaluate_mesh_convergence(base_result: Dict[str, Any], fine_result: Dict[str, Any], rms_limit_db: float,
                              max_abs_limit_db: float, complex_rms_limit: float = 0.02,
                              complex_max_limit: float = 0.05, phase_rms_limit_deg: float = 5.0,
                              phase_max_limit_deg: float = 15.0,
                              phase_floor_relative: float = MESH_CONVERGENCE_LEVEL_FLOOR,
                              db_floor_relative: float = MESH_CONVERGENCE_LEVEL_FLOOR) -> Dict[str, Any]:
    """Point-by-point comparison of two results (runs/quality.py:265-520).

    Per frequency with peak P = max |A| over both meshes: complex change |dA| / P; dB change
    of 20 log10(|A| + db_floor P); phase change where both |A| > phase_floor P."""
    policy = validate_mesh_convergence_policy({
        "rms_limit_db": rms_limit_db, "max_abs_limit_db": max_abs_limit_db,
        "complex_rms_limit": complex_rms_limit, "complex_max_limit": complex_max_limit,
        "phase_rms_limit_deg": phase_rms_limit_deg, "phase_max_limit_deg": phase_max_limit_deg,
        "phase_floor_relative": phase_floor_relative, "db_floor_relative": db_floor_relative})
    base_samples = base_result.get("samples", []) or []
    fine_samples = fine_result.get("samples", []) or []
    if not base_samples or not fine_samples:
        raise ValueError("Both base_result and fine_result must contain samples.")
    pairs = _matched(base_samples, fine_samples)
    base_amp = np.array([complex(_finite(a, "rcs_amp_real"), _finite(a, "rcs_amp_imag")) for a, _ in pairs])
    fine_amp = np.array([complex(_finite(b, "rcs_amp_real"), _finite(b, "rcs_amp_imag")) for _, b in pairs])
    unfloored = np.array([_sample_db(a) - _sample_db(b) for a, b in pairs], dtype=float)
    groups: Dict[float, List[int]] = {}
    for i, (a, _) in enumerate(pairs):
        groups.setdefault(round(float(a.get("frequency_ghz", 0.0)), 9), []).append(i)

    count = len(pairs)
    cerr = np.zeros(count)
    deltas = np.zeros(count)
    above = np.zeros(count, dtype=bool)
    phase_groups = []
    peaks = {}
    for freq, idx in sorted(groups.items()):
        idx = np.asarray(idx, dtype=int)
        ba, fa = np.abs(base_amp[idx]), np.abs(fine_amp[idx])
        peak = float(max(np.max(ba), np.max(fa)))
        peaks[str(freq)] = peak
        if not math.isfinite(peak) or peak <= 0.0:
            continue
        cerr[idx] = np.abs(base_amp[idx] - fine_amp[idx]) / peak
        floor = max(policy["db_floor_relative"] * peak, _TINY)
        deltas[idx] = 20.0 * (np.log10(ba + floor) - np.log10(fa + floor))
        above[idx] = np.maximum(ba, fa) >= policy["db_floor_relative"] * peak
        pmask = (ba > policy["phase_floor_relative"] * peak) & (fa > policy["phase_floor_relative"] * peak)
        if np.any(pmask):
            phase_groups.append(np.abs(np.degrees(np.angle(fine_amp[idx][pmask] *
                                                           np.conj(base_amp[idx][pmask])))))
    phase = np.concatenate(phase_groups) if phase_groups else np.zeros(0)
    absd = np.abs(deltas)
    rms_db = float(np.sqrt(np.mean(deltas * deltas)))
    max_abs_db = float(np.max(absd))
    crms = float(np.sqrt(np.mean(cerr ** 2)))
    cmax = float(np.max(cerr))
    prms = float(np.sqrt(np.mean(phase ** 2))) if phase.size else 0.0
    pmax = float(np.max(phase)) if phase.size else 0.0
    checks = ((rms_db, rms_limit_db, "RMS dB delta {v:.6g} exceeds limit {l:.6g}"),
              (max_abs_db, max_abs_limit_db, "Max |dB| delta {v:.6g} exceeds limit {l:.6g}"),
              (crms, complex_rms_limit, "RMS normalized complex-field delta {v:.6g} exceeds limit {l:.6g}"),
              (cmax, complex_max_limit, "Max normalized complex-field delta {v:.6g} exceeds limit {l:.6g}"),
              (prms, phase_rms_limit_deg, "RMS phase delta {v:.6g} deg exceeds limit {l:.6g} deg"),
              (pmax, phase_max_limit_deg, "Max phase delta {v:.6g} deg exceeds limit {l:.6g} deg"))
    violations = [msg.format(v=v, l=float(lim)) for v, lim, msg in checks if v > float(lim)]
    return {
        "passed": not violations, "sample_count": int(count), "rms_db": rms_db, "max_abs_db": max_abs_db,
        "mean_db": float(np.mean(deltas)), "median_abs_db": float(np.median(absd)),
        "db_metric": "20*log10(|A| + db_floor_relative*peak) per frequency",
        "db_sample_count_above_floor": int(np.count_nonzero(above)),
        "rms_db_unfloored": float(np.sqrt(np.mean(unfloored ** 2))),
        "max_abs_db_unfloored": float(np.max(np.abs(unfloored))),
        "complex_rms_normalized": crms, "complex_max_normalized": cmax,
        "complex_reference_peak_amplitude": float(max(peaks.values())),
        "complex_reference_peak_amplitude_by_frequency": peaks,
        "phase_sample_count": int(phase.size), "phase_rms_deg": prms, "phase_max_deg": pmax,
        "limits": {"rms_db": float(rms_limit_db), "max_abs_db": float(max_abs_limit_db),
                   "complex_rms_normalized": float(complex_rms_limit),
                   "complex_max_normalized": float(complex_max_limit),
                   "phase_rms_deg": float(phase_rms_limit_deg), "phase_max_deg": float(phase_max_limit_deg),
                   "phase_floor_relative": float(policy["phase_floor_relative"]),
                   "db_floor_relative": float(policy["db_floor_relative"])},
        "violations": violations,
        "reason": "; ".join(violations) if violations else "mesh convergence passed",
    }


def _channel(result: Dict[str, Any], export: str) -> Dict[str, Any]:
    value = {k: v for k, v in result.items() if k != "co_solved_samples"}
    value["samples"] = result["co_solved_samples"][export]
    return value


def _run_pair(solve: Callable[..., Dict[str, Any]], base_snapshot, fine_snapshot, kwargs, device):
    """Base and fine solves on two streams of one device, overlapping on the GPU."""
    import torch
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    results: Dict[str, Any] = {}
    errors: Dict[str, BaseException] = {}

    def run(name, snap):
        try:
            torch.cuda.set_device(dev)
            stream = torch.cuda.Stream(device=dev)
            with torch.cuda.stream(stream):
                results[name] = solve(geometry_snapshot=snap, device=dev, **kwargs)
            stream.synchronize()
        except BaseException as exc:   # surfaced below, base first
            errors[name] = exc

    threads = [threading.Thread(target=run, args=("base", base_snapshot)),
               threading.Thread(target=run, args=("fine", fine_snapshot))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for name in ("base", "fine"):
        if name in errors:
            raise errors[name]
    return results["base"], results["fine"]


def solve_monostatic_rcs_2d_certified(
        geometry_snapshot: Dict[str, Any], frequencies_ghz: List[float], elevations_deg: List[float],
        geometry_units: str = "inches", material_base_dir: Optional[str] = None,
        progress_callback: Optional[Callable[[int, int, str], None]] = None,
        quality_thresholds: Optional[Dict[str, Any]] = None,
        mesh_convergence_policy: Optional[Dict[str, Any]] = None, max_panels: int = MAX_PANELS_DEFAULT,
        mesh_reference_ghz: Optional[float] = None, rcs_normalization_mode: str = "physical",
        abort_event=None, solver_method: str = "auto", *, execution_options=None, precision: str = "double",
        device=None) -> Dict[str, Any]:
    """Certified co-polarized monostatic solve (twod/solver.py:3979-4005, 3685-3878).

    Both channels (VV <- TE, HH <- TM) must pass the mesh-convergence gate; the fine result is
    published with ``metadata["mesh_convergence"]`` per channel.  A failed gate raises
    ``ValueError("Certified 2-D mesh convergence failed: ...")``."""
    policy = validate_mesh_convergence_policy(mesh_convergence_policy)
    freqs = [float(f) for f in frequencies_ghz]
    if len(freqs) > 1:
        # each frequency is certified on its own and the results are merged in request order
        # (twod/solver.py:3096-3134)
        if len(set(freqs)) != len(freqs):
            raise ValueError("Duplicate frequencies are not supported in a co-polarized result grid.")
        parts = [solve_monostatic_rcs_2d_certified(
            geometry_snapshot, [f], elevations_deg, geometry_units, material_base_dir, None, quality_thresholds,
            mesh_convergence_policy, max_panels, mesh_reference_ghz, rcs_normalization_mode, abort_event,
            solver_method, execution_options=execution_options, precision=precision, device=device)
            for f in freqs]
        merged = {k: v for k, v in parts[0].items() if k not in ("samples", "co_solved_samples", "metadata")}
        merged["samples"] = [row for part in parts for row in part["samples"]]
        merged["co_solved_samples"] = {pol: [row for part in parts for row in part["co_solved_samples"][pol]]
                                       for pol in ("VV", "HH")}
        merged["metadata"] = {"frequency_metadata": [part["metadata"] for part in parts],
                              "mesh_convergence_certified": True, "certified_entry_point": True,
                              "published_mesh": "fine"}
        return merged
    # the certified entry always estimates the condition number, so its strict gate applies the
    # condition limit too (twod/solver.py:3734)
    kwargs = dict(frequencies_ghz=frequencies_ghz, elevations_deg=elevations_deg,
                  geometry_units=geometry_units, material_base_dir=material_base_dir,
                  quality_thresholds=quality_thresholds, strict_quality_gate=True, compute_condition_number=True,
                  max_panels=max_panels, mesh_reference_ghz=mesh_reference_ghz, rcs_normalization_mode=rcs_normalization_mode,
                  abort_event=abort_event, solver_method=solver_method, precision=precision,
                  execution_options=execution_options)
    fine_snapshot = certification_fine_snapshot(geometry_snapshot, policy["fine_factor"])
    base, fine = _run_pair(solve_monostatic_rcs_2d, geometry_snapshot, fine_snapshot, kwargs, device)
    if progress_callback is not None:
        try:
            progress_callback(2, 2, "Base and fine meshes solved")
        except Exception:
            pass
    base_panels = int(sum(base["metadata"].get("panel_counts", [0])))
    fine_panels = int(sum(fine["metadata"].get("panel_counts", [0])))
    if base_panels > 0 and fine_panels <= base_panels:
        raise ValueError("Certified 2-D mesh refinement failed: the fine solve used "
                         f"{fine_panels} panels versus {base_panels} on the base mesh. A mesh-convergence "
                         "certificate requires a genuinely refined discretization.")
    gates = {}
    limits = {k: v for k, v in policy.items() if k != "fine_factor"}
    for export, _internal in CO_POLARIZED_CHANNELS:
        gate = evaluate_mesh_convergence(_channel(base, export), _channel(fine, export), **limits)
        gate.update({"schema": "ghost.solver.mesh-convergence.v1", "fine_factor": policy["fine_factor"],
                     "published_mesh": "fine", "geometry_model": "piecewise_linear_input",
                     "geometry_approximation_certified": False, "refinement_kind": "mesh",
                     "base_panel_count": base_panels, "fine_panel_count": fine_panels,
                     "base_polynomial_degree": int(base["metadata"].get("polynomial_degree", 1)),
                     "fine_polynomial_degree": int(fine["metadata"].get("polynomial_degree", 1)),
                     "panel_refinement_ratio": float(fine_panels) / float(base_panels) if base_panels else
                     float("nan")})
        if not gate["passed"]:
            raise ValueError("Certified 2-D mesh convergence failed: " + str(gate["reason"]))
        gates[export] = gate
    result = fine
    meta = result.setdefault("metadata", {})
    meta["mesh_convergence"] = gates
    meta["mesh_convergence_certified"] = True
    meta["certified_entry_point"] = True
    meta["published_mesh"] = "fine"
    meta["certification_solve_order"] = ["base_TE", "base_TM", "fine_TE", "fine_TM"]
    quality_gate = meta.get("quality_gate")
    if isinstance(quality_gate, dict):
        quality_gate["mesh_convergence_certified"] = True
        quality_gate["certification_scope"] = "discrete_linear_system_and_mesh_convergence"
    return result
