"""Drop-in 2-D monostatic RCS entry points backed by the GRIM-2D HIP library.

Same call signatures, inputs and result layout as the reference's
``ghost_backend.twod.solver`` functions (tools/GHOST/ghost_backend/twod/solver.py):

* ``solve_monostatic_rcs_2d``                       (:3187-3255): VV <- TE, HH <- TM
* ``solve_monostatic_rcs_2d_single_polarization``   (:1984-2712)
* ``compute_boundary_densities``                    (:4150-4346)
* ``solve_bistatic_rcs_2d`` / ``..._single_polarization`` (:3262-3660): incidence x observation
  grids, the projection one complex GEMM on the library's FP64 MFMA kernel

for closed PEC contours in air with the plain dense linear-Galerkin combined
potential (``solver_method='direct'`` semantics).  Extra keyword ``precision``:
``'double'`` (complex128 LU, the reference's default) or ``'mixed'``
(complex64 LU + FP64 residual refinement, the reference's
``linear_precision('mixed')``).  ``device`` selects the GPU.
"""
from __future__ import annotations

import cmath
import math
from typing import Any, Callable, Dict, List, Optional

import numpy as np
import torch

from . import engine
from .mesh import C0, build_mesh, unit_scale

RCS_DB_FLOOR = 1e-12
AMPLITUDE_CONVENTION = "A_physical_asymptotic = +j * B_stored"
CO_POLARIZED_CHANNELS = (("VV", "TE"), ("HH", "TM"))
MAX_PANELS_DEFAULT = 20_000


def wavenumber(frequency_ghz: float) -> float:
    return 2.0 * math.pi * float(frequency_ghz) * 1e9 / C0


def rcs_db(sigma):
    sigma = np.asarray(sigma, float)
    return 10.0 * np.log10(np.maximum(sigma, RCS_DB_FLOOR))


def _validate(frequencies_ghz, elevations_deg):
    if not frequencies_ghz:
        raise ValueError("At least one frequency is required.")
    if not elevations_deg:
        raise ValueError("At least one elevation angle is required.")
    freqs = [float(f) for f in frequencies_ghz]
    angs = [float(a) for a in elevations_deg]
    if any((not math.isfinite(f)) or f <= 0.0 for f in freqs):
        raise ValueError("Frequencies must be positive finite GHz values.")
    if any(not math.isfinite(a) for a in angs):
        raise ValueError("Elevation angles must all be finite.")
    return freqs, angs


def _device(device):
    if device is None:
        if not torch.cuda.is_available():
            raise RuntimeError("The GRIM-2D solver needs a ROCm GPU (torch.cuda.is_available() is False).")
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def _systems(contour, k, z_te, z_tm, bu, bdn):
    """[TE, TM] systems (2, N, N) and right-hand sides (2, A, N): the PEC combined potential
    (TM Dirichlet rows -b_u, TE flux rows -b_dn), or, for contours with a surface impedance,
    the flux rows dn(u) + alpha u of both polarizations (engine.impedance_rows)."""
    pair = z_tm._base if z_tm._base is not None else torch.stack((z_te, z_tm))   # [TE, TM]
    if contour.mesh.impedance is None:
        return pair, torch.stack((-bdn, -bu))
    return engine.impedance_systems(contour, k, pair), engine.impedance_rhs(contour, k, bu, bdn)


def solve_frequency(mesh, frequency_ghz: float, angles_deg: List[float], pols=("TE", "TM"),
                    precision: str = "double", device=None, return_density: bool = False,
                    abort_event=None) -> Dict[str, Dict[str, Any]]:
    """Assemble both systems once, then factor, solve and project each requested polarization."""
    dev = _device(device)
    k = wavenumber(frequency_ghz)
    contour = engine.DeviceContour(mesh, dev)
    z_tm, z_te = engine.assemble(contour, k)
    angles = torch.tensor(angles_deg, dtype=torch.float64, device=dev)
    bu, bdn = engine.loads(contour, k, angles)
    if abort_event is not None and abort_event.is_set():
        raise InterruptedError("Solve canceled by user.")
    pair, rhs_all = _systems(contour, k, z_te, z_tm, bu, bdn)
    order = [p for p in ("TE", "TM") if p in pols]
    sel = [("TE", "TM").index(p) for p in order]
    systems = pair if len(order) == 2 else pair[sel[0]].unsqueeze(0)
    rhs = rhs_all if len(order) == 2 else rhs_all[sel[0]].unsqueeze(0)
    fac = engine.factor(systems.contiguous(), precision)
    x, info = engine.solve(fac, rhs)
    rel = engine.relative_residual(fac.z, x, rhs, info.get("residual"))
    out = {}
    for i, pol in enumerate(order):
        amp, sigma = engine.farfield(x[i], bu, bdn, k)
        res = {"amplitude": amp.cpu().numpy(), "sigma": sigma.cpu().numpy(),
               "residual": rel[i].cpu().numpy(), "info": {kk: v for kk, v in info.items() if kk != "residual"},
               "k": k, "n": mesh.n}
        if return_density:
            res["density"] = x[i].cpu().numpy().T      # (N, A) nodal densities
        out[pol] = res
    return out


def solve_bistatic_frequency(mesh, frequency_ghz: float, incidence_deg: List[float], observation_deg: List[float],
                             pols=("TE", "TM"), precision: str = "double", device=None,
                             abort_event=None) -> Dict[str, Dict[str, Any]]:
    """Both systems from one assembly; solve for every incidence, then project every solution
    on every observation direction (twod/solver.py:3476-3517 with observation angles)."""
    dev = _device(device)
    k = wavenumber(frequency_ghz)
    contour = engine.DeviceContour(mesh, dev)
    z_tm, z_te = engine.assemble(contour, k)
    inc = torch.tensor(incidence_deg, dtype=torch.float64, device=dev)
    obs = torch.tensor(observation_deg, dtype=torch.float64, device=dev)
    bu, bdn = engine.loads(contour, k, inc)
    if abort_event is not None and abort_event.is_set():
        raise InterruptedError("Solve canceled by user.")
    pair, rhs_all = _systems(contour, k, z_te, z_tm, bu, bdn)
    order = [p for p in ("TE", "TM") if p in pols]
    sel = [("TE", "TM").index(p) for p in order]
    systems = pair if len(order) == 2 else pair[sel[0]].unsqueeze(0)
    rhs = rhs_all if len(order) == 2 else rhs_all[sel[0]].unsqueeze(0)
    fac = engine.factor(systems.contiguous(), precision)
    x, info = engine.solve(fac, rhs)
    rel = engine.relative_residual(fac.z, x, rhs, info.get("residual"))
    bu_o, bdn_o = engine.loads(contour, k, obs)
    amp, sigma = engine.bistatic_farfield(x, bu_o, bdn_o, k)
    out = {}
    for i, pol in enumerate(order):
        out[pol] = {"amplitude": amp[i].cpu().numpy(), "sigma": sigma[i].cpu().numpy(),
                    "residual": rel[i].cpu().numpy(),
                    "info": {kk: v for kk, v in info.items() if kk != "residual"}, "k": k, "n": mesh.n}
    return out


def _bistatic_samples(frequency_ghz, incidence, observation, amp, sigma, residual):
    """Rows in request order: incidence, then observation (twod/solver.py:3519-3537)."""
    rows = []
    dbs = rcs_db(sigma)
    for i, a in enumerate(incidence):
        for j, o in enumerate(observation):
            z = complex(amp[i, j])
            rows.append({"frequency_ghz": float(frequency_ghz), "theta_inc_deg": float(a),
                         "theta_scat_deg": float(o), "rcs_linear": float(sigma[i, j]), "rcs_db": float(dbs[i, j]),
                         "rcs_amp_real": float(z.real), "rcs_amp_imag": float(z.imag),
                         "rcs_amp_phase_deg": float(math.degrees(cmath.phase(z))),
                         "linear_residual": float(residual[i])})
    return rows


def _samples(frequency_ghz, angles, amp, sigma, residual):
    rows = []
    dbs = rcs_db(sigma)
    for i, a in enumerate(angles):
        z = complex(amp[i])
        rows.append({
            "frequency_ghz": float(frequency_ghz),
            "theta_inc_deg": float(a),
            "theta_scat_deg": float(a),
            "rcs_linear": float(sigma[i]),
            "rcs_db": float(dbs[i]),
            "rcs_amp_real": float(z.real),
            "rcs_amp_imag": float(z.imag),
            "rcs_amp_phase_deg": float(math.degrees(cmath.phase(z))),
            "linear_residual": float(residual[i]),
        })
    return rows


def _normalize_pol(polarization: str) -> str:
    pol = str(polarization).strip().upper()
    if pol not in ("TE", "TM"):
        raise ValueError(f"Unsupported polarization '{polarization}'. Use TE or TM.")
    return pol


def solve_monostatic_rcs_2d_single_polarization(
        geometry_snapshot: Dict[str, Any], frequencies_ghz: List[float], elevations_deg: List[float],
        polarization: str, geometry_units: str = "inches", material_base_dir: Optional[str] = None,
        progress_callback: Optional[Callable[[int, int, str], None]] = None,
        quality_thresholds: Optional[Dict[str, Any]] = None, strict_quality_gate: bool = True,
        compute_condition_number: bool = False, max_panels: int = MAX_PANELS_DEFAULT,
        mesh_reference_ghz: Optional[float] = None, rcs_normalization_mode: str = "physical",
        cfie_alpha: float = 0.0, abort_event=None, solver_method: str = "auto", *,
        execution_options=None, precision: str = "double", device=None) -> Dict[str, Any]:
    freqs, angs = _validate(frequencies_ghz, elevations_deg)
    if float(cfie_alpha) != 0.0:
        raise ValueError("cfie_alpha is not implemented by any active 2-D formulation; use cfie_alpha=0.")
    if str(rcs_normalization_mode).strip().lower() != "physical":
        raise ValueError("Only the physical RCS normalization is supported.")
    pol = _normalize_pol(polarization)
    unit_scale(geometry_units)
    samples = []
    residuals = []
    total = len(freqs) * (len(angs) + 1)
    done = 0
    n_used = []
    for f in freqs:
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve canceled by user.")
        mesh = build_mesh(geometry_snapshot, geometry_units, mesh_reference_ghz or f, max_panels)
        res = solve_frequency(mesh, f, angs, (pol,), precision, device, abort_event=abort_event)[pol]
        samples += _samples(f, angs, res["amplitude"], res["sigma"], res["residual"])
        residuals.append(float(np.max(res["residual"])))
        n_used.append(mesh.n)
        done += len(angs) + 1
        if progress_callback is not None:
            try:
                progress_callback(done, total, f"Solved {f:g} GHz")
            except Exception:
                pass
    max_res = max(residuals)
    gate_limit = float((quality_thresholds or {}).get("max_linear_residual", 1e-6))
    passed = max_res <= gate_limit
    if strict_quality_gate and not passed:
        raise ValueError(f"Quality gate failed: linear residual {max_res:.3e} exceeds {gate_limit:.1e}.")
    metadata = {"solver_backend": "grim_20261007_amd (HIP, gfx950)", "precision": precision,
                "formulation": "2D combined regional potential S + gamma D (linear Galerkin)",
                "panel_counts": n_used, "residual_max": max_res,
                "quality_gate": {"passed": passed, "max_linear_residual": gate_limit},
                "solver_method": solver_method}
    return {"solver": "2d_bie_mom_rcs", "scattering_mode": "monostatic",
            "amplitude_convention": AMPLITUDE_CONVENTION, "amplitude_version": 2,
            "rcs_log_unit": "dBke", "rcs_linear_quantity": "sigma_2d", "polarization": pol,
            "samples": samples, "metadata": metadata}


def solve_monostatic_rcs_2d(
        geometry_snapshot: Dict[str, Any], frequencies_ghz: List[float], elevations_deg: List[float],
        geometry_units: str = "inches", material_base_dir: Optional[str] = None,
        progress_callback: Optional[Callable[[int, int, str], None]] = None,
        quality_thresholds: Optional[Dict[str, Any]] = None, strict_quality_gate: bool = True,
        compute_condition_number: bool = False, max_panels: int = MAX_PANELS_DEFAULT,
        mesh_reference_ghz: Optional[float] = None, rcs_normalization_mode: str = "physical",
        abort_event=None, solver_method: str = "auto", *, execution_options=None,
        precision: str = "double", device=None) -> Dict[str, Any]:
    """Co-polarized solve: both channels share one fused assembly per frequency."""
    freqs, angs = _validate(frequencies_ghz, elevations_deg)
    if str(rcs_normalization_mode).strip().lower() != "physical":
        raise ValueError("Only the physical RCS normalization is supported.")
    unit_scale(geometry_units)
    co = {"VV": [], "HH": []}
    flat = []
    residual_max = 0.0
    counts = []
    total = 2 * len(freqs) * (len(angs) + 1)
    done = 0
    for f in freqs:
        mesh = build_mesh(geometry_snapshot, geometry_units, mesh_reference_ghz or f, max_panels)
        counts.append(mesh.n)
        res = solve_frequency(mesh, f, angs, ("TE", "TM"), precision, device, abort_event=abort_event)
        per = {}
        for export, internal in CO_POLARIZED_CHANNELS:
            r = res[internal]
            rows = _samples(f, angs, r["amplitude"], r["sigma"], r["residual"])
            for row in rows:
                row["polarization"] = export
                row["polarization_internal"] = internal
            per[export] = rows
            co[export] += rows
            residual_max = max(residual_max, float(np.max(r["residual"])))
        for vv, hh in zip(per["VV"], per["HH"]):
            flat += [vv, hh]
        done += 2 * (len(angs) + 1)
        if progress_callback is not None:
            try:
                progress_callback(done, total, f"Solved {f:g} GHz")
            except Exception:
                pass
    gate_limit = float((quality_thresholds or {}).get("max_linear_residual", 1e-6))
    passed = residual_max <= gate_limit
    if strict_quality_gate and not passed:
        raise ValueError(f"Quality gate failed: linear residual {residual_max:.3e} exceeds {gate_limit:.1e}.")
    metadata = {"solver_backend": "grim_20261007_amd (HIP, gfx950)", "precision": precision,
                "panel_counts": counts, "residual_max": residual_max,
                "quality_gate": {"passed": passed, "max_linear_residual": gate_limit},
                "solver_method": solver_method}
    return {"solver": "2d_bie_mom_rcs", "scattering_mode": "monostatic",
            "amplitude_convention": AMPLITUDE_CONVENTION, "amplitude_version": 2,
            "rcs_log_unit": "dBke", "rcs_linear_quantity": "sigma_2d",
            "polarizations": ["VV", "HH"], "polarization_mapping": {"VV": "TE", "HH": "TM"},
            "samples": flat, "co_solved_samples": co, "metadata": metadata}


def _validate_bistatic(frequencies_ghz, incidence_angles_deg, observation_angles_deg):
    if not frequencies_ghz:
        raise ValueError("At least one frequency is required.")
    if not incidence_angles_deg:
        raise ValueError("At least one incidence angle is required.")
    if not observation_angles_deg:
        raise ValueError("At least one observation angle is required.")
    freqs = [float(f) for f in frequencies_ghz]
    inc = [float(a) for a in incidence_angles_deg]
    obs = [float(a) for a in observation_angles_deg]
    if any((not math.isfinite(f)) or f <= 0.0 for f in freqs):
        raise ValueError("Frequencies must be positive finite GHz values.")
    if any(not math.isfinite(a) for a in inc):
        raise ValueError("Incidence angles must all be finite.")
    if any(not math.isfinite(a) for a in obs):
        raise ValueError("Observation angles must all be finite.")
    return freqs, inc, obs


def _bistatic_run(geometry_snapshot, frequencies_ghz, incidence_angles_deg, observation_angles_deg, pols,
                  geometry_units, progress_callback, quality_thresholds, strict_quality_gate, max_panels,
                  mesh_reference_ghz, abort_event, precision, device):
    freqs, inc, obs = _validate_bistatic(frequencies_ghz, incidence_angles_deg, observation_angles_deg)
    unit_scale(geometry_units)
    rows = {p: [] for p in pols}
    counts, residual_max = [], 0.0
    total, done = len(freqs) * len(inc), 0
    for f in freqs:
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve canceled by user.")
        mesh = build_mesh(geometry_snapshot, geometry_units, mesh_reference_ghz or f, max_panels)
        counts.append(mesh.n)
        res = solve_bistatic_frequency(mesh, f, inc, obs, pols, precision, device, abort_event=abort_event)
        for p in pols:
            r = res[p]
            rows[p] += _bistatic_samples(f, inc, obs, r["amplitude"], r["sigma"], r["residual"])
            residual_max = max(residual_max, float(np.max(r["residual"])))
        done += len(inc)
        if progress_callback is not None:
            try:
                progress_callback(done, total, f"Bistatic {f:g} GHz")
            except Exception:
                pass
    gate_limit = float((quality_thresholds or {}).get("max_linear_residual", 1e-6))
    passed = residual_max <= gate_limit
    if strict_quality_gate and not passed:
        raise ValueError(f"Quality gate failed: linear residual {residual_max:.3e} exceeds {gate_limit:.1e}.")
    metadata = {"solver_backend": "grim_20261007_amd (HIP, gfx950)", "precision": precision,
                "formulation": "bistatic 2D BIE/MoM", "panel_count": max(counts), "panel_count_min": min(counts),
                "panel_count_max": max(counts), "panel_counts": counts, "residual_norm_max": residual_max,
                "quality_gate": {"passed": passed, "max_linear_residual": gate_limit}, "solver_method": "dense_lu",
                "amplitude_version": 2}
    return rows, metadata


def solve_bistatic_rcs_2d_single_polarization(
        geometry_snapshot: Dict[str, Any], frequencies_ghz: List[float], incidence_angles_deg: List[float],
        observation_angles_deg: List[float], polarization: str, geometry_units: str = "inches",
        material_base_dir: Optional[str] = None, progress_callback: Optional[Callable[[int, int, str], None]] = None,
        quality_thresholds: Optional[Dict[str, Any]] = None, strict_quality_gate: bool = True,
        compute_condition_number: bool = False, max_panels: int = MAX_PANELS_DEFAULT,
        mesh_reference_ghz: Optional[float] = None, cfie_alpha: float = 0.0, abort_event=None,
        solver_method: str = "auto", *, precision: str = "double", device=None) -> Dict[str, Any]:
    """One channel on an incidence x observation grid (twod/solver.py:3262-3600): samples in
    request order (frequency, incidence, observation), theta_scat_deg = observation angle."""
    if float(cfie_alpha) != 0.0:
        raise ValueError("cfie_alpha is not implemented by any active 2-D formulation; use cfie_alpha=0.")
    pol = _normalize_pol(polarization)
    rows, metadata = _bistatic_run(geometry_snapshot, frequencies_ghz, incidence_angles_deg,
                                   observation_angles_deg, (pol,), geometry_units, progress_callback,
                                   quality_thresholds, strict_quality_gate, max_panels, mesh_reference_ghz,
                                   abort_event, precision, device)
    metadata["solver_method_requested"] = str(solver_method)
    return {"solver": "2d_bie_mom_rcs", "scattering_mode": "bistatic", "amplitude_convention": AMPLITUDE_CONVENTION,
            "amplitude_version": 2, "polarization": pol, "polarization_export": pol, "samples": rows[pol],
            "metadata": metadata}


def solve_bistatic_rcs_2d(
        geometry_snapshot: Dict[str, Any], frequencies_ghz: List[float], incidence_angles_deg: List[float],
        observation_angles_deg: List[float], geometry_units: str = "inches",
        material_base_dir: Optional[str] = None, progress_callback: Optional[Callable[[int, int, str], None]] = None,
        quality_thresholds: Optional[Dict[str, Any]] = None, strict_quality_gate: bool = True,
        compute_condition_number: bool = False, max_panels: int = MAX_PANELS_DEFAULT,
        mesh_reference_ghz: Optional[float] = None, abort_event=None, *, precision: str = "double",
        device=None) -> Dict[str, Any]:
    """Both co-polarized bistatic channels, VV <- TE and HH <- TM, from one assembly per frequency
    (twod/solver.py:3603-3660); ``samples`` interleaves VV then HH per physical sample."""
    rows, metadata = _bistatic_run(geometry_snapshot, frequencies_ghz, incidence_angles_deg,
                                   observation_angles_deg, ("TE", "TM"), geometry_units, progress_callback,
                                   quality_thresholds, strict_quality_gate, max_panels, mesh_reference_ghz,
                                   abort_event, precision, device)
    co = {}
    for export, internal in CO_POLARIZED_CHANNELS:
        for row in rows[internal]:
            row["polarization"] = export
            row["polarization_internal"] = internal
        co[export] = rows[internal]
    flat = [r for pair in zip(co["VV"], co["HH"]) for r in pair]
    metadata["co_solve_shared_discretization"] = True
    return {"solver": "2d_bie_mom_rcs", "scattering_mode": "bistatic", "amplitude_convention": AMPLITUDE_CONVENTION,
            "amplitude_version": 2, "rcs_log_unit": "dBke", "rcs_linear_quantity": "sigma_2d",
            "polarizations": ["VV", "HH"], "polarization_mapping": {"VV": "TE", "HH": "TM"},
            "samples": flat, "co_solved_samples": co, "metadata": metadata}


def compute_boundary_densities(geometry_snapshot: Dict[str, Any], frequency_ghz: float, elevation_deg: float,
                               polarization: str, geometry_units: str = "inches",
                               material_base_dir: Optional[str] = None, max_panels: int = MAX_PANELS_DEFAULT,
                               mesh_reference_ghz: Optional[float] = None, *, precision: str = "double",
                               device=None) -> Dict[str, Any]:
    """Layer-potential density of one incidence: nodal values and element-centre samples."""
    pol = _normalize_pol(polarization)
    mesh = build_mesh(geometry_snapshot, geometry_units, mesh_reference_ghz or frequency_ghz, max_panels)
    res = solve_frequency(mesh, float(frequency_ghz), [float(elevation_deg)], (pol,), precision, device,
                          return_density=True)[pol]
    nodal = res["density"][:, 0]
    centre = 0.5 * (nodal + nodal[mesh.next])
    centers = 0.5 * (mesh.p0 + mesh.p1)
    normals = mesh.normals
    return {"polarization": pol, "frequency_ghz": float(frequency_ghz), "elevation_deg": float(elevation_deg),
            "centers_x": centers[:, 0].tolist(), "centers_y": centers[:, 1].tolist(),
            "normals_x": normals[:, 0].tolist(), "normals_y": normals[:, 1].tolist(),
            "lengths": mesh.lengths.tolist(),
            "density_real": centre.real.tolist(), "density_imag": centre.imag.tolist(),
            "density_abs": np.abs(centre).tolist(), "density_phase_deg": np.degrees(np.angle(centre)).tolist(),
            "nodal_density": nodal, "linear_residual": float(res["residual"][0])}
