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
import time
from typing import Any, Callable, Dict, List, Optional

import numpy as np
import torch

from . import engine
from .mesh import C0, SheetLayout, build_mesh, causal_index, unit_scale

RCS_DB_FLOOR = 1e-12
AMPLITUDE_CONVENTION = "A_physical_asymptotic = +j * B_stored"
CO_POLARIZED_CHANNELS = (("VV", "TE"), ("HH", "TM"))
MAX_PANELS_DEFAULT = 20_000
MEMORY_LIMIT_FRACTION = 0.9        # twod/solver.py:1416
FORMULATION = "2D combined regional potential S + gamma D (linear Galerkin)"
DENSITY_FORMULATION = ("Multi-region indirect (exterior potential density; closed contours include a DLP term)")
DIELECTRIC_DENSITY_FORMULATION = "Indirect dielectric (combined DLP + i*k0*SLP density)"
SOLVER_BACKEND = "grim_20261007_amd (HIP, gfx950)"
# twod/solver.py:953-1080
DEFAULT_QUALITY_THRESHOLDS = {"residual_norm_max": 1.0e-6, "constraint_residual_norm_max": 1.0e-8,
                              "condition_est_max": 1.0e6, "warnings_max": 10}


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


def _validate_cfie_alpha(value) -> float:
    """The exact disabled value only (twod/solver.py:814-836)."""
    try:
        alpha = float(value)
    except (TypeError, ValueError, OverflowError) as exc:
        raise ValueError("cfie_alpha is not implemented by any active 2-D formulation; "
                         "use the exact disabled value cfie_alpha=0.") from exc
    if not math.isfinite(alpha) or alpha != 0.0:
        raise ValueError("cfie_alpha is not implemented by any active 2-D formulation; "
                         "use the exact disabled value cfie_alpha=0. No unchanged field "
                         "was returned under a different solver setting.")
    return 0.0


def _device(device):
    if device is None:
        if not torch.cuda.is_available():
            raise RuntimeError("The GRIM-2D solver needs a ROCm GPU (torch.cuda.is_available() is False).")
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def evaluate_quality_gate(metadata: Dict[str, Any], thresholds: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
    """The reference's numeric quality gate on solver metadata (twod/solver.py:953-1100): residual,
    non-finite residuals, junction constraints, the condition estimate when it was computed, and
    the warning count, against the default or supplied thresholds (unknown keys are refused)."""
    merged = dict(DEFAULT_QUALITY_THRESHOLDS)
    if thresholds:
        supplied = dict(thresholds)
        unknown = sorted(set(supplied) - set(DEFAULT_QUALITY_THRESHOLDS))
        if unknown:
            raise ValueError("Unknown 2-D quality threshold field(s): " + ", ".join(str(k) for k in unknown))
        merged.update(supplied)
    residual_limit = float(merged["residual_norm_max"])
    constraint_limit = float(merged["constraint_residual_norm_max"])
    condition_limit = float(merged["condition_est_max"])
    try:
        warnings_float = float(merged["warnings_max"])
    except (TypeError, ValueError, OverflowError) as exc:
        raise ValueError("warnings_max must be a finite non-negative integer.") from exc
    if any(not math.isfinite(v) or v < 0.0 for v in (residual_limit, constraint_limit, condition_limit)):
        raise ValueError("2-D residual and condition quality thresholds must be finite non-negative values.")
    if not math.isfinite(warnings_float) or warnings_float < 0.0 or not warnings_float.is_integer():
        raise ValueError("warnings_max must be a finite non-negative integer.")
    warnings_limit = int(warnings_float)

    def as_float(raw):
        try:
            return float(raw) if raw is not None else float("nan")
        except (TypeError, ValueError):
            return float("nan")

    residual_value = as_float(metadata.get("residual_norm_max"))
    try:
        nonfinite = int(metadata.get("residual_nonfinite_count", 0))
    except (TypeError, ValueError, OverflowError):
        nonfinite = -1
    constraint_value = float(metadata.get("constraint_residual_norm_max", 0.0) or 0.0)
    condition_value = as_float(metadata.get("condition_est_max"))
    condition_computed = (bool(metadata.get("condition_est_computed")) if "condition_est_computed" in metadata
                          else math.isfinite(condition_value))
    warnings_count = len(list(metadata.get("warnings", []) or []))
    violations: List[str] = []
    if not math.isfinite(residual_value) or residual_value > residual_limit:
        violations.append(f"residual_norm_max={residual_value:.6g} exceeds limit {residual_limit:.6g}")
    if nonfinite > 0:
        violations.append(f"residual_nonfinite_count={nonfinite} must be zero")
    elif nonfinite < 0:
        violations.append("residual_nonfinite_count is missing a valid non-negative integer value")
    if bool(metadata.get("junction_constraints_applied", False)) and (
            not math.isfinite(constraint_value) or constraint_value > constraint_limit):
        violations.append(f"constraint_residual_norm_max={constraint_value:.6g} exceeds limit {constraint_limit:.6g}")
    if condition_computed and (not math.isfinite(condition_value) or condition_value > condition_limit):
        violations.append(f"condition_est_max={condition_value:.6g} exceeds limit {condition_limit:.6g}")
    if warnings_count > warnings_limit:
        violations.append(f"warnings_count={warnings_count} exceeds limit {warnings_limit}")
    return {
        "passed": not violations,
        "thresholds": {"residual_norm_max": residual_limit, "constraint_residual_norm_max": constraint_limit,
                       "condition_est_max": condition_limit, "warnings_max": warnings_limit},
        "values": {"residual_norm_max": residual_value, "residual_nonfinite_count": nonfinite,
                   "constraint_residual_norm_max": constraint_value, "condition_est_max": condition_value,
                   "condition_est_computed": condition_computed, "warnings_count": warnings_count},
        "violations": violations,
        "certification_scope": ("discrete_linear_system_residual_and_condition" if condition_computed
                                else "discrete_linear_system_residual_only"),
        "mesh_convergence_certified": bool(metadata.get("mesh_convergence_certified", False)),
        "reason": ("; ".join(violations) if violations else
                   "discrete linear-system quality thresholds satisfied; "
                   + ("condition number was not requested; " if not condition_computed else "")
                   + "mesh convergence is separately certified by the production workflow"),
    }


def _apply_quality_gate(metadata, quality_thresholds, strict_quality_gate):
    gate = evaluate_quality_gate(metadata, quality_thresholds)
    metadata["quality_gate"] = gate
    if strict_quality_gate and not gate["passed"]:
        raise ValueError(f"Quality gate failed: {gate.get('reason', 'quality gate failed')}")
    return gate


def _summarize_residuals(values):
    """Finite max and mean, and the count of non-finite residuals (twod/solver.py:457-468)."""
    r = np.asarray(values, dtype=float).reshape(-1)
    finite = r[np.isfinite(r)]
    if finite.size == 0:
        return float("nan"), float("nan"), int(r.size)
    return float(np.max(finite)), float(np.mean(finite)), int(r.size - finite.size)


def admit_device_memory(n: int, nangles: int, npols: int, precision: str, device, context: str,
                        condition: bool = False, degree: int = 1, sheet_elements: int = 0, extra: int = 0) -> float:
    """Refuse a solve whose device footprint exceeds 0.9 of the GPU memory it can use now (free
    memory plus what PyTorch's allocator holds unused), before anything is assembled, as the
    reference refuses host memory (twod/solver.py:1419-1554, 2342-2400).  Returns the estimate
    in GiB."""
    need = engine.request_device_bytes(n, nangles, npols, precision, degree=degree,
                                       sheet_elements=sheet_elements) + int(extra)
    if condition:   # equilibration scales, two probe columns and a 256 x 256 staging block per system
        need += 2 * (3 * 8 * n + 16 * 2 * 2 * n + 16 * 256 * 256)
    free, _total = torch.cuda.mem_get_info(device)
    available = free + torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    limit = MEMORY_LIMIT_FRACTION * available
    gib = 2.0 ** 30
    if need > limit:
        raise MemoryError(
            f"{context} requires an estimated {need / gib:.2f} GiB, but {available / gib:.2f} GiB is currently "
            f"available on the GPU; the safe allocation limit is {limit / gib:.2f} GiB. Planned system: {n} DOFs "
            "across 1 region(s). Reduce panel count or frequency, set an appropriate mesh_reference_ghz, or reduce "
            "the angle batch.")
    return need / gib


# The reference's execution-settings record (execution/options.py:16-34): every field with its
# default.  The dense GPU path honours basis_order and mesh_strategy 'global'; the CPU execution
# controls (threads, tiles, storage, RAM budget, batch size) have no meaning on it and are only
# validated.
EXECUTION_DEFAULTS = {
    "version": 1, "factorization": "dense", "mesh_strategy": "global", "basis_order": 1,
    "compressed_storage_mib": 2048, "ram_budget_gib": None, "temporary_directory": "",
    "assembly_threads": 1, "blas_threads": 1, "rhs_compression": "auto", "angle_batch_size": 256,
    "assembly_tile": 0, "far_quadrature_order": 0, "far_grading": True, "dense_residual_storage": "auto",
    "compressed_far_method": "full", "frequency_preconditioner": "off",
}


def execution_record(execution_options, scattering: str = "monostatic") -> dict:
    """The complete, validated settings record of one solve, with the reference's messages
    (execution/options.py:72-118, validate_for_run :422-428).  Factorization choices other than
    dense run the dense GPU LU (the hierarchical and compressed backends approximate it); mesh
    strategies other than 'global' (local and adaptive hp meshing) are not built here."""
    import os
    value = {} if execution_options is None else execution_options
    if not isinstance(value, dict) or set(value) - set(EXECUTION_DEFAULTS):
        raise ValueError("Execution settings must contain supported fields only.")
    out = dict(EXECUTION_DEFAULTS)
    out.update(value)
    if type(out["version"]) is not int or out["version"] != 1:
        raise ValueError("Unsupported execution settings version.")
    if out["factorization"] not in ("dense", "hierarchical", "auto", "compressed", "adaptive"):
        raise ValueError("Choose dense, hierarchical, auto, compressed, or adaptive factorization.")
    if type(out["basis_order"]) is not int or out["basis_order"] not in (1, 2, 3):
        raise ValueError("Boundary polynomial degree must be 1, 2 or 3.")
    if out["mesh_strategy"] not in ("global", "local", "adaptive"):
        raise ValueError("Mesh strategy must be global, local or adaptive.")
    if out["rhs_compression"] not in ("off", "auto", "on"):
        raise ValueError("RHS compression must be off, auto, or on.")
    if out["dense_residual_storage"] not in ("auto", "memory", "disk"):
        raise ValueError("Dense residual storage must be auto, memory, or disk.")
    if out["compressed_far_method"] not in ("full", "verified_cur"):
        raise ValueError("Compressed far method must be full or verified_cur (experimental).")
    if out["frequency_preconditioner"] not in ("off", "reuse"):
        raise ValueError("Frequency preconditioner must be off or reuse (experimental).")
    storage = out["compressed_storage_mib"]
    if type(storage) is not int or not (storage == 0 or 16 <= storage <= 1048576):
        raise ValueError("compressed_storage_mib must be 0 for automatic or an integer from 16 to 1048576.")
    for key, lower, upper in (("angle_batch_size", 1, 256), ("assembly_tile", 0, 65536),
                              ("far_quadrature_order", 0, 64)):
        number = out[key]
        if type(number) is not int or not lower <= number <= upper:
            raise ValueError(f"{key} must be an integer from {lower} to {upper}.")
    threads = out["assembly_threads"]
    if threads != "auto" and (type(threads) is not int or not 1 <= threads <= 1024):
        raise ValueError("Assembly threads must be auto or an integer from 1 to 1024.")
    threads = out["blas_threads"]
    if threads != "auto" and (type(threads) is not int or not 1 <= threads <= 1024):
        raise ValueError("BLAS threads must be auto or an integer from 1 to 1024.")
    ram = out["ram_budget_gib"]
    if ram is not None and (type(ram) not in (int, float) or not math.isfinite(ram) or ram <= 0):
        raise ValueError("RAM budget must be positive GiB or null for available memory.")
    if type(out["far_grading"]) is not bool:
        raise ValueError("Far grading must be true or false.")
    directory = out["temporary_directory"]
    if not isinstance(directory, str) or any(c in directory for c in "\r\n\x00"):
        raise ValueError("Temporary directory must be a path string.")
    if directory and not os.path.isabs(directory):
        raise ValueError("Temporary directory must be absolute or empty for the system temporary directory.")
    if (out["mesh_strategy"] == "adaptive" or out["basis_order"] > 1) and scattering != "monostatic":
        raise ValueError("Adaptive polynomial meshing supports 2D monostatic runs only.")
    if out["mesh_strategy"] != "global":
        raise ValueError(f"Mesh strategy '{out['mesh_strategy']}' (local / adaptive hp meshing) is not built "
                         "on the MI355X path; use 'global'.")
    if out["far_quadrature_order"] not in (0, 8):
        raise ValueError("A far quadrature order other than the W-calibrated rule is not built on the MI355X "
                         "path; use 0.")
    return out


def _execution_metadata(metadata: dict, settings: dict, started: float) -> None:
    """The effective settings record and the wall time of the call, as the reference's public
    entries record them (execution/options.py:651-666)."""
    metadata["execution_options"] = dict(settings, basis_order=metadata.get("polynomial_degree",
                                                                            settings["basis_order"]))
    metadata["execution_wall_seconds"] = time.perf_counter() - started


def basis_order(execution_options) -> int:
    """The Galerkin degree an execution-settings record asks for: 1 (default), 2 or 3."""
    return int(execution_record(execution_options)["basis_order"])


def _frequency_mesh(snapshot, units, f, freqs, mesh_reference_ghz, max_panels, material_base_dir, degree=1):
    """(mesh frequency, mesh) that frequency ``f`` solves on.  Without a reference frequency the mesh
    is built at f for f alone; with one, every frequency shares the mesh of the shortest wavelength
    over the sweep and the reference (twod/solver.py:2195-2215, geometry.py:2454-2471), whose junction
    grading and bound-surface-wave sizing see all of them.  Impedance laws are evaluated at f."""
    if mesh_reference_ghz:
        served = sorted(set(float(v) for v in freqs) | {float(mesh_reference_ghz)})
        mesh_ghz = max(served)
    else:
        served, mesh_ghz = [float(f)], float(f)
    mesh = build_mesh(snapshot, units, mesh_ghz, max_panels, material_base_dir=material_base_dir,
                      law_ghz=float(f), mesh_frequencies=served)
    if degree > 1 and isinstance(mesh, SheetLayout):
        raise ValueError("Polynomial bases of degree 2 and 3 are built for PEC contours only.")
    if degree > 1:
        if mesh.impedance is not None or mesh.medium is not None or mesh.regions is not None:
            raise ValueError("Polynomial bases of degree 2 and 3 are built for PEC contours only.")
        mesh.degree = int(degree)
    return mesh_ghz, mesh


def system_dofs(mesh) -> int:
    """Unknowns of the mesh's system: one per node, two (psi, phi) for a dielectric body, one per
    node and bordering region for a multi-region layout; for a sheet layout the nodes of its TM mesh
    (the TE mesh never has more)."""
    if isinstance(mesh, SheetLayout):
        return mesh.mesh("TM").n
    if mesh.regions is not None:
        return engine.region_plan(mesh).n_dof
    return 2 * mesh.n if mesh.medium is not None else mesh.dofs


def _systems(contour, k, angles, workspace):
    """[TE, TM] systems, right-hand sides (2, A, DOFs), the plain loads b_u, b_dn (A, N) and the map
    from a solution (A, DOFs) to the single-layer density (A, N) that radiates its far field:
    the PEC combined potential (TM Dirichlet rows -b_u, TE flux rows -b_dn); for contours with
    surface impedances the flux rows dn(u) + alpha u of both polarizations
    (engine.impedance_systems, which reads the near blocks the assembly left in ``workspace``);
    for a dielectric body the two-density system (engine.dielectric_systems); for a multi-region
    layout the regional system (engine.region_systems)."""
    bu, bdn = engine.loads(contour, k, angles)
    mesh = contour.mesh
    if mesh.regions is not None:
        plan = engine.region_plan(mesh)
        return (engine.region_systems(contour, k, workspace, plan), engine.region_rhs(plan, bu, bdn, k), bu, bdn,
                lambda x: engine.region_exterior_density(plan, mesh.n, x))
    if mesh.medium is not None:
        return (engine.dielectric_systems(contour, k, workspace), engine.dielectric_rhs(bu, bdn), bu, bdn,
                lambda x: engine.radiating_density(contour, k, x))
    z_tm, z_te = engine.assemble(contour, k, workspace)
    pair = z_tm._base                                                      # [TE, TM]
    if mesh.impedance is None:
        return pair, torch.stack((-bdn, -bu)), bu, bdn, lambda x: x
    return (engine.impedance_systems(contour, k, pair, workspace), engine.impedance_rhs(contour, k, angles),
            bu, bdn, lambda x: x)


SHEET_DENSITY_MESSAGE = (   # twod/solver.py:1636-1657
    "TYPE 1 free-floating sheets are not supported by compute_boundary_densities (the boundary-density export "
    "path).  They ARE supported for RCS: use solve_monostatic_rcs_2d or solve_bistatic_rcs_2d, which route sheet "
    "geometries to the dedicated sheet BIEs (_solve_tm_sheet / _solve_te_sheet / _solve_mixed_sheet_pec).")

SHEET_FORMULATIONS = {
    (True, "TM"): "2D sheet BIE (TM: single-layer representation)",
    (True, "TE"): "2D sheet BIE (TE: double-layer / hypersingular representation)",
    (False, "TM"): "2D mixed sheet+PEC BIE (TM: unified SLP representation)",
    (False, "TE"): "2D mixed sheet+PEC BIE (TE: unified DLP / hypersingular representation)",
}   # twod/solver.py:2409-2439


def _sheet_solve(layout: SheetLayout, pols, k: float, angles: torch.Tensor, precision: str, dev,
                 compute_condition_number: bool, abort_event, context: str, extra_angles: int = 0):
    """The requested polarizations of a sheet layout, each on its own mesh: the system S - c M (TM) or
    W - c M with pinned strip ends (TE), loads -b_u / b_dn (twod/solver.py:1746-1787).  When both
    meshes have the same node count the two systems are factored and solved as one batch of two.
    Yields (pol, sheet, density (A, N), b_u, b_dn, result fields)."""
    order = [p for p in ("TE", "TM") if p in pols]
    meshes = {p: layout.mesh(p) for p in order}
    if len(order) == 2 and meshes["TE"].n == meshes["TM"].n:
        groups = [tuple(order)]
    else:
        groups = [(p,) for p in order]
    for group in groups:
        n = meshes[group[0]].n
        ne = max(meshes[p].ne for p in group)
        admit_device_memory(n, int(angles.numel()) + extra_angles, len(group), precision, dev, context,
                            compute_condition_number, sheet_elements=ne)
        z = torch.empty((len(group), n, n), dtype=torch.complex128, device=dev)
        ws = engine.sheet_workspace(ne, dev)
        sheets, loads, rhs = [], [], []
        te, tm = meshes.get("TE"), meshes.get("TM")
        shared = (len(group) == 2 and np.array_equal(te.p0, tm.p0) and np.array_equal(te.p1, tm.p1)
                  and np.array_equal(te.elements, tm.elements))
        if shared:                              # one traversal for both (pinned rows: TE only)
            sheet = engine.DeviceSheet(te, dev)
            engine.sheet_systems(sheet, k, ws, te=z[0], tm=z[1])
        for i, pol in enumerate(group):
            sheet = engine.DeviceSheet(meshes[pol], dev)
            if not shared:
                engine.sheet_system(sheet, k, pol, ws, out=z[i])
            bu, bdn = engine.sheet_loads(sheet, k, angles)
            sheets.append(sheet)
            loads.append((bu, bdn))
            rhs.append(engine.sheet_rhs(sheet, pol, bu, bdn))
        del ws
        rhs = torch.stack(rhs)
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve canceled by user.")
        fac = engine.factor(z, precision)
        del z
        conds = methods = None
        if compute_condition_number:
            from . import condition
            conds, methods, replacement = condition.condition_estimates(fac)
            if replacement is not None:
                fac = replacement
        x, info = engine.solve(fac, rhs)
        rel = engine.relative_residual(fac.z, x, rhs, info.get("residual"))
        final = info.get("factor", fac)
        for i, pol in enumerate(group):
            part = engine.system(final, i) if final.parts is not None else final
            res = {"residual": rel[i].cpu().numpy(),
                   "info": {kk: v for kk, v in info.items() if kk not in ("residual", "factor")},
                   "k": k, "n": n, "precision_used": part.precision, "fallback_reason": part.fallback_reason,
                   "formulation": SHEET_FORMULATIONS[(layout.all_sheet, pol)]}
            if conds is not None:
                res["condition"] = float(conds[i])
                res["condition_method"] = methods[i]
            yield pol, sheets[i], x[i], loads[i][0], loads[i][1], res


def _solve_sheet_frequency(layout: SheetLayout, frequency_ghz: float, angles_deg: List[float], pols,
                           precision: str, device, return_density: bool, abort_event,
                           compute_condition_number: bool, context: str) -> Dict[str, Dict[str, Any]]:
    """Sheets alone or beside PEC contours (twod/solver.py:2409-2460): per polarization its own mesh
    and the single-layer (TM) or double-layer (TE) matched far field of the density
    (twod/operators.py:2925-3023)."""
    dev = _device(device)
    k = wavenumber(frequency_ghz)
    angles = torch.tensor(angles_deg, dtype=torch.float64, device=dev)
    out = {}
    for pol, _sheet, x, bu, bdn, res in _sheet_solve(layout, pols, k, angles, precision, dev,
                                                     compute_condition_number, abort_event, context):
        amp, sigma = engine.farfield_potential(x, bu if pol == "TM" else bdn, k)
        res.update({"amplitude": amp.cpu().numpy(), "sigma": sigma.cpu().numpy()})
        if return_density:
            res["density"] = x.cpu().numpy().T
        out[pol] = res
    return out


def _solve_sheet_bistatic_frequency(layout: SheetLayout, frequency_ghz: float, incidence_deg: List[float],
                                    observation_deg: List[float], pols, precision: str, device, abort_event,
                                    compute_condition_number: bool) -> Dict[str, Dict[str, Any]]:
    """Bistatic sheets: every incidence solved, each density projected on every observation
    direction with its polarization's potential."""
    dev = _device(device)
    k = wavenumber(frequency_ghz)
    inc = torch.tensor(incidence_deg, dtype=torch.float64, device=dev)
    obs = torch.tensor(observation_deg, dtype=torch.float64, device=dev)
    out = {}
    for pol, sheet, x, _bu, _bdn, res in _sheet_solve(layout, pols, k, inc, precision, dev, compute_condition_number,
                                                      abort_event, "The bistatic 2-D solve", len(observation_deg)):
        bu_o, bdn_o = engine.sheet_loads(sheet, k, obs)
        amp, sigma = engine.bistatic_farfield(x, bu_o if pol == "TM" else bdn_o, None, k)
        res.update({"amplitude": amp[0].cpu().numpy(), "sigma": sigma[0].cpu().numpy()})
        out[pol] = res
    return out


def solve_frequency(mesh, frequency_ghz: float, angles_deg: List[float], pols=("TE", "TM"),
                    precision: str = "double", device=None, return_density: bool = False,
                    abort_event=None, compute_condition_number: bool = False,
                    context: str = "The combined_potential 2-D solve") -> Dict[str, Dict[str, Any]]:
    """Assemble both systems once, then factor, solve and project each requested polarization.
    Per polarization: amplitude, sigma, relative residual per angle, refinement info, the
    condition estimate when requested, and the precision the factor ended in."""
    if isinstance(mesh, SheetLayout):
        return _solve_sheet_frequency(mesh, frequency_ghz, angles_deg, pols, precision, device, return_density,
                                      abort_event, compute_condition_number, context)
    dev = _device(device)
    k = wavenumber(frequency_ghz)
    order = [p for p in ("TE", "TM") if p in pols]
    admit_device_memory(system_dofs(mesh), len(angles_deg), len(order), precision, dev, context,
                        compute_condition_number, mesh.degree, extra=engine.lossy_workspace_bytes(mesh))
    contour = engine.DeviceContour(mesh, dev)
    ws = engine.assembly_workspace(mesh.n, dev, mesh.degree)
    angles = torch.tensor(angles_deg, dtype=torch.float64, device=dev)
    pair, rhs_all, bu, bdn, radiate = _systems(contour, k, angles, ws)
    del ws
    if abort_event is not None and abort_event.is_set():
        raise InterruptedError("Solve canceled by user.")
    sel = [("TE", "TM").index(p) for p in order]
    systems = pair if len(order) == 2 else pair[sel[0]].unsqueeze(0)
    rhs = rhs_all if len(order) == 2 else rhs_all[sel[0]].unsqueeze(0)
    fac = engine.factor(systems.contiguous(), precision)
    conds = methods = None
    if compute_condition_number:
        from . import condition
        conds, methods, replacement = condition.condition_estimates(fac)
        if replacement is not None:
            fac = replacement
    x, info = engine.solve(fac, rhs)
    rel = engine.relative_residual(fac.z, x, rhs, info.get("residual"))
    final = info.get("factor", fac)
    out = {}
    for i, pol in enumerate(order):
        part = engine.system(final, i) if final.parts is not None else final
        rho = radiate(x[i])
        amp, sigma = engine.farfield(rho, bu, bdn, k)
        res = {"amplitude": amp.cpu().numpy(), "sigma": sigma.cpu().numpy(),
               "residual": rel[i].cpu().numpy(),
               "info": {kk: v for kk, v in info.items() if kk not in ("residual", "factor")},
               "k": k, "n": mesh.n, "precision_used": part.precision,
               "fallback_reason": part.fallback_reason}
        if conds is not None:
            res["condition"] = float(conds[i])
            res["condition_method"] = methods[i]
        if return_density:
            # (N, A) nodal densities: the exterior one of a multi-region layout, psi of a dielectric body
            res["density"] = (rho if mesh.regions is not None else x[i, :, :mesh.dofs]).cpu().numpy().T
        out[pol] = res
    return out


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
    """TM, HH, H, HORIZONTAL -> TM; TE, VV, V, VERTICAL -> TE (twod/solver.py:395-422)."""
    pol = str(polarization or "").strip().upper()
    if pol in {"TM", "HH", "H", "HORIZONTAL"}:
        return "TM"
    if pol in {"TE", "VV", "V", "VERTICAL"}:
        return "TE"
    raise ValueError(f"Unsupported polarization '{polarization}'. Use TM/TE or VV/HH.")


class _Record:
    """What a run of solves reports in the reference's metadata layout."""

    def __init__(self, snapshot):
        self.snapshot = snapshot
        self.residuals: List[float] = []
        self.conditions: List[float] = []
        self.methods: List[str] = []
        self.panel_counts: List[int] = []
        self.lengths: List[float] = []
        self.mesh_ghz: List[float] = []
        self.reasons: List[str] = []
        self.precisions: List[str] = []
        self.columns = 0
        self.batches = 0
        self.factorizations = 0
        self.degrees: List[int] = []
        self.formulation = FORMULATION

    def add_mesh(self, mesh, mesh_ghz):
        self.panel_counts.append(mesh.n)
        self.degrees.append(int(getattr(mesh, "degree", 1)))
        self.lengths += [float(np.min(mesh.lengths)), float(np.max(mesh.lengths))]
        self.mesh_ghz.append(float(mesh_ghz))

    def add_solve(self, res, nangles):
        self.residuals += [float(v) for v in np.asarray(res["residual"]).reshape(-1)]
        if "condition" in res:
            self.conditions.append(res["condition"])
            if res["condition_method"] not in self.methods:
                self.methods.append(res["condition_method"])
        if res["fallback_reason"] and res["fallback_reason"] not in self.reasons:
            self.reasons.append(res["fallback_reason"])
        self.precisions.append(res["precision_used"])
        self.formulation = res.get("formulation", self.formulation)
        self.columns += nangles
        self.batches += 1

    def metadata(self, compute_condition_number, precision, solver_method):
        rmax, rmean, nonfinite = _summarize_residuals(self.residuals)
        ghz = self.mesh_ghz
        wl = [C0 / (g * 1e9) for g in ghz]
        mixed_solves = sum(p == "mixed" for p in self.precisions)
        return {
            "source_path": str(self.snapshot.get("source_path", "") or ""),
            "segment_count": int(len(self.snapshot.get("segments", []) or [])),
            "panel_count": max(self.panel_counts), "panel_count_min": min(self.panel_counts),
            "panel_count_max": max(self.panel_counts), "panel_counts": list(self.panel_counts),
            "panel_length_min_m": min(self.lengths), "panel_length_max_m": max(self.lengths),
            "mesh_reference_ghz": ghz[0] if len(set(round(v, 12) for v in ghz)) == 1 else None,
            "mesh_reference_ghz_min": min(ghz), "mesh_reference_ghz_max": max(ghz),
            "mesh_wavelength_m": wl[0] if len(set(round(v, 15) for v in wl)) == 1 else None,
            "mesh_wavelength_min_m": min(wl), "mesh_wavelength_max_m": max(wl),
            "mesh_max_refractive_index": 1.0, "mesh_material_flags": [],
            "polynomial_degree": self.degrees[0] if self.degrees else 1,
            "polynomial_degree_min": min(self.degrees) if self.degrees else 1,
            "polynomial_degree_max": max(self.degrees) if self.degrees else 1,
            "formulation": self.formulation, "solver_method": "dense_lu", "solver_method_requested": str(solver_method),
            "residual_norm_max": rmax, "residual_norm_mean": rmean, "residual_nonfinite_count": nonfinite,
            "constraint_residual_norm_max": 0.0, "constraint_residual_norm_mean": 0.0,
            "constraint_residual_applicable": False,
            "condition_est_max": float(np.max(self.conditions)) if self.conditions else float("nan"),
            "condition_est_mean": float(np.mean(self.conditions)) if self.conditions else float("nan"),
            "condition_est_computed": bool(compute_condition_number),
            "condition_estimator": (self.methods[0] if self.methods else "equilibrated_1norm_lu_onenormest")
            if compute_condition_number else "not_requested",
            "dense_condition_methods": sorted(self.methods),
            "warnings": [], "information": [], "warning_count": 0,
            "junction_constraints_applied": False, "junction_constraints": 0, "junction_nodes": 0,
            "linear_node_count": int(sum(self.panel_counts)), "linear_element_count": int(sum(self.panel_counts)),
            "precision": precision, "solver_backend": SOLVER_BACKEND,
            "linear_backend": "gpu_hip_mixed_lu" if mixed_solves == len(self.precisions) else
            ("gpu_hip" if mixed_solves == 0 else "mixed"),
            "dense_mixed_precision_solve_count": mixed_solves,
            "dense_gpu_solve_count": len(self.precisions), "dense_cpu_solve_count": 0,
            "dense_fallback_reasons": list(self.reasons), "dense_gpu_fallback_reasons": list(self.reasons),
            "dense_rhs_batch_count": self.batches, "dense_rhs_column_count": self.columns,
            "dense_largest_system": max(self.panel_counts),
            "amplitude_version": 2,
        }


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
    cfie_alpha = _validate_cfie_alpha(cfie_alpha)
    if str(rcs_normalization_mode).strip().lower() != "physical":
        raise ValueError("Only the physical RCS normalization is supported.")
    pol = _normalize_pol(polarization)
    unit_scale(geometry_units)
    started = time.perf_counter()
    settings = execution_record(execution_options)
    degree = settings["basis_order"]
    evaluate_quality_gate({}, quality_thresholds)            # refuse unknown thresholds up front
    samples = []
    record = _Record(geometry_snapshot)
    total = len(freqs) * (len(angs) + 1)
    done = 0
    for f in freqs:
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve canceled by user.")
        mesh_ghz, mesh = _frequency_mesh(geometry_snapshot, geometry_units, f, freqs, mesh_reference_ghz, max_panels,
                                         material_base_dir, degree)
        record.add_mesh(mesh, mesh_ghz)
        res = solve_frequency(mesh, f, angs, (pol,), precision, device, abort_event=abort_event,
                              compute_condition_number=compute_condition_number,
                              context=f"The combined_potential 2-D {pol} solve")[pol]
        samples += _samples(f, angs, res["amplitude"], res["sigma"], res["residual"])
        record.add_solve(res, len(angs))
        done += len(angs) + 1
        if progress_callback is not None:
            try:
                progress_callback(done, total, f"Solved {f:g} GHz")
            except Exception:
                pass
    metadata = record.metadata(compute_condition_number, precision, solver_method)
    alias = "HH" if pol == "TM" else "VV"
    metadata.update({"polarization_internal": pol, "polarization_user": pol, "polarization_aliases": [pol],
                     "polarization_export": pol, "polarization_export_alias": alias,
                     "rcs_normalization_mode": "physical", "cfie_alpha": cfie_alpha})
    _execution_metadata(metadata, settings, started)
    _apply_quality_gate(metadata, quality_thresholds, strict_quality_gate)
    return {"solver": "2d_bie_mom_rcs", "scattering_mode": "monostatic",
            "amplitude_convention": AMPLITUDE_CONVENTION, "amplitude_version": 2,
            "polarization": pol, "polarization_export": pol, "samples": samples, "metadata": metadata}


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
    started = time.perf_counter()
    settings = execution_record(execution_options)
    degree = settings["basis_order"]
    evaluate_quality_gate({}, quality_thresholds)
    co = {"VV": [], "HH": []}
    flat = []
    record = _Record(geometry_snapshot)
    total = 2 * len(freqs) * (len(angs) + 1)
    done = 0
    for f in freqs:
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve canceled by user.")
        mesh_ghz, mesh = _frequency_mesh(geometry_snapshot, geometry_units, f, freqs, mesh_reference_ghz, max_panels,
                                         material_base_dir, degree)
        record.add_mesh(mesh, mesh_ghz)
        res = solve_frequency(mesh, f, angs, ("TE", "TM"), precision, device, abort_event=abort_event,
                              compute_condition_number=compute_condition_number)
        per = {}
        for export, internal in CO_POLARIZED_CHANNELS:
            r = res[internal]
            record.add_solve(r, len(angs))
            rows = _samples(f, angs, r["amplitude"], r["sigma"], r["residual"])
            for row in rows:
                row["polarization"] = export
                row["polarization_internal"] = internal
            per[export] = rows
            co[export] += rows
        for vv, hh in zip(per["VV"], per["HH"]):
            flat += [vv, hh]
        done += 2 * (len(angs) + 1)
        if progress_callback is not None:
            try:
                progress_callback(done, total, f"Solved {f:g} GHz")
            except Exception:
                pass
    metadata = record.metadata(compute_condition_number, precision, solver_method)
    metadata.update({"polarizations": ["VV", "HH"], "co_solve_shared_discretization": True,
                     "rcs_normalization_mode": "physical"})
    _execution_metadata(metadata, settings, started)
    _apply_quality_gate(metadata, quality_thresholds, strict_quality_gate)
    return {"solver": "2d_bie_mom_rcs", "scattering_mode": "monostatic",
            "amplitude_convention": AMPLITUDE_CONVENTION, "amplitude_version": 2,
            "rcs_log_unit": "dBke", "rcs_linear_quantity": "sigma_2d",
            "polarizations": ["VV", "HH"], "polarization_mapping": {"VV": "TE", "HH": "TM"},
            "samples": flat, "co_solved_samples": co, "metadata": metadata}


def solve_bistatic_frequency(mesh, frequency_ghz: float, incidence_deg: List[float], observation_deg: List[float],
                             pols=("TE", "TM"), precision: str = "double", device=None,
                             abort_event=None, compute_condition_number: bool = False) -> Dict[str, Dict[str, Any]]:
    """Both systems from one assembly; solve for every incidence, then project every solution
    on every observation direction (twod/solver.py:3476-3517 with observation angles)."""
    if isinstance(mesh, SheetLayout):
        return _solve_sheet_bistatic_frequency(mesh, frequency_ghz, incidence_deg, observation_deg, pols, precision,
                                               device, abort_event, compute_condition_number)
    dev = _device(device)
    k = wavenumber(frequency_ghz)
    order = [p for p in ("TE", "TM") if p in pols]
    admit_device_memory(system_dofs(mesh), len(incidence_deg) + len(observation_deg), len(order), precision, dev,
                        "The bistatic 2-D solve", compute_condition_number, extra=engine.lossy_workspace_bytes(mesh))
    contour = engine.DeviceContour(mesh, dev)
    ws = engine.assembly_workspace(mesh.n, dev)
    inc = torch.tensor(incidence_deg, dtype=torch.float64, device=dev)
    obs = torch.tensor(observation_deg, dtype=torch.float64, device=dev)
    pair, rhs_all, bu, bdn, radiate = _systems(contour, k, inc, ws)
    del ws
    if abort_event is not None and abort_event.is_set():
        raise InterruptedError("Solve canceled by user.")
    sel = [("TE", "TM").index(p) for p in order]
    systems = pair if len(order) == 2 else pair[sel[0]].unsqueeze(0)
    rhs = rhs_all if len(order) == 2 else rhs_all[sel[0]].unsqueeze(0)
    fac = engine.factor(systems.contiguous(), precision)
    conds = methods = None
    if compute_condition_number:
        from . import condition
        conds, methods, replacement = condition.condition_estimates(fac)
        if replacement is not None:
            fac = replacement
    x, info = engine.solve(fac, rhs)
    rel = engine.relative_residual(fac.z, x, rhs, info.get("residual"))
    final = info.get("factor", fac)
    bu_o, bdn_o = engine.loads(contour, k, obs)
    rho = torch.stack([radiate(x[i]) for i in range(x.shape[0])])
    amp, sigma = engine.bistatic_farfield(rho, bu_o, bdn_o, k)
    out = {}
    for i, pol in enumerate(order):
        part = engine.system(final, i) if final.parts is not None else final
        out[pol] = {"amplitude": amp[i].cpu().numpy(), "sigma": sigma[i].cpu().numpy(),
                    "residual": rel[i].cpu().numpy(),
                    "info": {kk: v for kk, v in info.items() if kk not in ("residual", "factor")}, "k": k,
                    "n": mesh.n, "precision_used": part.precision, "fallback_reason": part.fallback_reason}
        if conds is not None:
            out[pol]["condition"] = float(conds[i])
            out[pol]["condition_method"] = methods[i]
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
                  mesh_reference_ghz, abort_event, precision, device, compute_condition_number=False,
                  material_base_dir=None):
    freqs, inc, obs = _validate_bistatic(frequencies_ghz, incidence_angles_deg, observation_angles_deg)
    unit_scale(geometry_units)
    evaluate_quality_gate({}, quality_thresholds)
    rows = {p: [] for p in pols}
    record = _Record(geometry_snapshot)
    total, done = len(freqs) * len(inc), 0
    for f in freqs:
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve canceled by user.")
        mesh_ghz, mesh = _frequency_mesh(geometry_snapshot, geometry_units, f, freqs, mesh_reference_ghz, max_panels,
                                         material_base_dir)
        record.add_mesh(mesh, mesh_ghz)
        res = solve_bistatic_frequency(mesh, f, inc, obs, pols, precision, device, abort_event=abort_event,
                                       compute_condition_number=compute_condition_number)
        for p in pols:
            r = res[p]
            rows[p] += _bistatic_samples(f, inc, obs, r["amplitude"], r["sigma"], r["residual"])
            record.add_solve(r, len(inc))
        done += len(inc)
        if progress_callback is not None:
            try:
                progress_callback(done, total, f"Bistatic {f:g} GHz")
            except Exception:
                pass
    metadata = record.metadata(compute_condition_number, precision, "dense_lu")
    metadata["formulation"] = "bistatic 2D BIE/MoM"
    _apply_quality_gate(metadata, quality_thresholds, strict_quality_gate)
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
    _validate_cfie_alpha(cfie_alpha)
    pol = _normalize_pol(polarization)
    rows, metadata = _bistatic_run(geometry_snapshot, frequencies_ghz, incidence_angles_deg,
                                   observation_angles_deg, (pol,), geometry_units, progress_callback,
                                   quality_thresholds, strict_quality_gate, max_panels, mesh_reference_ghz,
                                   abort_event, precision, device, compute_condition_number, material_base_dir)
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
                                   abort_event, precision, device, compute_condition_number, material_base_dir)
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
                               material_base_dir: Optional[str] = None, cfie_alpha: float = 0.0,
                               max_panels: int = MAX_PANELS_DEFAULT, abort_event=None, *,
                               mesh_reference_ghz: Optional[float] = None, precision: str = "double",
                               device=None) -> Dict[str, Any]:
    """Layer-potential density of one incidence at the element centres (twod/solver.py:4150-4346):
    the same arguments in the same positions, the same result keys.  ``mesh_reference_ghz``
    (keyword only) meshes at another frequency; ``precision`` and ``device`` choose the LU
    precision and the GPU."""

    def check_abort():
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Boundary-density calculation canceled by user.")

    check_abort()
    _validate_cfie_alpha(cfie_alpha)
    pol = _normalize_pol(polarization)
    unit_scale(geometry_units)
    frequency_ghz = float(frequency_ghz)
    if not math.isfinite(frequency_ghz) or frequency_ghz <= 0.0:
        raise ValueError("frequency_ghz must be a positive finite value.")
    mesh_ghz, mesh = _frequency_mesh(geometry_snapshot, geometry_units, frequency_ghz, [frequency_ghz],
                                     mesh_reference_ghz, max_panels, material_base_dir)
    if isinstance(mesh, SheetLayout):
        raise ValueError(SHEET_DENSITY_MESSAGE)
    check_abort()
    res = solve_frequency(mesh, frequency_ghz, [float(elevation_deg)], (pol,), precision, device,
                          return_density=True, abort_event=abort_event,
                          context="Boundary-density diagnostics")[pol]
    check_abort()
    nodal = res["density"][:, 0]
    if mesh.regions is not None:
        flags = sorted(r for r in mesh.region_media if r > 0)
        index = max([1.0] + [abs(causal_index(*mesh.region_media[r])) for r in flags])
    elif mesh.medium is not None:
        flags = [int(mesh.medium[0])]
        index = max(1.0, abs(causal_index(mesh.medium[1], mesh.medium[2])))
    else:
        flags, index = [], 1.0
    ids = mesh.elements
    density = 0.5 * (nodal[ids[:, 0]] + nodal[ids[:, 1]])      # linear basis at t = 1/2
    centers = 0.5 * (mesh.p0 + mesh.p1)
    normals = mesh.normals
    return {
        "quantity": "boundary_integral_layer_density",
        "is_physical_surface_current": False,
        "interpretation": ("Formulation-specific SLP/DLP representation density; do not interpret as "
                           "electric or magnetic surface current without a formulation- and "
                           "polarization-specific trace conversion."),
        "formulation": DENSITY_FORMULATION if mesh.medium is None else DIELECTRIC_DENSITY_FORMULATION,
        "frequency_ghz": frequency_ghz,
        "elevation_deg": float(elevation_deg),
        "polarization": pol,
        "coordinate_units": "meters",
        "length_units": "meters",
        "mesh_wavelength_m": float(C0 / (mesh_ghz * 1e9) / index),
        "mesh_max_refractive_index": float(index),
        "mesh_material_flags": flags,
        "element_count": int(mesh.n),
        "node_count": int(mesh.n),
        "centers_x": centers[:, 0].tolist(),
        "centers_y": centers[:, 1].tolist(),
        "normals_x": normals[:, 0].tolist(),
        "normals_y": normals[:, 1].tolist(),
        "lengths": mesh.lengths.tolist(),
        "density_real": np.real(density).tolist(),
        "density_imag": np.imag(density).tolist(),
        "density_abs": np.abs(density).tolist(),
        "density_phase_deg": np.degrees(np.angle(density)).tolist(),
        "amplitude_version": 2,
        "metadata": {"solver_backend": SOLVER_BACKEND, "precision": precision,
                     "precision_used": res["precision_used"], "fallback_reason": res["fallback_reason"],
                     "linear_residual": float(res["residual"][0])},
    }
