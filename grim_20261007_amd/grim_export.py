"""GRIM `.grim` archives of monostatic RCS results (formats 1 and 2).

The archive is the one GRIM and the reference's exporter exchange
(``tools/GHOST/ghost_backend/io/grim.py:1034-1170``, ``export_result_to_grim``;
layout notes in ``tools/GHOST/README.md:212-221``): a compressed NumPy archive
holding a rectangular grid over (azimuth, elevation, frequency, polarization)
with its axes and metadata.  A bistatic result becomes one archive per
incidence angle (``<path>_inc_<angle>.grim``), published together.

* Format 2 (the default) stores the solver's raw complex far-field amplitude
  once, as ``rcs_complex`` (complex128), framed by ``grim_format_version``,
  ``storage_kind = "complex"`` and ``storage_domain = "raw_field"``; readers
  derive power and phase under ``units.rcs_linear_quantity``.
* Format 1 stores ``rcs_power`` and ``rcs_phase`` (float32) and the raw
  ``rcs_amp_real`` / ``rcs_amp_imag`` (float64).

A co-polarized result (``polarizations == ["VV", "HH"]`` with matching
``co_solved_samples``) becomes one archive with both channels on the last
axis.  Every key, dtype and value follows the reference's writer, so the same
result dict gives the same arrays; ``tests/test_grim_export_cpu.py`` checks
that against archives the reference wrote.
"""
from __future__ import annotations

import json
import math
import os
import tempfile
from io import StringIO
from typing import Any, Dict, List, Optional

import numpy as np

C0 = 299_792_458.0
EPS = 1e-12
SOLVER_METADATA_SCHEMA = "ghost.solver_metadata.v1"
FORMAT_ENV = "GHOST_GRIM_FORMAT_VERSION"
DIAGNOSTIC_KEYS = ("linear_residual", "linear_backward_error", "constraint_residual",
                   "constraint_residual_norm", "condition_est")
TM_ALIASES = ["TM", "HH", "H", "HORIZONTAL"]
TE_ALIASES = ["TE", "VV", "V", "VERTICAL"]

def resolve_format_version(requested: Optional[int] = None) -> int:
    """1 or 2; None reads GHOST_GRIM_FORMAT_VERSION and defaults to 2 (io/grim.py:23-36)."""
    if requested is None:
        setting = os.environ.get(FORMAT_ENV, "").strip()
        return int(setting) if setting in ("1", "2") else 2
    if isinstance(requested, bool) or requested not in (1, 2):
        raise ValueError(f"GRIM format_version must be 1 or 2; got {requested!r}.")
    return int(requested)


# -- metadata envelope ---------------------------------------------------------
def json_safe(value: Any) -> Any:
    """Deterministic standard JSON data: sorted keys, marked non-finite floats and complex
    numbers, NumPy scalars and arrays as Python values (io/grim.py:39-81)."""
    if value is None or isinstance(value, (str, bool)):
        return value
    if isinstance(value, np.generic):
        return json_safe(value.item())
    if isinstance(value, int):
        return int(value)
    if isinstance(value, float):
        if math.isfinite(value):
            return float(value)
        label = "nan" if math.isnan(value) else ("infinity" if value > 0 else "-infinity")
        return {"__nonfinite_float__": label}
    if isinstance(value, complex):
        return {"__complex__": {"real": json_safe(float(value.real)), "imag": json_safe(float(value.imag))}}
    if isinstance(value, np.ndarray):
        return json_safe(value.tolist())
    if isinstance(value, dict):
        return {str(k): json_safe(v) for k, v in sorted(value.items(), key=lambda kv: str(kv[0]))}
    if isinstance(value, (list, tuple)):
        return [json_safe(v) for v in value]
    if isinstance(value, (set, frozenset)):
        items = [json_safe(v) for v in value]
        return sorted(items, key=lambda v: json.dumps(v, sort_keys=True, separators=(",", ":"), ensure_ascii=False))
    if isinstance(value, os.PathLike):
        return os.fspath(value)
    return str(value)


def solver_metadata_json(result: Dict[str, Any]) -> str:
    """The audit envelope stored beside every export (io/grim.py:84-149): result-level fields
    and one diagnostics row per sample, ordered by (frequency, incidence, scattering angle)."""
    samples = list(result.get("samples", []) or [])
    rows = [r for r in samples if any(k in r for k in DIAGNOSTIC_KEYS)]
    rows.sort(key=lambda r: tuple(float(r.get(k)) for k in ("frequency_ghz", "theta_inc_deg", "theta_scat_deg")))
    envelope = {
        "schema": SOLVER_METADATA_SCHEMA,
        "solver": result.get("solver", ""),
        "scattering_mode": result.get("scattering_mode", ""),
        "polarization": result.get("polarization", ""),
        "polarization_export": result.get("polarization_export", ""),
        "polarizations": result.get("polarizations", []) or [],
        "polarization_mapping": result.get("polarization_mapping", {}) or {},
        "rcs_log_unit": result.get("rcs_log_unit", ""),
        "rcs_linear_quantity": result.get("rcs_linear_quantity", ""),
        "amplitude_convention": result.get("amplitude_convention", ""),
        "amplitude_version": result.get("amplitude_version", 1),
        "metadata": result.get("metadata", {}) or {},
    }
    enc = json.JSONEncoder(sort_keys=True, separators=(",", ":"), ensure_ascii=False, allow_nan=False)
    out = StringIO()
    keys = sorted(list(envelope) + ["sample_diagnostics"])
    out.write("{")
    for i, key in enumerate(keys):
        if i:
            out.write(",")
        out.write(enc.encode(key) + ":")
        if key != "sample_diagnostics":
            out.write(enc.encode(json_safe(envelope[key])))
            continue
        out.write("[")
        for j, row in enumerate(rows):
            if j:
                out.write(",")
            entry = dict(frequency_ghz=row.get("frequency_ghz"), theta_inc_deg=row.get("theta_inc_deg"),
                         theta_scat_deg=row.get("theta_scat_deg"), polarization=row.get("polarization", ""))
            entry.update({k: row[k] for k in DIAGNOSTIC_KEYS if k in row})
            out.write(enc.encode(json_safe(entry)))
        out.write("]")
    out.write("}")
    return out.getvalue()


# -- grid ------------------------------------------------------------------------
def canonical_polarization(label: Optional[str]) -> str:
    text = str(label or "").strip().upper()
    if text in TM_ALIASES or not text:
        return "TM"
    if text in TE_ALIASES:
        return "TE"
    raise ValueError(f"Unsupported polarization '{label}'. Use TM/HH/H/HORIZONTAL or TE/VV/V/VERTICAL.")


def _finite(row: Dict[str, Any], key: str, index: int) -> float:
    if key not in row:
        raise ValueError(f"Invalid solver sample at index {index}: Solver sample is missing required field '{key}'.")
    try:
        value = float(row[key])
    except (TypeError, ValueError, OverflowError) as exc:
        raise ValueError(f"Invalid solver sample at index {index}: Solver sample field '{key}' must be a finite "
                         f"numeric value; got {row[key]!r}.") from exc
    if not math.isfinite(value):
        raise ValueError(f"Invalid solver sample at index {index}: Solver sample field '{key}' must be finite; "
                         f"got {row[key]!r}.")
    return value


def _columns(samples: List[Dict[str, Any]], quantity: str) -> np.ndarray:
    """(n, 5) theta_scat, frequency, linear, amp_real, amp_imag after the per-row checks of
    io/grim.py:232-292: finite fields, positive frequency, non-negative power, and power
    equal to |B|^2 / (4 k) (sigma_2d) or 4 pi |A|^2 (sigma_3d) to 2e-8."""
    out = np.empty((len(samples), 5))
    for i, row in enumerate(samples):
        theta, freq, lin, re, im = (_finite(row, k, i) for k in
                                    ("theta_scat_deg", "frequency_ghz", "rcs_linear", "rcs_amp_real", "rcs_amp_imag"))
        if freq <= 0.0:
            raise ValueError(f"Invalid solver sample at index {i}: frequency_ghz must be positive; got {freq!r}.")
        if lin < 0.0:
            raise ValueError(f"Invalid solver sample at index {i}: rcs_linear must be non-negative; got {lin!r}.")
        with np.errstate(over="ignore", invalid="ignore"):
            a2 = re * re + im * im
        if not math.isfinite(a2):
            raise ValueError(f"Invalid solver sample at index {i}: the finite complex amplitude is too large to "
                             "form finite physical power.")
        expected = a2 / (4.0 * (2.0 * math.pi * freq * 1.0e9 / C0)) if quantity == "sigma_2d" else 4.0 * math.pi * a2
        if not math.isfinite(expected):
            raise ValueError(f"Invalid solver sample at index {i}: normalized linear RCS is not finite.")
        if abs(lin - expected) > 2.0e-8 * max(expected, lin) + np.finfo(float).tiny:
            raise ValueError(f"Invalid solver sample at index {i}: rcs_linear {lin:.12g} is inconsistent with the "
                             f"complex amplitude under {quantity} normalization (expected {expected:.12g}).")
        out[i] = (theta, freq, lin, re, im)
    return out


def _channel_grid(samples: List[Dict[str, Any]], polarization: str, quantity: str):
    """Axes, float32 power, complex field and phase of one channel (io/grim.py:353-513)."""
    if not samples:
        raise ValueError("No samples available to export.")
    theta, freq, lin, re, im = _columns(samples, quantity).T
    az, fr = np.unique(theta), np.unique(freq)
    cells = np.searchsorted(az, theta) * len(fr) + np.searchsorted(fr, freq)
    unique_cells, first = np.unique(cells, return_index=True)
    with np.errstate(invalid="ignore"):
        phase = np.arctan2(im, re)
    phase[(re == 0.0) & (im == 0.0)] = 0.0
    repeat = np.ones(cells.size, bool)
    repeat[first] = False
    if np.any(repeat):
        # the first sample of a coordinate is stored; repeats must agree with it
        idx = np.flatnonzero(repeat)
        origin = first[np.searchsorted(unique_cells, cells[idx])]
        p_bad = np.abs(lin[origin].astype(np.float32).astype(np.float64) - lin[idx]) > EPS
        dphi = np.remainder(phase[idx] - phase[origin].astype(np.float32).astype(np.float64) + math.pi,
                            2.0 * math.pi) - math.pi
        bad = np.flatnonzero(p_bad | (np.abs(dphi) > 1e-5))
        if bad.size:
            r = idx[bad[0]]
            what = "sample" if p_bad[bad[0]] else "amplitude"
            raise ValueError(f"Duplicate {what} conflict at az={theta[r]}, el=0, f={freq[r]}, pol={polarization}.")
    shape = (len(az), 1, len(fr), 1)
    power = np.full(shape, np.nan, np.float32)
    power.flat[unique_cells] = lin[first]
    field = np.full(shape, complex(np.nan, np.nan), np.complex128)
    field.real.flat[unique_cells] = re[first]
    field.imag.flat[unique_cells] = im[first]
    missing = np.argwhere(~np.isfinite(power))
    if missing.size:
        raise ValueError("Solver samples do not form a complete rectangular (angle x frequency x polarization) "
                         f"grid: {len(missing)} cell(s) are missing; first missing index is "
                         f"{tuple(int(v) for v in missing[0])}.")
    return az, fr, power, field


def _units(quantity: str, log_unit: str) -> str:
    return json.dumps({"azimuth": "deg", "elevation": "deg", "frequency": "GHz",
                       "rcs_log_unit": str(log_unit), "rcs_linear_quantity": str(quantity)})


def _payload(channels, labels, alias_primary, aliases_json, quantity, log_unit, source_path, history):
    """One archive payload from per-channel grids (io/grim.py:353-592)."""
    az, fr = channels[0][0], channels[0][1]
    for c in channels[1:]:
        for name, a, b in (("azimuths", c[0], az), ("frequencies", c[1], fr)):
            if not np.array_equal(a, b):
                raise ValueError(f"VV and HH cannot share one GRIM file because their {name} axes differ.")
    power = np.concatenate([c[2] for c in channels], axis=-1)
    field = np.concatenate([c[3] for c in channels], axis=-1)
    if quantity == "sigma_2d":
        phase_reference = ("origin=(0,0), convention=exp(+jwt); stored complex field is the 2D layer-potential "
                           "bare-integral amplitude B. The coefficient in u_s~exp(-j(kr-pi/4))/sqrt(8*pi*k*r)*A "
                           "is A=j*B.")
        domain, convention = "2d_layer_potential_bare_integral_amplitude_B", "A_physical_asymptotic = +j * B_stored"
    else:
        phase_reference = "origin=(0,0), convention=exp(+jwt), monostatic far-field amplitude"
        domain, convention = "solver_raw_far_field_amplitude", "solver physical far-field amplitude"
    phase = np.angle(field)
    phase = np.where((field.real == 0.0) & (field.imag == 0.0), 0.0, phase).astype(np.float32)
    return {"azimuths": az, "elevations": np.asarray([0.0]), "frequencies": fr,
            "polarizations": np.asarray(labels, dtype=str), "polarization_alias_primary": alias_primary,
            "polarization_aliases_json": aliases_json, "power": power, "phase": phase, "field": field,
            "units": _units(quantity, log_unit), "phase_reference": phase_reference,
            "amplitude_convention": convention, "complex_field_domain": domain,
            "source_path": source_path, "history": history}


def _grim_path(path: str) -> str:
    return path if path.lower().endswith(".grim") else f"{path}.grim"


def incidence_suffix(theta_inc_deg: float) -> str:
    """File suffix of one incidence of a bistatic collection: inc_12p5, inc_m30 (io/grim.py:191-194)."""
    text = f"{theta_inc_deg:.6f}".rstrip("0").rstrip(".")
    return "inc_" + (text.replace("-", "m").replace(".", "p") or "0")


def _publish(entries, version: int, preserve_raw: bool) -> List[str]:
    """Write a bistatic collection as one unit (io/grim.py:894-1005): every archive is staged
    beside its target first; targets are then replaced, keeping backups until all are in place,
    so a failure restores the previous collection."""
    targets, seen = [], set()
    for _, path in entries:
        target = os.path.abspath(_grim_path(path))
        if os.path.normcase(target) in seen:
            raise ValueError(f"A GRIM output batch contains the duplicate target {target!r}.")
        if os.path.exists(target) and not os.path.isfile(target):
            raise ValueError(f"GRIM output target {target!r} exists but is not a regular file.")
        seen.add(os.path.normcase(target))
        targets.append(target)
    staged, backups, complete = [], {}, False
    try:
        for (payload, _), target in zip(entries, targets):
            os.makedirs(os.path.dirname(target), exist_ok=True)
            fd, stage = tempfile.mkstemp(prefix=f".{os.path.basename(target)}.", suffix=".stage.grim",
                                         dir=os.path.dirname(target))
            os.close(fd)
            staged.append((stage, target))
            _write(payload, stage, version, preserve_raw)
        for stage, target in staged:
            backup = None
            if os.path.exists(target):
                fd, backup = tempfile.mkstemp(prefix=f".{os.path.basename(target)}.", suffix=".backup",
                                              dir=os.path.dirname(target))
                os.close(fd)
                os.replace(target, backup)
            backups[target] = backup
            os.replace(stage, target)
        complete = True
        return targets
    except BaseException:
        for target, backup in reversed(list(backups.items())):
            if backup is None:
                if os.path.exists(target):
                    os.unlink(target)
            elif os.path.exists(backup):
                os.replace(backup, target)
        raise
    finally:
        for stage, _ in staged:
            if os.path.exists(stage):
                os.unlink(stage)
        if complete:
            for backup in backups.values():
                if backup is not None and os.path.exists(backup):
                    os.unlink(backup)


def _write(payload: Dict[str, Any], path: str, version: int, preserve_raw: bool) -> str:
    """Write the archive atomically (temporary file, fsync, rename), keys in the reference's order."""
    out = _grim_path(path)
    save = {k: payload[k] for k in ("azimuths", "elevations", "frequencies", "polarizations",
                                    "polarization_alias_primary", "polarization_aliases_json")}
    compact = version == 2 and preserve_raw
    if not compact:
        save["rcs_power"] = payload["power"]
        save["rcs_phase"] = payload["phase"]
    save.update(rcs_domain="power_phase", power_domain="linear_rcs", source_path=payload["source_path"],
                history=payload["history"], units=payload["units"], phase_reference=payload["phase_reference"],
                amplitude_convention=payload["amplitude_convention"], raw_complex_amplitude_preserved=preserve_raw)
    if preserve_raw:
        if compact:
            save.update(grim_format_version=np.int16(2), storage_kind="complex", storage_domain="raw_field",
                        rcs_complex=payload["field"])
        else:
            save["rcs_amp_real"] = payload["field"].real.copy()
            save["rcs_amp_imag"] = payload["field"].imag.copy()
        save["complex_field_domain"] = payload["complex_field_domain"]
    save["solver_metadata_json"] = payload["solver_metadata_json"]
    folder = os.path.dirname(os.path.abspath(out))
    os.makedirs(folder, exist_ok=True)
    fd, tmp = tempfile.mkstemp(prefix=f".{os.path.basename(out)}.", suffix=".tmp", dir=folder)
    os.close(fd)
    try:
        with open(tmp, "wb") as f:
            np.savez_compressed(f, **save)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, out)
    except BaseException:
        try:
            os.unlink(tmp)
        except FileNotFoundError:
            pass
        raise
    return out


def export_result_to_grim(result: Dict[str, Any], output_path: str, source_path: str = "", history: str = "",
                          preserve_raw_complex_amplitude: bool = True, format_version: Optional[int] = None
                          ) -> List[str]:
    """Write `result` (a solve_monostatic_rcs_2d result dict) as one `.grim` archive and
    return its absolute path in a list, as io/grim.py:1034-1090 does for monostatic results."""
    version = resolve_format_version(format_version)
    samples = result.get("samples", []) or []
    if not samples:
        raise ValueError("No solver samples were returned, nothing to export.")
    declared = [str(v).strip().upper() for v in list(result.get("polarizations", []) or [])]
    co = result.get("co_solved_samples", {}) or {}
    dual = declared == ["VV", "HH"] and isinstance(co, dict) and set(co) == {"VV", "HH"}
    if (declared or co) and not dual:
        raise ValueError("A declared co-polarized result must contain ordered polarizations [VV, HH] and exactly "
                         "matching co_solved_samples; single-channel fallback was refused.")
    mapping = result.get("polarization_mapping", {}) or {}
    if dual and set(mapping) != {"VV", "HH"}:
        raise ValueError("A co-polarized result must declare its VV/HH physical mapping.")
    bistatic = str(result.get("scattering_mode", "monostatic")).strip().lower() == "bistatic"
    quantity = str(result.get("rcs_linear_quantity", "sigma_2d")).strip().lower()
    if quantity not in ("sigma_2d", "sigma_3d"):
        raise ValueError("rcs_linear_quantity must be 'sigma_2d' or 'sigma_3d'.")
    log_unit = str(result.get("rcs_log_unit", "dBke"))
    label = None if dual else canonical_polarization(
        result.get("polarization_export") or result.get("polarization") or "TM")

    def build(channel_rows: Dict[str, List[Dict[str, Any]]], hist: str) -> Dict[str, Any]:
        if dual:
            internal = [str(mapping[c]).strip().upper() for c in ("VV", "HH")]
            if any(not lab for lab in internal):
                raise ValueError("Co-polarized GRIM mapping labels cannot be empty.")
            chans = [_channel_grid(channel_rows[c], c, quantity) for c in ("VV", "HH")]
            payload = _payload(chans, ("VV", "HH"), ",".join(internal), json.dumps(internal), quantity,
                               log_unit, source_path, hist)
        else:
            chans = [_channel_grid(channel_rows[label], label, quantity)]
            aliases = TM_ALIASES if label == "TM" else TE_ALIASES
            payload = _payload(chans, ("HH" if label == "TM" else "VV",), label, json.dumps(aliases), quantity,
                               log_unit, source_path, hist)
        if not np.all(np.isfinite(payload["field"].real) & np.isfinite(payload["field"].imag)):
            raise ValueError("Solver complex amplitudes are not finite after grid construction.")
        return payload

    channels = {c: list(co[c] or []) for c in ("VV", "HH")} if dual else {label: list(samples)}
    raw = bool(preserve_raw_complex_amplitude)
    if not bistatic:
        payload = build(channels, history)
        payload["solver_metadata_json"] = solver_metadata_json(result)
        return [os.path.abspath(_write(payload, output_path, version, raw))]
    # bistatic: one archive per incidence angle, published together (io/grim.py:1092-1170)
    groups = {c: _by_incidence(rows, c if dual else "") for c, rows in channels.items()}
    incidences = [set(g) for g in groups.values()]
    if any(s != incidences[0] for s in incidences[1:]):
        raise ValueError("VV and HH bistatic incidence axes differ and cannot share one GRIM collection.")
    root = _grim_path(output_path)[:-5]
    metadata = solver_metadata_json(result)
    entries = []
    for inc in sorted(incidences[0]):
        payload = build({c: g[inc] for c, g in groups.items()}, (history + f" | theta_inc_deg={inc:g}").strip(" |"))
        payload["solver_metadata_json"] = metadata
        entries.append((payload, f"{root}_{incidence_suffix(inc)}.grim"))
    return _publish(entries, version, raw)


def _by_incidence(rows: List[Dict[str, Any]], channel: str = "") -> Dict[float, List[Dict[str, Any]]]:
    """Rows grouped by theta_inc_deg, each group in its original order."""
    out: Dict[float, List[Dict[str, Any]]] = {}
    for i, row in enumerate(rows):
        try:
            inc = _finite(row, "theta_inc_deg", i)
        except ValueError as exc:
            raise ValueError(f"Invalid {channel + ' ' if channel else ''}bistatic solver sample at index {i}: "
                             f"{str(exc).split(': ', 1)[1]}") from exc
        out.setdefault(inc, []).append(row)
    return out


def load_grim(path: str) -> Dict[str, np.ndarray]:
    """Read a `.grim` archive into the format-1 key set (rcs_power, rcs_phase and, when the raw
    field is stored, rcs_amp_real / rcs_amp_imag), whichever format it was written in."""
    with np.load(path, allow_pickle=False) as data:
        arrays = {k: data[k] for k in data.files}
    if "grim_format_version" not in arrays:
        return arrays
    version = int(arrays.pop("grim_format_version"))
    if version != 2 or str(arrays.get("storage_kind")) != "complex" or str(arrays.get("storage_domain")) != "raw_field":
        raise ValueError(f"{path}: only format-2 raw-field complex archives are read here.")
    field = arrays.pop("rcs_complex")
    arrays.pop("storage_kind")
    arrays.pop("storage_domain")
    units = json.loads(str(arrays["units"]))
    if str(units.get("rcs_linear_quantity", "sigma_2d")).lower() == "sigma_2d":
        k = 2.0 * math.pi * np.asarray(arrays["frequencies"], float) * 1.0e9 / C0
        scale2 = (1.0 / (4.0 * k))[None, None, :, None]
    else:
        scale2 = 4.0 * math.pi
    arrays["rcs_power"] = ((field.real * field.real + field.imag * field.imag) * scale2).astype(np.float32)
    phase = np.angle(field)
    arrays["rcs_phase"] = np.where((field.real == 0.0) & (field.imag == 0.0), 0.0, phase).astype(np.float32)
    arrays["rcs_amp_real"] = field.real.copy()
    arrays["rcs_amp_imag"] = field.imag.copy()
    return arrays
