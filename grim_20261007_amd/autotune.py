"""Launch-shape autotuner for the HIP library.

Three launch shapes of the request path depend on the problem size and the device:

* far-pass source columns per wavefront (knob 0, keyed by n): few columns give
  more wavefronts (better fill of the 1024 SIMDs at small n), more columns
  give fewer chunk-boundary columns (shared between two wavefronts by atomic
  adds onto cleared entries) and longer register-carried walks;
* columns per wavefront of the factorization's 64-row triangular solves (knob 1,
  keyed by the columns of one call times its batch): more columns give more
  independent chains per wavefront but fewer wavefronts;
* columns per wavefront of the LU solve's 256-row block solves (knob 2, keyed by
  the right-hand-side columns times the batch): the same trade, against how many
  workgroups fit on the device at once.

``tune`` times every candidate on the current device with HIP events, on
synthetic star-shaped contours of the requested sizes, and returns a table;
``apply`` loads a table into the library (``grim_set_launch_shape``).  Shapes
never change results (the same operations compute every entry), so a tuned
server returns the same numbers as an untuned one.  Tables are plain JSON:
``save`` / ``load``.
"""
from __future__ import annotations

import json
import math
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np
import torch

from . import engine, native, workloads
from .mesh import build_mesh
from .solver import wavenumber

FAR_COLUMNS, TRSM_COLUMNS, SOLVE_COLUMNS = 0, 1, 2
# table measured by tune() on an MI355X (far pass n = 500-12000, factors of n = 2000, solves of
# 90-1440 angles)
MI355X_TABLE = Path(__file__).resolve().parent / "tuning_mi355x.json"
KNOB_NAMES = {FAR_COLUMNS: "far_columns", TRSM_COLUMNS: "trsm_columns", SOLVE_COLUMNS: "solve_columns"}
CHOICES = {FAR_COLUMNS: (1, 2, 4, 8, 16, 32), TRSM_COLUMNS: (1, 2, 4), SOLVE_COLUMNS: (1, 2, 4)}


def set_shape(knob: int, min_size: int, value: int) -> None:
    native.check("grim_set_launch_shape", native.lib().grim_set_launch_shape(int(knob), int(min_size), int(value)))


def clear() -> None:
    native.check("grim_clear_launch_shapes", native.lib().grim_clear_launch_shapes())


def shape(knob: int, size: int) -> int:
    """The table's value for `size` (0: the library's built-in rule)."""
    return int(native.lib().grim_launch_shape(int(knob), int(size)))


def _time(fn, repeats: int) -> float:
    fn()                                             # warm-up (and first-touch of buffers)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(repeats):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / repeats


def _contour(n: int, device) -> tuple:
    rng = np.random.default_rng(20261203)
    a, phi = workloads.star_shape(rng)
    pts = workloads.star_points(n, 0.5, a, phi)
    f = workloads.frequency_for_segments_per_wavelength(pts, 10.0)
    mesh = build_mesh(workloads._snapshot(pts, "tune"), "meters", f, max_panels=10 ** 6)
    return mesh, engine.DeviceContour(mesh, device), wavenumber(f)


def _boundaries(sizes: Sequence[int]) -> List[int]:
    """Entry i applies from the geometric midpoint between sizes[i-1] and sizes[i]."""
    out = [0]
    for lo, hi in zip(sizes[:-1], sizes[1:]):
        out.append(int(math.sqrt(lo * hi)))
    return out


def tune_far(sizes: Iterable[int], device=None, repeats: int = 5) -> Dict[int, Dict[int, float]]:
    """Far-pass time (ms) per candidate columns-per-wavefront at each n."""
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    native.ensure_device(dev.index or 0)
    stream = engine._stream(dev)
    times: Dict[int, Dict[int, float]] = {}
    for n in sorted(set(int(v) for v in sizes)):
        mesh, c, k = _contour(n, dev)
        ws = engine.assembly_workspace(c.n, dev)
        pair = torch.empty((2, c.n, c.n), dtype=torch.complex128, device=dev)
        args = (c.n, c.p0.data_ptr(), c.p1.data_ptr(), c.next.data_ptr(), c.prev.data_ptr(), c.fix_rows.data_ptr(),
                int(c.fix_rows.numel()), float(k), mesh.far_order(k), pair[1].data_ptr(), c.n, pair[0].data_ptr(),
                c.n, ws.data_ptr(), ws.numel(), engine.NEAR_PER_ELEMENT, stream)

        def run():
            native.check("grim_assemble_far", native.lib().grim_assemble_far(*args))

        times[n] = {}
        for value in CHOICES[FAR_COLUMNS]:
            set_shape(FAR_COLUMNS, 0, value)
            try:
                times[n][value] = _time(run, repeats)
            finally:
                set_shape(FAR_COLUMNS, 0, 0)
        del pair, ws
    return times


def _test_matrix(n: int, dev, g) -> torch.Tensor:
    z = torch.randn((2, n, n), dtype=torch.complex128, generator=g)
    z += n * torch.eye(n, dtype=torch.complex128)      # well conditioned: timing only
    return z.to(dev)


def tune_trsm(n: int = 2000, precisions: Sequence[str] = ("double", "mixed"), device=None, repeats: int = 3
              ) -> Dict[int, Dict[int, float]]:
    """Two-system LU factor time (ms, summed over `precisions`) per candidate columns-per-wavefront
    of the factorization's triangular solves.  One factor calls them with every column count from
    2 (n - 64) down, so the result is one table entry, under key 0."""
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    native.ensure_device(dev.index or 0)
    z = _test_matrix(n, dev, torch.Generator(device="cpu").manual_seed(11))
    times: Dict[int, Dict[int, float]] = {0: {}}
    for value in CHOICES[TRSM_COLUMNS]:
        set_shape(TRSM_COLUMNS, 0, value)
        try:
            times[0][value] = sum(_time(lambda: engine.factor(z, p), repeats) for p in precisions)
        finally:
            set_shape(TRSM_COLUMNS, 0, 0)
    return times


def tune_solve(column_counts: Iterable[int], n: int = 2000, device=None, repeats: int = 5
               ) -> Dict[int, Dict[int, float]]:
    """FP64 plus complex64 two-system LU solve time (ms) per candidate columns-per-wavefront of
    the block solves, for right-hand sides of A columns (the call then has 2 A columns: both
    polarizations)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    native.ensure_device(dev.index or 0)
    g = torch.Generator(device="cpu").manual_seed(11)
    z = _test_matrix(n, dev, g)
    facs = [engine.factor(z, p) for p in ("double", "mixed")]
    times: Dict[int, Dict[int, float]] = {}
    for na in sorted(set(int(v) for v in column_counts)):
        b = torch.randn((2, na, n), dtype=torch.complex128, generator=g).to(dev)
        rhss = [b.to(f.lu.dtype) for f in facs]          # each factor solves in its own precision
        cols = 2 * na
        times[cols] = {}
        for value in CHOICES[SOLVE_COLUMNS]:
            set_shape(SOLVE_COLUMNS, 0, value)
            try:
                times[cols][value] = sum(_time(lambda: engine.lu_solve_inplace(f, rhs.clone()), repeats)
                                         for f, rhs in zip(facs, rhss))
            finally:
                set_shape(SOLVE_COLUMNS, 0, 0)
    return times


def _table_from(times: Dict[int, Dict[int, float]]) -> List[List[int]]:
    sizes = sorted(times)
    best = [min(times[s], key=times[s].get) for s in sizes]
    entries = []
    for start, value in zip(_boundaries(sizes), best):
        if entries and entries[-1][1] == value:
            continue                                 # merge equal neighbours
        entries.append([start, value])
    return entries


def tune(far_sizes: Sequence[int] = (500, 1000, 2000, 4000, 8000),
         angle_counts: Sequence[int] = (90, 360, 720, 1440), factor_n: int = 2000, device=None,
         repeats: int = 5) -> dict:
    """Measure and return a launch-shape table for this device (does not apply it)."""
    measured = {"far_columns": tune_far(far_sizes, device, repeats),
                "trsm_columns": tune_trsm(factor_n, device=device, repeats=max(1, repeats // 2)),
                "solve_columns": tune_solve(angle_counts, device=device, repeats=repeats)}
    name = torch.cuda.get_device_name(torch.device(device) if device is not None else None)
    out = {"device": name, "measured_ms": {}}
    for key, times in measured.items():
        out[key] = _table_from(times)
        out["measured_ms"][key] = {str(n): {str(v): t for v, t in d.items()} for n, d in times.items()}
    return out


def apply(table: dict) -> None:
    """Replace the library's launch-shape table by `table` (from tune or load)."""
    clear()
    for knob, name in KNOB_NAMES.items():
        for start, value in table.get(name, []):
            set_shape(knob, int(start), int(value))


def save(table: dict, path) -> None:
    Path(path).write_text(json.dumps(table, indent=1, sort_keys=True))


def load(path) -> dict:
    return json.loads(Path(path).read_text())


def resolve(tuning: Optional[object]) -> Optional[dict]:
    """A table from a dict, a JSON path, "mi355x" (the shipped table), or None."""
    if tuning is None or isinstance(tuning, dict):
        return tuning
    if tuning == "mi355x":
        return load(MI355X_TABLE)
    return load(tuning)
