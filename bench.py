"""Benchmark: RCS solves/s and p99 latency, mixed precision vs plain FP64 on one MI355X.

Workload (the NACA 0012 configuration): NACA 0012 section, chord 1 m, closed
trailing edge, cosine-spaced stations, 1000 panels per surface (N = 2000
unknowns), frequency set so the longest panel is lambda/10, 720 monostatic
angles (0.5 degree steps), both polarizations.  One request is one
co-polarized solve on synthetic geometry: one fused TM+TE assembly, then for
both polarizations (batched) LU factorization, the 720-column solve and the
far-field projection -- 2 RCS solves.

Measured forms:
  * plain FP64 path: one request at a time, FP64 LU and the backward-error
    gate, per-stage HIP-event timers;
  * mixed form one request at a time (complex64 LU + FP64-residual
    refinement), per-stage timers and its accuracy against the FP64 path;
  * serving form (the headline value): --steps distinct requests (no cache
    reuse, so every request assembles and factors) with --workers requests in
    flight on their own streams, each in the precision the flop model picks
    for it (the server's default); forced mixed and forced FP64 serving of
    the same requests beside it;
  * a 32-request Poisson stream of the serving configuration (N up to about
    12 000, cache reuse): plain FP64 against the serving form.
Each request uploads its contour (a few tens of KB) as a serving system does;
everything else is produced and consumed on the GPU.

Prints ONE JSON line (N = 1).  --gpus > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6       # TFLOP/s, MI355X dense FP64 (vector and matrix)
PEAK_FP32 = 157.3      # TFLOP/s, MI355X dense FP32
PEAK_HBM = 8000.0      # GB/s
FLOPS_PER_EVAL = 125   # FP64 flops per kernel evaluation (distance, Hankel pair, projections, accumulation)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128, help="requests in the timed serving region")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workers", type=int, default=4, help="requests in flight in the serving form")
    ap.add_argument("--batch", type=int, default=4,
                    help="requests of equal size a serving worker runs as one batched factor + solve")
    ap.add_argument("--n-side", type=int, default=1000)
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sheet", action="store_true",
                    help="skip the sheet (TYPE 1) and lossy-coating request timings")
    ap.add_argument("--tuning", default="mi355x",
                    help="launch-shape table: 'mi355x' (shipped, measured by the autotuner), a JSON path, "
                         "'tune' (measure now) or 'none' (built-in rules)")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--workload", choices=("naca", "sweep", "stream", "large"), default="naca",
                    help="naca: the headline NACA 0012 N=2000 request (default); sweep: 64-frequency star-curve "
                         "sweep; stream: Poisson serving stream with cache reuse; large: one N=40000 request")
    ap.add_argument("--large-n", type=int, default=40000)
    ap.add_argument("--stream-count", type=int, default=64,
                    help="requests of the stream workload; the default run reports a 32-request stream (0: skip)")
    ap.add_argument("--stream-rate", type=float, default=200.0,
                    help="Poisson arrival rate (requests/s); the default saturates the server")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def copy_probe(mib: int):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE counters calibrate HBM traffic."""
    import torch
    n = mib * 2**20 // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "launches": 3}))


def workload(n_side: int, nangles: int):
    from grim_20261007_amd import workloads
    from grim_20261007_amd.mesh import build_mesh
    pts = workloads.naca0012_points(n_side)
    f = workloads.frequency_for_segments_per_wavelength(pts, 10.0)
    mesh = build_mesh(workloads._snapshot(pts, "naca0012"), "meters", f)
    angles = np.arange(nangles) * (360.0 / nangles)
    return pts, mesh, f, angles


class Timer:
    """HIP events on the stream the library launches on (torch's current stream)."""

    def __init__(self, torch):
        self.torch = torch
        self.acc = {}

    def run(self, name, fn):
        t = self.torch
        a = t.cuda.Event(enable_timing=True)
        b = t.cuda.Event(enable_timing=True)
        a.record()
        out = fn()
        b.record()
        self.acc.setdefault(name, []).append((a, b))
        return out

    def mean_ms(self, name):
        self.torch.cuda.synchronize()
        v = [a.elapsed_time(b) for a, b in self.acc.get(name, [])]
        return float(np.mean(v)) if v else float("nan")


def sheet_workload(n: int, nangles: int):
    """A 75 ohm resistive strip bent into a 120-degree circular arc of radius 0.5 m (TYPE 1 open
    contour), n panels of lambda/10."""
    from grim_20261007_amd.mesh import C0, build_mesh
    t = np.linspace(np.pi / 6, 5 * np.pi / 6, n + 1)
    pts = 0.5 * np.stack((np.cos(t), np.sin(t)), axis=1)
    chord = float(np.linalg.norm(pts[1] - pts[0]))
    f = C0 / (10.0 * chord) / 1e9
    seg = {"name": "arc", "seg_type": 1, "properties": ["1", "1", "1", "0", "0"],
           "point_pairs": [{"x1": float(a[0]), "y1": float(a[1]), "x2": float(b[0]), "y2": float(b[1])}
                           for a, b in zip(pts[:-1], pts[1:])]}
    snap = {"segments": [seg], "ibcs": [["1", "constant", "75", "0", "0", "0"]], "dielectrics": []}
    layout = build_mesh(snap, "meters", f)
    return layout, f, np.arange(nangles) * (360.0 / nangles)


def sheet_summary(n: int = 2000, nangles: int = 720, reps: int = 3):
    """One sheet request (TE + TM, each polarization on its own mesh, both systems factored and solved
    as one batch of two) timed stage by stage with HIP events."""
    import torch
    from grim_20261007_amd import engine
    from grim_20261007_amd.solver import wavenumber
    dev = torch.device("cuda", 0)
    layout, f, angles = sheet_workload(n, nangles)
    k = wavenumber(f)
    angles_t = torch.tensor(angles, dtype=torch.float64, device=dev)
    meshes = [layout.mesh(p) for p in ("TE", "TM")]
    assert np.array_equal(meshes[0].elements, meshes[1].elements)
    sheets = [engine.DeviceSheet(m, dev) for m in meshes]
    ws = engine.sheet_workspace(meshes[0].ne, dev)
    nn = meshes[0].n
    timer = Timer(torch)
    for rep in range(reps + 1):
        tm = timer.run if rep else (lambda name, fn: fn())
        z = torch.empty((2, nn, nn), dtype=torch.complex128, device=dev)
        rhs, proj = [], []
        # the arc's TE and TM meshes coincide: one traversal forms both systems
        tm("assemble", lambda: engine.sheet_systems(sheets[0], k, ws, te=z[0], tm=z[1]))
        for i, pol in enumerate(("TE", "TM")):
            bu, bdn = tm("loads", lambda: engine.sheet_loads(sheets[i], k, angles_t))
            rhs.append(engine.sheet_rhs(sheets[i], pol, bu, bdn))
            proj.append(bu if pol == "TM" else bdn)
        rhs = torch.stack(rhs)
        fac = tm("factor_double", lambda: engine.factor(z, "double", check=False))
        x, info = tm("solve_double", lambda: engine.solve(fac, rhs))
        for i in range(2):
            tm("farfield", lambda: engine.farfield_potential(x[i], proj[i], k))
    stage = {name: round(timer.mean_ms(name) * (2 if name in ("loads", "farfield") else 1), 3) for name in timer.acc}
    total = sum(stage.values())
    return {"workload": f"75 ohm resistive strip, 120-degree arc of radius 0.5 m, {meshes[0].ne} panels "
                        f"({nn} nodes, strip ends pinned in TE), f={f:.4f} GHz (panels lambda/10), {len(angles)} "
                        "monostatic angles, TE + TM", "request_ms": round(total, 3),
            "solves_per_s": round(2e3 / total, 3), "stage_ms": stage,
            "source": "HIP events; stage loads / farfield are both polarizations"}


def lossy_workload(n_out: int = 800, n_in: int = 700):
    """A PEC cylinder (radius 45 mm, n_in panels) under a 5 mm coating of the lossy medium of the
    reference's airfoil.geo (eps = 101.19 - 11.88j, mu = 3.45 - 3.47j; |n| = 22.3), outer radius 50 mm
    with n_out panels, at 3.4 GHz (outer panels about lambda_coating / 10)."""
    from grim_20261007_amd import workloads
    from grim_20261007_amd.mesh import build_mesh
    segs = []
    for i, (pts, seg_type) in enumerate(((workloads.circle_points(0.045, n_in), 4),
                                          (workloads.circle_points(0.05, n_out), 3))):
        seg = workloads._snapshot(pts, f"c{i}")["segments"][0]
        seg["seg_type"] = seg_type
        seg["properties"] = [str(seg_type), "1", "0", "1", "0"]
        segs.append(seg)
    snap = {"segments": segs, "ibcs": [], "dielectrics": [["1", "101.19", "-11.88", "3.45", "-3.47"]]}
    f = 3.4
    return build_mesh(snap, "meters", f), f


def lossy_summary(nangles: int = 720, reps: int = 2):
    """One request of the lossy coated cylinder (TE + TM regional systems, 720 angles), stage by stage
    with HIP events."""
    import torch
    from grim_20261007_amd import engine
    from grim_20261007_amd.solver import wavenumber
    dev = torch.device("cuda", 0)
    mesh, f = lossy_workload()
    k = wavenumber(f)
    angles_t = torch.tensor(np.arange(nangles) * (360.0 / nangles), dtype=torch.float64, device=dev)
    contour = engine.DeviceContour(mesh, dev)
    ws = engine.assembly_workspace(mesh.n, dev)
    plan = engine.region_plan(mesh)
    timer = Timer(torch)
    for rep in range(reps + 1):
        tm = timer.run if rep else (lambda name, fn: fn())
        pair = tm("assemble", lambda: engine.region_systems(contour, k, ws, plan))
        bu, bdn = tm("loads", lambda: engine.loads(contour, k, angles_t))
        rhs = engine.region_rhs(plan, bu, bdn)
        fac = tm("factor_double", lambda: engine.factor(pair, "double", check=False))
        x, info = tm("solve_double", lambda: engine.solve(fac, rhs))
        del pair, fac, x
    stage = {name: round(timer.mean_ms(name), 3) for name in timer.acc}
    total = sum(stage.values())
    return {"workload": f"PEC cylinder under a 5 mm coating of airfoil.geo's lossy medium (eps 101.19-11.88j, mu "
                        f"3.45-3.47j), {mesh.n} panels ({plan.n_dof} unknowns per polarization), f={f} GHz, "
                        f"{nangles} monostatic angles, TE + TM", "request_ms": round(total, 3),
            "solves_per_s": round(2e3 / total, 3), "stage_ms": stage,
            "source": "HIP events; assemble = air and lossy-region operators plus the routes, both polarizations"}


def request(engine, contour, k, angles_t, precision, timer=None, keep=False):
    """One co-polarized request; returns per-polarization (amplitude, sigma, residual, steps)."""
    import torch
    from grim_20261007_amd import native
    tm = timer.run if timer else (lambda name, fn: fn())
    n = contour.n
    pair_buf = torch.empty((2, n, n), dtype=torch.complex128, device=contour.device)
    z_te, z_tm = pair_buf[0], pair_buf[1]
    ws = request.ws
    st = torch.cuda.current_stream().cuda_stream
    args = lambda: (n, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(),
                    contour.prev.data_ptr(), contour.fix_rows.data_ptr(), int(contour.fix_rows.numel()), float(k),
                    contour.mesh.far_order(k), z_tm.data_ptr(), n, z_te.data_ptr(), n, ws.data_ptr(), ws.numel(),
                    engine.NEAR_PER_ELEMENT, st)
    native.check("far", tm("assemble_far", lambda: native.lib().grim_assemble_far(*args())))
    native.check("near", tm("assemble_near", lambda: native.lib().grim_assemble_near(*args())))
    bu, bdn = tm("loads", lambda: engine.loads(contour, k, angles_t))
    pair = z_tm._base                                   # [TE, TM]
    rhs = torch.stack((-bdn, -bu))
    fac = tm(f"factor_{precision}", lambda: engine.factor(pair, precision, check=False))
    x, info = tm(f"solve_{precision}", lambda: engine.solve(fac, rhs, gate=(precision == "double")))
    rel = tm("residual", lambda: engine.relative_residual(pair, x, rhs, info.get("residual")))
    out = {}
    for i, pol in enumerate(("TE", "TM")):
        amp, sig = tm("farfield", lambda: engine.farfield(x[i], bu, bdn, k))
        out[pol] = (amp, sig, rel[i], info.get("refinement_steps", 0))
    native.check("assembly status", native.lib().grim_workspace_status(ws.data_ptr(), st))
    return out


def far_pair_count(mesh):
    """Unordered far pairs (centre ratio >= 3, no shared node) -- the algorithmic count."""
    c = 0.5 * (mesh.p0 + mesh.p1)
    L = mesh.lengths
    n = mesh.n
    total = 0
    for s in range(0, n, 256):
        rows = np.arange(s, min(n, s + 256))
        d = np.linalg.norm(c[rows][:, None, :] - c[None, :, :], axis=2)
        far = d / np.maximum(np.maximum(L[rows][:, None], L[None, :]), 1e-12) >= 3.0
        for a in (rows, mesh.next[rows]):
            for b in (np.arange(n), mesh.next):
                far &= a[:, None] != b[None, :]
        total += int(np.sum(far))
    return total // 2


def cpu_baseline(mesh, f, angles, seconds):
    """Whole co-polarized requests on the host cores, timed end to end: the far pass with the
    reference's own far.c (built from its source into oracle/_ref, symmetric tiles on one thread
    per core, as its tile scheduler runs them), near pairs and mass terms with the oracle's
    restatement of the reference's NumPy rules, SciPy/LAPACK LU and solves of both systems, the
    matched far field.  Without oracle/_ref, the older bounded sample of the oracle alone."""
    from oracle import bem2d, ref_far
    if ref_far.available():
        om = bem2d.Mesh(mesh.nodes, mesh.elements, mesh.p0, mesh.p1, 0.5 * (mesh.p0 + mesh.p1), mesh.lengths,
                        mesh.normals)
        cores = len(os.sched_getaffinity(0))
        times, stages = [], {}
        t_start = time.perf_counter()
        while not times or (len(times) < 3 and time.perf_counter() - t_start < 0.5 * seconds):
            t0 = time.perf_counter()
            _, st = ref_far.full_request(om, f, angles, cores)
            times.append(time.perf_counter() - t0)
            for key, v in st.items():
                stages[key] = stages.get(key, 0.0) + v
        per = float(np.mean(times))
        stage_txt = ", ".join(f"{key} {v / len(times):.2f} s" for key, v in stages.items())
        return {"value": 2.0 / per, "unit": "RCS solves/s", "cores": cores, "kind": "reference",
                "extrapolated": False,
                "sample": (f"{len(times)} whole co-polarized requests (N={mesh.n}, {len(angles)} angles, TE + TM), "
                           f"{per:.2f} s each: far pass by the reference's own far.c on {cores} threads, near pairs "
                           f"by the oracle (one thread), LAPACK with {cores} threads; per request {stage_txt}")}
    return _cpu_baseline_sampled(mesh, f, angles, seconds)


def _cpu_baseline_sampled(mesh, f, angles, seconds):
    """The oracle (NumPy/SciPy restatement) on a bounded sample of one request, scaled to solves/s."""
    import scipy.linalg as sl
    from oracle import bem2d
    om = bem2d.Mesh(mesh.nodes, mesh.elements, mesh.p0, mesh.p1, 0.5 * (mesh.p0 + mesh.p1), mesh.lengths,
                    mesh.normals)
    k = bem2d.frequency_k(f)
    n = mesh.n
    q = bem2d.far_order(om, k)
    # far rows sample
    rows_done, t_far = 0, 0.0
    t0 = time.perf_counter()
    while rows_done < n and time.perf_counter() - t0 < 0.3 * seconds:
        rows = np.arange(rows_done, min(n, rows_done + 16))
        far = bem2d.far_mask(om, rows)
        bem2d._far_rows(om, k, rows, far, q)
        rows_done += len(rows)
    t_far = (time.perf_counter() - t0) * n / rows_done
    # near pairs sample
    near = np.argwhere(~bem2d.far_mask(om, np.arange(min(n, 64))))
    t0 = time.perf_counter()
    done = 0
    for e, g in near:
        bem2d.near_blocks(om, k, int(e), int(g))
        done += 1
        if time.perf_counter() - t0 > 0.2 * seconds:
            break
    per_pair = (time.perf_counter() - t0) / max(done, 1)
    total_near = len(near) * n / min(n, 64)
    t_near = per_pair * total_near
    # LU + 720-column solve + far field at the full size, one polarization
    rng = np.random.default_rng(0)
    z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + n * np.eye(n)
    t0 = time.perf_counter()
    lu = sl.lu_factor(z)
    t_lu = time.perf_counter() - t0
    b = rng.standard_normal((n, len(angles))) + 0j
    t0 = time.perf_counter()
    x = sl.lu_solve(lu, b)
    t_solve = time.perf_counter() - t0
    t0 = time.perf_counter()
    bem2d.farfield(om, k, angles, x)
    bem2d.rhs(om, k, angles, "TM")
    t_ff = time.perf_counter() - t0
    per_request = t_far + t_near + 2 * (t_lu + t_solve + t_ff)
    cores = len(os.sched_getaffinity(0))
    return {"value": 2.0 / per_request, "unit": "RCS solves/s", "cores": cores, "kind": "port", "extrapolated": True,
            "sample": (f"oracle far rows {rows_done}/{n} and {done}/{int(total_near)} near pairs timed and "
                       f"scaled; full-size LAPACK LU ({t_lu:.2f} s), {len(angles)}-column solve "
                       f"({t_solve:.2f} s) and far field; per co-polarized request {per_request:.1f} s")}


def far_traffic():
    """HBM bytes per far-pass launch from the committed rocprofv3 counter summary, or None."""
    import re
    path = os.path.join(ROOT, "profiles", "far_pass_naca0012_n2000.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def far_valu():
    """VALUBusy and VALUUtilization (percent) of the far kernel from the committed counter
    summary, or None."""
    import re
    path = os.path.join(ROOT, "profiles", "far_pass_naca0012_n2000.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    busy = re.search(r"valu_busy_percent:\s*([0-9.]+)", text)
    util = re.search(r"valu_utilization_percent:\s*([0-9.]+)", text)
    if not (busy and util):
        return None
    return {"valu_busy_percent": float(busy.group(1)), "valu_utilization_percent": float(util.group(1)),
            "source": "profiles/far_pass_naca0012_n2000.md (rocprofv3 --pmc VALUBusy VALUUtilization)"}


def serve(engine_mod, mesh, f, angles, precision, workers, nreq, warm, batch=1):
    """Closed-loop serving: nreq distinct requests (no cache reuse) with `workers` in flight, each
    worker running up to `batch` waiting requests as one batch.  Returns the results, the elapsed
    time, service times (ms) and the precision each request ran in ('auto' resolves per request
    by the flop model)."""
    from grim_20261007_amd import serving
    srv = serving.Server(workers=workers, batch=batch)
    try:
        serving.run_stream(srv, [serving.Request(mesh, f, angles, precision=precision, key=f"w{precision}{i}")
                                 for i in range(warm)])
        reqs = [serving.Request(mesh, f, angles, precision=precision, key=f"r{precision}{i}") for i in range(nreq)]
        res, elapsed = serving.run_stream(srv, reqs)
    finally:
        srv.close()
    svc = np.array([r.finished - r.started for r in res]) * 1e3
    chosen = {p: sum(r.precision == p for r in reqs) for p in ("double", "mixed")}
    return res, elapsed, svc, chosen


def stream_summary(workers: int, count: int = 32, seed: int = 11, rate_hz: float = 200.0, batch: int = 1):
    """The Poisson serving stream (star-shaped curves, N log-uniform in [200, 12000], 90 / 360 /
    1440 angles, 30 % reusing an earlier system): plain FP64 back to back against the serving
    form in the cost-model precision ('auto') and in forced mixed precision, on the same seeded
    requests and arrivals."""
    from grim_20261007_amd import engine, serving, workloads
    from grim_20261007_amd.mesh import build_mesh
    stream = workloads.serving_stream(count, seed=seed)
    meshes = {}
    for it in stream:
        if it["key"] not in meshes:
            meshes[it["key"]] = build_mesh(workloads._snapshot(it["points"], "star"), "meters", it["frequency_ghz"])
    ms = [meshes[it["key"]] for it in stream]
    freqs = [it["frequency_ghz"] for it in stream]
    angs = [it["angles"] for it in stream]
    _requests_plain(engine, ms[:1], freqs[:1], angs[:1], "double")
    t64, lat64 = _requests_plain(engine, ms, freqs, angs, "double")
    out = {"requests": len(stream), "n_min": min(m.n for m in ms), "n_max": max(m.n for m in ms),
           "reused": int(sum(it["reused"] for it in stream)), "arrival_rate_hz": rate_hz,
           "fp64_plain": {"value": round(2 * len(stream) / t64, 3), "p99_ms": round(float(np.percentile(lat64, 99)), 3),
                          "form": "back to back, no arrivals, no reuse"}}
    for precision in ("auto", "mixed"):
        srv = serving.Server(workers=workers, batch=batch)
        try:
            serving.run_stream(srv, [serving.Request(ms[0], freqs[0], angs[0], key="warm")])
            reqs = [serving.Request(m, f, a, precision=precision, key=it["key"])
                    for m, f, a, it in zip(ms, freqs, angs, stream)]
            res, tm = serving.run_stream(srv, reqs, rate_hz=rate_hz, rng=np.random.default_rng(seed))
            hits = srv.cache.hits
        finally:
            srv.close()
        lat = np.array([r.latency_s for r in res]) * 1e3
        out[f"{precision}_serving"] = {
            "value": round(2 * len(stream) / tm, 3), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "cache_hits": hits,
            "requests_mixed": int(sum(r.precision == "mixed" for r in reqs)),
            "vs_fp64_plain": round(t64 / tm, 3), "max_residual": max(max(r.residual.values()) for r in res)}
    return out


def _requests_plain(engine, meshes, freqs, angle_sets, precision):
    """One request at a time on the current stream, no reuse: (elapsed seconds, per-request ms)."""
    import torch
    from grim_20261007_amd import serving
    dev = torch.device("cuda", 0)
    lat, ws = [], {}
    t0 = time.perf_counter()
    for mesh, f, ang in zip(meshes, freqs, angle_sets):
        t1 = time.perf_counter()
        serving.execute(serving.Request(mesh, f, ang, precision=precision), dev, ws)
        torch.cuda.synchronize()
        lat.append(1e3 * (time.perf_counter() - t1))
    return time.perf_counter() - t0, np.array(lat)


def sweep_workload(args):
    """64-frequency sweep of one random star-shaped PEC curve (R = 0.5 m, seed 20261203), N from
    100 panels per wavelength of the longest panel (about 360-3600), 360 angles, TE + TM."""
    import torch
    from grim_20261007_amd import engine, serving, workloads
    from grim_20261007_amd.mesh import build_mesh
    cases = workloads.star_sweep()
    meshes, freqs, angle_sets = [], [], []
    for pts, f, ang in cases:
        meshes.append(build_mesh(workloads._snapshot(pts, "star"), "meters", f))
        freqs.append(f)
        angle_sets.append(ang)
    ns = [m.n for m in meshes]
    # plain FP64 path, one request at a time (warm once on the smallest)
    _requests_plain(engine, meshes[:1], freqs[:1], angle_sets[:1], "double")
    t64, lat64 = _requests_plain(engine, meshes, freqs, angle_sets, "double")
    # mixed-precision serving form
    srv = serving.Server(workers=args.workers)
    try:
        serving.run_stream(srv, [serving.Request(meshes[0], freqs[0], angle_sets[0], key="warm")])
        reqs = [serving.Request(m, f, a, precision="mixed", key=f"s{i}")
                for i, (m, f, a) in enumerate(zip(meshes, freqs, angle_sets))]
        res, tm = serving.run_stream(srv, reqs)
    finally:
        srv.close()
    lat = np.array([r.latency_s for r in res]) * 1e3
    solves = 2 * len(meshes)
    print(json.dumps({
        "workload": "star-curve frequency sweep", "requests": len(meshes), "n_min": min(ns), "n_max": max(ns),
        "angles": 360, "frequencies_ghz": [round(freqs[0], 4), round(freqs[-1], 4)],
        "mixed_serving": {"value": round(solves / tm, 3), "unit": "RCS solves/s", "workers": args.workers,
                          "p50_ms": round(float(np.percentile(lat, 50)), 3),
                          "p99_ms": round(float(np.percentile(lat, 99)), 3),
                          "max_residual": max(max(r.residual.values()) for r in res)},
        "fp64_plain": {"value": round(solves / t64, 3), "unit": "RCS solves/s",
                       "p50_ms": round(float(np.percentile(lat64, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat64, 99)), 3)},
        "speedup": round(t64 / tm, 3)}))


def stream_workload(args):
    """Poisson serving stream of star-shaped curves: N log-uniform in [200, 12000], angle sets of
    90 / 360 / 1440, 30 % of requests reusing an earlier (geometry, frequency) with new angles."""
    import torch
    from grim_20261007_amd import engine, serving, workloads
    from grim_20261007_amd.mesh import build_mesh
    stream = workloads.serving_stream(args.stream_count)
    meshes, cache = [], {}
    for item in stream:
        if item["key"] not in cache:
            cache[item["key"]] = build_mesh(workloads._snapshot(item["points"], "star"), "meters",
                                            item["frequency_ghz"])
        meshes.append(cache[item["key"]])
    freqs = [it["frequency_ghz"] for it in stream]
    angle_sets = [it["angles"] for it in stream]
    ns = [m.n for m in meshes]
    _requests_plain(engine, meshes[:1], freqs[:1], angle_sets[:1], "double")
    t64, lat64 = _requests_plain(engine, meshes, freqs, angle_sets, "double")
    forms = {}
    for precision in ("mixed", "auto"):
        srv = serving.Server(workers=args.workers)
        try:
            serving.run_stream(srv, [serving.Request(meshes[0], freqs[0], angle_sets[0], key="warm")])
            reqs = [serving.Request(m, f, a, precision=precision, key=it["key"])
                    for m, f, a, it in zip(meshes, freqs, angle_sets, stream)]
            res, tm = serving.run_stream(srv, reqs, rate_hz=args.stream_rate, rng=np.random.default_rng(11))
            forms[precision] = (res, tm, srv.cache.hits, sum(r.precision == "mixed" for r in reqs))
        finally:
            srv.close()
    res, tm, hits, _ = forms["mixed"]
    lat = np.array([r.latency_s for r in res]) * 1e3
    res_a, tm_a, hits_a, n_mixed = forms["auto"]
    lat_a = np.array([r.latency_s for r in res_a]) * 1e3
    solves = 2 * len(meshes)
    print(json.dumps({
        "workload": "Poisson serving stream", "requests": len(meshes), "n_min": min(ns), "n_max": max(ns),
        "reused": int(sum(it["reused"] for it in stream)), "cache_hits": hits,
        "arrival_rate_hz": args.stream_rate,
        "mixed_serving": {"value": round(solves / tm, 3), "unit": "RCS solves/s", "workers": args.workers,
                          "p50_ms": round(float(np.percentile(lat, 50)), 3),
                          "p99_ms": round(float(np.percentile(lat, 99)), 3),
                          "note": "value includes waiting for Poisson arrivals; latency is arrival to result",
                          "max_residual": max(max(r.residual.values()) for r in res)},
        "auto_serving": {"value": round(solves / tm_a, 3), "unit": "RCS solves/s", "workers": args.workers,
                         "p50_ms": round(float(np.percentile(lat_a, 50)), 3),
                         "p99_ms": round(float(np.percentile(lat_a, 99)), 3), "requests_mixed": int(n_mixed),
                         "cache_hits": hits_a, "max_residual": max(max(r.residual.values()) for r in res_a),
                         "note": "per request the cheaper form by the flop model (mixed iff angles < N/18)"},
        "fp64_plain": {"value": round(solves / t64, 3), "unit": "RCS solves/s",
                       "p50_ms": round(float(np.percentile(lat64, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat64, 99)), 3),
                       "note": "back-to-back, no arrivals, no reuse"}}))


def large_workload(args):
    """One elongated star-shaped body (R = 2 m, aspect 4:1), N = args.large_n, longest panel
    lambda/10, 1440 angles, TE + TM: complex64 LU + FP64 refinement against the FP64 LU."""
    import torch
    from grim_20261007_amd import engine, native, workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import wavenumber
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(20261203)
    a, phi = workloads.star_shape(rng)
    pts = workloads.star_points(args.large_n, 2.0, a, phi, aspect=4.0)
    f = workloads.frequency_for_segments_per_wavelength(pts, 10.0)
    mesh = build_mesh(workloads._snapshot(pts, "star"), "meters", f, max_panels=max(20000, args.large_n))
    n = mesh.n
    k = wavenumber(f)
    angles = torch.tensor(np.arange(1440) * 0.25, dtype=torch.float64, device=dev)
    contour = engine.DeviceContour(mesh, dev)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    out = {"workload": "large single request", "n": n, "angles": 1440, "frequency_ghz": f}
    e0, e1 = ev(), ev()
    e0.record()
    z_tm, z_te = engine.assemble(contour, k)
    e1.record()
    torch.cuda.synchronize()
    out["assemble_ms"] = e0.elapsed_time(e1)
    pair = z_tm._base
    bu, bdn = engine.loads(contour, k, angles)
    rhs = torch.stack((-bdn, -bu))
    lu_flops = 2 * 8.0 / 3.0 * n ** 3
    trsm_flops = 2 * 8.0 * n * n * 1440
    sig = {}
    for precision in ("mixed", "double"):
        e0, e1, e2 = ev(), ev(), ev()
        e0.record()
        fac = engine.factor(pair, precision)
        e1.record()
        x, info = engine.solve(fac, rhs, gate=(precision == "double"))
        e2.record()
        torch.cuda.synchronize()
        rel = engine.relative_residual(pair, x, rhs, info.get("residual"))
        amps = [engine.farfield(x[i], bu, bdn, k) for i in range(2)]
        sig[precision] = torch.stack([s for _, s in amps]).cpu().numpy()
        fm, sm = e0.elapsed_time(e1), e1.elapsed_time(e2)
        peak = PEAK_FP32 if precision == "mixed" else PEAK_FP64
        out[precision] = {"factor_ms": round(fm, 1), "solve_ms": round(sm, 1),
                          "lu_tflops": round(lu_flops / fm / 1e9, 2),
                          "lu_frac_of_peak": round(lu_flops / fm / 1e9 / peak, 4),
                          "refinement_steps": info.get("refinement_steps"),
                          "max_relative_residual": float(torch.max(rel)),
                          "solves_per_s": round(2000.0 / (out["assemble_ms"] + fm + sm), 4)}
        del fac, x, info
        torch.cuda.empty_cache()
    out["mixed_vs_fp64_max_db"] = float(np.max(np.abs(10 * np.log10(sig["mixed"]) - 10 * np.log10(sig["double"]))))
    out["trsm_flops_per_solve"] = trsm_flops
    print(json.dumps(out))


def apply_tuning(choice):
    """Load the launch-shape table the run uses; returns its entries for the JSON line."""
    from grim_20261007_amd import autotune, native
    native.ensure_device(0)
    if choice == "none":
        autotune.clear()
        return None
    table = autotune.tune() if choice == "tune" else autotune.resolve(choice)
    autotune.apply(table)
    return {name: table[name] for name in autotune.KNOB_NAMES.values() if name in table} | {"source": choice}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    shapes = apply_tuning(args.tuning)
    if args.workload != "naca":
        from grim_20261007_amd import native
        native.ensure_device(0)
        {"sweep": sweep_workload, "stream": stream_workload, "large": large_workload}[args.workload](args)
        return
    import torch
    from grim_20261007_amd import engine, native
    dev = torch.device("cuda", 0)
    pts, mesh, f, angles = workload(args.n_side, args.angles)
    from grim_20261007_amd.solver import wavenumber
    k = wavenumber(f)
    contour = engine.DeviceContour(mesh, dev)
    angles_t = torch.tensor(angles, dtype=torch.float64, device=dev)
    request.ws = engine.assembly_workspace(mesh.n, dev)
    # 1. one request at a time, per-stage HIP-event timers (plain FP64 path and mixed form)
    results, timers, lat = {}, {}, {}
    seq_steps = max(2, min(args.steps, 8))
    for precision in ("double", "mixed"):
        for _ in range(args.warmup):
            request(engine, contour, k, angles_t, precision)
        torch.cuda.synchronize()
        timer = Timer(torch)
        lat[precision] = []
        t_start = time.perf_counter()
        for _ in range(seq_steps):
            t0 = time.perf_counter()
            res = request(engine, contour, k, angles_t, precision, timer)
            torch.cuda.synchronize()
            lat[precision].append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        results[precision] = (res, time.perf_counter() - t_start)
        timers[precision] = timer
    # accuracy of the mixed form against the FP64 path on the same request
    db_err, amp_err, resid = 0.0, 0.0, 0.0
    steps_mixed = []
    for pol in ("TE", "TM"):
        a64, s64, r64, _ = results["double"][0][pol]
        a32, s32, r32, st = results["mixed"][0][pol]
        s64, s32 = s64.cpu().numpy(), s32.cpu().numpy()
        db_err = max(db_err, float(np.max(np.abs(10 * np.log10(s32) - 10 * np.log10(s64)))))
        amp_err = max(amp_err, float(torch.max(torch.abs(a32 - a64) / torch.abs(a64))))
        resid = max(resid, float(torch.max(r32)))
        steps_mixed.append(st)
    t64 = results["double"][1]
    plain64 = 2 * seq_steps / t64
    lat64 = np.array(lat["double"]) * 1e3
    lat32 = np.array(lat["mixed"]) * 1e3
    # 2. the serving form: args.steps distinct requests, args.workers in flight.  The headline is
    # the server's default, the cost-model precision per request ('auto'); forced mixed and
    # forced FP64 serving run on the same requests beside it.
    torch.cuda.synchronize()
    # each form warms its own server with two full rounds of batches (first-use allocations of
    # every worker's stream pool, a whole batch's systems included, fall outside the timed region)
    warm = max(args.warmup, 2 * args.workers * args.batch)
    # the first server of a process runs about 12 % slower whatever its form (measured: 253 against
    # 288 solves/s for the same requests in the next server), so one untimed server goes first
    bt = args.batch
    serve(engine, mesh, f, angles, "auto", args.workers, warm, args.workers, bt)
    res_a, t_a, svc_a, chosen = serve(engine, mesh, f, angles, "auto", args.workers, args.steps, warm, bt)
    res_m, t_m, svc_m, _ = serve(engine, mesh, f, angles, "mixed", args.workers, args.steps, warm, bt)
    res_d, t_d, svc_d, _ = serve(engine, mesh, f, angles, "double", args.workers, args.steps, warm, bt)
    value = 2 * args.steps / t_a
    # every served mixed result within the tolerance of the FP64 path
    ref_sig = {pol: results["double"][0][pol][1].cpu().numpy() for pol in ("TE", "TM")}
    serve_db = max(float(np.max(np.abs(10 * np.log10(r.sigma[pol]) - 10 * np.log10(ref_sig[pol]))))
                   for r in res_m + res_a for pol in ("TE", "TM"))
    serve_res = max(max(r.residual.values()) for r in res_m + res_a)
    # roofline of the dominant single launch (far-pair assembly pass)
    tm = timers["mixed"]
    far_ms = 0.5 * (tm.mean_ms("assemble_far") + timers["double"].mean_ms("assemble_far"))
    pairs = far_pair_count(mesh)
    q = mesh.far_order(k)
    flops = pairs * q * q * FLOPS_PER_EVAL
    achieved = flops / (far_ms * 1e-3) / 1e12
    # the refinement residual b - Z x (one batched FP64 complex GEMM, N x A x N per system)
    pair = torch.empty((2, mesh.n, mesh.n), dtype=torch.complex128, device=dev)
    pair.normal_()
    xs = torch.randn((2, len(angles), mesh.n), dtype=torch.complex128, device=dev)
    bs = torch.randn((2, len(angles), mesh.n), dtype=torch.complex128, device=dev)
    gt = Timer(torch)
    for i in range(6):
        (gt.run if i else (lambda name, fn: fn()))("residual_gemm", lambda: engine.residual(pair, xs, bs))
    gemm_ms = gt.mean_ms("residual_gemm")
    gemm_tf = 2 * 8.0 * mesh.n * mesh.n * len(angles) / (gemm_ms * 1e-3) / 1e12
    del pair, xs, bs
    stage_ms = {name: round(tm.mean_ms(name), 3) for name in tm.acc}
    stage64 = {name: round(timers["double"].mean_ms(name), 3) for name in timers["double"].acc}
    # assembly rate: TM + TE matrix entries per second of far + near assembly (FP64 path's timers)
    asm_ms = timers["double"].mean_ms("assemble_far") + timers["double"].mean_ms("assemble_near")
    assembly = {"entries_per_s": round(2.0 * mesh.n * mesh.n / (asm_ms * 1e-3), 1), "ms": round(asm_ms, 3),
                "far_evaluations_per_s": round(pairs * q * q / (far_ms * 1e-3), 1),
                "far_kernel_counters": far_valu()}
    # LU rate of the batched TE + TM factorization, (8/3) N^3 real flops per complex system
    lu_flops = 2 * (8.0 / 3.0) * mesh.n ** 3
    lu = {}
    for key, stage, peak, label in (("complex128", stage64.get("factor_double"), PEAK_FP64, "fp64"),
                                    ("complex64", stage_ms.get("factor_mixed"), PEAK_FP32, "fp32")):
        if stage:
            tf = lu_flops / (stage * 1e-3) / 1e12
            lu[key] = {"factor_ms": stage, "tflops": round(tf, 2), f"frac_of_{label}_peak": round(tf / peak, 4)}
    # the same factorization as the serving form runs it: the systems of `batch` requests in one
    # call (random systems of the same size; the pivot chain and the GEMM shapes are what count)
    if args.batch > 1:
        nsys = 2 * args.batch
        g = torch.Generator(device="cpu").manual_seed(5)
        zb = torch.randn((nsys, mesh.n, mesh.n), dtype=torch.complex128, generator=g).to(dev)
        lt = Timer(torch)
        for prec, peak, label in (("double", PEAK_FP64, "fp64"), ("mixed", PEAK_FP32, "fp32")):
            for i in range(3):
                (lt.run if i else (lambda name, fn: fn()))(prec, lambda: engine.factor(zb, prec, check=False))
            ms = lt.mean_ms(prec)
            tf = nsys * (8.0 / 3.0) * mesh.n ** 3 / (ms * 1e-3) / 1e12
            lu[("complex128" if prec == "double" else "complex64") + f"_batch{nsys}"] = {
                "factor_ms": round(ms, 3), "systems": nsys, "tflops": round(tf, 2),
                f"frac_of_{label}_peak": round(tf / peak, 4)}
        del zb
        torch.cuda.empty_cache()
    stream = stream_summary(args.workers, batch=args.batch) if args.stream_count > 0 else None
    sheet = sheet_summary() if not args.no_sheet else None
    lossy = lossy_summary() if not args.no_sheet else None
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(mesh, f, angles, args.cpu_seconds)
    plain32 = 2 * seq_steps / results["mixed"][1]
    if chosen["mixed"] == 0:
        dtype = "c128 (f64) LU + f64 backward-error gate"
    elif chosen["double"] == 0:
        dtype = "c64 LU + f64 refinement"
    else:
        dtype = "per request: c128 LU or c64 LU + f64 refinement"
    line = {
        "metric": "RCS solves/sec + p99 latency, mixed-prec vs FP64 GPU at <=0.05 dB RCS error",
        "value": round(value, 3), "unit": "RCS solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_a / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic NACA 0012 geometry",
        "config": {"workload": f"NACA 0012, chord 1 m, N={mesh.n} linear Galerkin unknowns, f={f:.4f} GHz "
                               f"(longest panel lambda/10), {len(angles)} monostatic angles, TE+TM co-polarized "
                               f"requests; serving form with {args.workers} workers in flight"
                               + (f", each running up to {args.batch} waiting requests as one batched factor + "
                                  "solve" if args.batch > 1 else "")
                               + " (distinct requests, no cache reuse), each request in the precision the flop "
                                 "model picks for it",
                   "n": mesh.n, "angles": len(angles), "frequency_ghz": f, "workers": args.workers,
                   "batch": args.batch,
                   "precision_choice": chosen},
        "p50_ms": round(float(np.percentile(svc_a, 50)), 3), "p99_ms": round(float(np.percentile(svc_a, 99)), 3),
        "serving_vs_plain_fp64": round(value / plain64, 3),
        "mixed_vs_fp64": {"serving": round(t_d / t_m, 3), "sequential": round(plain32 / plain64, 3),
                          "note": "same form both sides: mixed serving / FP64 serving, mixed one at a time / "
                                  "FP64 one at a time"},
        "angle_columns_per_s": round(value * len(angles), 1),
        "fp64_path": {"value": round(plain64, 3), "p50_ms": round(float(np.percentile(lat64, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat64, 99)), 3), "stage_ms": stage64,
                      "form": "one request at a time, FP64 LU + backward-error gate"},
        "fp64_serving": {"value": round(2 * args.steps / t_d, 3),
                         "p50_ms": round(float(np.percentile(svc_d, 50)), 3),
                         "p99_ms": round(float(np.percentile(svc_d, 99)), 3)},
        "mixed_serving": {"value": round(2 * args.steps / t_m, 3),
                          "p50_ms": round(float(np.percentile(svc_m, 50)), 3),
                          "p99_ms": round(float(np.percentile(svc_m, 99)), 3)},
        "mixed_sequential": {"value": round(2 * seq_steps / results["mixed"][1], 3),
                             "p50_ms": round(float(np.percentile(lat32, 50)), 3),
                             "p99_ms": round(float(np.percentile(lat32, 99)), 3), "stage_ms": stage_ms},
        "accuracy": {"max_rcs_db_error_vs_fp64": max(db_err, serve_db), "max_amplitude_rel_error": amp_err,
                     "max_refined_residual": max(resid, serve_res), "refinement_steps": steps_mixed},
        "roofline": {"kernel": "far pass (far_zero_all + far_kernel + far_fixup, grim_assemble_far)", "bound": "valu",
                     "achieved": round(achieved, 3), "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64, 4), "traffic": far_traffic(),
                     "model": f"{pairs} unordered far pairs x {q}x{q} points x {FLOPS_PER_EVAL} FP64 flops "
                              f"per launch; FP64 VALU-bound: peak is the FP64 vector rate, which equals the FP64 "
                              f"matrix rate on MI355X"},
        "roofline_fp64_gemm": {"kernel": "cgemm<double> (refinement residual b - Z x, 2 x N x A x N)",
                               "bound": "mfma", "achieved": round(gemm_tf, 2), "peak": PEAK_FP64,
                               "unit": "TFLOP/s", "frac": round(gemm_tf / PEAK_FP64, 4),
                               "ms": round(gemm_ms, 3)},
        "assembly": assembly,
        "lu_n2000": lu,
        "cpu_baseline": cpu,
        "stream": stream,
        "sheet_n2000": sheet,
        "lossy_coated_n1500": lossy,
        "launch_shapes": shapes,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
