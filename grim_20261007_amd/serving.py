"""Continuous-batching RCS request server on one GPU.

A request is (geometry, frequency, polarization set, angle set).  The server
admits requests by their HBM footprint while earlier ones are still running:
each worker thread owns a HIP stream and runs one request at a time on it, so
the kernels of several requests overlap on the device, and a new request is
admitted as soon as a worker slot and enough of the byte budget are free
(continuous batching, not a FIFO of whole batches).  Assembled systems and
their LU factors stay resident in a device cache keyed by (geometry digest,
wavenumber, precision) and are reused by later requests with new angle sets.

Per-request latency (arrival to completion) and throughput are recorded for
p50 / p99 reporting.  The memory estimate follows the reference's admission
idea (twod/solver.py:1474-1554): every dense buffer of the request is counted
before it is allocated.
"""
from __future__ import annotations

import hashlib
import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import autotune, engine
from .mesh import ContourMesh, SheetLayout
from .solver import _solve_sheet_frequency, system_dofs, wavenumber


@dataclass
class Request:
    mesh: ContourMesh
    frequency_ghz: float
    angles_deg: np.ndarray
    pols: Tuple[str, ...] = ("TE", "TM")
    precision: str = "auto"           # "auto" (choose_precision), "mixed" or "double"
    arrival: float = 0.0
    rid: int = 0
    key: Optional[str] = None          # cache key; None -> geometry digest


@dataclass
class Result:
    rid: int
    amplitude: Dict[str, np.ndarray]
    sigma: Dict[str, np.ndarray]
    residual: Dict[str, float]
    cached: bool
    latency_s: float
    started: float
    finished: float
    info: Dict[str, dict] = field(default_factory=dict)


def geometry_digest(mesh: ContourMesh) -> str:
    h = hashlib.sha256()
    for a in (mesh.p0, mesh.p1, mesh.next):
        h.update(np.ascontiguousarray(a).tobytes())
    if mesh.impedance is not None:
        h.update(np.ascontiguousarray(mesh.impedance).tobytes())
    if mesh.medium is not None:
        h.update(np.array([complex(mesh.medium[1]), complex(mesh.medium[2])]).tobytes())
    if mesh.regions is not None:
        h.update(np.ascontiguousarray(mesh.regions).tobytes())
        h.update(np.array([complex(v) for r in sorted(mesh.region_media) for v in mesh.region_media[r]]).tobytes())
    return h.hexdigest()[:24]


def choose_precision(n: int, nangles: int, corrections: Optional[int] = None) -> str:
    """Cheaper form for one request by the flop model (FP32 MFMA at twice the FP64 rate):
    plain  = (8/3) N^3 + 2 * 8 N^2 A                       (LU, solve, backward-error gate)
    mixed  = (8/3) N^3 / 2 + (s + 1) 8 N^2 A / 2 + s 8 N^2 A   (s corrections, FP64 residuals)
    so the mixed form pays only when A < N / 18 (s = 3).  s grows with N (the complex64
    factor's error does): 3 at N <= 2000 (measured on the NACA 0012 systems), one more per
    doubling (7 measured at N = 40 000)."""
    if corrections is None:
        corrections = 3 + max(0, int(np.ceil(np.log2(max(n, 1) / 2000.0))))
    n3, n2a = n ** 3, n * n * nangles
    plain = 8.0 / 3.0 * n3 + 16.0 * n2a
    mixed = 4.0 / 3.0 * n3 + 4.0 * (corrections + 1) * n2a + 8.0 * corrections * n2a
    return "mixed" if mixed < plain else "double"


def request_bytes(n: int, nangles: int, npols: int, precision: str, cached: bool, degree: int = 1) -> int:
    """Device bytes a request allocates (systems, factors, loads, RHS, solutions, workspace)."""
    return engine.request_device_bytes(n, nangles, npols, precision, cached, degree)


class SystemCache:
    """Device-resident assembled + factored systems, LRU within a byte budget."""

    def __init__(self, budget_bytes: int):
        self.budget = int(budget_bytes)
        self.used = 0
        self.items: "OrderedDict[tuple, dict]" = OrderedDict()
        self.lock = threading.Lock()
        self.hits = self.misses = 0

    def get(self, key):
        with self.lock:
            item = self.items.get(key)
            if item is not None:
                self.items.move_to_end(key)
                self.hits += 1
            else:
                self.misses += 1
            return item

    def put(self, key, item, nbytes):
        with self.lock:
            old = self.items.pop(key, None)           # a replaced entry (e.g. after a fallback)
            if old is not None:
                self.used -= old["bytes"]
            if nbytes > self.budget:
                return False
            while self.items and self.used + nbytes > self.budget:
                _, old = self.items.popitem(last=False)
                self.used -= old["bytes"]
            item["bytes"] = nbytes
            self.items[key] = item
            self.used += nbytes
            return True


def batch_signature(req: Request) -> Optional[tuple]:
    """Requests with equal signatures can share factor and solve launches: plain PEC contours
    (one unknown per node) of the same size, angle count and precision form.  None: the request
    runs on its own (impedance, dielectric and multi-region systems)."""
    mesh = req.mesh
    if (mesh.impedance is not None or mesh.medium is not None or mesh.regions is not None
            or mesh.degree != 1):
        return None
    precision = req.precision
    if precision == "auto":
        precision = choose_precision(mesh.n, len(req.angles_deg))
    return (mesh.n, len(req.angles_deg), precision)


class Server:
    """``workers`` threads, each with its own HIP stream; ``batch`` > 1 lets a worker take up to
    that many waiting requests of equal signature (``batch_signature``) and run their factor and
    solve as one batched launch sequence (``execute_batch``), which spreads the latency-bound
    panel chain of the LU over all of them."""

    def __init__(self, device=None, workers: int = 4, hbm_budget_bytes: Optional[int] = None,
                 cache_bytes: Optional[int] = None, tuning=None, batch: int = 4):
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        if batch < 1:
            raise ValueError("batch must be at least 1")
        self.batch = int(batch)
        self.batches: "deque" = deque(maxlen=4096)   # sizes of the latest request groups run, in order
        # launch-shape table from the autotuner (a dict or a JSON path); shapes never change results
        self.tuning = autotune.resolve(tuning)
        if self.tuning is not None:
            autotune.apply(self.tuning)
        total = torch.cuda.get_device_properties(self.device).total_memory
        self.budget = int(hbm_budget_bytes if hbm_budget_bytes is not None else 0.6 * total)
        self.cache = SystemCache(cache_bytes if cache_bytes is not None else int(0.25 * total))
        self.in_use = 0
        self.cv = threading.Condition()
        # single flight: a key being assembled and factored maps to an event that the request
        # doing the work sets once the factors are in the cache (or it failed)
        self.inflight: Dict[tuple, threading.Event] = {}
        self.inflight_lock = threading.Lock()
        self.waited = 0
        # waiting requests, in arrival order; a worker takes the head and, with batch > 1, the
        # waiting requests of the same signature behind it
        self.pending: "deque" = deque()
        self.pending_cv = threading.Condition()
        self.results: List[Result] = []
        self.workers = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(workers)]
        self.stop = False
        self.errors: List[BaseException] = []
        for t in self.workers:
            t.start()

    # -- admission -----------------------------------------------------------
    def _admit(self, nbytes: int):
        with self.cv:
            if nbytes > self.budget:
                raise MemoryError(f"request needs {nbytes / 2**30:.2f} GiB, budget {self.budget / 2**30:.2f} GiB")
            while self.in_use + nbytes > self.budget:
                self.cv.wait()
            self.in_use += nbytes

    def _release(self, nbytes: int):
        with self.cv:
            self.in_use -= nbytes
            self.cv.notify_all()

    def submit(self, req: Request) -> Future:
        fut: Future = Future()
        if req.arrival == 0.0:
            req.arrival = time.perf_counter()
        with self.pending_cv:
            self.pending.append((req, fut))
            self.pending_cv.notify()
        return fut

    def submit_many(self, reqs: List[Request]) -> List[Future]:
        """Submit several requests at once (they are queued together, so a batching worker
        sees all of them)."""
        futs = []
        now = time.perf_counter()
        with self.pending_cv:
            for req in reqs:
                fut: Future = Future()
                if req.arrival == 0.0:
                    req.arrival = now
                self.pending.append((req, fut))
                futs.append(fut)
            self.pending_cv.notify_all()
        return futs

    def close(self):
        with self.pending_cv:
            self.stop = True
            self.pending_cv.notify_all()
        for t in self.workers:
            t.join()

    def _take(self):
        """Next group of waiting requests (None once closed and drained): the head, plus with
        batch > 1 the waiting requests of its signature, up to ``batch`` and half the byte
        budget, in arrival order."""
        with self.pending_cv:
            while not self.pending and not self.stop:
                self.pending_cv.wait()
            if not self.pending:
                return None
            head = self.pending.popleft()
            group = [head]
            sig = batch_signature(head[0]) if self.batch > 1 else None
            if sig is not None:
                unit = request_bytes(sig[0], sig[1], 2, sig[2], False)
                keep = deque()
                while self.pending:
                    item = self.pending.popleft()
                    if (len(group) < self.batch and batch_signature(item[0]) == sig
                            and (len(group) + 1) * unit <= self.budget // 2):
                        group.append(item)
                    else:
                        keep.append(item)
                self.pending = keep
            return group

    # -- execution -----------------------------------------------------------
    def _worker(self, index: int):
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        ws = {}
        while True:
            group = self._take()
            if group is None:
                return
            with torch.cuda.stream(stream):
                if len(group) > 1:
                    self._run_group(group, ws)
                    continue
                req, fut = group[0]
                try:
                    self.batches.append(1)
                    res = self._run(req, ws)
                    fut.set_result(res)
                except BaseException as exc:   # surfaced through the future
                    fut.set_exception(exc)

    def _run_group(self, group, ws: dict):
        """Requests of one signature.  Those whose system is neither cached nor being built
        run as one batch (each one the single-flight owner of its key); the rest run on their
        own afterwards, as does every request of a batch that failed (redone request by
        request, so one bad system fails only its own request)."""
        owners, alone = [], []
        with self.inflight_lock:
            for req, fut in group:
                if req.precision == "auto":
                    req.precision = choose_precision(req.mesh.n, len(req.angles_deg))
                key = (req.key or geometry_digest(req.mesh), float(wavenumber(req.frequency_ghz)), req.precision)
                if self.cache.get(key) is None and key not in self.inflight:
                    self.inflight[key] = threading.Event()
                    owners.append((req, fut, key))
                else:
                    alone.append((req, fut))
        results = None
        if len(owners) > 1:
            reqs = [o[0] for o in owners]
            nbytes = sum(request_bytes(system_dofs(r.mesh), len(r.angles_deg), len(r.pols), r.precision, False)
                         for r in reqs)

            def store(i, item, nb):
                self.cache.put(owners[i][2], item, nb)

            try:
                self._admit(nbytes)
                try:
                    results = execute_batch(reqs, self.device, ws, store)
                finally:
                    self._release(nbytes)
            except Exception:
                results = None
        # the batch's factors are in the cache (waiters find them there) or it failed (the
        # requests are redone below, each its own owner again)
        with self.inflight_lock:
            for _, _, key in owners:
                ev = self.inflight.pop(key, None)
                if ev is not None:
                    ev.set()
        if results is not None:
            self.batches.append(len(owners))
            for (req, fut, key), res in zip(owners, results):
                fut.set_result(res)
        else:
            alone = [(r, f) for r, f, _ in owners] + alone
        for req, fut in alone:
            try:
                self.batches.append(1)
                fut.set_result(self._run(req, ws))
            except BaseException as exc:   # surfaced through the future
                fut.set_exception(exc)

    def _run_sheet(self, req: Request) -> Result:
        mesh = req.mesh
        tm = mesh.mesh("TM")
        nbytes = engine.request_device_bytes(tm.n, len(req.angles_deg), 1, "double", sheet_elements=tm.ne)
        self._admit(nbytes)
        try:
            return execute_sheet(req, self.device)
        finally:
            self._release(nbytes)

    def _run(self, req: Request, ws: dict) -> Result:
        mesh = req.mesh
        if isinstance(mesh, SheetLayout):
            return self._run_sheet(req)
        if req.precision == "auto":
            req.precision = choose_precision(mesh.n, len(req.angles_deg))
        k = wavenumber(req.frequency_ghz)
        key = (req.key or geometry_digest(mesh), float(k), req.precision)
        while True:
            with self.inflight_lock:
                cached = self.cache.get(key)
                pending = None if cached is not None else self.inflight.get(key)
                if pending is not None and pending.is_set():
                    pending = None      # its factors were stored and already evicted: redo them
                owner = cached is None and pending is None
                if owner:
                    self.inflight[key] = threading.Event()
            if pending is None:
                break
            # another worker is assembling and factoring this system: wait for its factors
            # instead of repeating the work, then look again (if it failed or the system did
            # not fit the cache, this request becomes the owner)
            self.waited += 1
            pending.wait()
        done = self.inflight.get(key) if owner else None

        def store(item, nb):
            if self.cache.put(key, item, nb) and done is not None:
                done.set()

        nbytes = request_bytes(system_dofs(mesh), len(req.angles_deg), len(req.pols), req.precision,
                               cached is not None, mesh.degree)
        if cached is None:
            nbytes += engine.lossy_workspace_bytes(mesh)
        try:
            self._admit(nbytes)
            try:
                return execute(req, self.device, ws, cached, store)
            finally:
                self._release(nbytes)
        finally:
            if owner:
                with self.inflight_lock:
                    self.inflight.pop(key, None)
                done.set()


def execute_sheet(req: Request, device) -> Result:
    """A sheet layout's request: each polarization on its own mesh, factored and solved alone (no
    factor cache)."""
    if req.precision == "auto":
        req.precision = choose_precision(system_dofs(req.mesh), len(req.angles_deg))
    started = time.perf_counter()
    res = _solve_sheet_frequency(req.mesh, req.frequency_ghz, list(map(float, req.angles_deg)), tuple(req.pols),
                                 req.precision, device, False, None, False, "The served sheet solve")
    finished = time.perf_counter()
    info = {pol: {"precision": r["precision_used"], "formulation": r["formulation"]} for pol, r in res.items()}
    return Result(req.rid, {p: r["amplitude"] for p, r in res.items()}, {p: r["sigma"] for p, r in res.items()},
                  {p: float(np.max(r["residual"])) for p, r in res.items()}, False,
                  finished - (req.arrival or started), started, finished, info)


def execute(req: Request, device, ws: dict, cached: Optional[dict] = None, store=None) -> Result:
    """Run one request on the current stream: assemble and factor (or reuse `cached`), then loads,
    solve and far field.  `store(item, nbytes)` receives the assembled + factored systems."""
    mesh = req.mesh
    if req.precision == "auto":
        req.precision = choose_precision(system_dofs(mesh), len(req.angles_deg))
    k = wavenumber(req.frequency_ghz)
    started = time.perf_counter()
    dev = torch.device(device)
    contour = cached["contour"] if cached else engine.DeviceContour(mesh, dev)
    plan = cached.get("plan") if cached else (engine.region_plan(mesh) if mesh.regions is not None else None)
    dofs = plan.n_dof if plan is not None else system_dofs(mesh)
    if cached is None:
        wkey = (mesh.n, mesh.degree)
        if wkey not in ws:
            ws[wkey] = engine.assembly_workspace(mesh.n, dev, mesh.degree)
        if plan is not None:
            pair = engine.region_systems(contour, k, ws[wkey], plan)
        elif mesh.medium is not None:
            pair = engine.dielectric_systems(contour, k, ws[wkey])
        else:
            z_tm, z_te = engine.assemble(contour, k, ws[wkey])
            pair = z_tm._base                                     # batch [TE, TM]
            if mesh.impedance is not None:
                engine.impedance_systems(contour, k, pair, ws[wkey])
        fac = engine.factor(pair, req.precision)
        if store is not None:
            sysbytes = 2 * (16 + (8 if req.precision == "mixed" else 16)) * dofs * dofs
            store({"contour": contour, "factor": fac, "plan": plan}, sysbytes)
    else:
        fac = cached["factor"]
        # the cached tensors are read on this worker's stream too
        cur = torch.cuda.current_stream()
        for t in (contour.p0, contour.p1, contour.next, contour.prev, contour.fix_rows,
                  fac.lu, fac.piv, fac.aux, fac.z):
            t.record_stream(cur)
    angles = torch.tensor(np.asarray(req.angles_deg, float), dtype=torch.float64, device=dev)
    bu, bdn = engine.loads(contour, k, angles)
    if plan is not None:
        rhs = engine.region_rhs(plan, bu, bdn, k)
    elif mesh.medium is not None:
        rhs = engine.dielectric_rhs(bu, bdn)
    elif mesh.impedance is not None:
        rhs = engine.impedance_rhs(contour, k, angles)
    else:
        rhs = torch.stack((-bdn, -bu))                            # [TE, TM]
    x, inf = engine.solve(fac, rhs, gate=(req.precision == "double"))
    if inf.get("factor") is not None and store is not None:
        # the mixed factorization fell back to FP64: later requests reuse the FP64 factors
        fac = inf.pop("factor")
        store({"contour": contour, "factor": fac, "plan": plan}, 2 * 32 * dofs * dofs)
    rel = engine.relative_residual(fac.z, x, rhs, inf.get("residual"))
    amp, sig, resid, info = {}, {}, {}, {}
    pending = []
    for i, pol in enumerate(("TE", "TM")):
        if pol not in req.pols:
            continue
        rho = (engine.region_exterior_density(plan, mesh.n, x[i]) if plan is not None
               else engine.radiating_density(contour, k, x[i]))
        a, s = engine.farfield(rho, bu, bdn, k)
        pending.append((pol, a, s, rel[i]))
        info[pol] = system_info(inf, fac, i)
    torch.cuda.current_stream().synchronize()
    for pol, a, s, r in pending:
        amp[pol] = a.cpu().numpy()
        sig[pol] = s.cpu().numpy()
        resid[pol] = float(r.max())
    finished = time.perf_counter()
    arrival = req.arrival or started
    return Result(req.rid, amp, sig, resid, cached is not None, finished - arrival, started, finished, info)


def system_info(inf: dict, fac, i: int) -> dict:
    """Accuracy statement of system i of a solve: refinement steps and error, the normwise
    backward error (the FP64 gate's metric) and the precision its factor ended in."""
    out = {kk: v for kk, v in inf.items()
           if kk not in ("residual", "factor", "backward_errors", "refinement_errors", "steps", "precisions")}
    if "backward_errors" in inf:
        out["backward_error"] = inf["backward_errors"][i]
    if "refinement_errors" in inf:
        out["refinement_error"] = inf["refinement_errors"][i]
    if "steps" in inf:
        out["refinement_steps"] = inf["steps"][i]
    out["precision_used"] = engine.system(fac, i).precision if fac.parts is not None else fac.precision
    return out


def execute_batch(reqs: List[Request], device, ws: dict, store=None) -> List[Result]:
    """Requests of one batch signature (plain PEC contours with equal N, angle count and
    precision), none of them cached: each is assembled into its two systems [TE, TM] of one
    (2B, N, N) batch, the batch is factored and solved in single batched launch sequences, and
    each request's far field comes from its own systems.  Every kernel works on each system of
    a batch with the same operations whatever the batch, and the FP64 gate and the mixed
    refinement decide per system, so each request gets, bit for bit, what ``execute`` gives it
    alone.  ``store(i, item, nbytes)`` receives request i's assembled + factored systems (views
    into the batch's storage)."""
    dev = torch.device(device)
    head = reqs[0]
    n, na, precision = head.mesh.n, len(head.angles_deg), head.precision
    for r in reqs:
        if batch_signature(r) != (n, na, precision) or r.precision != precision:
            raise ValueError("execute_batch needs requests of one signature")
    nb = len(reqs)
    started = time.perf_counter()
    if (n, 1) not in ws:
        ws[(n, 1)] = engine.assembly_workspace(n, dev)
    systems = torch.empty((2 * nb, n, n), dtype=torch.complex128, device=dev)
    rhs = torch.empty((2 * nb, na, n), dtype=torch.complex128, device=dev)
    contours, ks, loads = [], [], []
    for i, req in enumerate(reqs):
        k = wavenumber(req.frequency_ghz)
        contour = engine.DeviceContour(req.mesh, dev)
        engine.assemble(contour, k, ws[(n, 1)], out=systems[2 * i:2 * i + 2])
        angles = torch.tensor(np.asarray(req.angles_deg, float), dtype=torch.float64, device=dev)
        bu, bdn = engine.loads(contour, k, angles)
        torch.neg(bdn, out=rhs[2 * i])
        torch.neg(bu, out=rhs[2 * i + 1])
        contours.append(contour)
        ks.append(k)
        loads.append((bu, bdn))
    fac = engine.factor(systems, precision)
    sysbytes = 2 * (16 + (8 if precision == "mixed" else 16)) * n * n
    if store is not None:
        for i in range(nb):
            store(i, {"contour": contours[i], "factor": engine.systems(fac, 2 * i, 2 * i + 2), "plan": None},
                  sysbytes)
    x, inf = engine.solve(fac, rhs, gate=(precision == "double"))
    if inf.get("factor") is not None:
        # some systems fell back to the FP64 LU: later requests reuse those factors
        fac = inf.pop("factor")
        if store is not None:
            for i in range(nb):
                store(i, {"contour": contours[i], "factor": engine.systems(fac, 2 * i, 2 * i + 2), "plan": None},
                      2 * 32 * n * n)
    rel = engine.relative_residual(fac.z, x, rhs, inf.get("residual"))
    pending = []
    for i, req in enumerate(reqs):
        bu, bdn = loads[i]
        for j, pol in enumerate(("TE", "TM")):
            if pol not in req.pols:
                continue
            a, sg = engine.farfield(x[2 * i + j], bu, bdn, ks[i])
            pending.append((i, pol, a, sg, rel[2 * i + j], system_info(inf, fac, 2 * i + j)))
    torch.cuda.current_stream().synchronize()
    finished = time.perf_counter()
    out = [Result(r.rid, {}, {}, {}, False, finished - (r.arrival or started), started, finished, {})
           for r in reqs]
    for i, pol, a, sg, r, inf_i in pending:
        out[i].amplitude[pol] = a.cpu().numpy()
        out[i].sigma[pol] = sg.cpu().numpy()
        out[i].residual[pol] = float(r.max())
        out[i].info[pol] = inf_i
    return out


def run_stream(server: Server, requests: List[Request], rate_hz: Optional[float] = None,
               rng: Optional[np.random.Generator] = None) -> Tuple[List[Result], float]:
    """Submit requests with Poisson arrivals at rate_hz (None: all at once); wait for all."""
    rng = rng or np.random.default_rng(0)
    t0 = time.perf_counter()
    if not rate_hz:
        for req in requests:
            req.arrival = t0
        futs = server.submit_many(requests)
        results = [f.result() for f in futs]
        return results, time.perf_counter() - t0
    futs = []
    t_next = t0
    for req in requests:
        if rate_hz:
            t_next += rng.exponential(1.0 / rate_hz)
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        req.arrival = time.perf_counter()
        futs.append(server.submit(req))
    results = [f.result() for f in futs]
    return results, time.perf_counter() - t0


def latency_summary(results: List[Result]) -> dict:
    lat = np.array([r.latency_s for r in results]) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}
