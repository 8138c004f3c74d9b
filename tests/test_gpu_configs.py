"""The five workload configurations and the reference's cylinder gates on the GPU.

* the reference's PEC cylinder gates (tests/test_rcs_physics_regression.py:648-719) through
  the public single-polarization entry, both precision forms;
* C1: 400-chord circle, ka = 10, 360 angles, against the Mie series;
* C2: NACA 0012, N = 2000, 720 angles: FP64 and mixed forms, sampled matrix rows against the
  oracle, mixed within 0.05 dB of FP64, refined residuals <= 1e-10;
* C3: the 64-frequency star sweep through the server (mixed form), each request against its
  FP64 solve;
* C4: a seeded Poisson stream (N up to 11 493, 90 / 360 / 1440 angles, 8 of 20 requests
  reusing a cached system): alone == in the mix bit for bit, cached == fresh, mixed against FP64;
* C5: N = 40 000, 1440 angles: complex64 LU + refinement against the FP64 GPU LU of the same
  matrix.
"""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

J0_FIRST_ZERO = 2.404825557695773


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd import native
    native.ensure_device(0)
    return torch.device("cuda", 0)


def _db(sigma):
    return 10.0 * np.log10(np.asarray(sigma, float))


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_reference_cylinder_gate_ka1(pol, precision):
    """96-chord PEC circle, ka = 1, through the public entry: |dB| < 0.01 against Mie
    (tests/test_rcs_physics_regression.py:648-669)."""
    from grim_20261007_amd import workloads
    from grim_20261007_amd.solver import solve_monostatic_rcs_2d_single_polarization
    from oracle import mie
    _dev()
    f = workloads.frequency_for_ka(1.0, 0.1)
    res = solve_monostatic_rcs_2d_single_polarization(
        workloads.circle(0.1, 96), [f], [0.0], pol, geometry_units="meters", strict_quality_gate=False,
        max_panels=1000, precision=precision)
    got = res["samples"][0]["rcs_linear"]
    k = 2.0 * math.pi * f * 1e9 / mie.C0
    err = abs(10.0 * math.log10(got / mie.pec_sigma(0.1, k, pol)))
    print(f"ka=1 96 chords {pol} {precision}: |dB| vs Mie {err:.2e}", flush=True)
    assert err < 0.01


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_reference_cylinder_gate_first_j0_zero(pol):
    """64-chord circle at the first zero of J_0 (an interior resonance of the single-layer
    operators): complex amplitude within 5e-3 of Mie, residual below 1e-10
    (tests/test_rcs_physics_regression.py:671-719)."""
    from grim_20261007_amd import workloads
    from grim_20261007_amd.solver import solve_monostatic_rcs_2d_single_polarization
    from oracle import mie
    _dev()
    f = workloads.frequency_for_ka(J0_FIRST_ZERO, 0.1)
    res = solve_monostatic_rcs_2d_single_polarization(
        workloads.circle(0.1, 64), [f], [0.0], pol, geometry_units="meters", strict_quality_gate=False,
        max_panels=1000)
    s = res["samples"][0]
    k = 2.0 * math.pi * f * 1e9 / mie.C0
    want = mie.pec_backscatter_amplitude(0.1, k, pol)
    err = abs(complex(s["rcs_amp_real"], s["rcs_amp_imag"]) - want) / abs(want)
    print(f"J0 zero {pol}: amplitude vs Mie {err:.2e}, residual {s['linear_residual']:.1e}", flush=True)
    assert err < 5e-3
    assert s["linear_residual"] < 1e-10


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_c1_circle_against_mie(precision):
    """C1: 400-chord circle of radius 1 m, ka = 10, 360 angles.  Every angle's amplitude within
    6e-4 relative and 5e-3 dB of the Mie series (the 400-gon's discretisation error is 4.3e-4;
    the reference's own solve of this case lies at the same distance)."""
    from grim_20261007_amd import workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import solve_frequency, wavenumber
    from oracle import mie
    dev = _dev()
    f = workloads.frequency_for_ka(10.0, 1.0)
    mesh = build_mesh(workloads.circle(1.0, 400), "meters", f)
    angles = np.arange(360.0)
    res = solve_frequency(mesh, f, list(angles), ("TE", "TM"), precision, dev)
    k = wavenumber(f)
    for pol in ("TE", "TM"):
        want = mie.pec_backscatter_amplitude(1.0, k, pol)
        amp = res[pol]["amplitude"]
        rel = np.max(np.abs(amp - want)) / abs(want)
        db = np.max(np.abs(_db(res[pol]["sigma"]) - _db(mie.pec_sigma(1.0, k, pol))))
        print(f"C1 {pol} {precision}: amplitude vs Mie {rel:.2e}, dB {db:.2e}, "
              f"residual {np.max(res[pol]['residual']):.1e}", flush=True)
        assert rel < 6e-4 and db < 5e-3
        assert np.max(res[pol]["residual"]) < 1e-10


def test_c2_naca_720_angles_both_forms():
    """C2: NACA 0012, N = 2000, 720 angles, longest panel lambda / 10.  Ten matrix rows (both
    ends of the cosine-spaced contour, the leading edge, mid-chord) against the oracle; the mixed
    form within 0.05 dB (and 1e-9 relative amplitude) of the FP64 path at every angle, refined
    residuals <= 1e-10 in both forms."""
    from grim_20261007_amd import engine, workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import solve_frequency, wavenumber
    from oracle import bem2d
    dev = _dev()
    pts = workloads.naca0012_points(1000)
    f = workloads.frequency_for_segments_per_wavelength(pts, 10.0)
    snap = workloads._snapshot(pts, "naca0012")
    mesh = build_mesh(snap, "meters", f)
    assert mesh.n == 2000
    k = wavenumber(f)
    rows = np.array([0, 1, 2, 500, 999, 1000, 1001, 1500, 1998, 1999])
    om = bem2d.build_mesh(snap, "meters", f)
    assert np.array_equal(mesh.elements, om.elements)
    ref_tm, ref_te = bem2d.assemble(om, bem2d.frequency_k(f), node_rows=rows)
    z_tm, z_te = engine.assemble(engine.DeviceContour(mesh, dev), k)
    for pol, z, ref in (("TM", z_tm, ref_tm), ("TE", z_te, ref_te)):
        got = z.cpu().numpy().T[rows]
        err = np.max(np.abs(got - ref[rows]), axis=1) / np.max(np.abs(ref[rows]), axis=1)
        print(f"C2 {pol} rows vs oracle: max rel {err.max():.2e}", flush=True)
        assert err.max() < 1e-12
    del z_tm, z_te
    angles = list(np.arange(720) * 0.5)
    out = {p: solve_frequency(mesh, f, angles, ("TE", "TM"), p, dev) for p in ("double", "mixed")}
    for pol in ("TE", "TM"):
        a64, a32 = out["double"][pol]["amplitude"], out["mixed"][pol]["amplitude"]
        rel = np.max(np.abs(a32 - a64) / np.abs(a64))
        db = np.max(np.abs(_db(out["mixed"][pol]["sigma"]) - _db(out["double"][pol]["sigma"])))
        print(f"C2 {pol}: mixed vs FP64 amplitude {rel:.2e}, dB {db:.2e}, steps "
              f"{out['mixed'][pol]['info']['refinement_steps']}, residuals "
              f"{np.max(out['double'][pol]['residual']):.1e} / {np.max(out['mixed'][pol]['residual']):.1e}",
              flush=True)
        assert db <= 0.05 and rel < 1e-9
        for p in ("double", "mixed"):
            assert np.max(out[p][pol]["residual"]) <= 1e-10


def test_c3_frequency_sweep_through_server():
    """C3: the seeded 64-frequency star sweep (N = 362-3613, 360 angles) served in the mixed form
    with four requests in flight; each result within 0.05 dB and 1e-9 relative amplitude of its
    own FP64 solve, refined residual <= 1e-10."""
    from grim_20261007_amd import serving, workloads
    from grim_20261007_amd.mesh import build_mesh
    dev = _dev()
    cases = workloads.star_sweep()
    assert len(cases) == 64
    reqs = [serving.Request(build_mesh(workloads._snapshot(p, "star"), "meters", f), f, a, precision="mixed",
                            rid=i, key=f"sweep{i}") for i, (p, f, a) in enumerate(cases)]
    srv = serving.Server(workers=4)
    try:
        t0 = time.perf_counter()
        served, _ = serving.run_stream(srv, reqs)
        elapsed = time.perf_counter() - t0
    finally:
        srv.close()
    worst_db = worst_rel = worst_res = 0.0
    ws = {}
    for r, res in zip(reqs, served):
        ref = serving.execute(serving.Request(r.mesh, r.frequency_ghz, r.angles_deg, precision="double"), dev, ws)
        for pol in ("TE", "TM"):
            worst_db = max(worst_db, float(np.max(np.abs(_db(res.sigma[pol]) - _db(ref.sigma[pol])))))
            worst_rel = max(worst_rel, float(np.max(np.abs(res.amplitude[pol] - ref.amplitude[pol]) /
                                                    np.abs(ref.amplitude[pol]))))
            worst_res = max(worst_res, res.residual[pol])
    ns = [r.mesh.n for r in reqs]
    print(f"C3: {len(reqs)} requests N={min(ns)}-{max(ns)} served in {elapsed:.2f} s; vs FP64: dB {worst_db:.2e}, "
          f"amplitude {worst_rel:.2e}, residual {worst_res:.1e}", flush=True)
    assert worst_db <= 0.05 and worst_rel < 1e-9 and worst_res <= 1e-10


def _stream_requests(precision):
    from grim_20261007_amd import serving, workloads
    from grim_20261007_amd.mesh import build_mesh
    stream = workloads.serving_stream(20, seed=11)
    meshes = {}
    reqs = []
    for i, it in enumerate(stream):
        if it["key"] not in meshes:
            meshes[it["key"]] = build_mesh(workloads._snapshot(it["points"], "star"), "meters",
                                           it["frequency_ghz"])
        reqs.append(serving.Request(meshes[it["key"]], it["frequency_ghz"], it["angles"], precision=precision,
                                    rid=i, key=it["key"]))
    return stream, reqs


def test_c4_poisson_stream_alone_equals_mix():
    """C4: the seeded stream of 20 requests (N = 207-11 493; 90, 360 and 1440 angles; 8 reuse an
    earlier system with new angles) arriving as a Poisson process at 100/s.  Every result in the
    mix (four workers, device cache on) equals, bit for bit, the same request run alone with the
    cache off, so a cached system gives what a fresh one gives; every mixed result is within 0.05
    dB of the FP64 form, with refined residual <= 1e-10."""
    from grim_20261007_amd import serving
    _dev()
    stream, _ = _stream_requests("mixed")
    ns = [len(it["points"]) for it in stream]
    assert max(ns) > 11000 and {len(it["angles"]) for it in stream} == {90, 360, 1440}
    out = {}
    for label, precision, workers, cache in (("alone", "mixed", 1, 0), ("mix", "mixed", 4, None),
                                             ("fp64", "double", 1, 0)):
        _, reqs = _stream_requests(precision)
        srv = serving.Server(workers=workers, cache_bytes=cache)
        try:
            if label == "mix":
                res, elapsed = serving.run_stream(srv, reqs, rate_hz=100.0, rng=np.random.default_rng(5))
            else:
                res = [srv.submit(r).result() for r in reqs]
                elapsed = 0.0
            hits = srv.cache.hits
        finally:
            srv.close()
        out[label] = {r.rid: r for r in res}
        print(f"C4 {label}: {len(res)} requests, cache hits {hits}, {elapsed:.2f} s", flush=True)
    assert sum(r.cached for r in out["mix"].values()) == sum(it["reused"] for it in stream)
    worst_db = 0.0
    for rid, res in out["mix"].items():
        alone, ref = out["alone"][rid], out["fp64"][rid]
        for pol in ("TE", "TM"):
            assert np.array_equal(res.amplitude[pol], alone.amplitude[pol]), (rid, pol, res.cached)
            assert np.array_equal(res.sigma[pol], alone.sigma[pol]), (rid, pol)
            assert res.residual[pol] <= 1e-10
            worst_db = max(worst_db, float(np.max(np.abs(_db(res.sigma[pol]) - _db(ref.sigma[pol])))))
    print(f"C4: mix == alone bit for bit; mixed vs FP64 max dB {worst_db:.2e}", flush=True)
    assert worst_db <= 0.05


def test_c5_large_single_solve():
    """C5: elongated star (R = 2 m, aspect 4:1), N = 40 000, longest panel lambda / 10, 1440
    angles, TE + TM: complex64 LU + FP64 refinement against the FP64 GPU LU of the same matrix
    (its stated oracle).  Every RCS within 0.05 dB, refined residuals <= 1e-10 in both forms."""
    from grim_20261007_amd import engine, workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import wavenumber
    dev = _dev()
    rng = np.random.default_rng(20261203)
    a, phi = workloads.star_shape(rng)
    pts = workloads.star_points(40000, 2.0, a, phi, aspect=4.0)
    f = workloads.frequency_for_segments_per_wavelength(pts, 10.0)
    mesh = build_mesh(workloads._snapshot(pts, "star"), "meters", f, max_panels=40000)
    assert mesh.n == 40000
    k = wavenumber(f)
    contour = engine.DeviceContour(mesh, dev)
    z_tm, z_te = engine.assemble(contour, k)
    pair = z_tm._base
    angles = torch.tensor(np.arange(1440) * 0.25, dtype=torch.float64, device=dev)
    bu, bdn = engine.loads(contour, k, angles)
    rhs = torch.stack((-bdn, -bu))
    sig, amp = {}, {}
    for precision in ("mixed", "double"):
        t0 = time.perf_counter()
        fac = engine.factor(pair, precision)
        x, info = engine.solve(fac, rhs, gate=(precision == "double"))
        rel = engine.relative_residual(pair, x, rhs, info.get("residual"))
        out = [engine.farfield(x[i], bu, bdn, k) for i in range(2)]
        amp[precision] = torch.stack([o[0] for o in out]).cpu().numpy()
        sig[precision] = torch.stack([o[1] for o in out]).cpu().numpy()
        worst = float(torch.max(rel))
        print(f"C5 {precision}: {time.perf_counter() - t0:.1f} s, refinement steps {info.get('refinement_steps')}, "
              f"max relative residual {worst:.1e}", flush=True)
        assert worst <= 1e-10
        del fac, x, info, rel
        torch.cuda.empty_cache()
    db = float(np.max(np.abs(_db(sig["mixed"]) - _db(sig["double"]))))
    relamp = float(np.max(np.abs(amp["mixed"] - amp["double"]) / np.abs(amp["double"])))
    print(f"C5: mixed vs FP64 LU max dB {db:.2e}, amplitude {relamp:.2e}", flush=True)
    assert db <= 0.05 and relamp < 1e-8
