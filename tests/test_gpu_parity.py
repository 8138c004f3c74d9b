"""GPU parity of the HIP path against the oracle and the reference's golden vectors."""
import numpy as np
import pytest

from tests import golden_cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _mat(t):
    """Column-major device tensor -> numpy matrix."""
    return t.cpu().numpy().T


def _lib():
    from grim_20261007_amd import native
    native.ensure_device(0)
    return native


@pytest.mark.parametrize("prec,tol", [(64, 1e-13), (32, 2e-5)])
# k <= 256 runs on 32-row output tiles, longer products on 64-row tiles: both sides of that
# threshold, with row counts that are not multiples of either tile
@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (130, 77, 35), (200, 300, 129), (1, 5, 3), (97, 131, 256),
                                   (161, 70, 257), (300, 200, 300)])
def test_gemm_minus(prec, tol, m, n, k):
    dev = _dev()
    native = _lib()
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    dt = np.complex128 if prec == 64 else np.complex64
    a = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))).astype(dt)
    b = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(dt)
    c = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(dt)
    ta, tb, tc = (torch.as_tensor(np.ascontiguousarray(x.T), device=dev) for x in (a, b, c))
    code = native.lib().grim_gemm_minus(1, m, n, k, ta.data_ptr(), m, 0, tb.data_ptr(), k, 0, tc.data_ptr(), m, 0,
                                        prec, torch.cuda.current_stream().cuda_stream)
    assert code == 0
    want = c.astype(np.complex128) - a.astype(np.complex128) @ b.astype(np.complex128)
    got = _mat(tc)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < tol


@pytest.mark.parametrize("n", [1, 7, 64, 100, 257, 1100, 2100])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_lu_solve(n, precision):
    from grim_20261007_amd import engine
    dev = _dev()
    _lib()
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * np.sqrt(n) * np.eye(n) * (1 + 1j)
    b = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    z = torch.as_tensor(np.ascontiguousarray(a.T), device=dev)
    tb = torch.as_tensor(np.ascontiguousarray(b.T), device=dev)
    f = engine.factor(z, precision)
    x, info = engine.solve(f, tb)
    want = np.linalg.solve(a, b)
    err = np.max(np.abs(x.cpu().numpy().T - want)) / np.max(np.abs(want))
    print(f"lu n={n} {precision} err={err:.2e} steps={info.get('refinement_steps')}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_batched_lu_solve(precision):
    """Three systems factored and solved in the same launches equal three separate solves."""
    from grim_20261007_amd import engine
    dev = _dev()
    _lib()
    rng = np.random.default_rng(42)
    n, nb, na = 333, 3, 17
    a = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n)) + 30 * np.eye(n)
    b = rng.standard_normal((nb, n, na)) + 1j * rng.standard_normal((nb, n, na))
    z = torch.as_tensor(np.ascontiguousarray(np.transpose(a, (0, 2, 1))), device=dev)
    tb = torch.as_tensor(np.ascontiguousarray(np.transpose(b, (0, 2, 1))), device=dev)
    f = engine.factor(z, precision)
    x, info = engine.solve(f, tb)
    got = np.transpose(x.cpu().numpy(), (0, 2, 1))
    for i in range(nb):
        want = np.linalg.solve(a[i], b[i])
        assert np.max(np.abs(got[i] - want)) / np.max(np.abs(want)) < 1e-11
        single = engine.factor(z[i].contiguous(), precision)
        xi, _ = engine.solve(single, tb[i].contiguous())
        assert np.max(np.abs(xi.cpu().numpy().T - got[i])) <= 1e-13 * np.max(np.abs(want))


def test_refinement_is_per_system():
    """Mixed-precision refinement stops each system on its own error (linalg/refined_lu.py:41-61
    refines one matrix at a time): a batch of an easy and a harder system returns, bit for bit,
    what two separate calls return, the step count is the larger of the two, and the returned
    residual is each system's own final one."""
    from grim_20261007_amd import engine
    dev = _dev()
    _lib()
    rng = np.random.default_rng(77)
    n, na = 300, 9
    easy = 200.0 * np.eye(n) + rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    hard = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))       # Gaussian: kappa ~ 1e3
    a = np.stack([easy, hard])
    b = rng.standard_normal((2, n, na)) + 1j * rng.standard_normal((2, n, na))
    z = torch.as_tensor(np.ascontiguousarray(np.transpose(a, (0, 2, 1))), device=dev)
    tb = torch.as_tensor(np.ascontiguousarray(np.transpose(b, (0, 2, 1))), device=dev)
    x, info = engine.solve(engine.factor(z, "mixed"), tb)
    steps = []
    for i in range(2):
        xi, inf_i = engine.solve(engine.factor(z[i:i + 1].contiguous(), "mixed"), tb[i:i + 1].contiguous())
        steps.append(inf_i["refinement_steps"])
        assert torch.equal(xi[0].cpu(), x[i].cpu()), i
        assert torch.equal(inf_i["residual"][0].cpu(), info["residual"][i].cpu()), i
    print(f"refinement steps alone {steps}, batched {info['refinement_steps']}", flush=True)
    assert steps[0] < steps[1]                     # the systems do stop at different steps
    assert info["refinement_steps"] == max(steps)
    for i in range(2):
        want = np.linalg.solve(a[i], b[i])
        assert np.max(np.abs(x[i].cpu().numpy().T - want)) / np.max(np.abs(want)) < 1e-11


def test_lu_pivoting_needed():
    """A matrix with a zero leading entry needs row interchanges."""
    from grim_20261007_amd import engine
    dev = _dev()
    _lib()
    rng = np.random.default_rng(5)
    n = 300
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a[0, 0] = 0.0
    b = rng.standard_normal((n, 3)) + 0j
    f = engine.factor(torch.as_tensor(np.ascontiguousarray(a.T), device=dev), "double")
    x, _ = engine.solve(f, torch.as_tensor(np.ascontiguousarray(b.T), device=dev), gate=False)
    want = np.linalg.solve(a, b)
    assert np.max(np.abs(x.cpu().numpy().T - want)) / np.max(np.abs(want)) < 1e-10


@pytest.mark.parametrize("n", [300, 1500, 2500])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_lu_random_pivoting(n, precision):
    """Gaussian matrices pivot on almost every column: the factors reproduce P A = L U,
    the pivot sequence is LAPACK getrf's (|re| + |im| rule) and the solve is accurate."""
    import scipy.linalg
    from grim_20261007_amd import engine
    dev = _dev()
    _lib()
    rng = np.random.default_rng(1000 + n)
    a = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    b = rng.standard_normal((2, n, 4)) + 1j * rng.standard_normal((2, n, 4))
    z = torch.as_tensor(np.ascontiguousarray(np.transpose(a, (0, 2, 1))), device=dev)
    f = engine.factor(z, precision)
    x, _ = engine.solve(f, torch.as_tensor(np.ascontiguousarray(np.transpose(b, (0, 2, 1))), device=dev))
    piv = f.piv.cpu().numpy()
    lu = np.transpose(f.lu.cpu().numpy(), (0, 2, 1)).astype(np.complex128)
    for i in range(2):
        lu_ref, piv_ref = scipy.linalg.lu_factor(a[i].astype(np.complex64) if precision == "mixed" else a[i])
        agree = np.mean(piv[i] == piv_ref)
        # the first 64-column panel sees identical data; later FP64 pivots may differ by rounding
        # near ties, FP32 ones (different summation order from cgetrf, 24-bit mantissa) often do
        assert np.array_equal(piv[i, :64], piv_ref[:64])
        if precision == "double":
            assert agree > 0.95, agree
        # P A = L U with the returned pivots
        perm = np.arange(n)
        for c in range(n):
            perm[[c, piv[i, c]]] = perm[[piv[i, c], c]]
        low = np.tril(lu[i], -1) + np.eye(n)
        up = np.triu(lu[i])
        rel = np.max(np.abs(a[i][perm] - low @ up)) / np.max(np.abs(a[i]))
        tol = 1e-12 if precision == "double" else 2e-4
        want = np.linalg.solve(a[i], b[i])
        err = np.max(np.abs(x[i].cpu().numpy().T - want)) / np.max(np.abs(want))
        print(f"pivoting n={n} {precision} matrix {i}: pivots agree {agree:.4f}, |PA-LU| {rel:.2e}, "
              f"solve {err:.2e}", flush=True)
        assert rel < tol
        assert err < 1e-10


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_lu_tall_panels(precision):
    """N = 16500: the first panels are taller than one workgroup holds and run on the
    multi-workgroup leaf with device-wide barriers.  Checked by the first panel's pivots
    (LAPACK getrf on the 16500 x 64 panel alone picks the same rows), the backward error
    of the solve, and the status word."""
    import scipy.linalg
    from grim_20261007_amd import engine
    dev = _dev()
    _lib()
    n, nb = 16500, 3
    g = torch.Generator(device=dev).manual_seed(7)
    a = torch.randn((1, n, n), dtype=torch.complex128, generator=g, device=dev)   # column-major storage
    b = torch.randn((1, nb, n), dtype=torch.complex128, generator=g, device=dev)
    if precision == "mixed":
        # FP32 LU + refinement needs kappa * n * u32 < 1, which a Gaussian matrix of this size
        # does not give: use a well-conditioned matrix whose rows are shuffled, so that every
        # column still pivots
        a = 0.01 * a + 3.0 * torch.eye(n, dtype=torch.complex128, device=dev)
        perm = torch.randperm(n, generator=g, device=dev)
        a = a[:, :, perm].contiguous()
    f = engine.factor(a, precision)                                               # raises on a device error
    x, info = engine.solve(f, b, gate=False) if precision == "double" else engine.solve(f, b)
    panel = a[0, :64, :].cpu().numpy().T                                           # rows x 64 columns
    if precision == "mixed":
        panel = panel.astype(np.complex64)
    _, piv_ref = scipy.linalg.lu_factor(panel)
    assert np.array_equal(f.piv[0, :64].cpu().numpy(), piv_ref[:64])
    r = engine.residual(a, x, b)
    a_inf = float(torch.max(torch.sum(torch.abs(a[0]), dim=0)))                   # row sums of |A|
    berr = float(torch.max(torch.abs(r)) / (a_inf * torch.max(torch.abs(x)) + torch.max(torch.abs(b))))
    print(f"tall LU n={n} {precision}: backward error {berr:.2e} steps {info.get('refinement_steps')}", flush=True)
    assert berr < (1e-14 if precision == "double" else 1e-13)


@pytest.mark.parametrize("name", golden_cases.ALL)
def test_assembly_matches_reference(name):
    from grim_20261007_amd import engine
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import wavenumber
    dev = _dev()
    case = golden_cases.load(name)
    mesh = build_mesh(golden_cases.snapshot(case), "meters", float(case["frequency_ghz"]))
    assert np.array_equal(mesh.elements, case["elements"])
    k = wavenumber(float(case["frequency_ghz"]))
    contour = engine.DeviceContour(mesh, dev)
    z_tm, z_te = engine.assemble(contour, k)
    for pol, z in (("TM", _mat(z_tm)), ("TE", _mat(z_te))):
        if f"Z_{pol}" in case:
            ref = case[f"Z_{pol}"]
        else:
            rows = case["matrix_rows"]
            ref = case[f"Zrows_{pol}"]
            z = z[rows]
        err = np.max(np.abs(z - ref)) / np.max(np.abs(ref))
        print(f"{name} {pol} assembly max rel err {err:.2e}", flush=True)
        assert err < 1e-12


def test_assembly_matches_oracle_large_kr():
    """Circle R = 1 m, 400 panels, ka = 100: far-pair arguments k r from ~1 to 200 cover the
    series branch, every truncation length of the large-argument polynomials and the reduced
    sine/cosine.  The oracle evaluates the Hankel functions with SciPy."""
    from grim_20261007_amd import engine, workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import wavenumber
    from oracle import bem2d
    dev = _dev()
    _lib()
    f = workloads.frequency_for_ka(100.0, 1.0)
    snap = workloads.circle(1.0, 400)
    mesh = build_mesh(snap, "meters", f)
    om = bem2d.build_mesh(snap, "meters", f)
    assert np.array_equal(mesh.elements, om.elements)
    k = wavenumber(f)
    z_tm, z_te = engine.assemble(engine.DeviceContour(mesh, dev), k)
    ref_tm, ref_te = bem2d.assemble(om, bem2d.frequency_k(f))
    for pol, z, ref in (("TM", _mat(z_tm), ref_tm), ("TE", _mat(z_te), ref_te)):
        err = np.max(np.abs(z - ref)) / np.max(np.abs(ref))
        print(f"ka=100 {pol} assembly vs oracle max rel err {err:.2e}", flush=True)
        assert err < 1e-12


def _two_bodies():
    from grim_20261007_amd import workloads
    a = workloads.circle_points(0.5, 120) + np.array([-1.5, 0.0])
    b = workloads.star_points(150, 0.6, [0.05, 0.02, 0, 0, 0, 0, 0, 0], [0.3, 1.0, 0, 0, 0, 0, 0, 0]) + \
        np.array([1.5, 0.2])
    segs = workloads._snapshot(a, "left")["segments"] + workloads._snapshot(b, "right")["segments"]
    return {"segments": segs, "ibcs": [], "dielectrics": []}, workloads.frequency_for_ka(8.0, 0.5)


def test_two_contours_match_oracle():
    """Two closed bodies (a circle and a star, 270 nodes): contour starts inside a row block, far
    and near pairs between the bodies, and the full solve, against the oracle."""
    from grim_20261007_amd import engine
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import solve_frequency, wavenumber
    from oracle import bem2d
    dev = _dev()
    _lib()
    snap, f = _two_bodies()
    mesh = build_mesh(snap, "meters", f)
    om = bem2d.build_mesh(snap, "meters", f)
    assert np.array_equal(mesh.elements, om.elements)
    k = wavenumber(f)
    z_tm, z_te = engine.assemble(engine.DeviceContour(mesh, dev), k)
    ref_tm, ref_te = bem2d.assemble(om, bem2d.frequency_k(f))
    for pol, z, ref in (("TM", _mat(z_tm), ref_tm), ("TE", _mat(z_te), ref_te)):
        err = np.max(np.abs(z - ref)) / np.max(np.abs(ref))
        print(f"two bodies {pol} assembly vs oracle {err:.2e}", flush=True)
        assert err < 1e-12
    angles = np.arange(0.0, 360.0, 5.0)
    res = solve_frequency(mesh, f, list(angles), ("TE", "TM"), "double", dev)
    want = bem2d.solve_monostatic(om, f, angles)
    for pol in ("TE", "TM"):
        amp_ref = want[pol]["amplitude"]
        err = np.max(np.abs(res[pol]["amplitude"] - amp_ref)) / np.max(np.abs(amp_ref))
        print(f"two bodies {pol} amplitude vs oracle {err:.2e}", flush=True)
        assert err < 1e-10


def test_singular_system_is_reported():
    """A zero column gives a zero pivot: the status word turns it into LinAlgError."""
    from grim_20261007_amd import engine
    dev = _dev()
    _lib()
    rng = np.random.default_rng(3)
    n = 200
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a[:, 77] = 0.0
    for precision in ("double", "mixed"):
        with pytest.raises(np.linalg.LinAlgError):
            engine.factor(torch.as_tensor(np.ascontiguousarray(a.T), device=dev), precision)


@pytest.mark.parametrize("name", golden_cases.ALL)
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_solve_matches_reference(name, precision):
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import solve_frequency
    dev = _dev()
    case = golden_cases.load(name)
    f = float(case["frequency_ghz"])
    mesh = build_mesh(golden_cases.snapshot(case), "meters", f)
    angles = case["angles_deg"].tolist()
    res = solve_frequency(mesh, f, angles, ("TE", "TM"), precision, dev, return_density=True)
    for pol in ("TE", "TM"):
        r = res[pol]
        cols = case["density_columns"]
        xref = case[f"density_{pol}"]
        xerr = np.max(np.abs(r["density"][:, cols] - xref)) / np.max(np.abs(xref))
        aref = case[f"sample_amp_{pol}"]
        aerr = np.max(np.abs(r["amplitude"] - aref) / np.abs(aref))
        dberr = np.max(np.abs(10 * np.log10(r["sigma"]) - case[f"sample_db_{pol}"]))
        print(f"{name} {pol} {precision}: density {xerr:.2e} amplitude {aerr:.2e} dB {dberr:.2e} "
              f"residual {np.max(r['residual']):.2e}", flush=True)
        assert xerr < 1e-10
        assert aerr < 1e-10
        assert dberr < 1e-8
        assert np.max(r["residual"]) < 1e-10
