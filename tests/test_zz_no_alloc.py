"""grim_assemble_noalloc (static device scratch, no allocation) returns, bit for bit, what
grim_assemble with a caller-owned workspace returns: two calls in a row, the second smaller, so
the second runs on the first one's leftovers in the scratch; a request beyond the scratch is
refused before anything is launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(points, f):
    from grim_20261007_amd import engine, workloads
    from grim_20261007_amd.mesh import build_mesh
    from grim_20261007_amd.solver import wavenumber
    mesh = build_mesh(workloads._snapshot(points, "c"), "meters", f)
    return engine.DeviceContour(mesh, torch.device("cuda", 0)), wavenumber(f)


def _noalloc(contour, k):
    from grim_20261007_amd import engine, native
    n = contour.n
    pair = torch.empty((2, n, n), dtype=torch.complex128, device=contour.device)
    status = torch.ones(64, dtype=torch.uint8, device=contour.device)
    st = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    code = native.lib().grim_assemble_noalloc(
        n, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(), contour.prev.data_ptr(),
        contour.fix_rows.data_ptr(), int(contour.fix_rows.numel()), float(k), contour.mesh.far_order(k),
        pair[1].data_ptr(), n, pair[0].data_ptr(), n, engine.NEAR_PER_ELEMENT, status.data_ptr(), st)
    assert code == 0
    assert native.lib().grim_workspace_status(status.data_ptr(), st) == 0
    assert torch.cuda.memory_allocated() == before
    return pair


def test_noalloc_assembly_equals_workspace_assembly():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd import engine, native, workloads
    native.ensure_device(0)
    pts = workloads.naca0012_points(1000)
    cases = [(pts, workloads.frequency_for_segments_per_wavelength(pts, 10.0)),
             (workloads.circle_points(1.0, 400), workloads.frequency_for_ka(10.0, 1.0))]
    for points, f in cases:                     # N = 2000, then N = 400 on the same scratch
        contour, k = _case(points, f)
        print(f"no-alloc assembly N={contour.n}", flush=True)
        got = _noalloc(contour, k)
        z_tm, z_te = engine.assemble(contour, k)
        assert torch.equal(got[0], z_te) and torch.equal(got[1], z_tm)


def test_noalloc_refuses_beyond_scratch():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd import engine, native, workloads
    native.ensure_device(0)
    contour, k = _case(workloads.circle_points(1.0, 9975), workloads.frequency_for_ka(100.0, 1.0))
    n = contour.n
    status = torch.zeros(64, dtype=torch.uint8, device=contour.device)
    z = torch.zeros(8, dtype=torch.complex128, device=contour.device)   # never written: refused first
    code = native.lib().grim_assemble_noalloc(
        n, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(), contour.prev.data_ptr(),
        contour.fix_rows.data_ptr(), int(contour.fix_rows.numel()), float(k), 8, z.data_ptr(), n, z.data_ptr(), n,
        engine.NEAR_PER_ELEMENT, status.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert code == 1
    assert torch.count_nonzero(z).item() == 0
