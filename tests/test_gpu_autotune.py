"""Launch shapes on the GPU: every far-pass, triangular-solve and block-solve shape gives bit-identical
results; the autotuner measures a table and a server that applies it returns the same numbers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture
def clean_shapes():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from grim_20261007_amd import autotune, native
    native.ensure_device(0)
    autotune.clear()
    yield autotune
    autotune.clear()


def _system(n):
    from grim_20261007_amd import autotune, engine
    dev = torch.device("cuda", 0)
    mesh, contour, k = autotune._contour(n, dev)
    return mesh, contour, k, engine.assembly_workspace(contour.n, dev)


@pytest.mark.parametrize("n", [300, 1500])
def test_far_shapes_are_bit_identical(clean_shapes, n):
    from grim_20261007_amd import engine
    autotune = clean_shapes
    mesh, contour, k, ws = _system(n)
    ref = None
    for value in autotune.CHOICES[autotune.FAR_COLUMNS]:
        autotune.set_shape(autotune.FAR_COLUMNS, 0, value)
        print(f"n={n} far columns per wavefront {value}", flush=True)
        z_tm, z_te = engine.assemble(contour, k, ws)
        got = torch.stack((z_tm, z_te)).cpu()
        if ref is None:
            ref = got
        else:
            assert torch.equal(got, ref), value


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_trsm_shapes_are_bit_identical(clean_shapes, precision):
    """Factorization triangular-solve shapes (knob 1) give the same factors bit for bit."""
    from grim_20261007_amd import engine
    autotune = clean_shapes
    dev = torch.device("cuda", 0)
    g = torch.Generator(device="cpu").manual_seed(5)
    n = 700
    z = (torch.randn((2, n, n), dtype=torch.complex128, generator=g) + n * torch.eye(n, dtype=torch.complex128)).to(dev)
    ref = None
    for value in autotune.CHOICES[autotune.TRSM_COLUMNS]:
        autotune.set_shape(autotune.TRSM_COLUMNS, 0, value)
        print(f"{precision} factor: trsm columns per wavefront {value}", flush=True)
        f = engine.factor(z, precision)
        # factors, pivots and the row permutation (the first 4 n bytes of aux; the rest is scratch)
        got = (f.lu.cpu(), f.piv.cpu(), f.aux[:, :4 * n].cpu())
        if ref is None:
            ref = got
        else:
            assert all(torch.equal(x, y) for x, y in zip(got, ref)), value


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("n,na", [(300, 77), (700, 77), (300, 600)])
def test_solve_shapes_are_bit_identical(clean_shapes, precision, n, na):
    """Block-solve shapes (knob 2) give the same solutions bit for bit (n = 300: a ragged
    second 256-row block; 1 200 columns at one column per wavefront: more complex128
    workgroups than CUs, so the 4-column coupling chunks against the 16-column ones)."""
    from grim_20261007_amd import engine
    autotune = clean_shapes
    dev = torch.device("cuda", 0)
    g = torch.Generator(device="cpu").manual_seed(5)
    z = torch.randn((2, n, n), dtype=torch.complex128, generator=g) + n * torch.eye(n, dtype=torch.complex128)
    fac = engine.factor(z.to(dev), precision)
    b = torch.randn((2, na, n), dtype=torch.complex128, generator=g).to(dev).to(fac.lu.dtype)
    ref = None
    for value in autotune.CHOICES[autotune.SOLVE_COLUMNS]:
        autotune.set_shape(autotune.SOLVE_COLUMNS, 0, value)
        print(f"{precision} n={n} A={na} solve columns per wavefront {value}", flush=True)
        x = engine.lu_solve_inplace(fac, b.clone()).cpu()
        if ref is None:
            ref = x
            # tensors hold column-major matrices: Z = z[i].T, B = b[i].T, X = x[i].T
            want = torch.linalg.solve(z.transpose(1, 2), b.cpu().to(z.dtype).transpose(1, 2)).transpose(1, 2)
            tol = 1e-11 if precision == "double" else 1e-4
            assert float((x - want).abs().max() / want.abs().max()) < tol
        else:
            assert torch.equal(x, ref), value


def test_tuned_server_matches_untuned(clean_shapes):
    from grim_20261007_amd.serving import Request, Server
    from tests.test_gpu_serving import _requests
    autotune = clean_shapes
    table = autotune.tune(far_sizes=(200, 600), angle_counts=(40, 90), factor_n=600, repeats=2)
    print("table", table["far_columns"], table["trsm_columns"], table["solve_columns"], flush=True)
    for name in ("far_columns", "trsm_columns", "solve_columns"):
        assert table[name] and table[name][0][0] == 0
        assert all(v in autotune.CHOICES[k] for k, nm in autotune.KNOB_NAMES.items() if nm == name
                   for _, v in table[name])
    reqs = _requests()[:3]
    out = {}
    for label, tuning in (("plain", None), ("tuned", table)):
        autotune.clear()
        srv = Server(workers=2, tuning=tuning, cache_bytes=0)
        futs = [srv.submit(Request(r.mesh, r.frequency_ghz, r.angles_deg, precision="double", rid=r.rid, key=r.key))
                for r in reqs]
        out[label] = {f.result().rid: f.result() for f in futs}
        srv.close()
    assert autotune.shape(autotune.FAR_COLUMNS, 10 ** 6) == table["far_columns"][-1][1]
    for rid, res in out["tuned"].items():
        for pol in ("TE", "TM"):
            assert np.array_equal(res.amplitude[pol], out["plain"][rid].amplitude[pol])
