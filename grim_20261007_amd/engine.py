"""Device-side solve pipeline: assembly, factorization, many-angle solves and RCS.

Every numerical step runs in the HIP library through the C ABI; this module
only owns device buffers (PyTorch-ROCm tensors), streams and bookkeeping.
Matrices are stored column-major: a tensor ``t`` of shape (ncols, nrows)
holds the column-major matrix ``t.T``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import native
from .mesh import EPS, ContourMesh, causal_index

NEAR_PER_ELEMENT = 48
REFINE_TOL = 2e-12           # linalg/refined_lu.py:54
REFINE_MAX_STEPS = 10        # linalg/refined_lu.py:49
BACKWARD_ERROR_MAX = 1e-12   # twod/constants.py:24


def request_device_bytes(n: int, nangles: int, npols: int, precision: str, cached: bool = False,
                         degree: int = 1, sheet_elements: int = 0) -> int:
    """Device bytes one co-polarized request allocates: the two assembled systems and their
    factors (unless cached), the loads, right-hand sides, solutions and residuals, refinement
    and assembly workspace.  n counts unknowns; with degree-2 / degree-3 elements (n / degree
    elements) the assembly workspace holds three tables of element-pair blocks,
    48 (degree + 1)^2 (n / degree)^2 bytes.  A sheet mesh (sheet_elements > 0 panels) assembles its
    npols systems one after the other over one table of 64 ne^2 bytes of element-pair S blocks, each
    with its own loads."""
    z = 16 * n * n
    lu = (8 if precision == "mixed" else 16) * n * n
    vec = 16 * n * nangles
    if sheet_elements > 0:
        ne = sheet_elements
        work = 64 * ne * ne + NEAR_PER_ELEMENT * ne * 160 + 64 * ne
        refine = (24 * n * nangles) if precision == "mixed" else 0
        return npols * (z + lu + 2 * vec + vec * 3 + refine) + work
    if degree > 1:
        ne = n // degree
        work = 0 if cached else 48 * (degree + 1) ** 2 * ne * ne + NEAR_PER_ELEMENT * ne * 16 + 64 * ne
    else:
        work = 0 if cached else NEAR_PER_ELEMENT * n * 160 + 64 * n
    refine = (24 * n * nangles) if precision == "mixed" else 0
    per_pol = vec * 3 + refine                       # rhs, solution, residual
    systems = 0 if cached else 2 * (z + lu)          # both polarizations assembled together
    return systems + 2 * vec + npols * per_pol + work


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class DeviceContour:
    """Contour geometry resident on one GPU."""

    def __init__(self, mesh: ContourMesh, device: torch.device | str = "cuda"):
        self.mesh = mesh
        self.device = torch.device(device)
        native.ensure_device(self.device.index if self.device.index is not None else torch.cuda.current_device())
        self.n = mesh.n
        self.p0 = torch.as_tensor(np.ascontiguousarray(mesh.p0, np.float64), device=self.device)
        self.p1 = torch.as_tensor(np.ascontiguousarray(mesh.p1, np.float64), device=self.device)
        self.next = torch.as_tensor(np.ascontiguousarray(mesh.next, np.int32), device=self.device)
        self.prev = torch.as_tensor(np.ascontiguousarray(mesh.prev, np.int32), device=self.device)
        self.fix_rows = torch.as_tensor(mesh.fix_rows(), device=self.device)

    def matrix_bytes(self, precision: str = "double") -> int:
        return self.n * self.n * (16 if precision == "double" else 8)


def assembly_workspace(n: int, device, degree: int = 1) -> torch.Tensor:
    """Assembly workspace of an n-element contour with elements of the given degree."""
    if degree > 1:
        return poly_workspace(n, degree, device)
    nbytes = native.lib().grim_assemble_workspace_bytes(n, NEAR_PER_ELEMENT)
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device)


def assemble(contour: DeviceContour, k: float, workspace: torch.Tensor | None = None, check: bool = True,
             out: torch.Tensor | None = None):
    """TM and TE system matrices (complex128, column-major) from one fused traversal.

    Returns (z_tm, z_te), two views of one (2, N, N) tensor ``pair`` = [TE, TM]
    (also reachable as ``z_tm._base``) so that both can be factored and solved in
    batched launches.  ``out``, a contiguous (2, N, N) complex128 tensor (e.g. two systems of a
    larger batch), receives the pair instead of a new allocation."""
    if contour.mesh.degree > 1:
        return assemble_poly(contour, k, workspace, out)
    n = contour.n
    dev = contour.device
    if out is None:
        pair = torch.empty((2, n, n), dtype=torch.complex128, device=dev)
    else:
        if out.shape != (2, n, n) or out.dtype != torch.complex128 or not out.is_contiguous():
            raise ValueError("out must be a contiguous (2, N, N) complex128 tensor")
        pair = out
    z_te, z_tm = pair[0], pair[1]
    ws = workspace if workspace is not None else assembly_workspace(n, dev)
    st = _stream(dev)
    code = native.lib().grim_assemble(
        n, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(), contour.prev.data_ptr(),
        contour.fix_rows.data_ptr(), int(contour.fix_rows.numel()), float(k), contour.mesh.far_order(k),
        z_tm.data_ptr(), n, z_te.data_ptr(), n, ws.data_ptr(), ws.numel(), NEAR_PER_ELEMENT, st)
    native.check("grim_assemble", code)
    if check:
        native.check("grim_assemble (device)", native.lib().grim_workspace_status(ws.data_ptr(), st))
    return z_tm, z_te


# ----------------------------------------------------------------------------
# Degree-2 and degree-3 bases (twod/basis.py, twod/polynomial_quadrature.py)
# ----------------------------------------------------------------------------
POLY_NEAR_MESSAGE = "Polynomial near quadrature did not converge; refine the mesh."   # polynomial_quadrature.py:143
POLY_INTERSECT_MESSAGE = "Polynomial quadrature encountered intersecting integration points."   # :104


def poly_workspace(n: int, degree: int, device) -> torch.Tensor:
    nbytes = int(native.lib().grim_assemble_poly_workspace_bytes(n, degree, NEAR_PER_ELEMENT))
    if nbytes < 0:
        raise ValueError(f"no polynomial assembly for n={n}, degree={degree}")
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def assemble_poly(contour: DeviceContour, k: float, workspace: torch.Tensor | None = None,
                  out: torch.Tensor | None = None):
    """TM and TE systems of a contour with degree-2 or degree-3 elements (mesh.degree), N d x N d
    complex128 column-major; returns (z_tm, z_te) views of one (2, Nd, Nd) [TE, TM] tensor (also
    ``z_tm._base``), written into ``out`` when given.  Raises ValueError where the reference's
    polynomial near quadrature raises (non-convergence, coincident points)."""
    mesh = contour.mesh
    d = mesh.degree
    if d not in (2, 3):
        raise ValueError("assemble_poly needs a mesh of degree 2 or 3")
    n, m = contour.n, mesh.dofs
    dev = contour.device
    if out is None:
        pair = torch.empty((2, m, m), dtype=torch.complex128, device=dev)
    else:
        if out.shape != (2, m, m) or out.dtype != torch.complex128 or not out.is_contiguous():
            raise ValueError("out must be a contiguous (2, Nd, Nd) complex128 tensor")
        pair = out
    z_te, z_tm = pair[0], pair[1]
    ws = workspace if workspace is not None else poly_workspace(n, d, dev)
    st = _stream(dev)
    native.check("grim_assemble_poly", native.lib().grim_assemble_poly(
        n, d, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(), contour.prev.data_ptr(),
        float(k), mesh.far_order(k), z_tm.data_ptr(), m, z_te.data_ptr(), m, ws.data_ptr(), ws.numel(),
        NEAR_PER_ELEMENT, st))
    code = native.lib().grim_workspace_status(ws.data_ptr(), st)
    if code == 6:
        raise ValueError(POLY_NEAR_MESSAGE)
    if code == 10:
        raise ValueError(POLY_INTERSECT_MESSAGE)
    native.check("grim_assemble_poly (device)", code)
    return z_tm, z_te


_RULES: dict = {}


def _gauss_rule(order: int, device) -> tuple:
    key = (order, str(device))
    if key not in _RULES:
        x, w = np.polynomial.legendre.leggauss(order)
        _RULES[key] = (torch.tensor(0.5 * (x + 1.0), dtype=torch.float64, device=device),
                       torch.tensor(0.5 * w, dtype=torch.float64, device=device))
    return _RULES[key]


def loads_poly(contour: DeviceContour, k: float, angles_deg: torch.Tensor, want_u=True, want_dn=True):
    """Plane-wave loads of a degree-2 / degree-3 contour, (A, N d) complex128 (column-major
    N d x A): Gauss moments of the nodal basis (twod/basis.py:89-104, assembly/kernels.py:236-259)."""
    mesh = contour.mesh
    d = mesh.degree
    na = int(angles_deg.numel())
    dev = contour.device
    m = mesh.dofs
    order = mesh.load_order(k)
    qt, qw = _gauss_rule(order, dev)
    bu = torch.empty((na, m), dtype=torch.complex128, device=dev) if want_u else None
    bdn = torch.empty((na, m), dtype=torch.complex128, device=dev) if want_dn else None
    native.check("grim_loads_poly", native.lib().grim_loads_poly(
        contour.n, d, na, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.prev.data_ptr(), float(k),
        angles_deg.data_ptr(), order, qt.data_ptr(), qw.data_ptr(), native.ptr(bu), native.ptr(bdn), m,
        _stream(dev)))
    return bu, bdn


def loads(contour: DeviceContour, k: float, angles_deg: torch.Tensor, want_u=True, want_dn=True):
    """Nodal plane-wave loads (A, N) complex128 tensors (column-major N x A); (A, N d) for a contour
    of degree d > 1 (loads_poly)."""
    if contour.mesh.degree > 1:
        return loads_poly(contour, k, angles_deg, want_u, want_dn)
    na = int(angles_deg.numel())
    dev = contour.device
    bu = torch.empty((na, contour.n), dtype=torch.complex128, device=dev) if want_u else None
    bdn = torch.empty((na, contour.n), dtype=torch.complex128, device=dev) if want_dn else None
    code = native.lib().grim_loads(contour.n, na, contour.p0.data_ptr(), contour.p1.data_ptr(),
                                   contour.prev.data_ptr(), float(k), angles_deg.data_ptr(), native.ptr(bu),
                                   native.ptr(bdn), contour.n, _stream(dev))
    native.check("grim_loads", code)
    return bu, bdn


def farfield(x: torch.Tensor, bu: torch.Tensor, bdn: torch.Tensor, k: float):
    """Matched monostatic amplitude B (A,) and sigma_2D (A,)."""
    na, n = x.shape
    amp = torch.empty(na, dtype=torch.complex128, device=x.device)
    sigma = torch.empty(na, dtype=torch.float64, device=x.device)
    code = native.lib().grim_farfield(n, na, x.data_ptr(), n, bu.data_ptr(), bdn.data_ptr(), n, float(k),
                                      amp.data_ptr(), sigma.data_ptr(), _stream(x.device))
    native.check("grim_farfield", code)
    return amp, sigma


ETA0 = 376.730313668      # free-space wave impedance (twod/constants.py)


# ----------------------------------------------------------------------------
# Sheets (TYPE 1 open contours), alone or beside PEC contours (twod/formulations/sheet.py)
# ----------------------------------------------------------------------------
class DeviceSheet:
    """One polarization's sheet mesh resident on one GPU: element geometry, the near test's end ids,
    the node-to-element map and the pinned strip-end rows."""

    def __init__(self, mesh, device: torch.device | str = "cuda"):
        self.mesh = mesh
        self.device = torch.device(device)
        native.ensure_device(self.device.index if self.device.index is not None else torch.cuda.current_device())
        self.n, self.ne = mesh.n, mesh.ne
        dev = self.device
        self.p0 = torch.as_tensor(np.ascontiguousarray(mesh.p0, np.float64), device=dev)
        self.p1 = torch.as_tensor(np.ascontiguousarray(mesh.p1, np.float64), device=dev)
        self.next_ids = torch.as_tensor(mesh.next_ids(), device=dev)
        own_ptr, own_el, own_loc = mesh.owners()
        self.own_ptr = torch.as_tensor(own_ptr, device=dev)
        self.own_el = torch.as_tensor(own_el, device=dev)
        self.own_loc = torch.as_tensor(own_loc, device=dev)
        pinned = np.zeros(mesh.n, np.int32)
        pinned[mesh.endpoints] = 1
        self.pinned = torch.as_tensor(pinned, device=dev)
        self.endpoints = torch.as_tensor(mesh.endpoints, device=dev)


def sheet_workspace(ne: int, device) -> torch.Tensor:
    nbytes = int(native.lib().grim_sheet_workspace_bytes(ne, NEAR_PER_ELEMENT))
    if nbytes < 0:
        raise ValueError("Sheet meshes are limited to 65535 panels.")
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def sheet_coefficients(mesh, k: float, pol: str) -> np.ndarray:
    """Element mass coefficients: Zs / (j k eta0) (TM), j k Zs / eta0 (TE); 0 on PEC elements
    (twod/formulations/sheet.py:12-22)."""
    z = np.asarray(mesh.impedance, complex)
    return z / (1j * k * ETA0) if pol == "TM" else (1j * k / ETA0) * z


def sheet_system(sheet: DeviceSheet, k: float, pol: str, workspace: torch.Tensor | None = None,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """The (N, N) complex128 system (column-major) of one polarization: S - c M (TM) or W - c M with
    the strip-end rows pinned (TE)."""
    n = sheet.n
    z = torch.empty((n, n), dtype=torch.complex128, device=sheet.device) if out is None else out
    sheet_systems(sheet, k, workspace, **{pol.lower(): z})
    return z


def sheet_systems(sheet: DeviceSheet, k: float, workspace: torch.Tensor | None = None,
                  te: torch.Tensor | None = None, tm: torch.Tensor | None = None) -> None:
    """Both polarizations' systems from one traversal when their meshes coincide (a sheet whose TE
    chains need no flip and whose TM nodes are not split): one near pass and one table of S blocks
    feed both gathers.  Either output may be None; the pinned rows apply to TE."""
    n, ne, dev = sheet.n, sheet.ne, sheet.device
    ws = workspace if workspace is not None else sheet_workspace(ne, dev)
    c_te = torch.as_tensor(sheet_coefficients(sheet.mesh, k, "TE"), device=dev) if te is not None else None
    c_tm = torch.as_tensor(sheet_coefficients(sheet.mesh, k, "TM"), device=dev) if tm is not None else None
    st = _stream(dev)
    code = native.lib().grim_assemble_sheet(
        ne, n, sheet.p0.data_ptr(), sheet.p1.data_ptr(), sheet.next_ids.data_ptr(), sheet.own_ptr.data_ptr(),
        sheet.own_el.data_ptr(), sheet.own_loc.data_ptr(), float(k), sheet.mesh.far_order(k, False),
        sheet.mesh.far_order(k, True), native.ptr(c_tm), native.ptr(c_te),
        sheet.pinned.data_ptr() if te is not None else None, native.ptr(tm), n, native.ptr(te), n,
        ws.data_ptr(), ws.numel(), NEAR_PER_ELEMENT, st)
    native.check("grim_assemble_sheet", code)
    native.check("grim_assemble_sheet (device)", native.lib().grim_workspace_status(ws.data_ptr(), st))


def sheet_loads(sheet: DeviceSheet, k: float, angles_deg: torch.Tensor):
    """Plane-wave loads (A, N) of the sheet mesh's nodes: b_u and b_dn."""
    na = int(angles_deg.numel())
    bu = torch.empty((na, sheet.n), dtype=torch.complex128, device=sheet.device)
    bdn = torch.empty((na, sheet.n), dtype=torch.complex128, device=sheet.device)
    code = native.lib().grim_loads_sheet(sheet.ne, sheet.n, na, sheet.p0.data_ptr(), sheet.p1.data_ptr(),
                                         sheet.own_ptr.data_ptr(), sheet.own_el.data_ptr(),
                                         sheet.own_loc.data_ptr(), float(k), angles_deg.data_ptr(), bu.data_ptr(),
                                         bdn.data_ptr(), sheet.n, _stream(sheet.device))
    native.check("grim_loads_sheet", code)
    return bu, bdn


def sheet_rhs(sheet: DeviceSheet, pol: str, bu: torch.Tensor, bdn: torch.Tensor) -> torch.Tensor:
    """-b_u (TM); b_dn with the pinned strip-end rows zero (TE) (twod/formulations/sheet.py:33-39)."""
    if pol == "TM":
        return -bu
    rhs = bdn.clone()
    if sheet.endpoints.numel():
        rhs[:, sheet.endpoints] = 0
    return rhs


def farfield_potential(x: torch.Tensor, b: torch.Tensor, k: float):
    """Amplitude B (A,) = sum_p x[a, p] b[a, p] of one potential (b = b_u: single layer, b_dn: double
    layer) and sigma_2D (A,)."""
    na, n = x.shape
    amp = torch.empty(na, dtype=torch.complex128, device=x.device)
    sigma = torch.empty(na, dtype=torch.float64, device=x.device)
    code = native.lib().grim_farfield_potential(n, na, x.data_ptr(), n, b.data_ptr(), n, float(k), amp.data_ptr(),
                                                sigma.data_ptr(), _stream(x.device))
    native.check("grim_farfield_potential", code)
    return amp, sigma


@dataclass
class RobinPlan:
    """Coefficients of a contour's impedance rows at one wavenumber (host arrays, node order)."""
    alpha_te_e: np.ndarray      # (N,) element coefficients j k Zs / eta0 (0 on PEC elements)
    alpha_tm_e: np.ndarray      # (N,) element coefficients j k eta0 / Zs (0 on PEC elements)
    alpha_te: np.ndarray        # (N,) nodal: mean of the node's two elements
    alpha_tm: np.ndarray        # (N,) nodal, 0 on every node of a PEC element
    dirichlet: np.ndarray       # (N,) int32: TM rows with |alpha| <= EPS keep the Dirichlet row
    delta_te: np.ndarray        # (N,) (alpha_e(p) - alpha_e(prev p)) / 2
    delta_tm: np.ndarray        # (N,) the same in TM, 0 on Dirichlet rows
    flags: np.ndarray           # (N,) int32: bit 0 / bit 1 = TE / TM row with |alpha| > EPS
    walk: np.ndarray            # (64 B,) int32 walk blocks of the rows with delta != 0


def robin_plan(mesh: ContourMesh, k: float) -> RobinPlan:
    """Element and nodal Robin coefficients of the contour's surface impedances in air
    (twod/geometry.py:2488-2521: alpha_TE = j k Zs/eta0, alpha_TM = j k eta0/Zs; 0 on PEC),
    nodal values as the mean over the node's two elements with every node of a PEC element at 0
    in TM (twod/formulations/regions.py:66-119), and the walk of the rows whose two elements
    carry different coefficients (zones, tapers): per block, lane 0 the element before lane 1's
    row, lanes 1..63 consecutive rows along the contour."""
    z = np.asarray(mesh.impedance, complex)
    ibc = np.abs(z) > EPS
    safe = np.where(ibc, z, 1.0)
    a_te = np.where(ibc, 1j * k * safe / ETA0, 0.0).astype(np.complex128)
    a_tm = np.where(ibc, 1j * k * ETA0 / safe, 0.0).astype(np.complex128)
    prev = np.asarray(mesh.prev, np.int64)
    n_te = (a_te + a_te[prev]) / 2
    n_tm = (a_tm + a_tm[prev]) / 2
    pec_e = np.abs(a_tm) <= EPS
    n_tm = np.where(pec_e | pec_e[prev], 0.0, n_tm)
    dirichlet = np.abs(n_tm) <= EPS
    d_te = (a_te - a_te[prev]) / 2
    d_tm = np.where(dirichlet, 0.0, (a_tm - a_tm[prev]) / 2)
    flags = ((np.abs(n_te) > EPS).astype(np.int32) | 2 * (~dirichlet).astype(np.int32))
    need = (d_te != 0) | (d_tm != 0)
    blocks = []
    seen = np.zeros(mesh.n, bool)
    nxt = np.asarray(mesh.next, np.int64)
    for start in range(mesh.n):
        if seen[start]:
            continue
        cycle, e = [], start
        while not seen[e]:
            seen[e] = True
            cycle.append(e)
            e = int(nxt[e])
        flags_c = need[cycle]
        # rotate so that a run of needed rows never wraps past the end of the list
        if flags_c.all():
            order = cycle
        else:
            first_gap = int(np.argmin(flags_c))
            order = cycle[first_gap:] + cycle[:first_gap]
        run = []
        for e in order + [-1]:
            if e >= 0 and need[e]:
                run.append(e)
                continue
            for i in range(0, len(run), 63):
                part = run[i:i + 63]
                blocks.append([int(prev[part[0]])] + part + [-1] * (63 - len(part)))
            run = []
    walk = np.array(blocks, np.int32).reshape(-1) if blocks else np.zeros(0, np.int32)
    return RobinPlan(a_te, a_tm, n_te, n_tm, dirichlet.astype(np.int32), d_te.astype(np.complex128),
                     d_tm.astype(np.complex128), flags, walk)


def _c128(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, np.complex128), device=dev)


def impedance_systems(contour: DeviceContour, k: float, pair: torch.Tensor, workspace: torch.Tensor) -> torch.Tensor:
    """Turn the [TE, TM] PEC systems `pair` (2, N, N) in place into the systems of the contours'
    impedance laws: flux rows Z_TE + alpha_p Z_TM for both polarizations, TM Dirichlet rows
    unchanged (grim_robin_rows); then, on rows whose two elements carry different coefficients,
    the element-weighted alpha S and trace-jump mass (grim_robin_split, which reads the near
    blocks that `assemble` left in `workspace`)."""
    n, dev = contour.n, pair.device
    plan = robin_plan(contour.mesh, k)
    st = _stream(dev)
    ax, ay, dm = _c128(plan.alpha_te, dev), _c128(plan.alpha_tm, dev), torch.as_tensor(plan.dirichlet, device=dev)
    native.check("grim_robin_rows", native.lib().grim_robin_rows(n, n, pair[0].data_ptr(), n, pair[1].data_ptr(), n,
                                                                 ax.data_ptr(), ay.data_ptr(), dm.data_ptr(), st))
    nblocks = plan.walk.size // 64
    if nblocks:
        walk = torch.as_tensor(plan.walk, device=dev)
        cte, ctm = _c128(plan.delta_te, dev), _c128(plan.delta_tm, dev)
        flags = torch.as_tensor(plan.flags, device=dev)
        code = native.lib().grim_robin_split(
            n, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(), contour.prev.data_ptr(),
            float(k), contour.mesh.far_order(k), nblocks, walk.data_ptr(), cte.data_ptr(), ctm.data_ptr(),
            flags.data_ptr(), workspace.data_ptr(), workspace.numel(), NEAR_PER_ELEMENT, pair[0].data_ptr(), n,
            pair[1].data_ptr(), n, st)
        native.check("grim_robin_split", code)
    return pair


def impedance_rhs(contour: DeviceContour, k: float, angles: torch.Tensor) -> torch.Tensor:
    """Right-hand sides (2, A, N) [TE, TM] of impedance_systems: -(b_dn with each element's own
    Robin coefficient on its trace) on flux rows, -b_u on TM Dirichlet rows (grim_robin_rhs)."""
    n, na, dev = contour.n, int(angles.numel()), angles.device
    plan = robin_plan(contour.mesh, k)
    out = torch.empty((2, na, n), dtype=torch.complex128, device=dev)
    st = _stream(dev)
    zero = torch.zeros(n, dtype=torch.int32, device=dev)
    dirichlet = torch.as_tensor(plan.dirichlet, device=dev)
    for i, (alpha, dm) in enumerate(((_c128(plan.alpha_te_e, dev), zero), (_c128(plan.alpha_tm_e, dev), dirichlet))):
        code = native.lib().grim_robin_rhs(n, na, contour.p0.data_ptr(), contour.p1.data_ptr(),
                                           contour.prev.data_ptr(), float(k), angles.data_ptr(), alpha.data_ptr(),
                                           dm.data_ptr(), out[i].data_ptr(), n, st)
        native.check("grim_robin_rhs", code)
    return out


def medium_wavenumber(index: complex, k0: float, what: str):
    """k0 n for a medium of causal index n: a float for a lossless medium (the real-argument
    kernels), a complex number k0 n with Im <= 0 for a lossy one (grim_assemble_operators_lossy).
    Negative-index media (Re n <= 0) are refused."""
    tol = 64.0 * np.finfo(float).eps * max(1.0, abs(index))
    if not index.real > 0.0:
        raise ValueError(f"{what} has refractive index {index!r}: negative-index media are not evaluated by the GPU "
                         "path. Use a medium with a positive index or the reference's CPU path.")
    if abs(index.imag) <= tol:
        return float(k0 * index.real)
    return complex(k0 * index)


def is_lossy(k) -> bool:
    return isinstance(k, complex)


def lossy_workspace_bytes(mesh) -> int:
    """Device bytes of the complex-wavenumber operator pass a mesh needs beyond the real assembly:
    its two tables of element-pair blocks (128 n^2 bytes) and near lists for the largest lossy contour
    set (a lossy dielectric body, or the contours of a lossy region); 0 without lossy media."""
    if getattr(mesh, "medium", None) is not None:
        n = mesh.n if abs(causal_index(mesh.medium[1], mesh.medium[2]).imag) > 0 else 0
    elif getattr(mesh, "regions", None) is not None:
        n = 0
        for rid, (eps, mu) in mesh.region_media.items():
            if rid > 0 and abs(causal_index(eps, mu).imag) > 0:
                n = max(n, int(np.count_nonzero(np.any(np.asarray(mesh.regions) == rid, axis=1))))
    else:
        n = 0
    return 0 if n == 0 else 128 * n * n + NEAR_PER_ELEMENT * n * 160 + 64 * n


def lossy_workspace(n: int, device) -> torch.Tensor:
    nbytes = int(native.lib().grim_assemble_operators_lossy_workspace_bytes(n, NEAR_PER_ELEMENT))
    if nbytes < 0:
        raise ValueError("Lossy media are evaluated for contours of at most 65535 panels.")
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def assemble_lossy(contour: DeviceContour, k: complex, s=None, kp=None, d=None, w=None, ld: int | None = None,
                   mass_k: float = 0.0, mass_d: float = 0.0, coef_k=None, coef_d=None,
                   workspace: torch.Tensor | None = None) -> None:
    """S, K' + c_k M, D + c_d M and W of a contour set at complex k (pointers into column-major
    outputs with leading dimension ld; any may be None) from one traversal."""
    n, dev = contour.n, contour.device
    ws = workspace if workspace is not None else lossy_workspace(n, dev)
    st = _stream(dev)
    lib = native.lib()
    code = lib.grim_assemble_operators_lossy(
        n, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(), contour.prev.data_ptr(),
        float(k.real), float(k.imag), contour.mesh.far_order(k), s, kp, d, w, n if ld is None else ld,
        float(mass_k), float(mass_d), native.ptr(coef_k), native.ptr(coef_d), ws.data_ptr(), ws.numel(),
        NEAR_PER_ELEMENT, st)
    native.check("grim_assemble_operators_lossy", code)
    native.check("grim_assemble_operators_lossy (device)", lib.grim_workspace_status(ws.data_ptr(), st))


def dielectric_wavenumber(mesh: ContourMesh, k0: float):
    """Interior wavenumber k0 n of a homogeneous dielectric body (complex for a lossy medium)."""
    flag, eps, mu = mesh.medium
    return medium_wavenumber(causal_index(eps, mu), k0, f"Dielectric flag {flag} (eps={eps!r}, mu={mu!r})")


def dielectric_systems(contour: DeviceContour, k0: float, workspace: torch.Tensor) -> torch.Tensor:
    """[TE, TM] two-density systems (2, 2N, 2N) of a homogeneous dielectric body in air
    (twod/formulations/dielectric.py:17-81):
        [ eta Z_TM(k0)    -S(k1)            ]
        [ -eta Z_TE(k0)   factor (K'(k1) + M/2) ]
    with eta = j k0 (so that gamma = 1/eta in the PEC pair of k0), k1 = k0 n and factor =
    1/eps (TE) or 1/mu (TM).  The four blocks are assembled in place (grim_assemble at k0 into the
    left block column, grim_assemble_operators at k1 into the right one), the system is copied for the
    second polarization, and grim_scale_blocks applies each polarization's block factors."""
    flag, eps, mu = contour.mesh.medium
    k1 = dielectric_wavenumber(contour.mesh, k0)
    n, dev = contour.n, contour.device
    ld = 2 * n
    pair = torch.empty((2, ld, ld), dtype=torch.complex128, device=dev)
    base = pair[0].data_ptr()

    def block(r, c):          # column-major: entry (r, c) at r + c ld
        return base + 16 * (r + c * ld)

    st = _stream(dev)
    lib = native.lib()
    args = (n, contour.p0.data_ptr(), contour.p1.data_ptr(), contour.next.data_ptr(), contour.prev.data_ptr(),
            contour.fix_rows.data_ptr(), int(contour.fix_rows.numel()))
    native.check("grim_assemble", lib.grim_assemble(*args, float(k0), contour.mesh.far_order(k0), block(0, 0), ld,
                                                     block(n, 0), ld, workspace.data_ptr(), workspace.numel(),
                                                     NEAR_PER_ELEMENT, st))
    native.check("grim_assemble (device)", lib.grim_workspace_status(workspace.data_ptr(), st))
    if is_lossy(k1):
        assemble_lossy(contour, k1, s=block(0, n), kp=block(n, n), ld=ld, mass_k=0.5)
    else:
        native.check("grim_assemble_operators", lib.grim_assemble_operators(
            *args, float(k1), contour.mesh.far_order(k1), 1, block(0, n), ld, block(n, n), ld, 0.5, None,
            workspace.data_ptr(), workspace.numel(), NEAR_PER_ELEMENT, st))
        native.check("grim_assemble_operators (device)", lib.grim_workspace_status(workspace.data_ptr(), st))
    pair[1].copy_(pair[0])
    eta = 1j * float(k0)
    one = 1.0 + 0.0j
    for i, factor in enumerate((one / complex(eps), one / complex(mu))):
        scale = np.array([eta, -eta, -1.0, factor], dtype=np.complex128).view(np.float64)
        native.check("grim_scale_blocks", lib.grim_scale_blocks(n, pair[i].data_ptr(), ld,
                                                                scale.ctypes.data, st))
    return pair


def dielectric_rhs(bu: torch.Tensor, bdn: torch.Tensor) -> torch.Tensor:
    """Right-hand sides (2, A, 2N) [TE, TM] of dielectric_systems: [-b_u; b_dn]
    (twod/formulations/dielectric.py:84-91)."""
    both = torch.cat((-bu, bdn), dim=1)
    return torch.stack((both, both))


def radiating_density(contour: DeviceContour, k0: float, x: torch.Tensor) -> torch.Tensor:
    """The single-layer density whose matched far field (b_u + gamma b_dn weights, gamma = 1/(j k0))
    is the solution's: x for conductor contours; eta psi for a dielectric body, whose exterior
    field (D + eta S) psi has DLP density psi and SLP density eta psi (the reference's DLP plus
    second SLP potential, twod/solver.py:1672-1681)."""
    if contour.mesh.medium is None:
        return x
    return (1j * float(k0) * x[..., :contour.n]).contiguous()


# ----------------------------------------------------------------------------
# Multi-region layouts (twod/formulations/regions.py, combined_regions.py)
# ----------------------------------------------------------------------------
@dataclass
class RegionPlan:
    """Interfaces and unknowns of a multi-region layout (twod/formulations/regions.py:66-119):
    elements grouped by (minus, plus) region, interfaces in sorted key order, each interface's
    nodes in the reference's node order, and one block of unknowns per interface side that
    borders a region (minus first).  ``subs[rid]``: the region's elements (closed contours) and
    the map from mesh node to the region's local node."""
    ifaces: list
    region_ifaces: dict
    dofs: dict
    n_dof: int
    media: dict
    subs: dict


def region_plan(mesh: ContourMesh) -> RegionPlan:
    ref = np.arange(mesh.n) if mesh.ref_node is None else np.asarray(mesh.ref_node)
    groups = {}
    for e, (rm, rp) in enumerate(mesh.regions):
        groups.setdefault((int(rm), int(rp)), []).append(e)
    ifaces, region_ifaces, dofs, n_dof = [], {}, {}, 0
    for mi, ((rm, rp), eids) in enumerate(sorted(groups.items())):
        nodes = np.array(sorted(eids, key=lambda v: ref[v]), np.int64)   # element e starts at node e
        # one law per contour: node p's elements (p and prev[p]) share it
        zs = np.zeros(len(nodes), complex) if mesh.impedance is None else np.asarray(mesh.impedance)[nodes]
        ifaces.append({"r_m": rm, "r_p": rp, "eids": np.array(eids, np.int64), "nodes": nodes, "zs": zs})
        for side, rid in (("minus", rm), ("plus", rp)):
            if rid >= 0:
                region_ifaces.setdefault(rid, []).append(mi)
                dofs[mi, side] = (n_dof, len(nodes))
                n_dof += len(nodes)
    subs = {}
    for rid, members in region_ifaces.items():
        els = np.sort(np.concatenate([ifaces[mi]["eids"] for mi in members]))
        local = np.full(mesh.n, -1, np.int64)
        local[els] = np.arange(len(els))
        subs[rid] = {"elements": els, "local": local}
    return RegionPlan(ifaces, region_ifaces, dofs, n_dof, dict(mesh.region_media), subs)


def _region_wavenumber(media, rid, k0):
    eps, mu = media[rid]
    return medium_wavenumber(causal_index(eps, mu), k0, f"Region {rid} (eps={eps!r}, mu={mu!r})")


@dataclass
class Route:
    """dst[rows_dst, cols_dst] += weight * src[rows_src, cols_src] for one system (0 TE, 1 TM) or
    right-hand side.  src: (region, operator) with operator 0 S, 1 K' + c M, 2 D - c M, 3 W of the
    region's contours (local node numbers), or "bu" / "bdn" (mesh node numbers, all angles)."""
    pol: int
    rows_dst: np.ndarray
    cols_dst: np.ndarray | None
    src: tuple | str
    rows_src: np.ndarray
    cols_src: np.ndarray | None
    weight: complex


def region_mass_coefficients(plan: RegionPlan, rid: int) -> np.ndarray:
    """Per element of region rid's contours: -1/2 where the region borders the element's interface
    from its minus side, +1/2 from its plus side -- the mass of K' (regions.py:280-316: -M/2 + K' on
    the minus side, M/2 + K' on the plus side); D takes the opposite, its trace jump
    (combined_regions.py:139-141)."""
    local = plan.subs[rid]["local"]
    coef = np.empty(len(plan.subs[rid]["elements"]))
    for mi in plan.region_ifaces[rid]:
        ifc = plan.ifaces[mi]
        coef[local[ifc["eids"]]] = -0.5 if ifc["r_m"] == rid else 0.5
    return coef


def region_alpha(plan: RegionPlan, ifc: dict, pol: str, k_of) -> np.ndarray:
    """Robin coefficient of each node of a conductor interface, from its contour's surface impedance
    Zs and the region the interface borders (twod/geometry.py:2488-2521 with that region's
    eta = eta0 mu / n and k; twod/formulations/regions.py:80-89: negated where the region is on the
    plus side of a minus-side conductor); 0 on PEC nodes and on transmission interfaces."""
    zs = np.asarray(ifc.get("zs", np.zeros(len(ifc["nodes"]), complex)), complex)
    out = np.zeros(len(zs), complex)
    if ifc["r_m"] >= 0 and ifc["r_p"] >= 0:
        return out
    rid = ifc["r_p"] if ifc["r_m"] < 0 else ifc["r_m"]
    eps, mu = plan.media[rid]
    eta = ETA0 * complex(mu) / causal_index(eps, mu)
    k = k_of(rid)
    sign = -1.0 if ifc["r_m"] < 0 else 1.0
    for j, z in enumerate(zs):
        if abs(z) > EPS:
            out[j] = sign * (1j * k * eta / z if pol == "TM" else 1j * k * z / eta)
    return out


def _alpha_groups(alpha: np.ndarray):
    """(alpha value, node positions) per distinct value, in first-appearance order."""
    groups: dict = {}
    for j, a in enumerate(alpha):
        groups.setdefault(complex(a), []).append(j)
    return [(a, np.array(idx, np.int64)) for a, idx in groups.items()]


def region_routes(plan: RegionPlan, k_of) -> list:
    """The routes that compose both polarizations' regional systems (regions.py:239-333 base rows,
    combined_regions.py:43-62,116-146 D and W routes; ``k_of(rid)`` is the region's wavenumber).
    A conductor interface with a surface impedance (one law over the interface) carries the flux
    (Robin) row in both polarizations: K' + c M + alpha S, then -gamma W + gamma alpha (D - c M)."""
    routes = []

    def cols(mj, rid):
        src = plan.ifaces[mj]
        start, count = plan.dofs[mj, "minus" if src["r_m"] == rid else "plus"]
        return src["nodes"], np.arange(start, start + count)

    def inv_beta(ifc, pol):
        key = 1 if pol == "TM" else 0
        return complex(plan.media[ifc["r_m"]][key]) / complex(plan.media[ifc["r_p"]][key])

    def rows_of(mi, side):
        start, count = plan.dofs[mi, side]
        return np.arange(start, start + count)

    for i, pol in enumerate(("TE", "TM")):
        for mi, ifc in enumerate(plan.ifaces):
            rm, rp, nodes = ifc["r_m"], ifc["r_p"], ifc["nodes"]
            if rm < 0 or rp < 0:
                rid = rp if rm < 0 else rm
                local = plan.subs[rid]["local"]
                rows_all = rows_of(mi, "plus" if rm < 0 else "minus")
                for alpha, idx in _alpha_groups(region_alpha(plan, ifc, pol, k_of)):
                    rows, obs_nodes = rows_all[idx], local[nodes[idx]]
                    for mj in plan.region_ifaces[rid]:
                        src_nodes, cc = cols(mj, rid)
                        if alpha == 0:
                            routes.append(Route(i, rows, cc, (rid, 0 if pol == "TM" else 1), obs_nodes,
                                                local[src_nodes], 1.0))
                        else:
                            routes.append(Route(i, rows, cc, (rid, 1), obs_nodes, local[src_nodes], 1.0))
                            routes.append(Route(i, rows, cc, (rid, 0), obs_nodes, local[src_nodes], alpha))
            else:
                rs, rt = rows_of(mi, "minus"), rows_of(mi, "plus")
                for rid, sign, flux in ((rm, 1.0, 1.0), (rp, -1.0, inv_beta(ifc, pol))):
                    local = plan.subs[rid]["local"]
                    for mj in plan.region_ifaces[rid]:
                        src_nodes, cc = cols(mj, rid)
                        routes.append(Route(i, rs, cc, (rid, 1), local[nodes], local[src_nodes], sign * flux))
                        routes.append(Route(i, rt, cc, (rid, 0), local[nodes], local[src_nodes], sign))
        for rid, members in plan.region_ifaces.items():
            local = plan.subs[rid]["local"]
            for mi in members:
                side = "minus" if plan.ifaces[mi]["r_m"] == rid else "plus"
                gamma = (1.0 if side == "minus" else -1.0) / (1j * k_of(rid))
                src_nodes, cc = cols(mi, rid)
                for obs in members:
                    tgt = plan.ifaces[obs]
                    o_rows = local[tgt["nodes"]]
                    if tgt["r_m"] < 0 or tgt["r_p"] < 0:
                        rows_all = rows_of(obs, "plus" if tgt["r_m"] < 0 else "minus")
                        for alpha, idx in _alpha_groups(region_alpha(plan, tgt, pol, k_of)):
                            rows, orows = rows_all[idx], o_rows[idx]
                            if alpha != 0:
                                routes.append(Route(i, rows, cc, (rid, 3), orows, local[src_nodes], -gamma))
                                routes.append(Route(i, rows, cc, (rid, 2), orows, local[src_nodes], gamma * alpha))
                            elif pol == "TM":
                                routes.append(Route(i, rows, cc, (rid, 2), orows, local[src_nodes], gamma))
                            else:
                                routes.append(Route(i, rows, cc, (rid, 3), orows, local[src_nodes], -gamma))
                    else:
                        own = rid == tgt["r_m"]
                        routes.append(Route(i, rows_of(obs, "plus"), cc, (rid, 2), o_rows, local[src_nodes],
                                            gamma * (1.0 if own else -1.0)))
                        routes.append(Route(i, rows_of(obs, "minus"), cc, (rid, 3), o_rows, local[src_nodes],
                                            -gamma * (1.0 if own else -inv_beta(tgt, pol))))
    return routes


def region_load_routes(plan: RegionPlan, k_of=None) -> list:
    """Routes of both polarizations' right-hand sides (regions.py:557-590): on interfaces that
    border air, -b_dn (TE) or -b_u (TM) on conductor rows (-(b_dn + alpha b_u) on the Robin rows of
    an impedance conductor, which need ``k_of``); transmission rows with air on their minus side get
    -b_dn (flux) and -b_u (trace), with air on their plus side +(1/beta) b_dn and +b_u."""
    routes = []
    for i, pol in enumerate(("TE", "TM")):
        key = 1 if pol == "TM" else 0
        for mi, ifc in enumerate(plan.ifaces):
            rm, rp, nodes = ifc["r_m"], ifc["r_p"], ifc["nodes"]
            if rm < 0 or rp < 0:
                if (rp if rm < 0 else rm) == 0:
                    start, count = plan.dofs[mi, "plus" if rm < 0 else "minus"]
                    rows = np.arange(start, start + count)
                    zs = np.asarray(ifc.get("zs", np.zeros(count, complex)), complex)
                    if np.any(np.abs(zs) > EPS):
                        if k_of is None:
                            raise ValueError("The loads of an impedance conductor need the region wavenumbers.")
                        for alpha, idx in _alpha_groups(region_alpha(plan, ifc, pol, k_of)):
                            if alpha == 0:
                                routes.append(Route(i, rows[idx], None, "bu" if pol == "TM" else "bdn", nodes[idx],
                                                    None, -1.0))
                            else:
                                routes.append(Route(i, rows[idx], None, "bdn", nodes[idx], None, -1.0))
                                routes.append(Route(i, rows[idx], None, "bu", nodes[idx], None, -alpha))
                    else:
                        routes.append(Route(i, rows, None, "bu" if pol == "TM" else "bdn", nodes, None, -1.0))
                continue
            fl, tr = plan.dofs[mi, "minus"][0], plan.dofs[mi, "plus"][0]
            rs, rt = np.arange(fl, fl + len(nodes)), np.arange(tr, tr + len(nodes))
            if rm == 0:
                routes.append(Route(i, rs, None, "bdn", nodes, None, -1.0))
                routes.append(Route(i, rt, None, "bu", nodes, None, -1.0))
            if rp == 0:
                inv = complex(plan.media[rm][key]) / complex(plan.media[rp][key])
                routes.append(Route(i, rs, None, "bdn", nodes, None, inv))
                routes.append(Route(i, rt, None, "bu", nodes, None, 1.0))
    return routes


class _Router:
    """grim_route launches on torch's current stream.  The index and weight tensors they read are
    freed into torch's stream-ordered cache, so later work on that stream cannot overwrite them
    before the launches have read them."""

    def __init__(self, dev):
        self.dev, self.st = dev, _stream(dev)

    def __call__(self, dst, ld_dst, rows_dst, cols_dst, src, ld_src, rows_src, cols_src, weight):
        m, n = len(rows_src), len(cols_src)
        w = torch.as_tensor(np.full(m, complex(weight), np.complex128), device=self.dev)
        ts = [torch.as_tensor(np.ascontiguousarray(v, np.int32), device=self.dev)
              for v in (rows_src, cols_src, rows_dst, cols_dst)]
        code = native.lib().grim_route(m, n, src.data_ptr(), ld_src, ts[0].data_ptr(), ts[1].data_ptr(),
                                       dst.data_ptr(), ld_dst, ts[2].data_ptr(), ts[3].data_ptr(), w.data_ptr(),
                                       self.st)
        native.check("grim_route", code)


def region_systems(contour: DeviceContour, k0: float, workspace: torch.Tensor, plan: RegionPlan) -> torch.Tensor:
    """[TE, TM] systems (2, DOFs, DOFs) of a multi-region layout without surface impedances: per
    region, S, K' + c M, D - c M and W of the region's contours at its wavenumber
    (grim_assemble_operators, c from region_mass_coefficients), then the routes of region_routes
    (grim_route)."""
    mesh, dev = contour.mesh, contour.device
    nd = plan.n_dof
    pair = torch.zeros((2, nd, nd), dtype=torch.complex128, device=dev)
    route = _Router(dev)
    lib = native.lib()
    ops, ks = {}, {}
    for rid, sub in plan.subs.items():
        k = ks[rid] = _region_wavenumber(plan.media, rid, k0)
        els, local = sub["elements"], sub["local"]
        nr = len(els)
        rmesh = ContourMesh(mesh.p0[els], mesh.p0[els], mesh.p1[els], local[mesh.next[els]].astype(np.int32),
                            local[mesh.prev[els]].astype(np.int32))
        rc = DeviceContour(rmesh, dev)
        coef = region_mass_coefficients(plan, rid)
        c_k, c_d = torch.as_tensor(coef, device=dev), torch.as_tensor(-coef, device=dev)
        mats = torch.empty((4, nr, nr), dtype=torch.complex128, device=dev)     # S, K', D, W
        if is_lossy(k):         # on the current stream, as the routes
            assemble_lossy(rc, k, *(m.data_ptr() for m in mats), ld=nr, coef_k=c_k, coef_d=c_d)
            ops[rid] = mats
            continue
        args = (nr, rc.p0.data_ptr(), rc.p1.data_ptr(), rc.next.data_ptr(), rc.prev.data_ptr(),
                rc.fix_rows.data_ptr(), int(rc.fix_rows.numel()), float(k), rmesh.far_order(k))
        for which, a, b, c in ((1, mats[0], mats[1], c_k), (2, mats[2], mats[3], c_d)):
            # the workspace is sized for the whole mesh, so for any region's contours
            code = lib.grim_assemble_operators(*args, which, a.data_ptr(), nr, b.data_ptr(), nr, 0.0, c.data_ptr(),
                                               workspace.data_ptr(), workspace.numel(), NEAR_PER_ELEMENT, route.st)
            native.check("grim_assemble_operators", code)
            native.check("grim_assemble_operators (device)",
                         lib.grim_workspace_status(workspace.data_ptr(), route.st))
        ops[rid] = mats
    for r in region_routes(plan, ks.__getitem__):
        rid, op = r.src
        route(pair[r.pol], nd, r.rows_dst, r.cols_dst, ops[rid][op], len(plan.subs[rid]["elements"]), r.rows_src,
              r.cols_src, r.weight)
    return pair


def region_rhs(plan: RegionPlan, bu: torch.Tensor, bdn: torch.Tensor, k0: float | None = None) -> torch.Tensor:
    """Right-hand sides (2, A, DOFs) [TE, TM] of region_systems (routes of region_load_routes; k0, the
    air wavenumber, for the Robin loads of impedance conductors)."""
    dev = bu.device
    na, n = bu.shape
    out = torch.zeros((2, na, plan.n_dof), dtype=torch.complex128, device=dev)
    route = _Router(dev)
    acols = np.arange(na)
    loads = {"bu": bu, "bdn": bdn}
    k_of = None if k0 is None else (lambda rid: _region_wavenumber(plan.media, rid, k0))
    for r in region_load_routes(plan, k_of):
        route(out[r.pol], plan.n_dof, r.rows_dst, acols, loads[r.src], n, r.rows_src, acols, r.weight)
    return out


def region_exterior_density(plan: RegionPlan, n: int, x: torch.Tensor) -> torch.Tensor:
    """Nodal single-layer density (A, N) of the air-side unknowns of a solution x (A, DOFs), zero
    on nodes that do not border air (twod/formulations/regions.py:593-609)."""
    dev = x.device
    na = x.shape[0]
    out = torch.zeros((na, n), dtype=torch.complex128, device=dev)
    route = _Router(dev)
    acols = np.arange(na)
    xc = x.contiguous()
    for mi, ifc in enumerate(plan.ifaces):
        for side, rid in (("minus", ifc["r_m"]), ("plus", ifc["r_p"])):
            if rid == 0:
                start, count = plan.dofs[mi, side]
                route(out, n, ifc["nodes"], acols, xc, plan.n_dof, np.arange(start, start + count), acols, 1.0)
    return out


def bistatic_farfield(x: torch.Tensor, bu: torch.Tensor, bdn: torch.Tensor, k: float):
    """Bistatic amplitudes B[b, i, s] = sum_p x[b, i, p] (bu[s, p] + gamma bdn[s, p]), gamma = 1/(j k),
    for densities x (B, A_inc, N) and observation-direction loads (A_obs, N); with s = i this is
    the matched monostatic projection.  With bdn None, B[b, i, s] = sum_p x[b, i, p] bu[s, p]: the
    far field of the one potential whose loads bu holds.  One complex GEMM per batch on the
    library's FP64 MFMA kernel (the observation operator is shared: stride 0).  Returns
    (amplitude, sigma_2D), both (B, A_inc, A_obs)."""
    xb = _as_batch(x).contiguous()
    nb, ninc, n = xb.shape
    nobs = bu.shape[0]
    proj = bu if bdn is None else bu + bdn * complex(0.0, -1.0 / float(k))
    proj_t = proj.transpose(0, 1).contiguous()   # (N, A_obs)
    c = torch.zeros((nb, ninc, nobs), dtype=torch.complex128, device=x.device)
    code = native.lib().grim_gemm_minus(nb, nobs, ninc, n, proj_t.data_ptr(), nobs, 0, xb.data_ptr(), n, ninc * n,
                                        c.data_ptr(), nobs, ninc * nobs, 64, _stream(x.device))
    native.check("grim_gemm_minus", code)
    amp = -c
    sigma = (amp.real * amp.real + amp.imag * amp.imag) / (4.0 * float(k))
    return amp, sigma


@dataclass
class Factor:
    """LU factors of a batch of equally sized systems (column-major, batch first)."""
    precision: str              # "double" or "mixed"
    lu: torch.Tensor            # (B, N, N) complex128 or complex64
    piv: torch.Tensor           # (B, N) int32
    aux: torch.Tensor           # (B, aux_bytes) uint8: row permutation (int32 x N) + factorization scratch
    z: torch.Tensor | None      # (B, N, N) complex128 systems (for residuals)
    n: int
    batch: int
    fallback_reason: str = ""   # why a mixed factorization was replaced by the FP64 one
    parts: list | None = None   # per-system factors once some systems of a batch fell back
    norm_inf: torch.Tensor | None = None   # (B,) ||Z||_inf of the systems, once computed

    @property
    def prec(self) -> int:
        return 64 if self.precision == "double" else 32


def _as_batch(t: torch.Tensor) -> torch.Tensor:
    return t.unsqueeze(0) if t.dim() == 2 else t


def factor(z: torch.Tensor, precision: str = "double", keep_matrix: bool = True,
           status: torch.Tensor | None = None, check: bool = True) -> Factor:
    """Factor one system (N, N) or a batch (B, N, N) of complex128 systems."""
    zb = _as_batch(z)
    if not zb.is_contiguous():
        raise ValueError("systems must be contiguous")
    b, n = zb.shape[0], zb.shape[-1]
    dev = z.device
    st = _stream(dev)
    if precision == "double":
        lu = zb.clone()
        prec = 64
    elif precision == "mixed":
        lu = torch.empty((b, n, n), dtype=torch.complex64, device=dev)
        native.check("grim_to_c64", native.lib().grim_to_c64(b, n, n, zb.data_ptr(), n, n * n, lu.data_ptr(), n,
                                                             n * n, st))
        prec = 32
    else:
        raise ValueError("precision must be 'double' or 'mixed'")
    piv = torch.empty((b, n), dtype=torch.int32, device=dev)
    aux_bytes = int(native.lib().grim_lu_aux_bytes(n, prec))
    aux = torch.empty((b, aux_bytes), dtype=torch.uint8, device=dev)
    word = status if status is not None else torch.zeros(64, dtype=torch.uint8, device=dev)
    native.check("grim_lu_factor", native.lib().grim_lu_factor(b, n, lu.data_ptr(), n, n * n, prec, piv.data_ptr(),
                                                               n, aux.data_ptr(), aux_bytes, word.data_ptr(), st))
    if check:
        code = native.lib().grim_workspace_status(word.data_ptr(), st)
        if code == 8 and prec == 32 and b > 1:
            # only the systems whose complex64 LU is singular fall back (linalg/dense.py:86-91)
            del lu
            parts = [factor(zb[i:i + 1], "mixed", True, status, True) for i in range(b)]
            return Factor(precision, None, None, None, zb, n, b,
                          next((p.fallback_reason for p in parts if p.fallback_reason), ""), parts=parts)
        if code == 8 and prec == 32:
            del lu
            f = factor(zb, "double", keep_matrix, status, True)
            f.fallback_reason = MIXED_FALLBACK + "zero pivot in the single-precision LU."
            return f
        if code == 8:
            raise np.linalg.LinAlgError("zero pivot in the LU factorization")
        native.check("grim_lu_factor (device)", code)
    return Factor(precision, lu, piv, aux, zb if keep_matrix else None, n, b)


MIXED_FALLBACK = "Mixed precision fell back to double LU: "   # linalg/dense.py:90, 229


def system(f: Factor, i: int) -> Factor:
    """Factor of system i of a batch (views, no copies)."""
    if f.parts is not None:
        return f.parts[i]
    return Factor(f.precision, f.lu[i:i + 1], f.piv[i:i + 1], f.aux[i:i + 1],
                  None if f.z is None else f.z[i:i + 1], f.n, 1, f.fallback_reason)


def systems(f: Factor, i0: int, i1: int) -> Factor:
    """Factor of systems i0 .. i1-1 of a batch (views, no copies; they keep the whole batch's
    storage alive)."""
    if f.parts is not None:
        parts = f.parts[i0:i1]
        reason = next((p.fallback_reason for p in parts if p.fallback_reason), "")
        return Factor(f.precision, None, None, None, None if f.z is None else f.z[i0:i1], f.n, i1 - i0, reason,
                      parts=parts)
    return Factor(f.precision, f.lu[i0:i1], f.piv[i0:i1], f.aux[i0:i1], None if f.z is None else f.z[i0:i1],
                  f.n, i1 - i0, f.fallback_reason,
                  norm_inf=None if f.norm_inf is None else f.norm_inf[i0:i1])


def _solve_parts(parts, bb: torch.Tensor, gate: bool):
    """Solve each system with its own factor and merge the results into one batch."""
    xs, infos = [], []
    for i, part in enumerate(parts):
        xi, inf = solve(part, bb[i:i + 1], gate)
        xs.append(xi)
        infos.append(inf)
    new_parts = [inf.pop("factor", part) for inf, part in zip(infos, parts)]
    info = {"refinement_steps": max(inf.get("refinement_steps", 0) for inf in infos),
            "residual": torch.cat([inf["residual"] for inf in infos]),
            "precisions": [p.precision for p in new_parts]}
    reasons = [p.fallback_reason for p in new_parts if p.fallback_reason]
    if reasons:
        info["fallback_reason"] = reasons[0]
        info["fallback_reasons"] = reasons
    errs = [inf["refinement_error"] for inf in infos if "refinement_error" in inf]
    if errs:
        info["refinement_error"] = max(errs)
    # per-system statements, one entry per system of the batch
    for name, fallback in (("backward_errors", "backward_error"), ("refinement_errors", "refinement_error"),
                           ("steps", "refinement_steps")):
        vals = [inf[name][0] if name in inf else inf.get(fallback) for inf in infos]
        if any(v is not None for v in vals):
            info[name] = vals
    bes = [v for v in info.get("backward_errors", []) if v is not None]
    if bes:
        info["backward_error"] = max(bes)
    if any(a is not b for a, b in zip(new_parts, parts)):
        head = parts[0]
        info["factor"] = Factor("mixed", None, None, None, torch.cat([p.z for p in new_parts]), head.n,
                                len(new_parts), reasons[0] if reasons else "", parts=new_parts)
    return torch.cat(xs), info


def lu_solve_inplace(f: Factor, x: torch.Tensor) -> torch.Tensor:
    """x <- A^{-1} x for x of shape (B, A, N) in the factor's precision dtype."""
    b, na, n = x.shape
    st = _stream(x.device)
    nbytes = int(native.lib().grim_lu_solve_workspace_bytes(b, n, na, f.prec))
    work = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    native.check("grim_lu_solve", native.lib().grim_lu_solve(b, n, na, f.lu.data_ptr(), n, n * n, f.prec,
                                                             f.aux.data_ptr(), f.aux.shape[1], x.data_ptr(), n,
                                                             na * n, work.data_ptr(), nbytes, st))
    return x


def residual(z: torch.Tensor, x: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """r = b - Z x in FP64 (complex128); batched shapes (B, N, N), (B, A, N)."""
    zb, xb, bb = _as_batch(z), _as_batch(x), _as_batch(b)
    nb, na, n = bb.shape
    r = bb.clone()
    native.check("grim_gemm_minus", native.lib().grim_gemm_minus(nb, n, na, n, zb.data_ptr(), n, n * n,
                                                                 xb.data_ptr(), n, na * n, r.data_ptr(), n, na * n,
                                                                 64, _stream(z.device)))
    return r if b.dim() == 3 else r[0]


def solve(f: Factor, b: torch.Tensor, gate: bool = True):
    """x with Z x = b for every column: b is (A, N) (batch 1) or (B, A, N); returns (x, info)."""
    bb = _as_batch(b).contiguous()
    nb, na, n = bb.shape
    dev = b.device
    st = _stream(dev)
    info = {"refinement_steps": 0}
    if f.parts is not None:
        x, info = _solve_parts(f.parts, bb, gate)
        return (x if b.dim() == 3 else x[0]), info
    if f.precision == "double":
        x = lu_solve_inplace(f, bb.clone())
        if gate and f.z is not None:
            x, info = _backward_error_gate(f, x, bb)
        if f.fallback_reason:
            info["fallback_reason"] = f.fallback_reason
        return (x if b.dim() == 3 else x[0]), info
    x = torch.empty_like(bb)
    nbytes = native.lib().grim_refine_workspace_bytes(nb, n, na)
    ws = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
    steps = ctypes_int()
    err = ctypes_double()
    code = native.lib().grim_refined_solve(nb, n, na, f.z.data_ptr(), n, n * n, f.lu.data_ptr(), n, n * n,
                                           f.aux.data_ptr(), f.aux.shape[1], bb.data_ptr(), n, na * n,
                                           x.data_ptr(), n, na * n, REFINE_MAX_STEPS, REFINE_TOL, ws.data_ptr(),
                                           ws.numel(), steps, err, st)
    info["refinement_steps"] = int(steps.value)
    info["refinement_error"] = float(err.value)
    if code == 6:
        # RefinedLU raised; DenseFactor refactors in complex128 (linalg/dense.py:223-230).  Each
        # system of a batch falls back on its own, so a batched call still returns what
        # separate calls return.
        if nb > 1:
            x, info = _solve_parts([system(f, i) for i in range(nb)], bb, gate)
            return (x if b.dim() == 3 else x[0]), info
        why = ("Mixed-precision refinement exceeded its correction limit." if steps.value >= REFINE_MAX_STEPS
               else "Mixed-precision refinement stalled; double LU required.")
        fb = factor(f.z, "double")
        fb.fallback_reason = MIXED_FALLBACK + why
        x, info = solve(fb, bb, True)
        info["fallback_reason"] = fb.fallback_reason
        info["factor"] = fb
        return (x if b.dim() == 3 else x[0]), info
    native.check("grim_refined_solve", code)
    # the workspace starts with the final residual b - Z x of the returned x
    info["residual"] = ws[: 16 * nb * na * n].view(torch.complex128).view(nb, na, n)
    # the FP64 gate's acceptance metric for the refined answer (linalg/dense.py:257-268), from
    # that residual: a per-result statement of its accuracy next to the refinement error
    if f.norm_inf is None:
        f.norm_inf = norm_inf(f.z)
    errs = backward_errors(f.norm_inf, x, bb, info["residual"])
    info["backward_error"] = float(errs.max())
    info["backward_errors"] = [float(e) for e in errs]
    # the refinement's own measure per system: max over columns of max|r| / max|b| (max|b| = 0
    # counted as 1, as the refinement does)
    bmax = torch.amax(torch.abs(bb), dim=2)
    rmax = torch.amax(torch.abs(info["residual"]), dim=2)
    info["refinement_errors"] = [float(e) for e in
                                 torch.amax(rmax / torch.where(bmax > 0, bmax, torch.ones_like(bmax)), dim=1).cpu()]
    return (x if b.dim() == 3 else x[0]), info


def backward_errors(a_inf: torch.Tensor, x: torch.Tensor, b: torch.Tensor, r: torch.Tensor) -> np.ndarray:
    """Per system of a batch (B, A, N): max over columns of max|r| / (||Z||_inf max|x| + max|b|); a
    column with a zero denominator (b = 0, x = 0) counts 0, or infinity if its residual is not zero
    (linalg/dense.py:257-263)."""
    ri = torch.amax(torch.abs(r), dim=2)
    den = a_inf[:, None] * torch.amax(torch.abs(x), dim=2) + torch.amax(torch.abs(b), dim=2)
    pos = den > 0
    ratio = torch.where(pos, ri / torch.where(pos, den, torch.ones_like(den)),
                        torch.where(ri > 0, torch.full_like(ri, math.inf), torch.zeros_like(ri)))
    return torch.amax(ratio, dim=1).cpu().numpy()


def backward_error(a_inf: torch.Tensor, x: torch.Tensor, b: torch.Tensor, r: torch.Tensor) -> float:
    """max over systems and columns of max|r| / (||Z||_inf max|x| + max|b|), batched (B, A, N)."""
    return float(np.max(backward_errors(a_inf, x, b, r)))


def ctypes_int():
    import ctypes
    return ctypes.c_int(0)


def ctypes_double():
    import ctypes
    return ctypes.c_double(0.0)


def _backward_error_gate(f: Factor, x: torch.Tensor, b: torch.Tensor):
    """Normwise backward error <= 1e-12 with at most two corrections (linalg/dense.py:257-280).

    Each system of a batch is gated on its own, as the reference gates each matrix: a system
    within the limit takes no correction, a correction is kept only if it lowers that system's
    error, and its corrections run as batch-1 calls.  So a batch of several requests' systems
    returns, system by system, what separate calls return."""
    z = f.z
    if f.norm_inf is None:
        f.norm_inf = norm_inf(z)
    a_inf = f.norm_inf                                                  # (B,)
    r = residual(z, x, b)
    errs = backward_errors(a_inf, x, b, r)
    nb = errs.shape[0]
    steps = np.zeros(nb, dtype=np.int64)
    for i in range(nb):
        for _ in range(2):
            if errs[i] <= BACKWARD_ERROR_MAX:
                break
            fi = system(f, i)
            cand = x[i:i + 1] + lu_solve_inplace(fi, r[i:i + 1].clone())
            r2 = residual(z[i:i + 1], cand, b[i:i + 1])
            err2 = float(backward_errors(a_inf[i:i + 1], cand, b[i:i + 1], r2)[0])
            if err2 >= errs[i]:
                break
            x[i:i + 1] = cand
            r[i:i + 1] = r2
            errs[i] = err2
            steps[i] += 1
    bad = [float(e) for e in errs if not (math.isfinite(float(e)) and e <= BACKWARD_ERROR_MAX)]
    if bad:
        raise RuntimeError(f"normwise backward error {bad[0]} exceeds the release limit {BACKWARD_ERROR_MAX}.")
    return x, {"refinement_steps": int(steps.max()), "backward_error": float(errs.max()),
               "backward_errors": [float(e) for e in errs], "steps": [int(v) for v in steps], "residual": r}


def norm_inf(z: torch.Tensor) -> torch.Tensor:
    """||Z||_inf (max row sum of moduli) of each system of a batch (B, N, N), on the device."""
    zb = _as_batch(z)
    nb, n = zb.shape[0], zb.shape[-1]
    lib = native.lib()
    nbytes = int(lib.grim_norm_inf_workspace_bytes(nb, n))
    work = torch.empty(nbytes, dtype=torch.uint8, device=z.device)
    out = torch.empty(nb, dtype=torch.float64, device=z.device)
    native.check("grim_norm_inf", lib.grim_norm_inf(nb, n, zb.data_ptr(), n, n * n, work.data_ptr(), nbytes,
                                                    out.data_ptr(), _stream(z.device)))
    return out


def relative_residual(z: torch.Tensor, x: torch.Tensor, b: torch.Tensor, r: torch.Tensor | None = None):
    """||b - Z x||_2 / ||b||_2 per column (the reference's linear_residual); batched or not."""
    if r is None:
        r = residual(z, x, b)
    nr = torch.linalg.vector_norm(r, dim=-1)
    nb = torch.linalg.vector_norm(b, dim=-1)
    return nr / torch.where(nb <= 1e-12, torch.ones_like(nb), nb)
