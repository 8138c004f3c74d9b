"""CPU ORACLE -- test infrastructure only, never part of the product path.

The reference's own native far-pair quadrature (``ghost_far_block``,
tools/GHOST/ghost_backend/twod/assembly/native/far.c:100-257), compiled from its
source into ``oracle/_ref/libghost_far.so`` by ``oracle/Makefile``, driven the
way the reference drives it: a piecewise power-basis kernel table
(twod/assembly/kernels.py:312-394, restated in ``kernel_table``), element tiles
scheduled over a thread pool (twod/operators.py:1659-1711, ctypes releases the
GIL), far pairs only.  Used by ``bench.py``'s CPU baseline to time a full
request with the reference's fast far pass, and by the CPU tests to check the
oracle's far blocks against the reference's C code.
"""
from __future__ import annotations

import ctypes
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
from numpy.polynomial import Chebyshev, Polynomial
from scipy.fft import dct
from scipy.special import hankel2

from oracle import bem2d

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libghost_far.so")
TABLE_START = 5e-13          # kernels.py:325-334
TABLE_DEGREE = 12
TABLE_INTERVALS = 4096

_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB)
        d, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        lib.ghost_far_block.restype = i32
        lib.ghost_far_block.argtypes = [i64, i64, i32, i32, d, d, d, d, d, d, d, i32, i32, i32, i32,
                                        ctypes.c_double, ctypes.c_double, i32, d, i32, d, d, d, d]
        _lib = lib
    return _lib


def kernel_values(k: float, r):
    """(j/4) H0^(2)(kr) and (j/4) k H1^(2)(kr), stacked on the last axis (kernels.py:312-313)."""
    return np.stack((0.25j * hankel2(0, k * r), 0.25j * k * hankel2(1, k * r)), axis=-1)


def kernel_table(k: float, upper: float, degree: int = TABLE_DEGREE):
    """Interval bounds and coefficients coeff[interval][power][channel] (complex) of the
    piecewise polynomial in the normalized interval coordinate: intervals grow from 5e-13 m by
    x1.2 or +0.6/|k| (the smaller step) up to ``upper``; on each, the degree-12 Chebyshev
    interpolant at the first-kind nodes, converted to ascending powers."""
    bounds = [TABLE_START]
    while bounds[-1] < upper:
        bounds.append(min(upper, bounds[-1] * 1.2, bounds[-1] + 0.6 / abs(k)))
        if len(bounds) > TABLE_INTERVALS + 1:
            raise ValueError("kernel table interval budget exceeded")
    bounds = np.array(bounds)
    lo = bounds[:-1, None]
    width = np.diff(bounds)[:, None]
    n = degree + 1
    x = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    cheb = dct(kernel_values(k, lo + (x + 1) * 0.5 * width), type=2, axis=1) / n
    cheb[:, 0] *= 0.5
    conversion = np.zeros((n, n))
    for j in range(n):
        co = Chebyshev.basis(j, domain=[0, 1]).convert(kind=Polynomial).coef
        conversion[:len(co), j] = co
    power = np.einsum("pj,bjc->bpc", conversion, cheb)
    return bounds, np.ascontiguousarray(power)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def far_tile(mesh, k, table, rows, cols, q, w, phi, pts):
    """S, K' (observation-normal derivative) and D (source-normal derivative) blocks, each
    (2, 2, R, C), of element rows ``rows`` against element columns ``cols``, far pairs only."""
    lib = load()
    bounds, coeff = table
    mb, nb = len(rows), len(cols)
    far = np.ascontiguousarray(bem2d.far_mask(mesh, rows)[:, cols].astype(np.uint8))
    acc = [np.zeros((2, 2, mb, nb), np.complex128) for _ in range(3)]
    obs = np.ascontiguousarray(pts[rows])
    src = np.ascontiguousarray(pts[cols])
    onorm = np.ascontiguousarray(mesh.normal[rows])
    snorm = np.ascontiguousarray(mesh.normal[cols])
    code = lib.ghost_far_block(mb, nb, q, 2, _ptr(obs), _ptr(src), _ptr(w), _ptr(phi), _ptr(onorm), _ptr(snorm),
                               _ptr(far), 1, 1, 1, 1, float(k), 0.0, len(bounds) - 1, _ptr(bounds), TABLE_DEGREE,
                               _ptr(coeff), _ptr(acc[0]), _ptr(acc[1]), _ptr(acc[2]))
    if code != 0:
        raise RuntimeError(f"ghost_far_block returned {code}")
    scale = (mesh.length[rows][:, None] * mesh.length[cols][None, :])
    s, kp, dt = (a * scale for a in acc)
    return s, kp, np.swapaxes(dt, 0, 1)      # D[a][b] = acc_kt[b][a]


def far_matrices(mesh, k: float, threads: int | None = None, tile: int = 128):
    """The far-pair part of the TM and TE systems (no near pairs, no mass terms), assembled
    with the reference's far.c over a thread pool on symmetric tiles: each unordered pair of
    element tiles is evaluated once and also gives the transposed tile (S(f, e) = S(e, f)^T,
    K'(f, e) = D(e, f)^T, D(f, e) = K'(e, f)^T, W(f, e) = W(e, f)^T; twod/operators.py:2331-2332)."""
    threads = threads or len(os.sched_getaffinity(0))
    q = bem2d.far_order(mesh, k)
    t, w = bem2d.gauss(q)
    phi = np.ascontiguousarray(np.stack((1.0 - t, t), axis=1))
    pts = np.ascontiguousarray(mesh.p0[:, None, :] + t[None, :, None] * mesh.seg[:, None, :])
    centre = mesh.center
    radius = float(np.max(np.linalg.norm(centre - centre.mean(0), axis=1)))
    table = kernel_table(k, 1.01 * (2 * radius + 2 * float(np.max(mesh.length))))
    n = len(mesh.nodes)
    e = len(mesh.length)
    ids = mesh.elements
    gamma = 1.0 / (1j * k)
    z_tm = np.zeros((n, n), complex)
    z_te = np.zeros((n, n), complex)
    blocks = [np.arange(s, min(e, s + tile)) for s in range(0, e, tile)]
    pairs = [(i, j) for i in range(len(blocks)) for j in range(i, len(blocks))]

    def work(ij):
        rows, cols = blocks[ij[0]], blocks[ij[1]]
        return rows, cols, far_tile(mesh, k, table, rows, cols, q, w, phi, pts)

    def maue(rows, cols, s):
        nn = mesh.normal[rows] @ mesh.normal[cols].T
        lp = np.maximum(mesh.length[rows][:, None] * mesh.length[cols][None, :], bem2d.EPS * bem2d.EPS)
        ssum = s.sum(axis=(0, 1)) / lp
        return np.array([[-(k * k) * nn * s[a, b] + bem2d.TANGENT[a, b] * ssum for b in range(2)]
                         for a in range(2)])

    with ThreadPoolExecutor(threads) as pool:
        for rows, cols, (s, kp, d) in pool.map(work, pairs):
            wb = maue(rows, cols, s)
            mirror = rows[0] != cols[0]
            for a in range(2):
                for b in range(2):
                    r, c = ids[rows, a][:, None], ids[cols, b][None, :]
                    z_tm[r, c] += s[a, b] + gamma * d[a, b]
                    z_te[r, c] += kp[a, b] - gamma * wb[a, b]
                    if mirror:          # obs cols (local b), src rows (local a)
                        rt, ct = ids[cols, b][:, None], ids[rows, a][None, :]
                        z_tm[rt, ct] += (s[a, b] + gamma * kp[a, b]).T
                        z_te[rt, ct] += (d[a, b] - gamma * wb[a, b]).T
    return z_tm, z_te


def full_request(mesh, frequency_ghz: float, angles, threads: int | None = None):
    """One co-polarized request on the CPU: far pass with the reference's far.c, near pairs and
    mass terms with the oracle, LAPACK LU and solve of both systems, matched far field.
    Returns (amplitudes per polarization, seconds per stage)."""
    import time
    import scipy.linalg as sl
    k = bem2d.frequency_k(frequency_ghz)
    stages = {}
    t0 = time.perf_counter()
    z_tm, z_te = far_matrices(mesh, k, threads)
    stages["far"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    bem2d.add_near_and_mass(mesh, k, z_tm, z_te)
    stages["near"] = time.perf_counter() - t0
    out = {}
    for pol, z in (("TE", z_te), ("TM", z_tm)):
        t0 = time.perf_counter()
        lu = sl.lu_factor(z, overwrite_a=True, check_finite=False)
        stages["lu"] = stages.get("lu", 0.0) + time.perf_counter() - t0
        t0 = time.perf_counter()
        b = bem2d.rhs(mesh, k, angles, pol)
        x = sl.lu_solve(lu, b, check_finite=False)
        out[pol] = bem2d.farfield(mesh, k, angles, x)
        stages["solve_farfield"] = stages.get("solve_farfield", 0.0) + time.perf_counter() - t0
    return out, stages
