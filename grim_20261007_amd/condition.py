"""Equilibrated 1-norm condition estimate of factored systems.

The reference estimates kappa_1(R A C) = ||R A C||_1 ||C^-1 A^-1 R^-1||_1 for the row and column
equilibration R, C of the system matrix (twod/solver.py:528-579,
``_equilibrated_condition_from_lu``): the norm of the equilibrated matrix exactly, the norm of its
inverse with SciPy's block 1-norm estimator (``scipy.sparse.linalg.onenormest``, SciPy 1.15,
Higham and Tisseur, SIAM J. Matrix Anal. Appl. 21 (2000), Algorithm 2.4, t = 2 columns, at most 5
iterations) seeded with a fixed private seed (0x2D5EED, twod/solver.py:583-612) so that the
estimate is a deterministic function of the matrix.  The inverse is applied through the LU
factors: plain solves for FP64 factors, refined solves for mixed-precision ones
(linalg/dense.py:158-199).

Here every O(N^2) step runs in the HIP library: the equilibration (``grim_equilibration``), the
solves with A and with A^H (``grim_lu_solve_ex``, or ``grim_refined_solve_ex`` for mixed-precision
factors).  Between them the estimator's own steps act on N x 2 blocks on the host: the signs of
a block, its column norms and row maxima, and the choice of the next unit vectors.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import engine, native

ESTIMATE_SEED = 0x2D5EED      # twod/solver.py:588
ESTIMATE_COLUMNS = 2          # onenormest's t
ESTIMATE_ITERATIONS = 5       # onenormest's itmax
METHOD_LU = "equilibrated_1norm_lu_onenormest"              # linalg/dense.py:177
METHOD_REFINED = "equilibrated_1norm_refined_inverse"      # linalg/dense.py:170


def _signs(y: np.ndarray) -> np.ndarray:
    """Entrywise y / |y| with sign(0) = 1 (Higham and Tisseur's complex sign)."""
    s = np.array(y, copy=True)
    s[s == 0] = 1
    return s / np.abs(s)


def _parallel(v: np.ndarray, w: np.ndarray) -> bool:
    # columns with entries of modulus 1 are parallel when their dot product is n
    return np.dot(v, w) == v.shape[0]


def _needs_new_column(i: int, x: np.ndarray, old: np.ndarray | None = None) -> bool:
    v = x[:, i]
    if any(_parallel(v, x[:, j]) for j in range(i)):
        return True
    return old is not None and any(_parallel(v, w) for w in old.T)


def _random_column(i: int, x: np.ndarray, rng: np.random.RandomState) -> None:
    x[:, i] = rng.randint(0, 2, size=x.shape[0]) * 2 - 1


def block_onenorm_estimate(matmat, rmatmat, n: int, t: int = ESTIMATE_COLUMNS, itmax: int = ESTIMATE_ITERATIONS,
                           rng: np.random.RandomState | None = None) -> float:
    """Lower bound of ||B||_1 for an n x n operator given by B X (matmat) and B^H X (rmatmat):
    Higham and Tisseur's Algorithm 2.4 with the starting block, parallel-column tests and
    stopping rules SciPy's onenormest uses.  For n <= t the norm is formed exactly."""
    rng = rng if rng is not None else np.random.RandomState(ESTIMATE_SEED)
    if t >= n:
        cols = np.asarray(matmat(np.identity(n)))
        return float(np.max(np.abs(cols).sum(axis=0)))
    x = np.ones((n, t), dtype=float)
    for i in range(1, t):
        _random_column(i, x, rng)
    for i in range(t):
        while _needs_new_column(i, x):
            _random_column(i, x, rng)
    x /= float(n)
    visited = np.zeros(0, dtype=np.intp)
    est_old = 0.0
    s = np.zeros((n, t), dtype=float)
    k = 1
    order = None
    best = None
    while True:
        y = np.asarray(matmat(x))
        mags = np.abs(y).sum(axis=0)
        est = float(np.max(mags))
        best_j = int(np.argmax(mags))
        if est > est_old or k == 2:
            if k >= 2:
                best = order[best_j]
        if k >= 2 and est <= est_old:           # (1) no increase: keep the previous estimate
            est = est_old
            break
        est_old = est
        s_old = s
        if k > itmax:
            break
        s = _signs(y)
        if all(any(_parallel(v, w) for w in s_old.T) for v in s.T):   # (2) signs repeat
            break
        if t > 1:
            for i in range(t):
                while _needs_new_column(i, s, s_old):
                    _random_column(i, s, rng)
        z = np.asarray(rmatmat(s))                # (3)
        h = np.max(np.abs(z), axis=1)
        if k >= 2 and max(h) == h[best]:          # (4)
            break
        order = np.argsort(h)[::-1][:t + len(visited)].copy()
        if t > 1:                                 # (5) the most promising columns were all used
            if np.isin(order[:t], visited).all():
                break
            seen = np.isin(order, visited)
            order = np.concatenate((order[~seen], order[seen]))
        x = np.zeros((n, t), dtype=float)
        x[order[:t], np.arange(t)] = 1.0
        fresh = order[:t][~np.isin(order[:t], visited)]
        visited = np.concatenate((visited, fresh))
        k += 1
    return est


def equilibration(z: torch.Tensor):
    """Row scales, column scales and equilibrated 1-norms of a batch (B, N, N) (column-major)."""
    zb = engine._as_batch(z)
    nb, n = zb.shape[0], zb.shape[-1]
    rs = torch.empty((nb, n), dtype=torch.float64, device=z.device)
    cs = torch.empty((nb, n), dtype=torch.float64, device=z.device)
    norm = torch.empty(nb, dtype=torch.float64, device=z.device)
    native.check("grim_equilibration", native.lib().grim_equilibration(
        nb, n, zb.data_ptr(), n, n * n, rs.data_ptr(), cs.data_ptr(), norm.data_ptr(), engine._stream(z.device)))
    return rs, cs, norm


def _inverse_apply(f: engine.Factor, rhs: np.ndarray, trans: int) -> np.ndarray:
    """A^{-1} rhs (trans 0) or A^{-H} rhs (trans 2) for a single-system factor; rhs (N, t)."""
    dev = f.z.device if f.z is not None else f.lu.device
    n, t = rhs.shape
    b = torch.as_tensor(np.ascontiguousarray(rhs.T.astype(np.complex128)), device=dev)   # (t, N)
    st = engine._stream(dev)
    lib = native.lib()
    if f.precision == "double":
        nbytes = int(lib.grim_lu_solve_ex_workspace_bytes(1, n, t, 64, trans))
        work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        native.check("grim_lu_solve_ex", lib.grim_lu_solve_ex(1, n, t, f.lu.data_ptr(), n, n * n, 64,
                                                              f.aux.data_ptr(), f.aux.shape[1], trans, b.data_ptr(),
                                                              n, t * n, work.data_ptr(), nbytes, st))
        return b.cpu().numpy().T
    x = torch.empty_like(b)
    nbytes = int(lib.grim_refine_ex_workspace_bytes(1, n, t, trans))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    steps, err = engine.ctypes_int(), engine.ctypes_double()
    code = lib.grim_refined_solve_ex(1, n, t, f.z.data_ptr(), n, n * n, f.lu.data_ptr(), n, n * n, f.aux.data_ptr(),
                                     f.aux.shape[1], trans, b.data_ptr(), n, t * n, x.data_ptr(), n, t * n,
                                     engine.REFINE_MAX_STEPS, engine.REFINE_TOL, ws.data_ptr(), nbytes, steps, err,
                                     st)
    if code == 6:
        raise np.linalg.LinAlgError("Mixed-precision refinement stalled; double LU required.")
    native.check("grim_refined_solve_ex", code)
    return x.cpu().numpy().T


def system_condition(f: engine.Factor, scaling=None):
    """(estimate, method) for a single-system factor; a mixed factor whose refined inverse stalls
    is replaced by the FP64 one, as linalg/dense.py:166-191 does (the third value is that
    factor, or None)."""
    rs, cs, norm = scaling if scaling is not None else equilibration(f.z)
    r = rs[0].cpu().numpy()
    c = cs[0].cpu().numpy()
    norm_a = float(norm[0])
    n = f.n

    def matmat(x):
        return c[:, None] * _inverse_apply(f, r[:, None] * np.asarray(x, dtype=np.complex128), 0)

    def rmatmat(x):
        return r[:, None] * _inverse_apply(f, c[:, None] * np.asarray(x, dtype=np.complex128), 2)

    try:
        inverse = block_onenorm_estimate(matmat, rmatmat, n)
    except np.linalg.LinAlgError as exc:
        fb = engine.factor(f.z, "double")
        fb.fallback_reason = engine.MIXED_FALLBACK + str(exc)
        est, method, _ = system_condition(fb, (rs, cs, norm))
        return est, method, fb
    est = norm_a * inverse
    return (est if math.isfinite(est) else float("inf"),
            METHOD_LU if f.precision == "double" else METHOD_REFINED, None)


def condition_estimates(f: engine.Factor):
    """Per system of a factored batch: (estimates, methods, replacement factor or None)."""
    rs, cs, norm = equilibration(f.z)
    ests, methods, parts, changed = [], [], [], False
    for i in range(f.batch):
        part = engine.system(f, i)
        est, method, fb = system_condition(part, (rs[i:i + 1], cs[i:i + 1], norm[i:i + 1]))
        ests.append(est)
        methods.append(method)
        parts.append(fb if fb is not None else part)
        changed |= fb is not None
    replacement = None
    if changed:
        replacement = engine.Factor("mixed", None, None, None, f.z, f.n, f.batch,
                                    next((p.fallback_reason for p in parts if p.fallback_reason), ""), parts=parts)
    return ests, methods, replacement
