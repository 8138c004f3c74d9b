/* GRIM-2D: MI355X-native 2-D boundary-element RCS path (C ABI).
 *
 * Drop-in native seam for the reference's 2-D PEC solve path
 * (14cole/GRIM_20261007, tools/GHOST/ghost_backend).  Conventions follow the
 * reference's own native seam (twod/assembly/native/far.c, far.py):
 *   - int status return: 0 ok, 1 size/index out of range, 3 invalid arguments,
 *     5 HIP runtime failure, 6 adaptive near rule not converged / refinement
 *     stalled, 7 near-pair capacity exceeded, 8 zero pivot, 9 a device-wide
 *     barrier of the tall-panel LU did not complete (bounded wait expired),
 *     10 coincident integration points in a polynomial near block;
 *   - sizes are int64_t; complex arrays are interleaved double[2] (complex128)
 *     or float[2] (complex64), column-major, leading dimension in elements;
 *   - caller-owned outputs and workspace (all pointers are DEVICE pointers
 *     unless stated), no allocation inside the library, no internal threads;
 *   - `stream` is a hipStream_t (NULL = default stream); calls are
 *     asynchronous on that stream unless documented otherwise.
 *
 * Geometry of a set of closed contours with n elements and n nodes: element e
 * runs from node e (start point p0[e]) to node next[e] (end point p1[e]);
 * prev[next[e]] == e.  Stored normals are (-t_y, t_x) of the element tangent,
 * i.e. pointing into the conductor for a contour drawn with the reference's
 * TYPE 2 convention (twod/geometry.py:1009-1065, 1935).
 */
#ifndef GRIM2D_H
#define GRIM2D_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Library version (major * 10000 + minor * 100 + patch). */
int grim_version(void);

/* Load the quadrature rules into the current device's constant memory.
 * Must be called once per device before any other call on that device. */
int grim_init(int device);

/* Bytes of workspace grim_assemble needs for n elements and at most
 * near_per_element near pairs per element on average. */
int64_t grim_assemble_workspace_bytes(int64_t n, int64_t near_per_element);

/* Fused TM + TE system matrices of closed PEC contours in air, one traversal.
 * Replaces the reference's fused assembly (twod/operators.py:2087-2607 with
 * twod/assembly/native/far.c:100-207 ghost_far_block and far.c:333-363
 * ghost_scatter_tile, routed by twod/assembly/scatter.py:303-368 and
 * twod/formulations/combined_regions.py:149-213):
 *   z_tm = S + gamma (D + M/2),  z_te = -M/2 + K' - gamma W,  gamma = 1/(j k).
 * k: real wavenumber (rad/m).  far_order: Gauss points per axis of far pairs,
 * 8 or 16 (the reference's W-calibrated far rule: 8 when 1e-4 <= k Lmax <= 3
 * and Lmin >= 1e-3 Lmax, else 16; twod/operators.py:1509-1527).  fix_rows: the nfix
 * rows r with r % 64 == 0 or prev[r] != r - 1.  Outputs complex128 n x n.
 * Asynchronous; kernel-side failures are reported by grim_workspace_status. */
int grim_assemble(int64_t n, const double* p0, const double* p1, const int32_t* next, const int32_t* prev,
                  const int32_t* fix_rows, int64_t nfix, double k, int far_order, double* z_tm, int64_t ld_tm,
                  double* z_te, int64_t ld_te, void* workspace, int64_t workspace_bytes,
                  int64_t near_per_element, void* stream);

/* grim_assemble without a workspace argument, for callers that cannot allocate: code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation.  No counterpart in the
 * reference.  It allocates nothing: the near-pair lists and blocks that grim_assemble keeps in
 * its workspace live in a static device array of 64 MiB, one per device the library runs on,
 * held until the process ends.  That array must hold
 * n * 8 + near_per_element * n * 140 bytes (each of six arrays rounded up to 256 bytes):
 * n <= 9974 at 48 near pairs per element, 14952 at 32, 29852 at 16; larger requests return 1
 * and launch nothing.  status: caller-owned device word (>= 4 bytes), cleared and then set by
 * the kernels; read it with grim_workspace_status(status, stream).  Same results as
 * grim_assemble, bit for bit.  Every call on one GPU uses that GPU's one array: calls that
 * overlap on one GPU (on different streams) race. */
int grim_assemble_noalloc(int64_t n, const double* p0, const double* p1, const int32_t* next, const int32_t* prev,
                          const int32_t* fix_rows, int64_t nfix, double k, int far_order, double* z_tm,
                          int64_t ld_tm, double* z_te, int64_t ld_te, int64_t near_per_element, void* status,
                          void* stream);

/* The two passes of grim_assemble, for callers that time or overlap them:
 * grim_assemble_far writes every entry of z_tm / z_te (far pairs; zero for
 * near pairs) and must run first; grim_assemble_near adds the near-pair
 * blocks and the mass terms.  Same arguments as grim_assemble. */
int grim_assemble_far(int64_t n, const double* p0, const double* p1, const int32_t* next, const int32_t* prev,
                      const int32_t* fix_rows, int64_t nfix, double k, int far_order, double* z_tm, int64_t ld_tm,
                      double* z_te, int64_t ld_te, void* workspace, int64_t workspace_bytes,
                      int64_t near_per_element, void* stream);
int grim_assemble_near(int64_t n, const double* p0, const double* p1, const int32_t* next, const int32_t* prev,
                       const int32_t* fix_rows, int64_t nfix, double k, int far_order, double* z_tm, int64_t ld_tm,
                       double* z_te, int64_t ld_te, void* workspace, int64_t workspace_bytes,
                       int64_t near_per_element, void* stream);

/* The operators themselves, from the same traversal and near rules as grim_assemble, for
 * formulations that weight them per region and interface: which = 1 gives a = S and
 * b = K' (observation-normal derivative) + c M; which = 2 gives a = D + c M and b = W (Maue),
 * with c = mass_coef[e] for each element e (device double[n]) or `mass` when mass_coef is NULL.
 * Replaces twod/operators.py:2087-2607 (_assemble_linear_operator_matrices[_multi] with single-,
 * double-layer and hypersingular outputs) as used by twod/formulations/dielectric.py:64-69 (S, K'
 * + M/2 at the interior wavenumber) and twod/formulations/regions.py:239-333,
 * combined_regions.py:116-146 (S, K', D, W of each region, the trace-jump mass in D).  Same
 * arguments as grim_assemble otherwise; a and b are complex128 n x n, column-major, and may be
 * blocks of larger matrices through ld_a, ld_b. */
int grim_assemble_operators(int64_t n, const double* p0, const double* p1, const int32_t* next,
                            const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k, int far_order,
                            int which, double* a, int64_t ld_a, double* b, int64_t ld_b, double mass,
                            const double* mass_coef, void* workspace, int64_t workspace_bytes,
                            int64_t near_per_element, void* stream);

/* dst[rows_dst[i], cols_dst[j]] += weight[i] src[rows_src[i], cols_src[j]] for i < m, j < n
 * (complex128, column-major; weight complex128 [m]; index arrays int32 device arrays; a
 * negative destination index skips its row or column).  Composes a regional system, its loads
 * and the exterior density from assembled operator blocks, as the reference's SystemScatter
 * routes tiles (twod/assembly/scatter.py:65-116, formulations/regions.py:239-333, 557-609). */
int grim_route(int64_t m, int64_t n, const double* src, int64_t ld_src, const int32_t* rows_src,
               const int32_t* cols_src, double* dst, int64_t ld_dst, const int32_t* rows_dst,
               const int32_t* cols_dst, const double* weight, void* stream);

/* Scales the four n x n blocks of a complex128 2n x 2n column-major matrix in place:
 * block (i, j) (rows i n .. i n + n - 1, columns j n ..) by scale[i + 2 j] (complex128 host
 * array of 4).  Forms the two-density dielectric system from its assembled operators: eta Z_TM,
 * -eta Z_TE (exterior, eta = j k0), -S1 and factor (K'1 + M/2) (interior), as
 * twod/formulations/dielectric.py:17-81 scales them. */
int grim_scale_blocks(int64_t n, double* z, int64_t ld, const double* scale, void* stream);

/* Synchronises `stream` and returns the kernel status of the last
 * grim_assemble / grim_lu_factor that used `workspace` (0 ok, else a status
 * code as above). */
int grim_workspace_status(void* workspace, void* stream);

/* Degree-2 and degree-3 Galerkin bases (execution option basis_order 2 / 3).  The contour's n
 * elements keep their end nodes 0 .. n-1 (element e from node e to next[e]); each element adds
 * degree - 1 interior nodes n + (degree - 1) e + m at (m + 1) / degree along it, so a system has
 * n * degree unknowns (twod/basis.py:47-63).  grim_assemble_poly writes the TM and TE systems
 * (complex128 n*degree x n*degree, column-major): far pairs with the Q x Q Gauss rule, Q = far_order
 * (9 inside the W-calibrated range, else 16; twod/operators.py:1502-1527, 2335-2518 with width
 * degree + 1), every other pair with the adaptive polynomial near rule (orders 20 / 36, then 72 / 144
 * for a self pair or at depth 24, otherwise bisection; t^4 corner rule on self and touching pairs,
 * exact log moments on self blocks; twod/polynomial_quadrature.py:23-156), the Maue W of each
 * block and the element mass.  Replaces the polynomial branch of _assemble_multi and
 * _near_pair_blocks (twod/operators.py:2053-2067).  The workspace holds the near-pair lists and
 * three tables of n x n element blocks (48 (degree + 1)^2 n^2 bytes).  grim_workspace_status
 * afterwards returns 6 where the reference raises "Polynomial near quadrature did not converge"
 * and 10 where it raises "Polynomial quadrature encountered intersecting integration points". */
int64_t grim_assemble_poly_workspace_bytes(int64_t n, int degree, int64_t near_per_element);
int grim_assemble_poly(int64_t n, int degree, const double* p0, const double* p1, const int32_t* next,
                       const int32_t* prev, double k, int far_order, double* z_tm, int64_t ld_tm, double* z_te,
                       int64_t ld_te, void* workspace, int64_t workspace_bytes, int64_t near_per_element,
                       void* stream);

/* Plane-wave loads of a degree-2 / degree-3 contour: Gauss moments of the nodal basis with the
 * `order`-point rule rule_t / rule_w (device arrays, nodes and weights on [0, 1]), the reference's
 * order being max(12, degree + 3 + ceil(|k| L_max)) (twod/basis.py:89-104, used by
 * twod/assembly/kernels.py:236-259).  Outputs complex128 (n * degree) x na as grim_loads. */
int grim_loads_poly(int64_t n, int degree, int64_t na, const double* p0, const double* p1, const int32_t* prev,
                    double k, const double* angles_deg, int64_t order, const double* rule_t, const double* rule_w,
                    double* bu, double* bdn, int64_t ld, void* stream);

/* Plane-wave loads of all nodes for na incidence angles (degrees, device
 * array): bu = int phi_i exp(j k d.r) ds, bdn = int phi_i (j k d.n) exp(j k d.r) ds,
 * d = (cos phi, sin phi).  Replaces twod/assembly/kernels.py:222-294
 * (incident_loads).  RHS of the TM system is -bu, of the TE system -bdn.
 * Either output may be NULL.  complex128 n x na, leading dimension ld. */
int grim_loads(int64_t n, int64_t na, const double* p0, const double* p1, const int32_t* prev, double k,
               const double* angles_deg, double* bu, double* bdn, int64_t ld, void* stream);

/* Matched monostatic far field B[a] = sum_p x[p,a] (bu[p,a] + gamma bdn[p,a])
 * and sigma_2D[a] = |B[a]|^2 / (4 k).  Replaces twod/assembly/kernels.py:162-219
 * (farfield, matched projection reusing the RHS loads) and
 * twod/solver.py:838-859 (_rcs_sigma_from_amp).  amp: complex128 [na];
 * sigma may be NULL. */
int grim_farfield(int64_t n, int64_t na, const double* x, int64_t ldx, const double* bu, const double* bdn,
                  int64_t ldb, double k, double* amp, double* sigma, void* stream);

/* Far field of one potential: B[a] = sum_p x[p,a] b[p,a] with b the plane-wave moments of that
 * potential (bu for a single layer, bdn for a double layer), sigma[a] = |B[a]|^2 / (4k).  Replaces
 * twod/operators.py:2925-3023 (_farfield_linear_density_many, matched projection, potential 'SLP' or
 * 'DLP') as the sheet solves use it (twod/solver.py:1780-1787).  amp: complex128 [na]; sigma may be
 * NULL. */
int grim_farfield_potential(int64_t n, int64_t na, const double* x, int64_t ldx, const double* b, int64_t ldb,
                            double k, double* amp, double* sigma, void* stream);

/* Sheet (TYPE 1 open contour) systems, alone or beside PEC contours, linear Galerkin bases on a
 * general node map (twod/formulations/sheet.py:6-30, twod/solver.py:1746-1787):
 *   z_tm = S - diag-block mass with element coefficient coef_tm (Zs / (j k eta0) on sheets, 0 on PEC)
 *   z_te = W - mass with coef_te (j k Zs / eta0), rows with pinned[p] != 0 replaced by the identity
 *          row (the strip ends, twod/solver.py:1683-1736)
 * where S is the single layer (far rule far_order_tm, 8 or 16) and W the Maue hypersingular form of
 * the element-pair S blocks (twod/operators.py:2799-2858; far rule far_order_te, 8 or 16).  Elements
 * are given by p0, p1 (ne x 2 float64); next_ids[e] is the element starting at element e's end node,
 * or an id >= ne of that node alone when none does (the ids of the near test); nodes own their
 * elements through own_ptr (nn + 1) / own_el / own_loc (local 0 = start, 1 = end); coef_* are
 * complex128 [ne]; pinned int32 [nn] or NULL.  Either output may be NULL (that polarization is not
 * formed).  Status as grim_assemble (grim_workspace_status on the workspace afterwards).
 * ne, nn <= 65535. */
int64_t grim_sheet_workspace_bytes(int64_t ne, int64_t near_per_element);
int grim_assemble_sheet(int64_t ne, int64_t nn, const double* p0, const double* p1, const int32_t* next_ids,
                        const int32_t* own_ptr, const int32_t* own_el, const int32_t* own_loc, double k,
                        int far_order_tm, int far_order_te, const double* coef_tm, const double* coef_te,
                        const int32_t* pinned, double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te,
                        void* workspace, int64_t workspace_bytes, int64_t near_per_element, void* stream);

/* Plane-wave loads of a sheet mesh's nodes (as grim_loads, over the node map of
 * grim_assemble_sheet): the TM load is -bu, the TE load bdn with pinned rows zero
 * (twod/formulations/sheet.py:33-39).  complex128 nn x na, leading dimension ld. */
int grim_loads_sheet(int64_t ne, int64_t nn, int64_t na, const double* p0, const double* p1, const int32_t* own_ptr,
                     const int32_t* own_el, const int32_t* own_loc, double k, const double* angles_deg, double* bu,
                     double* bdn, int64_t ld, void* stream);

/* Operators of closed contours at a complex wavenumber k = k_re + j k_im (k_re > 0, k_im <= 0: a
 * lossy medium under e^{+j omega t}), in the layout and with the mass terms of
 * grim_assemble_operators, all four from one traversal: s = S, kp = K' + c_k M, d = D + c_d M,
 * w = the Maue W (any may be NULL), c_k / c_d from coef_k / coef_d per element (float64 [n]) or the
 * scalars mass_k / mass_d.  Replaces twod/operators.py:2087-2607 at complex k with the complex
 * Hankel functions of twod/special.py:220-264 (SciPy's hankel2 in the reference).  n <= 65535.
 * Status as grim_assemble (grim_workspace_status on the workspace afterwards). */
int64_t grim_assemble_operators_lossy_workspace_bytes(int64_t n, int64_t near_per_element);
int grim_assemble_operators_lossy(int64_t n, const double* p0, const double* p1, const int32_t* next,
                                  const int32_t* prev, double k_re, double k_im, int far_order, double* s, double* kp,
                                  double* d, double* w, int64_t ld, double mass_k, double mass_d,
                                  const double* coef_k, const double* coef_d, void* workspace,
                                  int64_t workspace_bytes, int64_t near_per_element, void* stream);

/* (j/4) H0^(2)(z) and (j/4) H1^(2)(z) at m complex points z (device arrays: z complex128 [m], out
 * float64 [4 m] = g re, g im, h1 re, h1 im), the device routine the lossy operators use; Re z > 0,
 * Im z <= 0.  Replaces twod/special.py:220-264 (scipy.special.hankel2). */
int grim_hankel2_probe(int64_t m, const double* z, double* out, void* stream);

/* Batched dense LU.  All dense calls take `batch` equally strided matrices
 * (strides in complex elements for matrices/vectors, in elements for int32
 * arrays, in bytes for aux), e.g. the TE and TM systems of one request.
 *
 * Bytes of factor-side auxiliary storage per matrix: the row permutation the
 * pivots compose to (4 n bytes) and the interchange list of the panel being
 * factored (1 KiB), each rounded up to 256 bytes; precision 64 or 32. */
int64_t grim_lu_aux_bytes(int64_t n, int precision);

/* In-place LU with partial pivoting, P A = L U (LAPACK pivot rule: first row of
 * largest |re| + |im|), complex128 (precision 64) or complex64 (32), n <= 65536.
 * piv: int32 [n] per matrix (absolute row exchanged with row c, LAPACK getrf
 * order, 0-based); aux: grim_lu_aux_bytes per matrix, stride a multiple of 256.
 * status: device uint32 word (read with grim_workspace_status).  Replaces
 * scipy.linalg.lu_factor as used by linalg/dense.py:114-155 (zgetrf) and
 * linalg/refined_lu.py:26-39 (cgetrf on the complex64 copy). */
int grim_lu_factor(int64_t batch, int64_t n, void* a, int64_t lda, int64_t stride_a, int precision, int32_t* piv,
                   int64_t stride_piv, void* aux, int64_t stride_aux, void* status, void* stream);

/* Bytes of workspace grim_lu_solve needs. */
int64_t grim_lu_solve_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs, int precision);

/* Solve A X = B in place for nrhs columns with the factors and aux of
 * grim_lu_factor.  Replaces scipy.linalg.lu_solve (linalg/dense.py:239-301). */
int grim_lu_solve(int64_t batch, int64_t n, int64_t nrhs, const void* lu, int64_t ldlu, int64_t stride_lu,
                  int precision, const void* aux, int64_t stride_aux, void* b, int64_t ldb, int64_t stride_b,
                  void* work, int64_t work_bytes, void* stream);

/* Bytes of workspace grim_lu_solve_ex needs (trans 0: as grim_lu_solve; trans 2: also one
 * 256 x 256 staging block per matrix). */
int64_t grim_lu_solve_ex_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs, int precision, int trans);

/* grim_lu_solve with a transpose option: trans 0 solves A X = B, trans 2 solves A^H X = B
 * (conjugate transpose) with the same factors.  Replaces scipy.linalg.lu_solve(..., trans=2)
 * as twod/solver.py:546-565 (_equilibrated_condition_from_lu, the adjoint products of the
 * condition estimate) and linalg/refined_lu.py:41-61 (RefinedLU.solve with trans) call it. */
int grim_lu_solve_ex(int64_t batch, int64_t n, int64_t nrhs, const void* lu, int64_t ldlu, int64_t stride_lu,
                     int precision, const void* aux, int64_t stride_aux, int trans, void* b, int64_t ldb,
                     int64_t stride_b, void* work, int64_t work_bytes, void* stream);

/* C -= op(A) B, op(A) = A (trans_a 0, A m x k) or A^H (trans_a 2, A stored k x m, lda >= k):
 * the FP64 residual b - Z^H x of linalg/refined_lu.py:50-51 with trans=2. */
int grim_gemm_minus_ex(int64_t batch, int64_t m, int64_t n, int64_t k, int trans_a, const void* a, int64_t lda,
                       int64_t stride_a, const void* b, int64_t ldb, int64_t stride_b, void* c, int64_t ldc,
                       int64_t stride_c, int precision, void* stream);

/* Row/column equilibration of complex128 matrices and the equilibrated 1-norm: row_scale[i] =
 * max_j |z_ij| (1 if 0), col_scale[j] = max_i |z_ij| / row_scale[i] (1 if 0), norm =
 * max_j sum_i |z_ij| / row_scale[i] / col_scale[j].  row_scale, col_scale: double [n] per
 * matrix (stride n); norm: double [batch].  Replaces _equilibrated_scaling_and_norm_1
 * (twod/solver.py:472-525). */
int grim_equilibration(int64_t batch, int64_t n, const double* z, int64_t ldz, int64_t stride_z, double* row_scale,
                       double* col_scale, double* norm, void* stream);

/* Infinity norms max_i sum_j |z_ij| of complex128 matrices into out[batch] (device), with a
 * caller-owned workspace of grim_norm_inf_workspace_bytes: the matrix norm of the normwise
 * backward-error gate (linalg/workspace.py:19-98, matrix_inf_norm / checked_row_norms, as
 * linalg/dense.py:36 and :257-280 use it).  Deterministic (fixed summation order). */
int64_t grim_norm_inf_workspace_bytes(int64_t batch, int64_t n);
int grim_norm_inf(int64_t batch, int64_t n, const double* z, int64_t ldz, int64_t stride_z, void* work,
                  int64_t work_bytes, double* out, void* stream);

/* C -= A B (complex; precision 64 or 32); the FP64 refinement residual. */
int grim_gemm_minus(int64_t batch, int64_t m, int64_t n, int64_t k, const void* a, int64_t lda, int64_t stride_a,
                    const void* b, int64_t ldb, int64_t stride_b, void* c, int64_t ldc, int64_t stride_c,
                    int precision, void* stream);

/* Bytes of workspace grim_refined_solve needs. */
int64_t grim_refine_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs);

/* Mixed-precision solve (linalg/refined_lu.py:41-61): x = LU32^{-1} b, then
 * r = b - Z x in FP64, x += LU32^{-1} r until max_col(|r|_inf / |b|_inf) <= tol,
 * for each matrix of the batch on its own: a matrix within tol takes no more
 * corrections while the others continue, so a batched call returns what
 * separate calls return (at most max_steps corrections; a step that does not
 * reduce a matrix's error below 0.98 of its previous one returns 6).
 * steps_out: corrections of the slowest matrix (max_steps when a matrix used every
 * correction without reaching tol, as distinct from a stall); error_out: largest final error.
 * z: complex128 systems; lu/aux: complex64 factors of the same matrices.
 * Synchronises the stream once per step.  steps_out / error_out: host
 * pointers (may be NULL).  On GRIM_OK the workspace starts with the final
 * residual b - Z x (complex128, leading dimension n, stride n * nrhs per
 * matrix), the residual linalg/refined_lu.py:41-61 returns with
 * return_residual=True. */
int grim_refined_solve(int64_t batch, int64_t n, int64_t nrhs, const double* z, int64_t ldz, int64_t stride_z,
                       const float* lu, int64_t ldlu, int64_t stride_lu, const void* aux, int64_t stride_aux,
                       const double* b, int64_t ldb, int64_t stride_b, double* x, int64_t ldx, int64_t stride_x,
                       int max_steps, double tol, void* workspace, int64_t workspace_bytes, int* steps_out,
                       double* error_out, void* stream);

/* Bytes of workspace grim_refined_solve_ex needs (trans 0: as grim_refined_solve). */
int64_t grim_refine_ex_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs, int trans);

/* grim_refined_solve for op(Z) x = b, op = identity (trans 0) or conjugate transpose (trans 2):
 * linalg/refined_lu.py:41-61 with its trans argument (the refined inverse of the condition
 * estimate of mixed-precision factors, linalg/dense.py:166-172). */
int grim_refined_solve_ex(int64_t batch, int64_t n, int64_t nrhs, const double* z, int64_t ldz, int64_t stride_z,
                          const float* lu, int64_t ldlu, int64_t stride_lu, const void* aux, int64_t stride_aux,
                          int trans, const double* b, int64_t ldb, int64_t stride_b, double* x, int64_t ldx,
                          int64_t stride_x, int max_steps, double tol, void* workspace, int64_t workspace_bytes,
                          int* steps_out, double* error_out, void* stream);

/* complex128 -> complex64 copy (m x n) and complex128 copy, batched. */
int grim_to_c64(int64_t batch, int64_t m, int64_t n, const double* src, int64_t lds, int64_t stride_src, float* dst,
                int64_t ldd, int64_t stride_dst, void* stream);
int grim_copy_c128(int64_t batch, int64_t m, int64_t n, const double* src, int64_t lds, int64_t stride_src,
                   double* dst, int64_t ldd, int64_t stride_dst, void* stream);

/* Impedance (Robin) rows of closed contours with a surface impedance that is
 * constant on each contour, in place: x <- x + alpha_x[r] y on every row r;
 * y <- y on rows with dirichlet[r] != 0 and y <- x + alpha_y[r] y elsewhere
 * (both right-hand sides from the original x, y).  x, y: complex128 rows x
 * cols, column-major (cols <= 65535); alpha_x, alpha_y: complex128 [rows];
 * dirichlet: int32 [rows].  With x = Z_TE, y = Z_TM (or x = b_dn, y = b_u),
 * alpha = j k Zs / eta (TE) and j k eta / Zs (TM), and dirichlet marking TM
 * rows of PEC contours, this forms the flux rows dn(u) + alpha u of the
 * combined potential: Z_TE + alpha Z_TM.  Replaces the per-node routing of
 * twod/formulations/combined_regions.py:43-62 and the Robin system of
 * twod/formulations/robin.py:25-83 for such laws (Robin coefficient:
 * twod/geometry.py:2488-2521). */
int grim_robin_rows(int64_t rows, int64_t cols, double* x, int64_t ldx, double* y, int64_t ldy,
                    const double* alpha_x, const double* alpha_y, const int32_t* dirichlet, void* stream);

/* Impedance rows of a law that varies along a contour (PEC and impedance zones,
 * tapers, tables), in place, after grim_robin_rows with the nodal coefficients
 * alpha_p = (alpha_e(p) + alpha_e(prev p)) / 2: for each row p listed in `walk`
 *   Z_TE[p, :] += c_te[p] ((S0 - S1)[p, :] + (gamma/2) m_p)   (mass part if flags[p] & 1)
 *   Z_TM[p, :] += c_tm[p] ((S0 - S1)[p, :] + (gamma/2) m_p)   (mass part if flags[p] & 2)
 * with S0 (S1) the rows of S that element p (prev p) contributes through its
 * basis function at node p, m_p = (L_p/6 at next p, (L_p - L_prev p)/3 at p,
 * -L_prev p/6 at prev p), and c = (alpha_e(p) - alpha_e(prev p)) / 2.  With
 * these the rows carry the element-weighted alpha S and trace-jump mass of the
 * reference: twod/formulations/regions.py:274-304 (single-layer observation
 * coefficients), combined_regions.py:43-83 (routing, _trace_jump_mass),
 * regions.py:66-119 (element and nodal alpha).  walk: int32 [64 nblocks]:
 * per block, lane 0 the element before lane 1's row, lanes 1..63 rows p whose
 * previous element is the lane before's, -1 for none (entries in [-1, n)).
 * c_te, c_tm: complex128 [n]; flags: int32 [n].  The far rule is far_order;
 * near pairs are read from `workspace`, which must hold the near lists and
 * blocks of a grim_assemble call on the same contour and wavenumber (same
 * near_per_element), unchanged since. */
int grim_robin_split(int64_t n, const double* p0, const double* p1, const int32_t* next, const int32_t* prev,
                     double k, int far_order, int64_t nblocks, const int32_t* walk, const double* c_te,
                     const double* c_tm, const int32_t* flags, const void* workspace, int64_t workspace_bytes,
                     int64_t near_per_element, double* z_te, int64_t ld_te, double* z_tm, int64_t ld_tm,
                     void* stream);

/* Right-hand sides of one polarization's impedance rows: out[p, a] = -b_u[p, a]
 * on rows with dirichlet[p] != 0, else -sum_e i_e(p, a) (j k d_a.n_e + alpha_e[e])
 * over the two elements of node p -- the derivative load with each element's own
 * Robin coefficient.  Replaces twod/formulations/regions.py:557-590 (rhs_many)
 * with twod/assembly/kernels.py:222-294 (incident_loads, observation
 * coefficients).  alpha_e: complex128 [n]; out: complex128 n x na (ld >= n). */
int grim_robin_rhs(int64_t n, int64_t na, const double* p0, const double* p1, const int32_t* prev, double k,
                   const double* angles_deg, const double* alpha_e, const int32_t* dirichlet, double* out,
                   int64_t ld, void* stream);

/* Launch shapes (no counterpart in the reference; for a project autotuner and
 * callers that pin their own shapes).  Knob 0: far-pass source columns per
 * wavefront (1, 2, 4, 8, 16 or 32), keyed by n; knob 1: columns per wavefront
 * (1, 2 or 4) of the factorization's 64-row triangular solves, keyed by the
 * columns of one call times its batch; knob 2: columns per wavefront (1, 2 or
 * 4) of grim_lu_solve's 256-row block solves, keyed by the right-hand-side
 * columns times the batch.  An entry (min_size, value) applies to every size >= min_size up
 * to the next entry; value 0 removes the entry; without an entry the library's
 * built-in rule applies.  Shapes never change results: every matrix entry and
 * solution component is computed by the same operations.  Process-wide; set
 * them before launching work (calls are serialised by a lock, but a launch
 * that runs concurrently may see either shape).  grim_launch_shape returns the
 * table's value for a size (0 = built-in rule).  Status: 3 for an unknown knob
 * or value, 1 when a knob already has 16 entries. */
int grim_set_launch_shape(int knob, int64_t min_size, int value);
int grim_clear_launch_shapes(void);
int grim_launch_shape(int knob, int64_t size);

/* Per-column 2-norm of an m x n complex128 matrix into out[n] (device). */
int grim_column_norms(int64_t m, int64_t n, const double* v, int64_t ld, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
