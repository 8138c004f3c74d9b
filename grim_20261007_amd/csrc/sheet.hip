// Resistive and reactive sheets (TYPE 1 open contours), alone or beside PEC contours, with linear
// Galerkin bases on general node maps:
//   TM  S  - (Zs / (j k eta0)) M                      (single-layer representation)
//   TE  W  - (j k Zs / eta0) M,  strip-end rows pinned (double-layer / hypersingular representation)
// with Zs = 0 on PEC elements, W the Maue form of S, and the loads -b_u (TM), b_dn (TE).
//
// Replaces the reference's sheet assembly and loads (tools/GHOST/ghost_backend/twod/formulations/
// sheet.py:6-39): S from _assemble_linear_operator_matrices without the double layer, W from
// _assemble_linear_hypersingular_matrix (twod/operators.py:2799-2858: W-only traversal, the Maue
// identity on each element-pair S block), the element mass (twod/assembly/mass.py:6-24) and the
// open-edge rows of _geometric_sheet_endpoint_nodes (twod/solver.py:1683-1736).
//
// Layout: element e runs from its local node 0 to its local node 1; a node may have any number of
// elements (open strips have end nodes with one, junctions more), listed in CSR form:
// own_el[own_ptr[p] .. own_ptr[p + 1]) with local indices own_loc.  The pair classification takes
// the linear near pass's id convention: element e's start has id e, its end id nid[e] (the element
// starting there, or an id of its own >= ne where none does); the test is geometric apart from these
// ids, so both passes see exactly the same far pairs.  The element-pair S blocks live
// in one table, block (e, f) at (f ne + e) * 4 complex entries (row a = observation basis, column
// b = source basis, scaled by both lengths): far pairs by the Q x Q Gauss rule below, near pairs
// copied from the linear near pass (self series, touching corner rule, fixed and adaptive boxes).
// A gather then forms every node entry from the blocks of the two nodes' elements in a fixed order
// (no atomics) and adds the mass.
#include "common.h"

// ---------------------------------------------------------------------------
// Far pairs: one thread per unordered element pair e < f (e = lane of the block row, f =
// blockIdx.y) writes block (e, f) and its transpose, block (f, e).  Pairs that are not far are
// left to the near pass, which stores both orders (the test is symmetric, so the two passes agree
// on every pair).
// ---------------------------------------------------------------------------
template <int Q>
__global__ void __launch_bounds__(64) sheet_far(int ne, const double* __restrict__ p0, const double* __restrict__ p1,
                                                const int* __restrict__ nid, double k, double* __restrict__ bs) {
    __shared__ double st[Q], sw[Q];
    const int lane = threadIdx.x;
    if (lane < Q) {
        st[lane] = Q == 8 ? GRIM_GL8_T[lane] : GRIM_GL16_T[lane];
        sw[lane] = Q == 8 ? GRIM_GL8_W[lane] : GRIM_GL16_W[lane];
    }
    __syncthreads();
    const int e = blockIdx.x * 64 + lane, f = blockIdx.y;
    if (e >= ne || e >= f) return;
    const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
    if (!is_far(o, e, nid[e], s, f, nid[f])) return;   // the near pass lists exactly the others
    cplx S[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) S[a][b] = {0.0, 0.0};
    for (int i = 0; i < Q; ++i) {
        const double ti = st[i];
        const double xo = fma(ti, o.sx, o.x0), yo = fma(ti, o.sy, o.y0);
        cplx t0 = {0.0, 0.0}, t1 = {0.0, 0.0};
        for (int j = 0; j < Q; ++j) {
            const double tj = st[j], wj = sw[j];
            const double dx = xo - fma(tj, s.sx, s.x0), dy = yo - fma(tj, s.sy, s.y0);
            double x = k * sqrt(dx * dx + dy * dy);
            if (x < 1e-12) x = 1e-12;
            const HK v = hankel_quarter_v(x);
            const double w0 = wj * (1.0 - tj), w1 = wj * tj;
            t0.re = fma(w0, v.gr, t0.re);
            t0.im = fma(w0, v.gi, t0.im);
            t1.re = fma(w1, v.gr, t1.re);
            t1.im = fma(w1, v.gi, t1.im);
        }
        const double c0 = sw[i] * (1.0 - ti), c1 = sw[i] * ti;
        S[0][0].re = fma(c0, t0.re, S[0][0].re);
        S[0][0].im = fma(c0, t0.im, S[0][0].im);
        S[0][1].re = fma(c0, t1.re, S[0][1].re);
        S[0][1].im = fma(c0, t1.im, S[0][1].im);
        S[1][0].re = fma(c1, t0.re, S[1][0].re);
        S[1][0].im = fma(c1, t0.im, S[1][0].im);
        S[1][1].re = fma(c1, t1.re, S[1][1].re);
        S[1][1].im = fma(c1, t1.im, S[1][1].im);
    }
    const double scale = o.len * s.len;
    double* out = bs + 8 * ((int64_t)f * ne + e);
    double* mirror = bs + 8 * ((int64_t)e * ne + f);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            out[2 * (2 * a + b)] = S[a][b].re * scale;
            out[2 * (2 * a + b) + 1] = S[a][b].im * scale;
            mirror[2 * (2 * b + a)] = S[a][b].re * scale;
            mirror[2 * (2 * b + a) + 1] = S[a][b].im * scale;
        }
}

// Near pairs: the S part (first 8 values) of each listed pair's linear near block into the table.
__global__ void __launch_bounds__(256) sheet_near_store(int ne, const int* __restrict__ offsets, int capacity,
                                                        const int* __restrict__ pair_e,
                                                        const int* __restrict__ pair_f,
                                                        const double* __restrict__ blocks, double* __restrict__ bs) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= offsets[ne] || offsets[ne] > capacity) return;   // overflow: the lists were not written
    const int e = pair_e[idx], f = pair_f[idx];
    const double* src = blocks + 16 * (int64_t)idx;
    double* out = bs + 8 * ((int64_t)f * ne + e);
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = src[i];
}

// ---------------------------------------------------------------------------
// Node gather: Z[p, q] = sum over the elements (e, a) of p and (f, b) of q of the block entry
// (S, or the Maue W = -k^2 (n_e . n_f) S + s_a s_b sum(S) / (L_e L_f), s = (-1, +1)), minus
// coef_e L_e m_ab on e == f (m = [[1/3, 1/6], [1/6, 1/3]]); pinned rows p hold the identity row.
// ---------------------------------------------------------------------------
template <bool HYPER>
__global__ void __launch_bounds__(256) sheet_gather(int nn, int ne, const double* __restrict__ p0,
                                                    const double* __restrict__ p1, const int* __restrict__ own_ptr,
                                                    const int* __restrict__ own_el, const int* __restrict__ own_loc,
                                                    const double* __restrict__ bs, double k,
                                                    const double* __restrict__ coef, const int* __restrict__ pinned,
                                                    double* __restrict__ z, int64_t ld) {
    const int p = blockIdx.x * 256 + threadIdx.x, q = blockIdx.y;
    if (p >= nn) return;
    double zr = 0.0, zi = 0.0;
    if (pinned && pinned[p]) {
        zr = p == q ? 1.0 : 0.0;
    } else {
        const double kk = -k * k;
        for (int r = own_ptr[p]; r < own_ptr[p + 1]; ++r) {
            const int e = own_el[r], a = own_loc[r];
            const Elem o = load_elem(p0, p1, e);
            for (int c = own_ptr[q]; c < own_ptr[q + 1]; ++c) {
                const int f = own_el[c], b = own_loc[c];
                const double* B = bs + 8 * ((int64_t)f * ne + e);
                const int ab = 2 * a + b;
                double vr, vi;
                if (HYPER) {
                    const Elem s = load_elem(p0, p1, f);
                    double denom = o.len * s.len;
                    if (denom < GRIM_EPS * GRIM_EPS) denom = GRIM_EPS * GRIM_EPS;
                    const double sr = (B[0] + B[2]) + (B[4] + B[6]), si = (B[1] + B[3]) + (B[5] + B[7]);
                    const double tsign = a == b ? 1.0 : -1.0;
                    const double nn_ = kk * (o.nx * s.nx + o.ny * s.ny);
                    vr = nn_ * B[2 * ab] + tsign * (sr / denom);
                    vi = nn_ * B[2 * ab + 1] + tsign * (si / denom);
                } else {
                    vr = B[2 * ab];
                    vi = B[2 * ab + 1];
                }
                if (e == f) {
                    const double m = o.len * (a == b ? 1.0 / 3.0 : 1.0 / 6.0);
                    vr -= coef[2 * e] * m;
                    vi -= coef[2 * e + 1] * m;
                }
                zr += vr;
                zi += vi;
            }
        }
    }
    const int64_t at = 2 * ((int64_t)p + (int64_t)q * ld);
    z[at] = zr;
    z[at + 1] = zi;
}

// ---------------------------------------------------------------------------
// Plane-wave loads: b_u[p, a] = sum over the elements (e, l) of p of the exact moment i_l of
// element e, b_dn the same times j k n_e.d (twod/assembly/kernels.py:20-39, 260-285).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) sheet_loads(int nn, int na, const double* __restrict__ p0,
                                                   const double* __restrict__ p1, const int* __restrict__ own_ptr,
                                                   const int* __restrict__ own_el, const int* __restrict__ own_loc,
                                                   double k, const double* __restrict__ angles_deg,
                                                   double* __restrict__ bu, double* __restrict__ bdn, int64_t ld) {
    const int p = blockIdx.x * 256 + threadIdx.x, a = blockIdx.y;
    if (p >= nn) return;
    const double phi = angles_deg[a] * (GRIM_PI / 180.0);
    double dy, dx;
    sincos(phi, &dy, &dx);
    cplx u = {0.0, 0.0}, v = {0.0, 0.0};
    for (int r = own_ptr[p]; r < own_ptr[p + 1]; ++r) {
        const Elem E = load_elem(p0, p1, own_el[r]);
        cplx i0, i1, dn;
        elem_moments(E, k, dx, dy, i0, i1, dn);
        const cplx m = own_loc[r] == 0 ? i0 : i1;
        u = cadd(u, m);
        v = cadd(v, cmul(m, dn));
    }
    const int64_t at = 2 * ((int64_t)p + (int64_t)a * ld);
    if (bu) { bu[at] = u.re; bu[at + 1] = u.im; }
    if (bdn) { bdn[at] = v.re; bdn[at + 1] = v.im; }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
extern "C" hipError_t grim_launch_sheet_far(int q, int ne, const double* p0, const double* p1, const int* nid,
                                            double k, double* bs, hipStream_t st) {
    const dim3 g((ne + 63) / 64, ne);
    if (q == 16) sheet_far<16><<<g, 64, 0, st>>>(ne, p0, p1, nid, k, bs);
    else sheet_far<8><<<g, 64, 0, st>>>(ne, p0, p1, nid, k, bs);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_sheet_near_store(int ne, const int* offsets, int capacity, const int* pair_e,
                                                   const int* pair_f, const double* blocks, double* bs,
                                                   hipStream_t st) {
    sheet_near_store<<<(capacity + 255) / 256, 256, 0, st>>>(ne, offsets, capacity, pair_e, pair_f, blocks, bs);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_sheet_gather(int hyper, int nn, int ne, const double* p0, const double* p1,
                                               const int* own_ptr, const int* own_el, const int* own_loc,
                                               const double* bs, double k, const double* coef, const int* pinned,
                                               double* z, int64_t ld, hipStream_t st) {
    const dim3 g((nn + 255) / 256, nn);
    if (hyper)
        sheet_gather<true><<<g, 256, 0, st>>>(nn, ne, p0, p1, own_ptr, own_el, own_loc, bs, k, coef, pinned, z, ld);
    else
        sheet_gather<false><<<g, 256, 0, st>>>(nn, ne, p0, p1, own_ptr, own_el, own_loc, bs, k, coef, pinned, z, ld);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_sheet_loads(int nn, int na, const double* p0, const double* p1, const int* own_ptr,
                                              const int* own_el, const int* own_loc, double k,
                                              const double* angles_deg, double* bu, double* bdn, int64_t ld,
                                              hipStream_t st) {
    const dim3 g((nn + 255) / 256, na);
    sheet_loads<<<g, 256, 0, st>>>(nn, na, p0, p1, own_ptr, own_el, own_loc, k, angles_deg, bu, bdn, ld);
    return hipGetLastError();
}
