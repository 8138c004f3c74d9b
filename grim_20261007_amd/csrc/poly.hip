// Degree-2 and degree-3 Galerkin assembly and loads for closed PEC contours
// (TM: S + gamma (D + M/2), TE: -M/2 + K' - gamma W, gamma = 1/(jk)).
//
// Replaces the reference's polynomial branch of the fill: the far pass of
// _assemble_multi with width = degree + 1 (tools/GHOST/ghost_backend/twod/operators.py:2335-2518,
// W-calibrated far rule :1502-1527), the polynomial near blocks (twod/polynomial_quadrature.py:23-156:
// adaptive Gauss orders 20 / 36, then 72 / 144 on a self pair or at the depth limit, otherwise
// bisection of the longer interval; t^4 corner rule on self and touching pairs, exact log moments on
// the self block), the Maue W (polynomial_quadrature.py:23-26), the element mass (twod/basis.py:66-76)
// and the Gauss plane-wave moments (basis.py:89-104, assembly/kernels.py:236-294).
//
// Layout: element e runs from node e to node next[e] (local nodes 0 and 1); its interior nodes
// (local 2 .. d) are n + (d - 1) e + 0 .. d - 2, the reference's enrichment order (basis.py:47-63).
// Element-pair blocks are kept in three tables of n x n blocks, block (e, f) at (f n + e) W^2:
// S(e, f), K'(e, f) and the finished TE block K'(e, f) - gamma W(e, f), each W x W complex,
// row a = observation basis, column b = source basis, scaled by the element lengths.  A gather
// then forms every node entry from at most four blocks in a fixed order (no atomics), with
// D(e, f) = K'(f, e)^T.
#include "common.h"
#include "poly_rules.h"

#define POLY_RTOL 1.0e-9       // NEAR_PAIR_QUADRATURE_RTOL, twod/operators.py:1168
#define POLY_MAX_DEPTH 24      // 2 * NEAR_PAIR_QUADRATURE_MAX_DEPTH (operators.py:1169), polynomial_quadrature.py:128
#define POLY_STACK 32

template <int W> struct Poly;
template <> struct Poly<3> {
    static __device__ __forceinline__ double c(int m, int l) { return GRIM_P2_C[m][l]; }
    static __device__ __forceinline__ double der(int i, int l) { return GRIM_P2_DER[i][l]; }
    static __device__ __forceinline__ double mass(int a, int b) { return GRIM_P2_MASS[a][b]; }
    static __device__ __forceinline__ double logm(int a, int b) { return GRIM_P2_LOG[a][b]; }
};
template <> struct Poly<4> {
    static __device__ __forceinline__ double c(int m, int l) { return GRIM_P3_C[m][l]; }
    static __device__ __forceinline__ double der(int i, int l) { return GRIM_P3_DER[i][l]; }
    static __device__ __forceinline__ double mass(int a, int b) { return GRIM_P3_MASS[a][b]; }
    static __device__ __forceinline__ double logm(int a, int b) { return GRIM_P3_LOG[a][b]; }
};

// Nodal basis values at x (Horner in the monomial coefficients, basis.py:28-32).
template <int W>
__device__ __forceinline__ void basis_at(double x, double (&phi)[W]) {
#pragma unroll
    for (int l = 0; l < W; ++l) {
        double v = Poly<W>::c(W - 1, l);
#pragma unroll
        for (int m = W - 2; m >= 0; --m) v = fma(v, x, Poly<W>::c(m, l));
        phi[l] = v;
    }
}

// Gauss-Legendre rules on [0, 1] by order.
__device__ __forceinline__ const double* rule_t(int order) {
    switch (order) {
        case 9: return GRIM_GL9_T;
        case 16: return GRIM_GL16_T;
        case 20: return GRIM_GL20_T;
        case 36: return GRIM_GL36_T;
        case 72: return GRIM_GL72_T;
        default: return GRIM_GL144_T;
    }
}
__device__ __forceinline__ const double* rule_w(int order) {
    switch (order) {
        case 9: return GRIM_GL9_W;
        case 16: return GRIM_GL16_W;
        case 20: return GRIM_GL20_W;
        case 36: return GRIM_GL36_W;
        case 72: return GRIM_GL72_W;
        default: return GRIM_GL144_W;
    }
}

// Maue W and the finished TE block K' - gamma W of one element pair from its scaled S and K'
// (entry a W + b), gamma = (0, -1/k): W = -k^2 (n_o . n_s) S + DER^T S DER / (L_o L_s).
template <int W>
__device__ __forceinline__ cplx te_entry(const double* S, const double* K, int a, int b, double k, double nn,
                                         double lolen) {
    double wr = 0.0, wi = 0.0;
#pragma unroll
    for (int i = 0; i < W; ++i)
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const double c = Poly<W>::der(i, a) * Poly<W>::der(j, b);
            wr = fma(c, S[2 * (i * W + j)], wr);
            wi = fma(c, S[2 * (i * W + j) + 1], wi);
        }
    const double kk = -k * k * nn;
    wr = kk * S[2 * (a * W + b)] + wr / lolen;
    wi = kk * S[2 * (a * W + b) + 1] + wi / lolen;
    const double g = -1.0 / k;
    return {K[2 * (a * W + b)] + g * wi, K[2 * (a * W + b) + 1] - g * wr};
}

// ---------------------------------------------------------------------------
// Far pairs: one thread per ordered element pair (observation e = lane of the block row,
// source f = blockIdx.y), Q x Q Gauss rule (far.c / _far_tile semantics with width W).
// ---------------------------------------------------------------------------
template <int W, int Q>
__global__ void __launch_bounds__(64) poly_far(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                               const int* __restrict__ next, double k, double* __restrict__ bs,
                                               double* __restrict__ bk, double* __restrict__ bte) {
    __shared__ double st[Q], swp[Q][W];   // nodes t_j and weights w_j phi_b(t_j)
    const int lane = threadIdx.x;
    if (lane < Q) {
        const double t = rule_t(Q)[lane], w = rule_w(Q)[lane];
        double phi[W];
        basis_at<W>(t, phi);
        st[lane] = t;
#pragma unroll
        for (int b = 0; b < W; ++b) swp[lane][b] = w * phi[b];
    }
    __syncthreads();
    const int e = blockIdx.x * 64 + lane, f = blockIdx.y;
    if (e >= n) return;
    const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
    if (!is_far(o, e, next[e], s, f, next[f])) return;   // near pairs: poly_near
    cplx S[W][W], K[W][W];
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
        for (int b = 0; b < W; ++b) S[a][b] = K[a][b] = {0.0, 0.0};
    for (int i = 0; i < Q; ++i) {
        const double ti = st[i];
        const double xo = fma(ti, o.sx, o.x0), yo = fma(ti, o.sy, o.y0);
        cplx ts[W], tk[W];
#pragma unroll
        for (int b = 0; b < W; ++b) ts[b] = tk[b] = {0.0, 0.0};
        for (int j = 0; j < Q; ++j) {
            const double tj = st[j];
            const double dx = xo - fma(tj, s.sx, s.x0), dy = yo - fma(tj, s.sy, s.y0);
            const double r = sqrt(dx * dx + dy * dy);
            const HK v = hankel_quarter_v(k * r);
            const double po = (dx * o.nx + dy * o.ny) / r;
            const double kr_ = -k * v.hr * po, ki_ = -k * v.hi * po;
#pragma unroll
            for (int b = 0; b < W; ++b) {
                const double wp = swp[j][b];
                ts[b].re = fma(wp, v.gr, ts[b].re);
                ts[b].im = fma(wp, v.gi, ts[b].im);
                tk[b].re = fma(wp, kr_, tk[b].re);
                tk[b].im = fma(wp, ki_, tk[b].im);
            }
        }
#pragma unroll
        for (int a = 0; a < W; ++a) {
            const double c = swp[i][a];
#pragma unroll
            for (int b = 0; b < W; ++b) {
                S[a][b].re = fma(c, ts[b].re, S[a][b].re);
                S[a][b].im = fma(c, ts[b].im, S[a][b].im);
                K[a][b].re = fma(c, tk[b].re, K[a][b].re);
                K[a][b].im = fma(c, tk[b].im, K[a][b].im);
            }
        }
    }
    const double scale = o.len * s.len;
    double Sv[2 * W * W], Kv[2 * W * W];
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
        for (int b = 0; b < W; ++b) {
            Sv[2 * (a * W + b)] = S[a][b].re * scale;
            Sv[2 * (a * W + b) + 1] = S[a][b].im * scale;
            Kv[2 * (a * W + b)] = K[a][b].re * scale;
            Kv[2 * (a * W + b) + 1] = K[a][b].im * scale;
        }
    const int64_t base = 2 * ((int64_t)f * n + e) * (W * W);
    const double nn = o.nx * s.nx + o.ny * s.ny;
    double lolen = o.len * s.len;
    if (lolen < GRIM_EPS * GRIM_EPS) lolen = GRIM_EPS * GRIM_EPS;
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const int m = a * W + b;
            bs[base + 2 * m] = Sv[2 * m];
            bs[base + 2 * m + 1] = Sv[2 * m + 1];
            bk[base + 2 * m] = Kv[2 * m];
            bk[base + 2 * m + 1] = Kv[2 * m + 1];
            const cplx t = te_entry<W>(Sv, Kv, a, b, k, nn, lolen);
            bte[base + 2 * m] = t.re;
            bte[base + 2 * m + 1] = t.im;
        }
}

// ---------------------------------------------------------------------------
// Near pairs: one wavefront per listed (non-far) ordered pair, the reference's adaptive rule run
// depth first with an explicit stack.  Block values are spread over the lanes: lane i holds entry
// i of (S re, S im interleaved, W^2 entries) then (K' likewise), 4 W^2 values.
// ---------------------------------------------------------------------------
// Shared interval ends (polynomial_quadrature.py:35-55): bit 0 = observation side at its
// interval start, bit 1 = source side at its start, -1 none; the last matching combination wins.
__device__ __forceinline__ int shared_ends(const Elem& o, int e, int ne, const Elem& s, int f, int nf, double oa,
                                           double ob, double sa, double sb) {
    int shared = -1;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const double to = a == 0 ? oa : ob, ts = b == 0 ? sa : sb;
            const double dx = (o.x0 + to * o.sx) - (s.x0 + ts * s.sx);
            const double dy = (o.y0 + to * o.sy) - (s.y0 + ts * s.sy);
            const double d = sqrt(dx * dx + dy * dy);
            const int nodo = to == 0.0 ? e : (to == 1.0 ? ne : -1);
            const int nods = ts == 0.0 ? f : (ts == 1.0 ? nf : -1);
            if (d <= 1e-12 || (nodo >= 0 && nodo == nods && d <= 2.0e-9)) shared = (a == 0 ? 1 : 0) | (b == 0 ? 2 : 0);
        }
    return shared;
}

template <int W>
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Modulus of this lane's complex entry (lanes 2m, 2m + 1 hold entry m), 0 beyond nv.
__device__ __forceinline__ double entry_mod(double v, int lane, int nv) {
    const double sq = v * v;
    const double other = __shfl_xor(sq, 1, 64);
    return lane < nv ? sqrt(sq + other) : 0.0;
}

// One polynomial block (polynomial_quadrature.py:58-120) over the intervals [oa, ob] x [sa, sb] with
// the Gauss rule of `order`: returns this lane's entry (see above).  `same`: a self pair on equal
// intervals; `shared`: shared_ends of a distinct pair.  red: 4 W^2 x 65 doubles of LDS.
template <int W>
__device__ double poly_block(const Elem& o, const Elem& s, double k, double oa, double ob, double sa, double sb,
                             int order, bool same, int shared, double (*red)[65], unsigned* status) {
    constexpr int NV = 4 * W * W;
    const int lane = threadIdx.x;
    const double* T = rule_t(order);
    const double* Wt = rule_w(order);
    const bool singular = same || shared >= 0;
    const bool exact_log = same && oa == 0.0 && ob == 1.0 && sa == 0.0 && sb == 1.0;
    const int sq = order * order;
    const int count = singular ? 2 * sq : sq;
    const double ho = ob - oa, hs = sb - sa;
    // corner geometry of a shared pair
    double orx = 0.0, ory = 0.0, dox = 0.0, doy = 0.0, dsx = 0.0, dsy = 0.0;
    const bool os = (shared & 1) != 0, ss = (shared & 2) != 0;
    if (!same && shared >= 0) {
        const double to = os ? oa : ob, ts = ss ? sa : sb;
        const double eox = o.x0 + to * o.sx, eoy = o.y0 + to * o.sy;
        const double esx = s.x0 + ts * s.sx, esy = s.y0 + ts * s.sy;
        orx = eox - esx;
        ory = eoy - esy;
        const double big = fmax(fmax(sqrt(eox * eox + eoy * eoy), sqrt(esx * esx + esy * esy)), fmax(o.len, s.len));
        if (sqrt(orx * orx + ory * ory) <= 64.0 * 2.220446049250313e-16 * big) { orx = 0.0; ory = 0.0; }
        dox = ho * o.sx * (os ? 1.0 : -1.0);
        doy = ho * o.sy * (os ? 1.0 : -1.0);
        dsx = hs * s.sx * (ss ? 1.0 : -1.0);
        dsy = hs * s.sy * (ss ? 1.0 : -1.0);
    }
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    bool bad = false;
    for (int pt = lane; pt < count; pt += 64) {
        double x, y, wgt;
        if (singular) {
            const int half = pt >= sq ? 1 : 0;
            const int r = pt - half * sq;
            const int i = r / order, j = r - (r / order) * order;
            const double t = T[i], t2 = t * t, t4 = t2 * t2;
            const double v = T[j];
            wgt = 4.0 * (t4 * t2 * t) * Wt[i] * Wt[j];
            x = half ? t4 * v : t4;
            y = half ? t4 : t4 * v;
        } else {
            const int i = pt / order, j = pt - (pt / order) * order;
            x = T[i];
            y = T[j];
            wgt = Wt[i] * Wt[j];
        }
        double dx, dy;
        if (same) {
            const double dd = ho * (x - y);
            dx = dd * o.sx;
            dy = dd * o.sy;
        } else if (shared >= 0) {
            dx = orx + x * dox - y * dsx;
            dy = ory + x * doy - y * dsy;
            if (!os) x = 1.0 - x;
            if (!ss) y = 1.0 - y;
        }
        x = oa + ho * x;
        y = sa + hs * y;
        if (!same && shared < 0) {
            dx = (o.x0 - s.x0) + x * o.sx - y * s.sx;
            dy = (o.y0 - s.y0) + x * o.sy - y * s.sy;
        }
        const double dist = sqrt(dx * dx + dy * dy);
        if (!(dist > 0.0)) {
            bad = true;
            continue;
        }
        const HK v = hankel_quarter_v(k * dist);
        double gr = v.gr;
        const double gi = v.gi;
        if (exact_log) gr -= log(fabs(x - y)) * (0.5 / GRIM_PI);
        double kr_ = 0.0, ki_ = 0.0;
        if (!same) {
            const double po = (dx * o.nx + dy * o.ny) / dist;
            kr_ = -k * v.hr * po;
            ki_ = -k * v.hi * po;
        }
        double po_[W], ps_[W];
        basis_at<W>(x, po_);
        basis_at<W>(y, ps_);
#pragma unroll
        for (int a = 0; a < W; ++a) {
            const double ca = wgt * po_[a];
#pragma unroll
            for (int b = 0; b < W; ++b) {
                const double c = ca * ps_[b];
                const int m = 2 * (a * W + b);
                acc[m] = fma(c, gr, acc[m]);
                acc[m + 1] = fma(c, gi, acc[m + 1]);
                acc[2 * W * W + m] = fma(c, kr_, acc[2 * W * W + m]);
                acc[2 * W * W + m + 1] = fma(c, ki_, acc[2 * W * W + m + 1]);
            }
        }
    }
    if (bad) atomicOr(status, (unsigned)GRIM_DEV_POLY_INTERSECT);
    // lane-ordered reduction through LDS: entry i is summed by lane i over lanes 0..63
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][lane] = acc[i];
    __syncthreads();
    double sum = 0.0;
    if (lane < NV)
        for (int l = 0; l < 64; ++l) sum += red[lane][l];
    __syncthreads();
    if (exact_log && lane < 2 * W * W && (lane & 1) == 0) {
        const int m = lane >> 1;
        sum += Poly<W>::logm(m / W, m % W) * (0.5 / GRIM_PI);
    }
    return sum * (o.len * s.len * ho * hs);
}

template <int W>
__global__ void __launch_bounds__(64) poly_near(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                const int* __restrict__ next, const int* __restrict__ offsets,
                                                const int* __restrict__ pair_e, const int* __restrict__ pair_f,
                                                int capacity, double k, double* __restrict__ bs,
                                                double* __restrict__ bk, double* __restrict__ bte,
                                                unsigned* __restrict__ status) {
    constexpr int NV = 4 * W * W;
    __shared__ double red[NV][65];
    __shared__ double stk[POLY_STACK][4];
    __shared__ int sdep[POLY_STACK];
    __shared__ double fin[NV];
    const int idx = blockIdx.x;
    // an overflowing list was not written (near_list flags it): nothing to read then
    if (idx >= offsets[n] || offsets[n] > capacity) return;
    const int lane = threadIdx.x;
    const int e = pair_e[idx], f = pair_f[idx];
    if ((unsigned)e >= (unsigned)n || (unsigned)f >= (unsigned)n) return;
    const int ne = next[e], nf = next[f];
    const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
    const bool self = e == f;
    const double floor_scale = o.len * s.len * 1e-10;
    double total = 0.0;
    bool failed = false;
    int top = 0;
    if (lane == 0) {
        stk[0][0] = 0.0; stk[0][1] = 1.0; stk[0][2] = 0.0; stk[0][3] = 1.0;
        sdep[0] = 0;
    }
    top = 1;
    __syncthreads();
    while (top > 0) {
        --top;
        const double oa = stk[top][0], ob = stk[top][1], sa = stk[top][2], sb = stk[top][3];
        const int depth = sdep[top];
        __syncthreads();
        const bool same = self && oa == sa && ob == sb;
        const int shared = same ? -1 : shared_ends(o, e, ne, s, f, nf, oa, ob, sa, sb);
        const double low = poly_block<W>(o, s, k, oa, ob, sa, sb, 20, same, shared, red, status);
        const double high = poly_block<W>(o, s, k, oa, ob, sa, sb, 36, same, shared, red, status);
        const double scale = fmax(wave_max<W>(entry_mod(high, lane, NV)), floor_scale);
        double err = wave_max<W>(entry_mod(low - high, lane, NV));
        if (err <= POLY_RTOL * scale) {
            total += high;
            continue;
        }
        if (self || depth >= POLY_MAX_DEPTH) {
            double finest = poly_block<W>(o, s, k, oa, ob, sa, sb, 72, same, shared, red, status);
            err = wave_max<W>(entry_mod(high - finest, lane, NV));
            if (err > POLY_RTOL * scale) {
                const double extra = poly_block<W>(o, s, k, oa, ob, sa, sb, 144, same, shared, red, status);
                err = wave_max<W>(entry_mod(finest - extra, lane, NV));
                finest = extra;
            }
            if (err > POLY_RTOL * scale) failed = true;
            total += finest;
            continue;
        }
        if (top + 2 > POLY_STACK) {   // cannot happen below the depth limit; reported, never silent
            failed = true;
            total += high;
            continue;
        }
        // bisect the physically longer interval; the left child is taken first
        if (lane == 0) {
            if ((ob - oa) * o.len >= (sb - sa) * s.len) {
                const double mid = 0.5 * (oa + ob);
                stk[top][0] = mid; stk[top][1] = ob; stk[top][2] = sa; stk[top][3] = sb;
                stk[top + 1][0] = oa; stk[top + 1][1] = mid; stk[top + 1][2] = sa; stk[top + 1][3] = sb;
            } else {
                const double mid = 0.5 * (sa + sb);
                stk[top][0] = oa; stk[top][1] = ob; stk[top][2] = mid; stk[top][3] = sb;
                stk[top + 1][0] = oa; stk[top + 1][1] = ob; stk[top + 1][2] = sa; stk[top + 1][3] = mid;
            }
            sdep[top] = depth + 1;
            sdep[top + 1] = depth + 1;
        }
        top += 2;
        __syncthreads();
    }
    if (failed && lane == 0) atomicOr(status, (unsigned)GRIM_DEV_POLY_NEAR);
    if (lane < NV) fin[lane] = total;
    __syncthreads();
    const int64_t base = 2 * ((int64_t)f * n + e) * (W * W);
    if (lane < 2 * W * W) {
        bs[base + lane] = fin[lane];
        bk[base + lane] = fin[2 * W * W + lane];
    }
    if (lane < W * W) {
        const double nn = o.nx * s.nx + o.ny * s.ny;
        double lolen = o.len * s.len;
        if (lolen < GRIM_EPS * GRIM_EPS) lolen = GRIM_EPS * GRIM_EPS;
        const cplx t = te_entry<W>(fin, fin + 2 * W * W, lane / W, lane % W, k, nn, lolen);
        bte[base + 2 * lane] = t.re;
        bte[base + 2 * lane + 1] = t.im;
    }
}

// ---------------------------------------------------------------------------
// Node gather: Z_TM[p, q] and Z_TE[p, q] from the blocks of the (at most two) elements of each
// node, in a fixed order, with the element mass on e == f.
// ---------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ int node_owners(int p, int n, const int* __restrict__ prev, int (&el)[2], int (&loc)[2]) {
    if (p < n) {
        el[0] = p; loc[0] = 0;
        el[1] = prev[p]; loc[1] = 1;
        return 2;
    }
    const int r = p - n;
    el[0] = r / (W - 2);
    loc[0] = 2 + r % (W - 2);
    return 1;
}

template <int W>
__global__ void __launch_bounds__(256) poly_gather(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                   const int* __restrict__ prev, double k,
                                                   const double* __restrict__ bs, const double* __restrict__ bk,
                                                   const double* __restrict__ bte, double* __restrict__ ztm,
                                                   int64_t ldtm, double* __restrict__ zte, int64_t ldte) {
    const int nodes = n * (W - 1);
    const int p = blockIdx.x * 256 + threadIdx.x, q = blockIdx.y;
    if (p >= nodes) return;
    int re_[2], ra[2], ce[2], cb[2];
    const int nr = node_owners<W>(p, n, prev, re_, ra);
    const int nc = node_owners<W>(q, n, prev, ce, cb);
    const double g = -1.0 / k;   // gamma = (0, g)
    double tmr = 0.0, tmi = 0.0, ter = 0.0, tei = 0.0;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const int e = re_[r], a = ra[r], f = ce[c], b = cb[c];
            const int64_t ef = 2 * (((int64_t)f * n + e) * (W * W) + a * W + b);
            const int64_t fe = 2 * (((int64_t)e * n + f) * (W * W) + b * W + a);
            const double sr = bs[ef], si = bs[ef + 1];
            const double dr = bk[fe], di = bk[fe + 1];
            tmr += sr - g * di;
            tmi += si + g * dr;
            ter += bte[ef];
            tei += bte[ef + 1];
            if (e == f) {
                const double ex = p0[2 * e] - p1[2 * e], ey = p0[2 * e + 1] - p1[2 * e + 1];
                const double m = sqrt(ex * ex + ey * ey) * Poly<W>::mass(a, b);
                tmi += 0.5 * g * m;
                ter -= 0.5 * m;
            }
        }
    const int64_t at = 2 * ((int64_t)p + (int64_t)q * ldtm);
    ztm[at] = tmr;
    ztm[at + 1] = tmi;
    const int64_t at2 = 2 * ((int64_t)p + (int64_t)q * ldte);
    zte[at2] = ter;
    zte[at2 + 1] = tei;
}

// ---------------------------------------------------------------------------
// Plane-wave loads: b_u[p, a] = sum over the owners (e, l) of p of the Gauss moment
// L_e e^{j k c_e.d} sum_i (w_i / 2) e^{j (t_i - 1/2) k (p1 - p0).d} phi_l(t_i)  (basis.py:89-104),
// b_dn the same times j k n_e.d; d = (cos phi, sin phi), t_i / w_i the Gauss rule on [-1, 1]
// mapped to [0, 1] (w_i / 2 = the mapped weight).
// ---------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(256) poly_loads(int n, int na, const double* __restrict__ p0,
                                                  const double* __restrict__ p1, const int* __restrict__ prev,
                                                  double k, const double* __restrict__ angles_deg, int order,
                                                  const double* __restrict__ qt, const double* __restrict__ qw,
                                                  double* __restrict__ bu, double* __restrict__ bdn, int64_t ld) {
    const int nodes = n * (W - 1);
    const int p = blockIdx.x * 256 + threadIdx.x, a = blockIdx.y;
    if (p >= nodes) return;
    const double phi_a = angles_deg[a] * (GRIM_PI / 180.0);
    double dy, dx;
    sincos(phi_a, &dy, &dx);
    int el[2], loc[2];
    const int no = node_owners<W>(p, n, prev, el, loc);
    double ur = 0.0, ui = 0.0, vr = 0.0, vi = 0.0;
    for (int r = 0; r < no; ++r) {
        const Elem E = load_elem(p0, p1, el[r]);
        const double z = k * (E.sx * dx + E.sy * dy);
        double mr = 0.0, mi = 0.0;
        for (int i = 0; i < order; ++i) {
            double phi[W];
            basis_at<W>(qt[i], phi);
            double sn, cs;
            sincos((qt[i] - 0.5) * z, &sn, &cs);
            const double c = qw[i] * phi[loc[r]];
            mr = fma(c, cs, mr);
            mi = fma(c, sn, mi);
        }
        double sb, cb;
        sincos(k * (E.cx * dx + E.cy * dy), &sb, &cb);
        const double br = E.len * cb, bi = E.len * sb;
        const double m_re = br * mr - bi * mi, m_im = br * mi + bi * mr;
        ur += m_re;
        ui += m_im;
        const double dn = k * (E.nx * dx + E.ny * dy);   // j k n.d = (0, dn)
        vr += -dn * m_im;
        vi += dn * m_re;
    }
    const int64_t at = 2 * ((int64_t)p + (int64_t)a * ld);
    if (bu) { bu[at] = ur; bu[at + 1] = ui; }
    if (bdn) { bdn[at] = vr; bdn[at + 1] = vi; }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <int W>
static hipError_t launch_poly(int n, int q, const double* p0, const double* p1, const int* next, const int* prev,
                              double k, const int* offsets, int capacity, const int* pair_e, const int* pair_f,
                              double* bs, double* bk, double* bte, unsigned* status, double* ztm, int64_t ldtm,
                              double* zte, int64_t ldte, hipStream_t st) {
    const dim3 gf((n + 63) / 64, n);
    if (q == 9) poly_far<W, 9><<<gf, 64, 0, st>>>(n, p0, p1, next, k, bs, bk, bte);
    else poly_far<W, 16><<<gf, 64, 0, st>>>(n, p0, p1, next, k, bs, bk, bte);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    poly_near<W><<<capacity, 64, 0, st>>>(n, p0, p1, next, offsets, pair_e, pair_f, capacity, k, bs, bk, bte,
                                          status);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    const int nodes = n * (W - 1);
    poly_gather<W><<<dim3((nodes + 255) / 256, nodes), 256, 0, st>>>(n, p0, p1, prev, k, bs, bk, bte, ztm, ldtm,
                                                                     zte, ldte);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_poly(int degree, int n, int q, const double* p0, const double* p1, const int* next,
                                       const int* prev, double k, const int* offsets, int capacity,
                                       const int* pair_e, const int* pair_f, double* bs, double* bk, double* bte,
                                       unsigned* status, double* ztm, int64_t ldtm, double* zte, int64_t ldte,
                                       hipStream_t st) {
    if (degree == 2)
        return launch_poly<3>(n, q, p0, p1, next, prev, k, offsets, capacity, pair_e, pair_f, bs, bk, bte, status,
                              ztm, ldtm, zte, ldte, st);
    return launch_poly<4>(n, q, p0, p1, next, prev, k, offsets, capacity, pair_e, pair_f, bs, bk, bte, status, ztm,
                          ldtm, zte, ldte, st);
}

extern "C" hipError_t grim_launch_poly_loads(int degree, int n, int na, const double* p0, const double* p1,
                                             const int* prev, double k, const double* angles_deg, int order,
                                             const double* qt, const double* qw, double* bu, double* bdn, int64_t ld,
                                             hipStream_t st) {
    const int nodes = n * degree;
    const dim3 g((nodes + 255) / 256, na);
    if (degree == 2) poly_loads<3><<<g, 256, 0, st>>>(n, na, p0, p1, prev, k, angles_deg, order, qt, qw, bu, bdn, ld);
    else poly_loads<4><<<g, 256, 0, st>>>(n, na, p0, p1, prev, k, angles_deg, order, qt, qw, bu, bdn, ld);
    return hipGetLastError();
}
