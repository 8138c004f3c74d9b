// Single layer S, K', D and the Maue W of closed contours at a complex wavenumber k = k0 n (a lossy
// medium, n = n' - j n'', n' > 0, n'' > 0 under e^{+j omega t}): the operators grim_assemble_operators
// forms for a lossless medium, with the complex-argument Hankel functions of the second kind.
//
// Replaces, for complex k, the reference's fused operator traversal (tools/GHOST/ghost_backend/
// twod/operators.py:2087-2607: far tiles, near pairs, scatter), its complex Hankel evaluation
// (twod/special.py:220-264, twod/operators.py:898-945: G = (j/4) H0^(2)(k r), dG = (j/4) k H1^(2)(k r),
// k r floored at 1e-12) and the exact self series at complex z = k L (operators.py:418-501).
//
// Complex Hankel functions (hankel2_quarter_c): |z| < 2, the power series of J0, J1, Y0, Y1
// (DLMF 10.2.2, 10.8.1; 8e-15 relative to SciPy's Amos-based hankel2 over the fourth quadrant);
// 2 <= |z| < 20, the integral form of DLMF 10.9.E11 on the 32-point generalized Gauss-Laguerre rule
// of gen_lossy_rules.py (4e-15); |z| >= 20, 20 terms of the asymptotic expansion (DLMF 10.17.4).
// All are written for Re z > 0, Im z <= 0.
//
// Layout: the near-pair lists and linear near blocks of the real path (near_list), then two tables of
// element-pair blocks, TS (S) and TK (K', observation normal) of n x n blocks of 2 x 2 complex, block
// (e, f) at (f n + e).  The far pass evaluates each unordered far pair once and writes (e, f) and
// (f, e) of both tables; the near blocks fill the rest.  A gather forms the node entries of the
// closed contours (node p = start of element p and end of element prev[p]) in a fixed order, with
// D(e, f) = K'(f, e)^T and the mass terms of grim_assemble_operators.
#include "common.h"
#include "lossy_rules.h"

__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {   // Smith's division
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

__device__ __forceinline__ cplx csqrt_p(cplx w) {   // principal square root
    const double r = hypot(w.re, w.im);
    if (w.re >= 0.0) {
        const double s = sqrt(0.5 * (r + w.re));
        return {s, s > 0.0 ? 0.5 * w.im / s : 0.0};
    }
    const double t = sqrt(0.5 * (r - w.re));
    return {0.5 * fabs(w.im) / t, w.im >= 0.0 ? t : -t};
}

__device__ __forceinline__ cplx clog_p(cplx z) { return {log(hypot(z.re, z.im)), atan2(z.im, z.re)}; }

__device__ __forceinline__ cplx cexp_mi(cplx z) {   // exp(-i z) = e^{Im z} (cos Re z - i sin Re z)
    double s, c;
    sincos(z.re, &s, &c);
    const double m = exp(z.im);
    return {m * c, -m * s};
}

struct HC {
    cplx g, h1;   // (j/4) H0^(2)(z), (j/4) H1^(2)(z)
};

__device__ __noinline__ HC hankel2_quarter_c(cplx z) {
    const double az = hypot(z.re, z.im);
    cplx H0, H1;
    if (az < 2.0) {
        const cplx y = cscale(cmul(z, z), -0.25);
        cplx t = {1.0, 0.0}, j0 = {0.0, 0.0}, j1s = {0.0, 0.0}, s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
        double harm = 0.0;
        for (int k = 0; k < 24; ++k) {
            if (k > 0) {
                t = cscale(cmul(t, y), 1.0 / ((double)k * (double)k));
                harm += 1.0 / k;
            }
            const double inv = 1.0 / (k + 1);
            j0 = cadd(j0, t);
            j1s = cadd(j1s, cscale(t, inv));
            s0 = cadd(s0, cscale(t, harm));
            s1 = cadd(s1, cscale(t, (2.0 * harm + inv) * inv));
        }
        const cplx j1 = cmul(j1s, cscale(z, 0.5));
        cplx L = clog_p(cscale(z, 0.5));
        L.re += 0.5772156649015329;
        const double tp = 2.0 / GRIM_PI;
        const cplx y0 = cscale(csub(cmul(L, j0), s0), tp);
        cplx y1 = cscale(cmul(L, j1), tp);
        y1 = csub(y1, cdiv(cplx{tp, 0.0}, z));
        y1 = csub(y1, cscale(cmul(z, s1), 0.5 / GRIM_PI));
        H0 = {j0.re + y0.im, j0.im - y0.re};   // J - i Y
        H1 = {j1.re + y1.im, j1.im - y1.re};
    } else if (az >= 20.0) {
        const cplx u = cdiv(cplx{0.0, -1.0}, z);   // -i / z
        cplx zk = {1.0, 0.0}, s0 = {1.0, 0.0}, s1 = {1.0, 0.0};
        for (int k = 1; k < 20; ++k) {
            zk = cmul(zk, u);
            s0.re = fma(GRIM_HASYM0[k], zk.re, s0.re);
            s0.im = fma(GRIM_HASYM0[k], zk.im, s0.im);
            s1.re = fma(GRIM_HASYM1[k], zk.re, s1.re);
            s1.im = fma(GRIM_HASYM1[k], zk.im, s1.im);
        }
        const cplx pre = csqrt_p(cdiv(cplx{2.0 / GRIM_PI, 0.0}, z));
        const cplx e0 = cmul(pre, cexp_mi(z));
        const cplx r4 = {GRIM_SQRT_HALF, GRIM_SQRT_HALF}, r34 = {-GRIM_SQRT_HALF, GRIM_SQRT_HALF};
        H0 = cmul(cmul(e0, r4), s0);
        H1 = cmul(cmul(e0, r34), s1);
    } else {
        const cplx c = cdiv(cplx{0.0, -0.5}, z);   // -i / (2 z)
        cplx i0 = {0.0, 0.0}, i1 = {0.0, 0.0};
        for (int q = 0; q < 32; ++q) {
            const double x = GRIM_LAG32_X[q], wq = GRIM_LAG32_W[q];
            const cplx w = {fma(x, c.re, 1.0), x * c.im};
            const cplx sw = csqrt_p(w);                 // w^(1/2); w^(-1/2) = conj(sw) / |w|
            const double inv = wq / hypot(w.re, w.im);
            i0.re = fma(inv, sw.re, i0.re);
            i0.im = fma(-inv, sw.im, i0.im);
            i1.re = fma(wq * x, sw.re, i1.re);
            i1.im = fma(wq * x, sw.im, i1.im);
        }
        const cplx pre = csqrt_p(cdiv(cplx{2.0 / GRIM_PI, 0.0}, z));
        const cplx e0 = cmul(pre, cexp_mi(z));
        const cplx r4 = {GRIM_SQRT_HALF, GRIM_SQRT_HALF}, r34 = {-GRIM_SQRT_HALF, GRIM_SQRT_HALF};
        const double rpi = 0.56418958354775628695;   // 1 / Gamma(1/2); 1 / Gamma(3/2) = 2 / sqrt(pi)
        H0 = cscale(cmul(cmul(e0, r4), i0), rpi);
        H1 = cscale(cmul(cmul(e0, r34), i1), 2.0 * rpi);
    }
    return HC{{-0.25 * H0.im, 0.25 * H0.re}, {-0.25 * H1.im, 0.25 * H1.re}};
}

// g = (j/4) H0^(2)(k r), h = (j/4) k H1^(2)(k r); r floored at GRIM_EPS, |k r| at 1e-12.
__device__ __forceinline__ void green_c(cplx k, double r, cplx& g, cplx& h) {
    const double rr = r > GRIM_EPS ? r : GRIM_EPS;
    cplx z = cscale(k, rr);
    if (hypot(z.re, z.im) <= 1e-12) z = {1e-12, 0.0};
    const HC v = hankel2_quarter_c(z);
    g = v.g;
    h = cmul(v.h1, k);
}

// Exact linear-Galerkin self S block at complex z = k L, |z| <= 8 (operators.py:418-501): the series of
// the real case with complex z, log and alpha.
__device__ bool self_block_c(double len, cplx k, cplx (&S)[2][2]) {
    const cplx z = cscale(k, len);
    const double az = hypot(z.re, z.im);
    if (!(az <= 8.0) || az <= 1e-30) return false;
    const double cd[4] = {2.0 / 3.0, -1.0, 0.0, 1.0 / 3.0};
    const double co[4] = {1.0 / 3.0, 0.0, 0.0, -1.0 / 3.0};
    const double tp = 2.0 / GRIM_PI;
    cplx lt = clog_p(cscale(z, 0.5));
    lt.re += 0.5772156649015329;
    const cplx zq = cscale(cmul(z, z), 0.25);
    cplx bd = {0, 0}, bo = {0, 0}, alpha = {1.0, 0.0};
    double harmonic = 0.0;
    for (int m = 0;;) {
        const int p = 2 * m;
        // a_m = alpha (1 - j tp (lt - H_m))
        const cplx f = {1.0 + tp * lt.im, -tp * (lt.re - harmonic)};
        const cplx am = cmul(alpha, f);
        double md = 0, mo = 0, ld = 0, lo = 0;
        for (int q = 0; q < 4; ++q) {
            const double dd = (double)(p + q + 1);
            md += cd[q] / dd;
            mo += co[q] / dd;
            ld -= cd[q] / (dd * dd);
            lo -= co[q] / (dd * dd);
        }
        // a_m moment - j tp alpha log_moment
        bd.re += am.re * md + tp * alpha.im * ld;
        bd.im += am.im * md - tp * alpha.re * ld;
        bo.re += am.re * mo + tp * alpha.im * lo;
        bo.im += am.im * mo - tp * alpha.re * lo;
        ++m;
        alpha = cmul(alpha, cscale(zq, -1.0 / ((double)m * (double)m)));
        harmonic += 1.0 / m;
        if (m > 60) return false;
        const double mag = hypot(bd.re, bd.im);
        if (hypot(alpha.re, alpha.im) < 1e-18 * (mag > 1.0 ? mag : 1.0)) break;
    }
    const double c = 0.25 * len * len;   // 0.25 j L^2 [[bd, bo], [bo, bd]]
    const cplx d = {-c * bd.im, c * bd.re}, o = {-c * bo.im, c * bo.re};
    S[0][0] = d; S[0][1] = o; S[1][0] = o; S[1][1] = d;
    return true;
}

struct SKc {
    double v[16];   // S[a][b] re, im (8) then K'[a][b] re, im (8)
};

__device__ __forceinline__ void acc_point(double (&acc)[16], double c0a, double c1a, double ps0, double ps1,
                                          const cplx& g, const cplx& dk) {
    const double po[2] = {c0a, c1a}, ps[2] = {ps0, ps1};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const double c = po[a] * ps[b];
            acc[2 * (2 * a + b)] += c * g.re;
            acc[2 * (2 * a + b) + 1] += c * g.im;
            acc[8 + 2 * (2 * a + b)] += c * dk.re;
            acc[8 + 2 * (2 * a + b) + 1] += c * dk.im;
        }
}

// Tensor Gauss 16 x 16 box on sub-intervals, wave-parallel (the real path's box16 at complex k).
__device__ void box16_c(const Elem& o, const Elem& s, cplx k, double oa, double ob, double sa, double sb, SKc& out) {
    const int lane = threadIdx.x & 63;
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0;
    const double ho = ob - oa, hs = sb - sa;
    for (int pt = lane; pt < 256; pt += 64) {
        const int i = pt >> 4, j = pt & 15;
        const double xo = oa + ho * GRIM_GL16_T[i], xs = sa + hs * GRIM_GL16_T[j];
        const double w = GRIM_GL16_W[i] * GRIM_GL16_W[j];
        const double dx = (o.x0 + xo * o.sx) - (s.x0 + xs * s.sx);
        const double dy = (o.y0 + xo * o.sy) - (s.y0 + xs * s.sy);
        const double dist = sqrt(dx * dx + dy * dy);
        cplx g, h;
        green_c(k, dist, g, h);
        const double safe = dist > GRIM_EPS ? dist : GRIM_EPS;
        const double proj = (dx * o.nx + dy * o.ny) / safe;
        cplx dk = {-h.re * proj, -h.im * proj};
        if (dist <= GRIM_EPS) dk = {0.0, 0.0};
        acc_point(acc, w * (1.0 - xo), w * xo, 1.0 - xs, xs, g, dk);
    }
    const double scale = o.len * ho * s.len * hs;
#pragma unroll
    for (int i = 0; i < 16; ++i) out.v[i] = wave_sum(acc[i]) * scale;
}

// The graded t^4 two-triangle corner rule of the real path (operators.py:549-757), its 800 points
// formed from the 20-point Gauss rule exactly as grim_init forms them for the constant table.
__device__ void touching16_c(const Elem& o, const Elem& s, cplx k, bool obs_start, bool src_start, SKc& out) {
    const int lane = threadIdx.x & 63;
    const double cox = obs_start ? o.x0 : o.x0 + o.sx, coy = obs_start ? o.y0 : o.y0 + o.sy;
    const double csx = src_start ? s.x0 : s.x0 + s.sx, csy = src_start ? s.y0 : s.y0 + s.sy;
    const double dox = obs_start ? o.sx : -o.sx, doy = obs_start ? o.sy : -o.sy;
    const double dsx = src_start ? s.sx : -s.sx, dsy = src_start ? s.sy : -s.sy;
    double orx = cox - csx, ory = coy - csy;
    const double eps = 2.220446049250313e-16;
    const double big = fmax(fmax(sqrt(cox * cox + coy * coy), sqrt(csx * csx + csy * csy)), fmax(o.len, s.len));
    if (sqrt(orx * orx + ory * ory) <= 64.0 * eps * big) { orx = 0.0; ory = 0.0; }
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0;
    for (int pt = lane; pt < 800; pt += 64) {
        const int a = pt % 400, i = a / 20, j = a % 20;
        const double t = GRIM_GL20_T[i], v = GRIM_GL20_T[j];
        const double u = t * t * t * t;
        const double w = 4.0 * (t * t * t * t * t * t * t) * GRIM_GL20_W[i] * GRIM_GL20_W[j];
        const double x = pt < 400 ? u : u * v, y = pt < 400 ? u * v : u;
        const double dx = orx + x * dox - y * dsx;
        const double dy = ory + x * doy - y * dsy;
        const double xo = obs_start ? x : 1.0 - x, xs = src_start ? y : 1.0 - y;
        const double dist = sqrt(dx * dx + dy * dy);
        const double safe = dist > GRIM_EPS ? dist : GRIM_EPS;
        cplx g, h;
        green_c(k, safe, g, h);
        const double proj = (dx * o.nx + dy * o.ny) / safe;
        cplx dk = {-h.re * proj, -h.im * proj};
        if (dist <= GRIM_EPS) dk = {0.0, 0.0};
        acc_point(acc, w * (1.0 - xo), w * xo, 1.0 - xs, xs, g, dk);
    }
    const double scale = o.len * s.len;
#pragma unroll
    for (int i = 0; i < 16; ++i) out.v[i] = wave_sum(acc[i]) * scale;
}

// Touching test and adaptive criterion of the real path (geometric, wavenumber-free).
__device__ __forceinline__ bool touching_c(const Elem& o, int e, int ne, const Elem& s, int f, int nf,
                                           bool& obs_start, bool& src_start) {
    const double ox[2] = {o.x0, o.x0 + o.sx}, oy[2] = {o.y0, o.y0 + o.sy};
    const double sx[2] = {s.x0, s.x0 + s.sx}, sy[2] = {s.y0, s.y0 + s.sy};
    const int on[2] = {e, ne}, sn[2] = {f, nf};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const double dx = ox[a] - sx[b], dy = oy[a] - sy[b];
            const double d = sqrt(dx * dx + dy * dy);
            if (d <= 1e-9 || (on[a] == sn[b] && d <= 2e-9)) {
                obs_start = a == 0;
                src_start = b == 0;
                return true;
            }
        }
    return false;
}

__device__ __forceinline__ double point_seg_sq_c(double x, double y, double ax, double ay, double bx, double by) {
    const double ex = bx - ax, ey = by - ay;
    const double den = ex * ex + ey * ey;
    double t = den > 0 ? ((x - ax) * ex + (y - ay) * ey) / den : 0.0;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    const double dx = x - ax - t * ex, dy = y - ay - t * ey;
    return dx * dx + dy * dy;
}

// separation.py:18-78
__device__ bool requires_adaptive_c(const Elem& o, const Elem& s) {
    const double lo = o.len, ls = s.len;
    const double scale = lo > ls ? lo : ls;
    const double cdx = o.cx - s.cx, cdy = o.cy - s.cy;
    const double distance = sqrt(cdx * cdx + cdy * cdy);
    if (distance < 0.75 * scale) return true;
    if (distance >= 0.5 * (lo + ls) + 0.5 * scale) return false;
    const double p0x = o.x0, p0y = o.y0, p1x = o.x0 + o.sx, p1y = o.y0 + o.sy;
    const double q0x = s.x0, q0y = s.y0, q1x = s.x0 + s.sx, q1y = s.y0 + s.sy;
    if (distance < 0.5 * (lo + ls) + 0.25 * scale) {
        double sq = point_seg_sq_c(p0x, p0y, q0x, q0y, q1x, q1y);
        sq = fmin(sq, point_seg_sq_c(p1x, p1y, q0x, q0y, q1x, q1y));
        sq = fmin(sq, point_seg_sq_c(q0x, q0y, p0x, p0y, p1x, p1y));
        sq = fmin(sq, point_seg_sq_c(q1x, q1y, p0x, p0y, p1x, p1y));
        const double ux = p1x - p0x, uy = p1y - p0y, vx = q1x - q0x, vy = q1y - q0y;
        const double wx = q0x - p0x, wy = q0y - p0y;
        const double den = ux * vy - uy * vx;
        double segd = sqrt(fmax(sq, 0.0));
        if (den != 0) {
            const double t = (wx * vy - wy * vx) / den, sp = (wx * uy - wy * ux) / den;
            if (t >= 0 && t <= 1 && sp >= 0 && sp <= 1) segd = 0.0;
        }
        if (segd < 0.25 * scale) return true;
    }
    double best = 1e300;
    const double X[4] = {p0x, p1x, q0x, q1x}, Y[4] = {p0y, p1y, q0y, q1y};
    const double AX[4] = {q0x, q0x, p0x, p0x}, AY[4] = {q0y, q0y, p0y, p0y};
    const double BX[4] = {q1x, q1x, p1x, p1x}, BY[4] = {q1y, q1y, p1y, p1y};
    for (int i = 0; i < 4; ++i) {
        const double ex = BX[i] - AX[i], ey = BY[i] - AY[i];
        const double den = ex * ex + ey * ey;
        const double t = den > 0 ? ((X[i] - AX[i]) * ex + (Y[i] - AY[i]) * ey) / den : -1.0;
        if (t > 0.0 && t < 1.0) {
            const double dx = X[i] - AX[i] - t * ex, dy = Y[i] - AY[i] - t * ey;
            best = fmin(best, sqrt(dx * dx + dy * dy));
        }
    }
    return best < 0.5 * scale;
}

// Near blocks: 0 self, 1 touching, 2 fixed, 3 adaptive (integrated by lossy_near_adaptive).
__global__ void __launch_bounds__(64) lossy_near_blocks(int n, const double* __restrict__ p0,
                                                        const double* __restrict__ p1, const int* __restrict__ next,
                                                        const int* __restrict__ offsets,
                                                        const int* __restrict__ pair_e,
                                                        const int* __restrict__ pair_f, int capacity, double kr,
                                                        double ki, double* __restrict__ blocks,
                                                        int* __restrict__ kind, unsigned* __restrict__ status) {
    const int idx = blockIdx.x;
    if (idx >= offsets[n] || offsets[n] > capacity) return;   // overflow: the lists were not written
    const cplx k = {kr, ki};
    const int e = pair_e[idx], f = pair_f[idx];
    const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
    SKc r;
    int cls;
    bool os = false, ss = false;
    if (e == f) {
        cls = 0;
        cplx S[2][2];
        if (!self_block_c(o.len, k, S)) {
            if (threadIdx.x == 0) atomicOr(status, (unsigned)GRIM_DEV_SELF_SERIES);
            for (int i = 0; i < 16; ++i) r.v[i] = 0.0;
        } else {
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    r.v[2 * (2 * a + b)] = S[a][b].re;
                    r.v[2 * (2 * a + b) + 1] = S[a][b].im;
                }
            for (int i = 8; i < 16; ++i) r.v[i] = 0.0;
        }
    } else if (touching_c(o, e, next[e], s, f, next[f], os, ss)) {
        cls = 1;
        touching16_c(o, s, k, os, ss, r);
    } else if (requires_adaptive_c(o, s)) {
        if (threadIdx.x == 0) kind[idx] = 3;
        return;
    } else {
        cls = 2;
        box16_c(o, s, k, 0.0, 1.0, 0.0, 1.0, r);
    }
    if (threadIdx.x < 16) blocks[16 * (int64_t)idx + threadIdx.x] = r.v[threadIdx.x];
    if (threadIdx.x == 0) kind[idx] = cls;
}

// Recursive four-way bisection, rtol 1e-9, depth 12 (operators.py:1172-1292), as the real path's
// near_adaptive: one workgroup of four wavefronts per adaptive pair, the four child boxes at once.
#define GRIM_LOSSY_STACK 48
__global__ void __launch_bounds__(256) lossy_near_adaptive(int n, const double* __restrict__ p0,
                                                           const double* __restrict__ p1,
                                                           const int* __restrict__ offsets,
                                                           const int* __restrict__ pair_e,
                                                           const int* __restrict__ pair_f, int capacity, double kr,
                                                           double ki, double* __restrict__ blocks,
                                                           const int* __restrict__ kind,
                                                           unsigned* __restrict__ status) {
    __shared__ double st_iv[GRIM_LOSSY_STACK][4];
    __shared__ double st_blk[GRIM_LOSSY_STACK][16];
    __shared__ int st_depth[GRIM_LOSSY_STACK];
    __shared__ double kid_blk[4][16];
    const cplx k = {kr, ki};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int listed = offsets[n];
    if (listed > capacity) return;                          // overflow: the lists were not written
    for (int idx = blockIdx.x; idx < listed; idx += gridDim.x) {
        if (kind[idx] != 3) continue;                      // uniform over the workgroup
        const int e = pair_e[idx], f = pair_f[idx];
        const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
        const double floor_ = 2.220446049250313e-16 * fmax(1.0, o.len * s.len);
        if (wave == 0) {
            SKc root;
            box16_c(o, s, k, 0.0, 1.0, 0.0, 1.0, root);
            if (lane == 0) {
                st_iv[0][0] = 0.0; st_iv[0][1] = 1.0; st_iv[0][2] = 0.0; st_iv[0][3] = 1.0;
                st_depth[0] = 0;
            }
            if (lane < 16) st_blk[0][lane] = root.v[lane];
        }
        int top = 1;
        __syncthreads();
        double total[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) total[i] = 0.0;
        bool failed = false;
        while (top > 0) {
            --top;
            const double oa = st_iv[top][0], ob = st_iv[top][1], sa = st_iv[top][2], sb = st_iv[top][3];
            const int depth = st_depth[top];
            double coarse[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) coarse[i] = st_blk[top][i];
            const double om = 0.5 * (oa + ob), sm = 0.5 * (sa + sb);
            const double iv[4][4] = {{oa, om, sa, sm}, {oa, om, sm, sb}, {om, ob, sa, sm}, {om, ob, sm, sb}};
            {
                SKc kid;
                box16_c(o, s, k, iv[wave][0], iv[wave][1], iv[wave][2], iv[wave][3], kid);
                if (lane < 16) kid_blk[wave][lane] = kid.v[lane];
            }
            __syncthreads();
            double err = 0.0;
            for (int part = 0; part < 2; ++part) {
                double diff2 = 0.0, denom = 0.0;
                for (int c = 0; c < 4; ++c) {
                    double n2 = 0.0;
                    for (int i = 0; i < 8; ++i) n2 += kid_blk[c][8 * part + i] * kid_blk[c][8 * part + i];
                    denom += sqrt(n2);
                }
                for (int i = 0; i < 8; ++i) {
                    const double fine = kid_blk[0][8 * part + i] + kid_blk[1][8 * part + i] +
                                        kid_blk[2][8 * part + i] + kid_blk[3][8 * part + i];
                    const double d = fine - coarse[8 * part + i];
                    diff2 += d * d;
                }
                err = fmax(err, sqrt(diff2) / fmax(denom, floor_));
            }
            if (err <= 1e-9 || depth >= 12 || top + 4 > GRIM_LOSSY_STACK) {
                if (err > 1e-9) failed = true;
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int i = 0; i < 16; ++i) total[i] += kid_blk[c][i];
            } else {
                for (int c = 3; c >= 0; --c) {
                    if (tid == 0) {
                        for (int j = 0; j < 4; ++j) st_iv[top][j] = iv[c][j];
                        st_depth[top] = depth + 1;
                    }
                    if (tid < 16) st_blk[top][tid] = kid_blk[c][tid];
                    ++top;
                }
            }
            __syncthreads();
        }
        if (failed && tid == 0) atomicOr(status, (unsigned)GRIM_DEV_ADAPTIVE_DEPTH);
        if (tid < 16) blocks[16 * (int64_t)idx + tid] = total[tid];
        __syncthreads();                                   // the stack is reused by the next pair
    }
}

// Far pairs: one thread per unordered far pair e < f (e = lane of the block row, f = blockIdx.y): the
// Q x Q rule gives S(e, f), K'(e, f) (normal of e) and K'(f, e) (normal of f); writes (e, f) and
// (f, e) of both tables.
template <int Q>
__global__ void __launch_bounds__(64) lossy_far(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                const int* __restrict__ next, double kr, double ki,
                                                double* __restrict__ ts, double* __restrict__ tk) {
    __shared__ double st[Q], sw[Q];
    const int lane = threadIdx.x;
    if (lane < Q) {
        st[lane] = Q == 8 ? GRIM_GL8_T[lane] : GRIM_GL16_T[lane];
        sw[lane] = Q == 8 ? GRIM_GL8_W[lane] : GRIM_GL16_W[lane];
    }
    __syncthreads();
    const int e = blockIdx.x * 64 + lane, f = blockIdx.y;
    if (e >= n || e >= f) return;
    const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
    if (!is_far(o, e, next[e], s, f, next[f])) return;   // the near pass lists exactly the others
    const cplx k = {kr, ki};
    const int64_t ef = 8 * ((int64_t)f * n + e), fe = 8 * ((int64_t)e * n + f);
    // Negligible kernels (twod/operators.py:1355-1363, 2373-2379): with lower = centre distance less
    // half of each length, -Im(k) lower >= 40 bounds |G| below e^{-40} of the near scale; the
    // reference neither evaluates nor scatters such pairs, so their blocks are zero.
    if (ki < 0.0) {
        const double cdx = o.cx - s.cx, cdy = o.cy - s.cy;
        const double lower = sqrt(cdx * cdx + cdy * cdy) - 0.5 * (o.len + s.len);
        if (-ki * lower >= 40.0) {
#pragma unroll
            for (int i = 0; i < 8; ++i) ts[ef + i] = ts[fe + i] = tk[ef + i] = tk[fe + i] = 0.0;
            return;
        }
    }
    cplx S[2][2], Ke[2][2], Kf[2][2];   // Kf[b][a]: observation basis b of f, source basis a of e
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) S[a][b] = Ke[a][b] = Kf[a][b] = {0.0, 0.0};
    for (int i = 0; i < Q; ++i) {
        const double ti = st[i], wi = sw[i];
        const double xo = fma(ti, o.sx, o.x0), yo = fma(ti, o.sy, o.y0);
        for (int j = 0; j < Q; ++j) {
            const double tj = st[j], w = wi * sw[j];
            const double dx = xo - fma(tj, s.sx, s.x0), dy = yo - fma(tj, s.sy, s.y0);
            const double r = sqrt(dx * dx + dy * dy);
            cplx g, h;
            green_c(k, r, g, h);
            const double pe = (dx * o.nx + dy * o.ny) / r, pf = -(dx * s.nx + dy * s.ny) / r;
            const cplx de = {-h.re * pe, -h.im * pe}, df = {-h.re * pf, -h.im * pf};
            const double po[2] = {w * (1.0 - ti), w * ti}, ps[2] = {1.0 - tj, tj};
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const double c = po[a] * ps[b];
                    S[a][b].re = fma(c, g.re, S[a][b].re);
                    S[a][b].im = fma(c, g.im, S[a][b].im);
                    Ke[a][b].re = fma(c, de.re, Ke[a][b].re);
                    Ke[a][b].im = fma(c, de.im, Ke[a][b].im);
                    Kf[b][a].re = fma(c, df.re, Kf[b][a].re);
                    Kf[b][a].im = fma(c, df.im, Kf[b][a].im);
                }
        }
    }
    const double scale = o.len * s.len;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int m = 2 * (2 * a + b), t = 2 * (2 * b + a);
            ts[ef + m] = S[a][b].re * scale;
            ts[ef + m + 1] = S[a][b].im * scale;
            ts[fe + t] = S[a][b].re * scale;
            ts[fe + t + 1] = S[a][b].im * scale;
            tk[ef + m] = Ke[a][b].re * scale;
            tk[ef + m + 1] = Ke[a][b].im * scale;
            tk[fe + m] = Kf[a][b].re * scale;
            tk[fe + m + 1] = Kf[a][b].im * scale;
        }
}

// Near pairs: S and K' parts of each listed pair's block into the tables.
__global__ void __launch_bounds__(256) lossy_near_store(int n, const int* __restrict__ offsets, int capacity,
                                                        const int* __restrict__ pair_e,
                                                        const int* __restrict__ pair_f,
                                                        const double* __restrict__ blocks, double* __restrict__ ts,
                                                        double* __restrict__ tk) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= offsets[n] || offsets[n] > capacity) return;   // overflow: the lists were not written
    const int e = pair_e[idx], f = pair_f[idx];
    const double* src = blocks + 16 * (int64_t)idx;
    const int64_t at = 8 * ((int64_t)f * n + e);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        ts[at + i] = src[i];
        tk[at + i] = src[8 + i];
    }
}

// Node gather of closed contours (node p: start of element p, local 0, and end of element prev[p],
// local 1): s = S, kp = K' + c_k M, d = D + c_d M with D(e, f) = K'(f, e)^T, w = the Maue W
// -k^2 (n_e . n_f) S + s_a s_b sum(S) / (L_e L_f); c from coef_* per element, else the scalars.
// Any output may be null.  Every entry has one writer.
__global__ void __launch_bounds__(256) lossy_gather(int n, const double* __restrict__ p0,
                                                    const double* __restrict__ p1, const int* __restrict__ prev,
                                                    double kr, double ki, const double* __restrict__ ts,
                                                    const double* __restrict__ tk, double mass_k, double mass_d,
                                                    const double* __restrict__ coef_k,
                                                    const double* __restrict__ coef_d, double* __restrict__ s_out,
                                                    double* __restrict__ k_out, double* __restrict__ d_out,
                                                    double* __restrict__ w_out, int64_t ld) {
    const int p = blockIdx.x * 256 + threadIdx.x, q = blockIdx.y;
    if (p >= n) return;
    const cplx k2 = cmul(cplx{kr, ki}, cplx{kr, ki});
    cplx sa = {0, 0}, ka = {0, 0}, da = {0, 0}, wa = {0, 0};
    const int re_[2] = {p, prev[p]}, ce[2] = {q, prev[q]};
    for (int r = 0; r < 2; ++r) {
        const int e = re_[r], a = r;
        const Elem o = load_elem(p0, p1, e);
        for (int c = 0; c < 2; ++c) {
            const int f = ce[c], b = c;
            const int64_t ef = 8 * ((int64_t)f * n + e), fe = 8 * ((int64_t)e * n + f);
            const int m = 2 * (2 * a + b), t = 2 * (2 * b + a);
            sa.re += ts[ef + m];
            sa.im += ts[ef + m + 1];
            ka.re += tk[ef + m];
            ka.im += tk[ef + m + 1];
            da.re += tk[fe + t];
            da.im += tk[fe + t + 1];
            if (w_out) {
                const Elem s = load_elem(p0, p1, f);
                double denom = o.len * s.len;
                if (denom < GRIM_EPS * GRIM_EPS) denom = GRIM_EPS * GRIM_EPS;
                const double sr = (ts[ef] + ts[ef + 2]) + (ts[ef + 4] + ts[ef + 6]);
                const double si = (ts[ef + 1] + ts[ef + 3]) + (ts[ef + 5] + ts[ef + 7]);
                const double nn = o.nx * s.nx + o.ny * s.ny;
                const cplx kk = {-k2.re * nn, -k2.im * nn};
                const cplx v = cmul(kk, cplx{ts[ef + m], ts[ef + m + 1]});
                const double tsign = a == b ? 1.0 : -1.0;
                wa.re += v.re + tsign * (sr / denom);
                wa.im += v.im + tsign * (si / denom);
            }
            if (e == f) {
                const double mm = o.len * (a == b ? 1.0 / 3.0 : 1.0 / 6.0);
                ka.re += (coef_k ? coef_k[e] : mass_k) * mm;
                da.re += (coef_d ? coef_d[e] : mass_d) * mm;
            }
        }
    }
    const int64_t at = 2 * ((int64_t)p + (int64_t)q * ld);
    if (s_out) { s_out[at] = sa.re; s_out[at + 1] = sa.im; }
    if (k_out) { k_out[at] = ka.re; k_out[at + 1] = ka.im; }
    if (d_out) { d_out[at] = da.re; d_out[at + 1] = da.im; }
    if (w_out) { w_out[at] = wa.re; w_out[at + 1] = wa.im; }
}

// Complex Hankel probe: out[i] = (j/4) H0^(2)(z_i), (j/4) H1^(2)(z_i) (4 doubles per point).
__global__ void __launch_bounds__(256) hankel2_probe(int m, const double* __restrict__ z, double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const HC v = hankel2_quarter_c(cplx{z[2 * i], z[2 * i + 1]});
    out[4 * i] = v.g.re;
    out[4 * i + 1] = v.g.im;
    out[4 * i + 2] = v.h1.re;
    out[4 * i + 3] = v.h1.im;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
extern "C" hipError_t grim_launch_lossy(int q, int n, const double* p0, const double* p1, const int* next,
                                        const int* prev, double kr, double ki, const int* offsets, int capacity,
                                        const int* pair_e, const int* pair_f, double* blocks, int* kind,
                                        unsigned* status, double* ts, double* tk, double mass_k, double mass_d,
                                        const double* coef_k, const double* coef_d, double* s_out, double* k_out,
                                        double* d_out, double* w_out, int64_t ld, hipStream_t st) {
    lossy_near_blocks<<<capacity, 64, 0, st>>>(n, p0, p1, next, offsets, pair_e, pair_f, capacity, kr, ki, blocks,
                                                kind, status);
    lossy_near_adaptive<<<capacity < 1024 ? capacity : 1024, 256, 0, st>>>(n, p0, p1, offsets, pair_e, pair_f,
                                                                           capacity, kr, ki, blocks, kind, status);
    const dim3 gf((n + 63) / 64, n);
    if (q == 16) lossy_far<16><<<gf, 64, 0, st>>>(n, p0, p1, next, kr, ki, ts, tk);
    else lossy_far<8><<<gf, 64, 0, st>>>(n, p0, p1, next, kr, ki, ts, tk);
    lossy_near_store<<<(capacity + 255) / 256, 256, 0, st>>>(n, offsets, capacity, pair_e, pair_f, blocks, ts, tk);
    lossy_gather<<<dim3((n + 255) / 256, n), 256, 0, st>>>(n, p0, p1, prev, kr, ki, ts, tk, mass_k, mass_d, coef_k,
                                                           coef_d, s_out, k_out, d_out, w_out, ld);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_hankel2_probe(int m, const double* z, double* out, hipStream_t st) {
    hankel2_probe<<<(m + 255) / 256, 256, 0, st>>>(m, z, out);
    return hipGetLastError();
}
