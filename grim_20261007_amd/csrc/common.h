// Shared device helpers of the GRIM-2D HIP library (gfx950, FP64 VALU).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hankel_coeffs.h"

#define GRIM_EPS 1e-12
#define GRIM_PI 3.14159265358979323846
#define GRIM_SQRT_HALF 0.70710678118654752440

// Status codes of the C ABI (same style as the reference's native seam:
// 0 ok, 1 index/size out of range, 3 invalid arguments; plus HIP failures).
enum {
    GRIM_OK = 0,
    GRIM_ERR_RANGE = 1,
    GRIM_ERR_ARGS = 3,
    GRIM_ERR_HIP = 5,
    GRIM_ERR_NOT_CONVERGED = 6,
    GRIM_ERR_CAPACITY = 7,
    GRIM_ERR_SINGULAR = 8,
    GRIM_ERR_TIMEOUT = 9,
    GRIM_ERR_GEOMETRY = 10,
};

// Device status word bits (written by kernels, read by the host once per call).
enum {
    GRIM_DEV_ADAPTIVE_DEPTH = 1,
    GRIM_DEV_NEAR_CAPACITY = 2,
    GRIM_DEV_SELF_SERIES = 4,
    GRIM_DEV_ZERO_PIVOT = 8,
    GRIM_DEV_BARRIER_TIMEOUT = 16,
    GRIM_DEV_POLY_NEAR = 32,        // polynomial near block did not converge (refine the mesh)
    GRIM_DEV_POLY_INTERSECT = 64,   // polynomial quadrature met coincident integration points
};

// Launch-shape knobs (grim_set_launch_shape): far-pass source columns per
// wavefront, keyed by n; columns per wavefront of the factorization's 64-row
// triangular solves, keyed by the total columns of the call; columns per
// wavefront of the LU solve's 256-row block solves, keyed by the total
// right-hand-side columns.  Shapes change only how work is spread over
// wavefronts: every entry is computed by the same operations.
enum {
    GRIM_SHAPE_FAR_COLUMNS = 0,
    GRIM_SHAPE_TRSM_COLUMNS = 1,
    GRIM_SHAPE_SOLVE_COLUMNS = 2,
    GRIM_SHAPE_KNOBS = 3
};
// The table's value for `size` (largest entry with min_size <= size), 0 = built-in rule.
extern "C" int grim_shape_value(int knob, int64_t size);

struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }

template <int N>
__device__ __forceinline__ double horner(const double (&c)[N], double t) {
    double r = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) r = fma(r, t, c[i]);
    return r;
}

// (j/4) H0^(2)(x) and (j/4) H1^(2)(x) for real x > 0 (x = k r, floored at 1e-12).
// Returns g = (Y0/4, J0/4) and h1 = (Y1/4, J1/4); the caller multiplies h1 by k.
// Small x: Taylor series of the entire parts (DLMF 10.8.1); x >= 6: modulus /
// phase form with fitted P_nu, Q_nu (DLMF 10.17.3), 12 terms on [6, 20) and 6
// terms from 20 on, the common case of far pairs.  Coefficients generated by
// gen_hankel_coeffs.py (max relative error 3e-15 for x up to 50; beyond, the
// eps * x conditioning of the phase dominates).
struct HK { double gr, gi, hr, hi; };
__device__ __noinline__ HK hankel_quarter_v(double x) {
    cplx g, h1;
    if (x < GRIM_HANKEL_XS) {
        const double y = 0.25 * x * x;
        const double L = log(0.5 * x);
        const double j0v = horner(GRIM_HK_A0, y);
        const double j1v = 0.5 * x * horner(GRIM_HK_A1, y);
        const double y0v = (2.0 / GRIM_PI) * L * j0v + horner(GRIM_HK_B0, y);
        const double y1v = (2.0 / GRIM_PI) * L * j1v - 2.0 / (GRIM_PI * x) + 0.5 * x * horner(GRIM_HK_B1, y);
        g = {0.25 * y0v, 0.25 * j0v};
        h1 = {0.25 * y1v, 0.25 * j1v};
    } else {
        // 1/sqrt(x) by the hardware estimate and two Newton steps; 1/x = its square
        double rs = __builtin_amdgcn_rsq(x);
        rs = rs * fma(-0.5 * x, rs * rs, 1.5);
        rs = rs * fma(-0.5 * x, rs * rs, 1.5);
        const double inv = rs * rs;
        const double amp = (0.7978845608028654 * GRIM_SQRT_HALF) * rs;
        double p0, q0, p1, q1;
        if (x < GRIM_HANKEL_XM) {
            const double r = GRIM_HANKEL_XS * inv;
            const double u = fma(r * r, GRIM_HANKEL_MID_SCALE, GRIM_HANKEL_MID_SHIFT);
            p0 = horner(GRIM_HK_P0M, u);
            q0 = r * horner(GRIM_HK_Q0M, u);
            p1 = horner(GRIM_HK_P1M, u);
            q1 = r * horner(GRIM_HK_Q1M, u);
        } else {
            const double r = GRIM_HANKEL_XM * inv;
            const double w = r * r;
            p0 = horner(GRIM_HK_P0, w);
            q0 = r * horner(GRIM_HK_Q0, w);
            p1 = horner(GRIM_HK_P1, w);
            q1 = r * horner(GRIM_HK_Q1, w);
        }
        // cos x + sin x and cos x - sin x: x = n pi/2 + t, |t| <= pi/4 (two-part
        // Cody-Waite reduction with FMA, exact to an ulp of t for n < 2^30), then
        // Taylor polynomials of sin t and cos t (remainders below 7e-17)
        double cps, cms;
        if (x < 1.0e9) {
            const double n = rint(x * 0.6366197723675814);
            double t = fma(-n, 1.5707963267948966, x);
            t = fma(-n, 6.123233995736766e-17, t);
            const double t2 = t * t;
            double ps = -7.647163731819816e-13;
            ps = fma(ps, t2, 1.6059043836821613e-10);
            ps = fma(ps, t2, -2.505210838544172e-08);
            ps = fma(ps, t2, 2.7557319223985893e-06);
            ps = fma(ps, t2, -1.984126984126984e-04);
            ps = fma(ps, t2, 8.333333333333333e-03);
            ps = fma(ps, t2, -1.6666666666666666e-01);
            const double sn = fma(ps * t2, t, t);
            double pc = 4.779477332387385e-14;
            pc = fma(pc, t2, -1.1470745597729725e-11);
            pc = fma(pc, t2, 2.08767569878681e-09);
            pc = fma(pc, t2, -2.755731922398589e-07);
            pc = fma(pc, t2, 2.48015873015873e-05);
            pc = fma(pc, t2, -1.388888888888889e-03);
            pc = fma(pc, t2, 4.1666666666666664e-02);
            const double cs = fma(fma(pc, t2, -0.5), t2, 1.0);
            const int q = (int)n & 3;
            const double a = cs + sn, b = cs - sn;
            cps = (q & 1) ? b : a;
            cms = (q & 1) ? -a : b;
            if (q & 2) { cps = -cps; cms = -cms; }
        } else {
            double s0, c0;
            sincos(x, &s0, &c0);
            cps = c0 + s0;
            cms = c0 - s0;
        }
        // H0 = amp (P0 - j Q0) (u0 + j v0), u0 = c + s, v0 = c - s
        const double u0 = cps, v0 = cms;
        const double h0re = amp * (p0 * u0 + q0 * v0), h0im = amp * (p0 * v0 - q0 * u0);
        // H1 = amp (P1 - j Q1) (u1 + j v1), u1 = s - c, v1 = c + s
        const double u1 = -cms, v1 = cps;
        const double h1re = amp * (p1 * u1 + q1 * v1), h1im = amp * (p1 * v1 - q1 * u1);
        g = {-0.25 * h0im, 0.25 * h0re};
        h1 = {-0.25 * h1im, 0.25 * h1re};
    }
    return HK{g.re, g.im, h1.re, h1.im};
}
// Two arguments at once, for the far quadrature: the same operations as hankel_quarter_v on each
// argument (so the same values, bit for bit), written as two independent chains that the
// scheduler interleaves; the range forms are selected per argument instead of branched.
struct HK2 { HK a, b; };

__device__ __forceinline__ void hk_phase(double x, double& cps, double& cms) {
    const double n = rint(x * 0.6366197723675814);
    double t = fma(-n, 1.5707963267948966, x);
    t = fma(-n, 6.123233995736766e-17, t);
    const double t2 = t * t;
    double ps = -7.647163731819816e-13;
    ps = fma(ps, t2, 1.6059043836821613e-10);
    ps = fma(ps, t2, -2.505210838544172e-08);
    ps = fma(ps, t2, 2.7557319223985893e-06);
    ps = fma(ps, t2, -1.984126984126984e-04);
    ps = fma(ps, t2, 8.333333333333333e-03);
    ps = fma(ps, t2, -1.6666666666666666e-01);
    const double sn = fma(ps * t2, t, t);
    double pc = 4.779477332387385e-14;
    pc = fma(pc, t2, -1.1470745597729725e-11);
    pc = fma(pc, t2, 2.08767569878681e-09);
    pc = fma(pc, t2, -2.755731922398589e-07);
    pc = fma(pc, t2, 2.48015873015873e-05);
    pc = fma(pc, t2, -1.388888888888889e-03);
    pc = fma(pc, t2, 4.1666666666666664e-02);
    const double cs = fma(fma(pc, t2, -0.5), t2, 1.0);
    const int q = (int)n & 3;
    const double a = cs + sn, b = cs - sn;
    cps = (q & 1) ? b : a;
    cms = (q & 1) ? -a : b;
    if (q & 2) { cps = -cps; cms = -cms; }
}

__device__ __forceinline__ HK hk_combine(double amp, double p0, double q0, double p1, double q1, double cps,
                                         double cms) {
    const double u0 = cps, v0 = cms;
    const double h0re = amp * (p0 * u0 + q0 * v0), h0im = amp * (p0 * v0 - q0 * u0);
    const double u1 = -cms, v1 = cps;
    const double h1re = amp * (p1 * u1 + q1 * v1), h1im = amp * (p1 * v1 - q1 * u1);
    return HK{-0.25 * h0im, 0.25 * h0re, -0.25 * h1im, 0.25 * h1re};
}

__device__ __forceinline__ HK hk_small(double x) {
    const double y = 0.25 * x * x;
    const double L = log(0.5 * x);
    const double j0v = horner(GRIM_HK_A0, y);
    const double j1v = 0.5 * x * horner(GRIM_HK_A1, y);
    const double y0v = (2.0 / GRIM_PI) * L * j0v + horner(GRIM_HK_B0, y);
    const double y1v = (2.0 / GRIM_PI) * L * j1v - 2.0 / (GRIM_PI * x) + 0.5 * x * horner(GRIM_HK_B1, y);
    return HK{0.25 * y0v, 0.25 * j0v, 0.25 * y1v, 0.25 * j1v};
}

__device__ __forceinline__ HK hk_select(bool c, const HK& u, const HK& v) {
    return HK{c ? u.gr : v.gr, c ? u.gi : v.gi, c ? u.hr : v.hr, c ? u.hi : v.hi};
}

__device__ __noinline__ HK2 hankel_quarter_v2(double xa, double xb) {
    double rsa = __builtin_amdgcn_rsq(xa), rsb = __builtin_amdgcn_rsq(xb);
    rsa = rsa * fma(-0.5 * xa, rsa * rsa, 1.5);
    rsb = rsb * fma(-0.5 * xb, rsb * rsb, 1.5);
    rsa = rsa * fma(-0.5 * xa, rsa * rsa, 1.5);
    rsb = rsb * fma(-0.5 * xb, rsb * rsb, 1.5);
    const double inva = rsa * rsa, invb = rsb * rsb;
    const double ampa = (0.7978845608028654 * GRIM_SQRT_HALF) * rsa;
    const double ampb = (0.7978845608028654 * GRIM_SQRT_HALF) * rsb;
    // x >= 20 form for both
    const double ra = GRIM_HANKEL_XM * inva, rb = GRIM_HANKEL_XM * invb;
    const double wa = ra * ra, wb = rb * rb;
    double p0a = horner(GRIM_HK_P0, wa), p0b = horner(GRIM_HK_P0, wb);
    double q0a = ra * horner(GRIM_HK_Q0, wa), q0b = rb * horner(GRIM_HK_Q0, wb);
    double p1a = horner(GRIM_HK_P1, wa), p1b = horner(GRIM_HK_P1, wb);
    double q1a = ra * horner(GRIM_HK_Q1, wa), q1b = rb * horner(GRIM_HK_Q1, wb);
    if (xa < GRIM_HANKEL_XM || xb < GRIM_HANKEL_XM) {
        const double sa = GRIM_HANKEL_XS * inva, sb = GRIM_HANKEL_XS * invb;
        const double ua = fma(sa * sa, GRIM_HANKEL_MID_SCALE, GRIM_HANKEL_MID_SHIFT);
        const double ub = fma(sb * sb, GRIM_HANKEL_MID_SCALE, GRIM_HANKEL_MID_SHIFT);
        const double mp0a = horner(GRIM_HK_P0M, ua), mp0b = horner(GRIM_HK_P0M, ub);
        const double mq0a = sa * horner(GRIM_HK_Q0M, ua), mq0b = sb * horner(GRIM_HK_Q0M, ub);
        const double mp1a = horner(GRIM_HK_P1M, ua), mp1b = horner(GRIM_HK_P1M, ub);
        const double mq1a = sa * horner(GRIM_HK_Q1M, ua), mq1b = sb * horner(GRIM_HK_Q1M, ub);
        const bool ma = xa < GRIM_HANKEL_XM, mb = xb < GRIM_HANKEL_XM;
        p0a = ma ? mp0a : p0a; q0a = ma ? mq0a : q0a; p1a = ma ? mp1a : p1a; q1a = ma ? mq1a : q1a;
        p0b = mb ? mp0b : p0b; q0b = mb ? mq0b : q0b; p1b = mb ? mp1b : p1b; q1b = mb ? mq1b : q1b;
    }
    double cpsa, cmsa, cpsb, cmsb;
    hk_phase(xa, cpsa, cmsa);
    hk_phase(xb, cpsb, cmsb);
    if (xa >= 1.0e9 || xb >= 1.0e9) {
        double s0, c0;
        if (xa >= 1.0e9) { sincos(xa, &s0, &c0); cpsa = c0 + s0; cmsa = c0 - s0; }
        if (xb >= 1.0e9) { sincos(xb, &s0, &c0); cpsb = c0 + s0; cmsb = c0 - s0; }
    }
    HK a = hk_combine(ampa, p0a, q0a, p1a, q1a, cpsa, cmsa);
    HK b = hk_combine(ampb, p0b, q0b, p1b, q1b, cpsb, cmsb);
    if (xa < GRIM_HANKEL_XS || xb < GRIM_HANKEL_XS) {
        if (xa < GRIM_HANKEL_XS) a = hk_small(xa);
        if (xb < GRIM_HANKEL_XS) b = hk_small(xb);
    }
    return HK2{a, b};
}

// green_kernels for two distances at once (see hankel_quarter_v2).
__device__ __forceinline__ void green_kernels2(double k, double ra, double rb, cplx& ga, cplx& ha, cplx& gb,
                                               cplx& hb) {
    double xa = k * (ra > GRIM_EPS ? ra : GRIM_EPS), xb = k * (rb > GRIM_EPS ? rb : GRIM_EPS);
    if (xa < 1e-12) xa = 1e-12;
    if (xb < 1e-12) xb = 1e-12;
    const HK2 v = hankel_quarter_v2(xa, xb);
    ga = {v.a.gr, v.a.gi};
    ha = cscale(cplx{v.a.hr, v.a.hi}, k);
    gb = {v.b.gr, v.b.gi};
    hb = cscale(cplx{v.b.hr, v.b.hi}, k);
}

__device__ __forceinline__ void hankel_quarter(double x, cplx& g, cplx& h1) {
    const HK v = hankel_quarter_v(x);
    g = {v.gr, v.gi};
    h1 = {v.hr, v.hi};
}

// Kernels at distance r for real k: g = (j/4) H0(k r), h = (j/4) k H1(k r).
__device__ __forceinline__ void green_kernels(double k, double r, cplx& g, cplx& h) {
    double rr = r > GRIM_EPS ? r : GRIM_EPS;
    double x = k * rr;
    if (x < 1e-12) x = 1e-12;
    cplx h1;
    hankel_quarter(x, g, h1);
    h = cscale(h1, k);
}

// Geometry of one straight element (element e runs from node e to node next[e]).
struct Elem {
    double x0, y0, sx, sy;   // start point and p1 - p0
    double len, nx, ny;      // length and stored normal (-t_y, t_x)
    double cx, cy;           // centre 0.5 (p0 + p1)
};

__device__ __forceinline__ Elem load_elem(const double* p0, const double* p1, int e) {
    Elem E;
    E.x0 = p0[2 * e];
    E.y0 = p0[2 * e + 1];
    const double x1 = p1[2 * e], y1 = p1[2 * e + 1];
    E.sx = x1 - E.x0;
    E.sy = y1 - E.y0;
    E.len = sqrt(E.sx * E.sx + E.sy * E.sy);
    const double tx = E.sx / E.len, ty = E.sy / E.len;
    E.nx = -ty;
    E.ny = tx;
    E.cx = 0.5 * (E.x0 + x1);
    E.cy = 0.5 * (E.y0 + y1);
    return E;
}

// Far pair test of the fused traversal: centre distance / max length >= 3,
// no shared mesh node, not the same element.
__device__ __forceinline__ bool is_far(const Elem& a, int ea, int na, const Elem& b, int eb, int nb) {
    if (ea == eb) return false;
    if (ea == eb || ea == nb || na == eb || na == nb) return false;
    const double dx = a.cx - b.cx, dy = a.cy - b.cy;
    const double dist = sqrt(dx * dx + dy * dy);
    double scale = a.len > b.len ? a.len : b.len;
    if (scale < GRIM_EPS) scale = GRIM_EPS;
    return dist / scale >= 3.0;
}

// Plane-wave moments of straight elements
__device__ __forceinline__ void sinc_d(double u, double& f, double& fp) {
    if (fabs(u) < 0.01) {
        const double v2 = u * u;
        f = 1 + v2 * (-1.0 / 6 + v2 * (1.0 / 120 + v2 * (-1.0 / 5040 + v2 / 362880)));
        fp = u * (-1.0 / 6 + v2 * (1.0 / 60 + v2 * (-1.0 / 1680 + v2 * (1.0 / 90720 - v2 / 7983360))));
    } else {
        double s, c;
        sincos(u, &s, &c);
        f = s / u;
        fp = (u * c - s) / (2 * u * u);
    }
}

// Moments of element e for direction (dx, dy): i0, i1 and j k d.n (kernels.py:20-39, 260-285).
__device__ __forceinline__ void elem_moments(const Elem& E, double k, double dx, double dy, cplx& i0, cplx& i1,
                                             cplx& dn) {
    const double z = k * (E.sx * dx + E.sy * dy);
    double f, fp;
    sinc_d(0.5 * z, f, fp);
    double s, c;
    sincos(k * (E.cx * dx + E.cy * dy), &s, &c);
    const cplx ph = {E.len * c, E.len * s};
    i0 = cmul(ph, cplx{0.5 * f, fp});
    i1 = cmul(ph, cplx{0.5 * f, -fp});
    dn = {0.0, k * (E.nx * dx + E.ny * dy)};
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
