// Impedance-matrix fill, plane-wave loads and far-field projection for closed
// PEC contours (TM: S + gamma (D + M/2), TE: -M/2 + K' - gamma W, gamma = 1/(jk)).
//
// Replaces the reference's CPU fill: the far pass of _assemble_multi with its
// native block quadrature (tools/GHOST/ghost_backend/twod/operators.py:2335-2518,
// twod/assembly/native/far.c:100-207), the near-pair rules
// (operators.py:418-501, 549-757, 1172-1292, 1971-2085), the routed scatter
// (assembly/scatter.py:303-368, formulations/combined_regions.py:149-213), the
// plane-wave loads (assembly/kernels.py:222-294) and the matched far field
// (assembly/kernels.py:162-219).
//
// Layout: element e runs from node e to node next[e]; every node starts exactly
// one element and ends exactly one (closed contours).  Matrices are column-major
// complex128 (interleaved re, im), Z[p + q * ld].
#include "common.h"

__constant__ double c_gl8_t[8], c_gl8_w[8], c_gl16_t[16], c_gl16_w[16];
__constant__ double c_touch_x[800], c_touch_y[800], c_touch_w[800];

template <int Q> struct Rule;
template <> struct Rule<8> {
    static __device__ __forceinline__ double t(int i) { return c_gl8_t[i]; }
    static __device__ __forceinline__ double w(int i) { return c_gl8_w[i]; }
};
template <> struct Rule<16> {
    static __device__ __forceinline__ double t(int i) { return c_gl16_t[i]; }
    static __device__ __forceinline__ double w(int i) { return c_gl16_w[i]; }
};

struct Blk {   // TM and TE contributions of one element pair, [obs local][src local]
    cplx tm[2][2], te[2][2];
};

__device__ __forceinline__ void zero_blk(Blk& b) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            b.tm[a][c] = {0.0, 0.0};
            b.te[a][c] = {0.0, 0.0};
        }
}

// W from S (Maue identity, operators.py:1295-1321, 1943-1967): S already scaled.
__device__ __forceinline__ void combine(const cplx (&S)[2][2], const cplx (&K)[2][2], const cplx (&D)[2][2],
                                        const Elem& o, const Elem& s, double k, Blk& out) {
    const double g = -1.0 / k;   // gamma = 1/(jk) = (0, -1/k)
    const double nn = o.nx * s.nx + o.ny * s.ny;
    double denom = o.len * s.len;
    if (denom < GRIM_EPS * GRIM_EPS) denom = GRIM_EPS * GRIM_EPS;
    const cplx sum = cadd(cadd(S[0][0], S[0][1]), cadd(S[1][0], S[1][1]));
    const cplx q = {sum.re / denom, sum.im / denom};
    const double kk = -k * k * nn;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const double tsign = (a == b) ? 1.0 : -1.0;
            const cplx w = {kk * S[a][b].re + tsign * q.re, kk * S[a][b].im + tsign * q.im};
            // gamma * D = (0, g) * D = (-g D.im, g D.re)
            out.tm[a][b] = {S[a][b].re - g * D[a][b].im, S[a][b].im + g * D[a][b].re};
            out.te[a][b] = {K[a][b].re + g * w.im, K[a][b].im - g * w.re};
        }
}

// Far blocks of one unordered element pair, both orientations:
//   tm  = S + gamma D,        te  = K' - gamma W            (obs e, src f)
//   tmt = (S + gamma K')^T,   tet = (D - gamma W)^T         (obs f, src e)
// since S(f,e) = S(e,f)^T, K'(f,e) = D(e,f)^T, D(f,e) = K'(e,f)^T, W(f,e) = W(e,f)^T.
struct Blk2 {
    cplx tm[2][2], te[2][2], tmt[2][2], tet[2][2];
};

__device__ __forceinline__ void zero_blk2(Blk2& b) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) b.tm[a][c] = b.te[a][c] = b.tmt[a][c] = b.tet[a][c] = {0.0, 0.0};
}

// Far element-pair quadrature with the Q x Q Gauss rule (far.c:100-191 semantics,
// kernels evaluated directly instead of through an interpolation table).  MODE 0 gives the
// combined potential; MODE 1 the operators S and K' (tm = S, te = K', tmt = S^T, tet = D^T,
// the K' of the reversed pair); MODE 2 the operators D and W (tm = D, te = W, tmt = K'^T,
// the D of the reversed pair, tet = W^T).
template <int Q, int MODE = 0>
__device__ __forceinline__ void far_pair(const Elem& o, const Elem& s, double k, Blk2& out) {
    cplx S[2][2], K[2][2], D[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) S[a][b] = K[a][b] = D[a][b] = {0.0, 0.0};
    for (int qi = 0; qi < Q; ++qi) {
        const double ti = Rule<Q>::t(qi), wi = Rule<Q>::w(qi);
        const double xo = fma(ti, o.sx, o.x0), yo = fma(ti, o.sy, o.y0);
        cplx ps0 = {0, 0}, ps1 = {0, 0}, pk0 = {0, 0}, pk1 = {0, 0}, pd0 = {0, 0}, pd1 = {0, 0};
        // two source points per step (Q is even): the pair's Hankel evaluations run as two
        // independent chains; the sums take the points in order, as one point per step would
        // unrolled for Q = 8 (fastest measured); kept rolled for Q = 16, which would spill
        constexpr int QJ_UNROLL = Q == 8 ? Q / 2 : 1;
#pragma unroll QJ_UNROLL
        for (int qj = 0; qj < Q; qj += 2) {
            double po[2], pso[2], r[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const double tj = Rule<Q>::t(qj + u);
                const double dx = xo - fma(tj, s.sx, s.x0);
                const double dy = yo - fma(tj, s.sy, s.y0);
                // r and 1/r from the hardware 1/sqrt estimate: two Newton steps for 1/r, one
                // residual correction for r (r is the Hankel phase argument, so it is kept
                // to the rounding of a correctly rounded sqrt)
                const double r2 = fmax(fma(dx, dx, dy * dy), GRIM_EPS * GRIM_EPS);
                double iv = __builtin_amdgcn_rsq(r2);
                iv = iv * fma(-0.5 * r2, iv * iv, 1.5);
                iv = iv * fma(-0.5 * r2, iv * iv, 1.5);
                double rr = r2 * iv;
                r[u] = fma(0.5 * iv, fma(-rr, rr, r2), rr);
                po[u] = (dx * o.nx + dy * o.ny) * iv;
                pso[u] = (dx * s.nx + dy * s.ny) * iv;
            }
            cplx g[2], h[2];
            green_kernels2(k, r[0], r[1], g[0], h[0], g[1], h[1]);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const double tj = Rule<Q>::t(qj + u), wj = Rule<Q>::w(qj + u);
                const double w0 = wj * (1.0 - tj), w1 = wj * tj;
                ps0.re = fma(w0, g[u].re, ps0.re); ps0.im = fma(w0, g[u].im, ps0.im);
                ps1.re = fma(w1, g[u].re, ps1.re); ps1.im = fma(w1, g[u].im, ps1.im);
                const double kr_ = -h[u].re * po[u], ki_ = -h[u].im * po[u];
                pk0.re = fma(w0, kr_, pk0.re); pk0.im = fma(w0, ki_, pk0.im);
                pk1.re = fma(w1, kr_, pk1.re); pk1.im = fma(w1, ki_, pk1.im);
                const double dr_ = h[u].re * pso[u], di_ = h[u].im * pso[u];
                pd0.re = fma(w0, dr_, pd0.re); pd0.im = fma(w0, di_, pd0.im);
                pd1.re = fma(w1, dr_, pd1.re); pd1.im = fma(w1, di_, pd1.im);
            }
        }
        const double wo0 = wi * (1.0 - ti), wo1 = wi * ti;
        S[0][0] = cadd(S[0][0], cscale(ps0, wo0)); S[0][1] = cadd(S[0][1], cscale(ps1, wo0));
        S[1][0] = cadd(S[1][0], cscale(ps0, wo1)); S[1][1] = cadd(S[1][1], cscale(ps1, wo1));
        K[0][0] = cadd(K[0][0], cscale(pk0, wo0)); K[0][1] = cadd(K[0][1], cscale(pk1, wo0));
        K[1][0] = cadd(K[1][0], cscale(pk0, wo1)); K[1][1] = cadd(K[1][1], cscale(pk1, wo1));
        D[0][0] = cadd(D[0][0], cscale(pd0, wo0)); D[0][1] = cadd(D[0][1], cscale(pd1, wo0));
        D[1][0] = cadd(D[1][0], cscale(pd0, wo1)); D[1][1] = cadd(D[1][1], cscale(pd1, wo1));
    }
    const double scale = o.len * s.len;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            S[a][b] = cscale(S[a][b], scale);
            K[a][b] = cscale(K[a][b], scale);
            D[a][b] = cscale(D[a][b], scale);
        }
    if (MODE == 1) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                out.tm[a][b] = S[a][b];
                out.te[a][b] = K[a][b];
                out.tmt[b][a] = S[a][b];
                out.tet[b][a] = D[a][b];
            }
        return;
    }
    const double g = -1.0 / k;   // gamma = (0, g)
    const double nn = o.nx * s.nx + o.ny * s.ny;
    double denom = scale < GRIM_EPS * GRIM_EPS ? GRIM_EPS * GRIM_EPS : scale;
    const cplx sum = cadd(cadd(S[0][0], S[0][1]), cadd(S[1][0], S[1][1]));
    const cplx q = {sum.re / denom, sum.im / denom};
    const double kk = -k * k * nn;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const double tsign = (a == b) ? 1.0 : -1.0;
            const cplx w = {kk * S[a][b].re + tsign * q.re, kk * S[a][b].im + tsign * q.im};
            if (MODE == 2) {
                out.tm[a][b] = D[a][b];
                out.te[a][b] = w;
                out.tmt[b][a] = K[a][b];
                out.tet[b][a] = w;
                continue;
            }
            out.tm[a][b] = {S[a][b].re - g * D[a][b].im, S[a][b].im + g * D[a][b].re};
            out.te[a][b] = {K[a][b].re + g * w.im, K[a][b].im - g * w.re};
            out.tmt[b][a] = {S[a][b].re - g * K[a][b].im, S[a][b].im + g * K[a][b].re};
            out.tet[b][a] = {D[a][b].re + g * w.im, D[a][b].im - g * w.re};
        }
}

template <int Q, int MODE>
__device__ __forceinline__ void far_or_zero(const Elem& o, int e, int ne, const Elem& s, int f, int nf,
                                            double k, bool active, Blk2& out) {
    if (active && is_far(o, e, ne, s, f, nf)) far_pair<Q, MODE>(o, s, k, out);
    else zero_blk2(out);
}

__device__ __forceinline__ void store_add(double* z, int64_t ld, int p, int q, cplx v, bool add) {
    double* t = z + 2 * ((int64_t)p + (int64_t)q * ld);
    if (add) { t[0] += v.re; t[1] += v.im; }
    else { t[0] = v.re; t[1] = v.im; }
}

// Far pass over unordered element pairs.  A wavefront owns 64 consecutive
// observation elements (lanes, block I) and walks a chunk of `chunk` source
// columns (inside column block J >= I) in order.  Node entry Z[p, q] gathers
//   (blk(p, prev q)[0][1] + blk(prev p, prev q)[1][1])      source-side part P1
// + (blk(p, q)[0][0] + blk(prev p, q)[1][0])                  own part P2
// (always in this association, so no entry's bits depend on the chunk): the
// observation-side neighbour comes from lane - 1 by a shuffle, and P1 of the
// next column is formed from the current column's blocks and carried along the
// walk.  For J > I the same evaluations also give the transposed tile (rows in
// the chunk, columns in block I), written with the roles exchanged.
// A chunk whose first column continues the previous chunk of its row block
// does not evaluate that column's source-side neighbour again: the previous
// chunk adds P1 after its last column and this chunk adds P2, both by atomic
// adds onto entries far_zero cleared.  Two addends onto zero give the same sum
// in either order, so every entry is still written deterministically.  Only the
// first chunk of a row block and columns that start a contour evaluate their
// neighbour.  Entries whose observation-side neighbour is not lane - 1 (lane 0,
// contour starts) are completed by far_fixup.  MODE selects the two outputs
// (see far_pair).
__device__ __forceinline__ cplx shfl_up1(cplx v) {
    return {__shfl_up(v.re, 1, 64), __shfl_up(v.im, 1, 64)};
}

__device__ __forceinline__ void atomic_add_entry(double* z, int64_t ld, int p, int q, cplx v) {
    double* t = z + 2 * ((int64_t)p + (int64_t)q * ld);
    atomicAdd(t, v.re);
    atomicAdd(t + 1, v.im);
}

// Column c (a multiple of `chunk`, continuing a contour) receives two atomic partial sums in
// the row blocks whose chunk boundary it is: rows below 64 ceil(c / 64) of column c (direct
// tiles) and columns below 64 floor(c / 64) of row c (mirror tiles).  One workgroup per c.
__global__ void __launch_bounds__(256) far_zero(int n, int chunk, const int* __restrict__ prev,
                                                double* __restrict__ ztm, int64_t ldtm, double* __restrict__ zte,
                                                int64_t ldte) {
    const int c = (blockIdx.x + 1) * chunk;
    if (c >= n || prev[c] != c - 1) return;
    const int rows = min(n, 64 * ((c + 63) / 64)), cols = 64 * (c / 64);
    for (int r = threadIdx.x; r < rows; r += 256) {
        double* a = ztm + 2 * ((int64_t)r + (int64_t)c * ldtm);
        double* b = zte + 2 * ((int64_t)r + (int64_t)c * ldte);
        a[0] = 0.0; a[1] = 0.0; b[0] = 0.0; b[1] = 0.0;
    }
    for (int r = threadIdx.x; r < cols; r += 256) {
        double* a = ztm + 2 * ((int64_t)c + (int64_t)r * ldtm);
        double* b = zte + 2 * ((int64_t)c + (int64_t)r * ldte);
        a[0] = 0.0; a[1] = 0.0; b[0] = 0.0; b[1] = 0.0;
    }
}

// Both outputs cleared whole, column by column (coalesced): cheaper than far_zero's strided
// mirror rows once half or more of the columns are chunk boundaries.
__global__ void __launch_bounds__(256) far_zero_all(int n, double* __restrict__ ztm, int64_t ldtm,
                                                    double* __restrict__ zte, int64_t ldte) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= n) return;
    double* a = ztm + 2 * ((int64_t)r + (int64_t)c * ldtm);
    double* b = zte + 2 * ((int64_t)r + (int64_t)c * ldte);
    a[0] = 0.0; a[1] = 0.0; b[0] = 0.0; b[1] = 0.0;
}

template <int Q, int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) far_kernel(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                 const int* __restrict__ next, const int* __restrict__ prev,
                                                 double k, int chunk, double* __restrict__ ztm, int64_t ldtm,
                                                 double* __restrict__ zte, int64_t ldte) {
    const int lane = threadIdx.x;
    // tile = (row block I, column chunk j) with j * chunk / 64 >= I, enumerated row block by row block
    const int nc = (n + chunk - 1) / chunk, cpb = 64 / chunk, nb = (n + 63) / 64;
    int2 tile = {0, 0};
    for (int I = 0, idx = blockIdx.x; I < nb; ++I) {
        const int cnt = nc - I * cpb;
        if (idx < cnt) { tile.x = I; tile.y = I * cpb + idx; break; }
        idx -= cnt;
    }
    const int e = tile.x * 64 + lane;
    const bool active = e < n;
    const int ee = active ? e : n - 1;
    const Elem o = load_elem(p0, p1, ee);
    const int ne = next[ee];
    const bool rowok = active && lane > 0 && prev[ee] == ee - 1;
    const int c0 = tile.y * chunk;
    const int c1 = min(n, c0 + chunk);
    const bool mirror = (c0 / 64) > tile.x;
    const bool first_chunk = c0 == tile.x * 64;
    // P1 of the current column: direct TM / TE (column f, rows e) and mirror TM / TE (row f, columns e)
    cplx qtm = {0, 0}, qte = {0, 0}, rtm = {0, 0}, rte = {0, 0};
    bool shared_first = false;   // this chunk's first column gets its P1 from the previous chunk
    for (int f = c0; f < c1; ++f) {
        const int pf = prev[f];
        if (f == c0 || pf != f - 1) {
            if (f == c0 && !first_chunk && pf == f - 1) {
                shared_first = true;
                qtm = qte = rtm = rte = {0.0, 0.0};
            } else {
                Blk2 b;
                far_or_zero<Q, MODE>(o, ee, ne, load_elem(p0, p1, pf), pf, next[pf], k, active, b);
                const cplx utm = shfl_up1(b.tm[1][1]), ute = shfl_up1(b.te[1][1]);
                const cplx vtm = shfl_up1(b.tmt[1][1]), vte = shfl_up1(b.tet[1][1]);
                qtm = rowok ? cadd(b.tm[0][1], utm) : b.tm[0][1];
                qte = rowok ? cadd(b.te[0][1], ute) : b.te[0][1];
                rtm = rowok ? cadd(b.tmt[1][0], vtm) : b.tmt[1][0];
                rte = rowok ? cadd(b.tet[1][0], vte) : b.tet[1][0];
            }
        }
        Blk2 b;
        far_or_zero<Q, MODE>(o, ee, ne, load_elem(p0, p1, f), f, next[f], k, active, b);
        // own parts P2: direct column f from blk(e, f)[0][0] + blk(e - 1, f)[1][0], mirror row f from
        // blk(e, f).tmt[0][0] + blk(e - 1, f).tmt[0][1]
        const cplx u0 = shfl_up1(b.tm[1][0]), u1 = shfl_up1(b.te[1][0]);
        const cplx u2 = shfl_up1(b.tmt[0][1]), u3 = shfl_up1(b.tet[0][1]);
        const cplx otm = rowok ? cadd(b.tm[0][0], u0) : b.tm[0][0];
        const cplx ote = rowok ? cadd(b.te[0][0], u1) : b.te[0][0];
        const cplx mtm = rowok ? cadd(b.tmt[0][0], u2) : b.tmt[0][0];
        const cplx mte = rowok ? cadd(b.tet[0][0], u3) : b.tet[0][0];
        if (active) {
            if (f == c0 && shared_first) {
                atomic_add_entry(ztm, ldtm, e, f, otm);
                atomic_add_entry(zte, ldte, e, f, ote);
                if (mirror) {
                    atomic_add_entry(ztm, ldtm, f, e, mtm);
                    atomic_add_entry(zte, ldte, f, e, mte);
                }
            } else {
                store_add(ztm, ldtm, e, f, cadd(qtm, otm), false);
                store_add(zte, ldte, e, f, cadd(qte, ote), false);
                if (mirror) {
                    store_add(ztm, ldtm, f, e, cadd(rtm, mtm), false);
                    store_add(zte, ldte, f, e, cadd(rte, mte), false);
                }
            }
        }
        // P1 of column f + 1 from this column's blocks
        const cplx w0 = shfl_up1(b.tm[1][1]), w1 = shfl_up1(b.te[1][1]);
        const cplx w2 = shfl_up1(b.tmt[1][1]), w3 = shfl_up1(b.tet[1][1]);
        qtm = rowok ? cadd(b.tm[0][1], w0) : b.tm[0][1];
        qte = rowok ? cadd(b.te[0][1], w1) : b.te[0][1];
        rtm = rowok ? cadd(b.tmt[1][0], w2) : b.tmt[1][0];
        rte = rowok ? cadd(b.tet[1][0], w3) : b.tet[1][0];
    }
    // the next chunk of this row block takes its first column's P1 from here
    if (active && c1 < n && prev[c1] == c1 - 1) {
        atomic_add_entry(ztm, ldtm, e, c1, qtm);
        atomic_add_entry(zte, ldte, e, c1, qte);
        if ((c1 / 64) > tile.x) {
            atomic_add_entry(ztm, ldtm, c1, e, rtm);
            atomic_add_entry(zte, ldte, c1, e, rte);
        }
    }
}

// Completes the entries far_kernel leaves short, for each listed index c
// (c % 64 == 0 or prev[c] != c - 1), with e' = prev(c):
//   row c    (entries written by a direct tile, block(q) >= block(c)):
//       Z[c, q] += blk(e', prev q).tm[1][1] + blk(e', q).tm[1][0]
//   column c (entries written by a mirror tile,  block(q) >  block(c)):
//       Z[q, c] += blk(e', prev q).tmt[1][1] + blk(e', q).tmt[0][1]
template <int Q, int MODE>
__global__ void __launch_bounds__(64) far_fixup(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                const int* __restrict__ next, const int* __restrict__ prev,
                                                double k, const int* __restrict__ rows, double* __restrict__ ztm,
                                                int64_t ldtm, double* __restrict__ zte, int64_t ldte) {
    // one wavefront per 63 consecutive q: lane l >= 1 owns q = qb + l - 1 and evaluates
    // blk(e', q); lane 0 evaluates blk(e', prev qb), the neighbour of lane 1.  Each lane
    // takes blk(e', prev q) from lane - 1 (prev q == q - 1) or, at a contour start inside
    // the wavefront, evaluates it itself.
    const int c = rows[blockIdx.y];
    const int lane = threadIdx.x;
    const int qb = blockIdx.x * 63;
    const int q = qb + lane - 1;
    const int e = prev[c];
    const Elem o = load_elem(p0, p1, e);
    const int ne = next[e];
    const int qc = q < n ? q : n - 1;
    const int col = lane == 0 ? prev[qb] : qc;
    Blk2 b;
    far_or_zero<Q, MODE>(o, e, ne, load_elem(p0, p1, col), col, next[col], k, lane == 0 || q < n, b);
    cplx ltm, lte, lmt, lme;
    ltm.re = __shfl_up(b.tm[1][1].re, 1, 64); ltm.im = __shfl_up(b.tm[1][1].im, 1, 64);
    lte.re = __shfl_up(b.te[1][1].re, 1, 64); lte.im = __shfl_up(b.te[1][1].im, 1, 64);
    lmt.re = __shfl_up(b.tmt[1][1].re, 1, 64); lmt.im = __shfl_up(b.tmt[1][1].im, 1, 64);
    lme.re = __shfl_up(b.tet[1][1].re, 1, 64); lme.im = __shfl_up(b.tet[1][1].im, 1, 64);
    const int pq = prev[qc];
    if (lane >= 2 && q < n && pq != q - 1) {
        Blk2 b1;
        far_or_zero<Q, MODE>(o, e, ne, load_elem(p0, p1, pq), pq, next[pq], k, true, b1);
        ltm = b1.tm[1][1]; lte = b1.te[1][1]; lmt = b1.tmt[1][1]; lme = b1.tet[1][1];
    }
    if (lane == 0 || q >= n) return;
    if (q / 64 >= c / 64) {
        store_add(ztm, ldtm, c, q, cadd(ltm, b.tm[1][0]), true);
        store_add(zte, ldte, c, q, cadd(lte, b.te[1][0]), true);
    }
    if (q / 64 > c / 64) {
        store_add(ztm, ldtm, q, c, cadd(lmt, b.tmt[0][1]), true);
        store_add(zte, ldte, q, c, cadd(lme, b.tet[0][1]), true);
    }
}

// ---------------------------------------------------------------------------
// Near pairs
// ---------------------------------------------------------------------------
// One wavefront per observation element: count (pass 0) or list (pass 1) the
// source elements that are not far, in ascending order.
__global__ void __launch_bounds__(64) near_list(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                const int* __restrict__ next, int pass, int* __restrict__ counts,
                                                const int* __restrict__ offsets, int capacity,
                                                int* __restrict__ pair_e, int* __restrict__ pair_f,
                                                unsigned* __restrict__ status) {
    const int e = blockIdx.x;
    const int lane = threadIdx.x;
    const Elem o = load_elem(p0, p1, e);
    const int ne = next[e];
    int base = pass ? offsets[e] : 0;
    int total = 0;
    if (pass && offsets[n] > capacity) {
        if (lane == 0 && e == 0) atomicOr(status, (unsigned)GRIM_DEV_NEAR_CAPACITY);
        return;
    }
    for (int f0 = 0; f0 < n; f0 += 64) {
        const int f = f0 + lane;
        bool near = false;
        if (f < n) {
            const Elem s = load_elem(p0, p1, f);
            near = !is_far(o, e, ne, s, f, next[f]);
        }
        const unsigned long long mask = __ballot(near);
        if (pass && near) {
            const int idx = base + __popcll(mask & ((1ull << lane) - 1ull));
            pair_e[idx] = e;
            pair_f[idx] = f;
        }
        base += __popcll(mask);
        total += __popcll(mask);
    }
    if (!pass && lane == 0) counts[e] = total;
}

// Exclusive scan of n counts into offsets[0..n] (single workgroup).
__global__ void __launch_bounds__(1024) scan_counts(int n, const int* __restrict__ counts, int* __restrict__ offsets) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int lo = min(n, t * per), hi = min(n, lo + per);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t ? part[t - 1] : 0;
    for (int i = lo; i < hi; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    if (t == 1023) offsets[n] = part[1023];
}

// Touching test of one ordered pair (operators.py:1912-1940): the first endpoint
// pair within the 1e-9 m snap, or one mesh node within two snap widths.
__device__ __forceinline__ bool touching(const Elem& o, int e, int ne, const Elem& s, int f, int nf,
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

__device__ __forceinline__ double point_seg_sq(double x, double y, double ax, double ay, double bx, double by) {
    const double ex = bx - ax, ey = by - ay;
    const double den = ex * ex + ey * ey;
    double t = den > 0 ? ((x - ax) * ex + (y - ay) * ey) / den : 0.0;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    const double dx = x - ax - t * ex, dy = y - ay - t * ey;
    return dx * dx + dy * dy;
}

// separation.py:18-78
__device__ bool requires_adaptive(const Elem& o, const Elem& s) {
    const double lo = o.len, ls = s.len;
    const double scale = lo > ls ? lo : ls;
    const double cdx = o.cx - s.cx, cdy = o.cy - s.cy;
    const double distance = sqrt(cdx * cdx + cdy * cdy);
    if (distance < 0.75 * scale) return true;
    if (distance >= 0.5 * (lo + ls) + 0.5 * scale) return false;
    const double p0x = o.x0, p0y = o.y0, p1x = o.x0 + o.sx, p1y = o.y0 + o.sy;
    const double q0x = s.x0, q0y = s.y0, q1x = s.x0 + s.sx, q1y = s.y0 + s.sy;
    if (distance < 0.5 * (lo + ls) + 0.25 * scale) {
        double sq = point_seg_sq(p0x, p0y, q0x, q0y, q1x, q1y);
        sq = fmin(sq, point_seg_sq(p1x, p1y, q0x, q0y, q1x, q1y));
        sq = fmin(sq, point_seg_sq(q0x, q0y, p0x, p0y, p1x, p1y));
        sq = fmin(sq, point_seg_sq(q1x, q1y, p0x, p0y, p1x, p1y));
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

// Exact linear-Galerkin self S block for |k L| <= 8 (operators.py:418-501).
__device__ bool self_block(double len, double k, cplx (&S)[2][2]) {
    const double z = k * len;
    if (!(z <= 8.0) || z <= 1e-30) return false;
    const double cd[4] = {2.0 / 3.0, -1.0, 0.0, 1.0 / 3.0};
    const double co[4] = {1.0 / 3.0, 0.0, 0.0, -1.0 / 3.0};
    const double tp = 2.0 / GRIM_PI;
    const double lt = log(0.5 * z) + 0.5772156649015329;
    const double zq = 0.25 * z * z;
    cplx bd = {0, 0}, bo = {0, 0};
    double alpha = 1.0, harmonic = 0.0;
    for (int m = 0;; ) {
        const int p = 2 * m;
        const cplx am = {alpha, -alpha * tp * (lt - harmonic)};
        double md = 0, mo = 0, ld = 0, lo = 0;
        for (int q = 0; q < 4; ++q) {
            const double dd = (double)(p + q + 1);
            md += cd[q] / dd; mo += co[q] / dd;
            ld -= cd[q] / (dd * dd); lo -= co[q] / (dd * dd);
        }
        // a_m * moment - j (2/pi) alpha log_moment
        bd.re += am.re * md; bd.im += am.im * md - tp * alpha * ld;
        bo.re += am.re * mo; bo.im += am.im * mo - tp * alpha * lo;
        ++m;
        alpha *= -zq / ((double)m * (double)m);
        harmonic += 1.0 / m;
        if (m > 60) return false;
        const double mag = sqrt(bd.re * bd.re + bd.im * bd.im);
        if (fabs(alpha) < 1e-18 * (mag > 1.0 ? mag : 1.0)) break;
    }
    // 0.25 j L^2 [[bd, bo], [bo, bd]]
    const double c = 0.25 * len * len;
    const cplx d = {-c * bd.im, c * bd.re}, f = {-c * bo.im, c * bo.re};
    S[0][0] = d; S[0][1] = f; S[1][0] = f; S[1][1] = d;
    return true;
}

struct SK {
    double v[16];   // S[a][b] re,im (8) then K'[a][b] re,im (8)
};

// Tensor Gauss 16 x 16 box on sub-intervals; wave-parallel, result in every lane.
__device__ void box16(const Elem& o, const Elem& s, double k, double oa, double ob, double sa, double sb, SK& out) {
    const int lane = threadIdx.x & 63;
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0;
    const double ho = ob - oa, hs = sb - sa;
    for (int pt = lane; pt < 256; pt += 64) {
        const int i = pt >> 4, j = pt & 15;
        const double xo = oa + ho * c_gl16_t[i], xs = sa + hs * c_gl16_t[j];
        const double w = c_gl16_w[i] * c_gl16_w[j];
        const double dx = (o.x0 + xo * o.sx) - (s.x0 + xs * s.sx);
        const double dy = (o.y0 + xo * o.sy) - (s.y0 + xs * s.sy);
        const double dist = sqrt(dx * dx + dy * dy);
        cplx g, h;
        green_kernels(k, dist, g, h);
        const double safe = dist > GRIM_EPS ? dist : GRIM_EPS;
        const double proj = (dx * o.nx + dy * o.ny) / safe;
        cplx dk = {-h.re * proj, -h.im * proj};
        if (dist <= GRIM_EPS) dk = {0.0, 0.0};
        const double po[2] = {1.0 - xo, xo}, ps[2] = {1.0 - xs, xs};
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const double c = w * po[a] * ps[b];
                acc[2 * (2 * a + b)] += c * g.re;
                acc[2 * (2 * a + b) + 1] += c * g.im;
                acc[8 + 2 * (2 * a + b)] += c * dk.re;
                acc[8 + 2 * (2 * a + b) + 1] += c * dk.im;
            }
    }
    const double scale = o.len * ho * s.len * hs;
#pragma unroll
    for (int i = 0; i < 16; ++i) out.v[i] = wave_sum(acc[i]) * scale;
}

// Graded t^4 two-triangle corner rule (operators.py:549-757, polynomial_quadrature.py:319-335).
__device__ void touching16(const Elem& o, const Elem& s, double k, bool obs_start, bool src_start, SK& out) {
    const int lane = threadIdx.x & 63;
    const double cox = obs_start ? o.x0 : o.x0 + o.sx, coy = obs_start ? o.y0 : o.y0 + o.sy;
    const double csx = src_start ? s.x0 : s.x0 + s.sx, csy = src_start ? s.y0 : s.y0 + s.sy;
    const double dox = obs_start ? o.sx : -o.sx, doy = obs_start ? o.sy : -o.sy;
    const double dsx = src_start ? s.sx : -s.sx, dsy = src_start ? s.sy : -s.sy;
    double orx = cox - csx, ory = coy - csy;
    const double eps = 2.220446049250313e-16;
    double big = fmax(fmax(sqrt(cox * cox + coy * coy), sqrt(csx * csx + csy * csy)), fmax(o.len, s.len));
    if (sqrt(orx * orx + ory * ory) <= 64.0 * eps * big) { orx = 0.0; ory = 0.0; }
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0;
    for (int pt = lane; pt < 800; pt += 64) {
        const double x = c_touch_x[pt], y = c_touch_y[pt], w = c_touch_w[pt];
        const double dx = orx + x * dox - y * dsx;
        const double dy = ory + x * doy - y * dsy;
        const double xo = obs_start ? x : 1.0 - x, xs = src_start ? y : 1.0 - y;
        const double dist = sqrt(dx * dx + dy * dy);
        const double safe = dist > GRIM_EPS ? dist : GRIM_EPS;
        cplx g, h;
        green_kernels(k, safe, g, h);
        const double proj = (dx * o.nx + dy * o.ny) / safe;
        cplx dk = {-h.re * proj, -h.im * proj};
        if (dist <= GRIM_EPS) dk = {0.0, 0.0};
        const double po[2] = {1.0 - xo, xo}, ps[2] = {1.0 - xs, xs};
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const double c = w * po[a] * ps[b];
                acc[2 * (2 * a + b)] += c * g.re;
                acc[2 * (2 * a + b) + 1] += c * g.im;
                acc[8 + 2 * (2 * a + b)] += c * dk.re;
                acc[8 + 2 * (2 * a + b) + 1] += c * dk.im;
            }
    }
    const double scale = o.len * s.len;
#pragma unroll
    for (int i = 0; i < 16; ++i) out.v[i] = wave_sum(acc[i]) * scale;
}

// Pair classes: 0 self, 1 touching, 2 fixed, 3 adaptive.
__global__ void __launch_bounds__(64) near_blocks(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                  const int* __restrict__ next, const int* __restrict__ offsets,
                                                  const int* __restrict__ pair_e, const int* __restrict__ pair_f,
                                                  int capacity, double k, double* __restrict__ blocks,
                                                  int* __restrict__ kind, unsigned* __restrict__ status) {
    const int idx = blockIdx.x;
    if (idx >= offsets[n] || offsets[n] > capacity) return;   // overflow: the lists were not written
    const int e = pair_e[idx], f = pair_f[idx];
    const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
    SK r;
    int cls;
    bool os = false, ss = false;
    if (e == f) {
        cls = 0;
        cplx S[2][2];
        if (!self_block(o.len, k, S)) {
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
    } else if (touching(o, e, next[e], s, f, next[f], os, ss)) {
        cls = 1;
        touching16(o, s, k, os, ss, r);
    } else if (requires_adaptive(o, s)) {
        cls = 3;   // integrated by near_adaptive
        if (threadIdx.x == 0) kind[idx] = cls;
        return;
    } else {
        cls = 2;
        box16(o, s, k, 0.0, 1.0, 0.0, 1.0, r);
    }
    if (threadIdx.x < 16) blocks[16 * (int64_t)idx + threadIdx.x] = r.v[threadIdx.x];
    if (threadIdx.x == 0) kind[idx] = cls;
}

// Recursive four-way bisection with rtol 1e-9 and depth 12 (operators.py:1172-1292),
// run depth first with an explicit stack.  One workgroup of four wavefronts per adaptive pair,
// a grid-stride loop over the listed pairs: the four child boxes of a node are integrated at
// once, one per wavefront (each with the wave-parallel 16 x 16 rule), so a pair's recursion
// takes one box time per node instead of four.  Every box and every sum is formed in the same
// order as by a single wavefront.
#define GRIM_ADAPT_STACK 48
__global__ void __launch_bounds__(256) near_adaptive(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                     const int* __restrict__ offsets, const int* __restrict__ pair_e,
                                                     const int* __restrict__ pair_f, int capacity, double k,
                                                     double* __restrict__ blocks, const int* __restrict__ kind,
                                                     unsigned* __restrict__ status) {
    __shared__ double st_iv[GRIM_ADAPT_STACK][4];
    __shared__ double st_blk[GRIM_ADAPT_STACK][16];
    __shared__ int st_depth[GRIM_ADAPT_STACK];
    __shared__ double kid_blk[4][16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int listed = offsets[n];
    if (listed > capacity) return;                          // overflow: the lists were not written
    for (int idx = blockIdx.x; idx < listed; idx += gridDim.x) {
        if (kind[idx] != 3) continue;                      // uniform over the workgroup
        const int e = pair_e[idx], f = pair_f[idx];
        const Elem o = load_elem(p0, p1, e), s = load_elem(p0, p1, f);
        const double floor_ = 2.220446049250313e-16 * fmax(1.0, o.len * s.len);
        if (wave == 0) {
            SK root;
            box16(o, s, k, 0.0, 1.0, 0.0, 1.0, root);
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
                SK kid;
                box16(o, s, k, iv[wave][0], iv[wave][1], iv[wave][2], iv[wave][3], kid);
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
            if (err <= 1e-9 || depth >= 12 || top + 4 > GRIM_ADAPT_STACK) {
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

__device__ __forceinline__ int find_pair(const int* __restrict__ offsets, const int* __restrict__ pair_f, int row, int col) {
    int lo = offsets[row], hi = offsets[row + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = pair_f[mid];
        if (v == col) return mid;
        if (v < col) lo = mid + 1; else hi = mid;
    }
    return -1;
}

__device__ __forceinline__ void add_entry(double* z, int64_t ld, int p, int q, cplx v) {
    double* t = z + 2 * ((int64_t)p + (int64_t)q * ld);
    t[0] += v.re;
    t[1] += v.im;
}

// Row-owner gather of near blocks and mass terms: thread p adds everything of node row p.
// mode 1: the outputs are S and K' + c M; mode 2: D + c M and W; c = mass_coef[e] per element
// (or mass if mass_coef is null).  No other mass or jump terms in these modes.
__global__ void __launch_bounds__(256) near_gather(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                   const int* __restrict__ next, const int* __restrict__ prev,
                                                   const int* __restrict__ offsets, const int* __restrict__ pair_f,
                                                   int capacity, const double* __restrict__ blocks, double k, int mode,
                                                   double mass, const double* __restrict__ mass_coef,
                                                   double* __restrict__ ztm, int64_t ldtm,
                                                   double* __restrict__ zte, int64_t ldte) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || offsets[n] > capacity) return;           // overflow: the lists were not written
    const double g = -1.0 / k;
    const int elems[2] = {p, prev[p]};
    for (int loc = 0; loc < 2; ++loc) {
        const int e = elems[loc];
        const Elem o = load_elem(p0, p1, e);
        for (int idx = offsets[e]; idx < offsets[e + 1]; ++idx) {
            const int f = pair_f[idx];
            const Elem s = load_elem(p0, p1, f);
            const double* B = blocks + 16 * (int64_t)idx;
            const int rev = find_pair(offsets, pair_f, f, e);
            const double* R = blocks + 16 * (int64_t)(rev < 0 ? idx : rev);
            cplx S[2][2], K[2][2], D[2][2];
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    S[a][b] = {B[2 * (2 * a + b)], B[2 * (2 * a + b) + 1]};
                    K[a][b] = {B[8 + 2 * (2 * a + b)], B[8 + 2 * (2 * a + b) + 1]};
                    // D(e, f) = K'(f, e)^T
                    D[a][b] = rev < 0 ? cplx{0.0, 0.0} : cplx{R[8 + 2 * (2 * b + a)], R[8 + 2 * (2 * b + a) + 1]};
                }
            Blk blk;
            if (mode == 1) {
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        blk.tm[a][b] = S[a][b];
                        blk.te[a][b] = K[a][b];
                    }
            } else if (mode == 2) {
                // D and the Maue W of combine(), without gamma
                const double nn = o.nx * s.nx + o.ny * s.ny;
                double denom = o.len * s.len;
                if (denom < GRIM_EPS * GRIM_EPS) denom = GRIM_EPS * GRIM_EPS;
                const cplx sum = cadd(cadd(S[0][0], S[0][1]), cadd(S[1][0], S[1][1]));
                const cplx q = {sum.re / denom, sum.im / denom};
                const double kk = -k * k * nn;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        const double tsign = (a == b) ? 1.0 : -1.0;
                        blk.tm[a][b] = D[a][b];
                        blk.te[a][b] = {kk * S[a][b].re + tsign * q.re, kk * S[a][b].im + tsign * q.im};
                    }
            } else {
                combine(S, K, D, o, s, k, blk);
            }
            const int cols[2] = {f, next[f]};
            for (int b = 0; b < 2; ++b) {
                add_entry(ztm, ldtm, p, cols[b], blk.tm[loc][b]);
                add_entry(zte, ldte, p, cols[b], blk.te[loc][b]);
            }
        }
        // mass of element e: L [[1/3, 1/6], [1/6, 1/3]]; TM + gamma/2 M, TE - M/2
        const int cols[2] = {e, next[e]};
        const double c = mass_coef ? mass_coef[e] : mass;
        for (int b = 0; b < 2; ++b) {
            const double m = o.len * (loc == b ? 1.0 / 3.0 : 1.0 / 6.0);
            if (mode == 1) {
                add_entry(zte, ldte, p, cols[b], cplx{c * m, 0.0});
            } else if (mode == 2) {
                add_entry(ztm, ldtm, p, cols[b], cplx{c * m, 0.0});
            } else {
                add_entry(ztm, ldtm, p, cols[b], cplx{0.0, 0.5 * m * g});
                add_entry(zte, ldte, p, cols[b], cplx{-0.5 * m, 0.0});
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Plane-wave loads and far field
// ---------------------------------------------------------------------------
// bu[p, a] = i0(p) + i1(prev p); bdn likewise with j k d.n.  angles in degrees.
__global__ void __launch_bounds__(256) loads_kernel(int n, int na, const double* __restrict__ p0,
                                                    const double* __restrict__ p1, const int* __restrict__ prev,
                                                    double k, const double* __restrict__ angles_deg,
                                                    double* __restrict__ bu, double* __restrict__ bdn, int64_t ld) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (p >= n) return;
    const double phi = angles_deg[a] * (GRIM_PI / 180.0);
    double dy, dx;
    sincos(phi, &dy, &dx);
    const Elem A = load_elem(p0, p1, p), B = load_elem(p0, p1, prev[p]);
    cplx i0, i1, dnA, j0_, j1_, dnB;
    elem_moments(A, k, dx, dy, i0, i1, dnA);
    elem_moments(B, k, dx, dy, j0_, j1_, dnB);
    const cplx u = cadd(i0, j1_);
    const cplx v = cadd(cmul(i0, dnA), cmul(j1_, dnB));
    const int64_t at = 2 * ((int64_t)p + (int64_t)a * ld);
    if (bu) { bu[at] = u.re; bu[at + 1] = u.im; }
    if (bdn) { bdn[at] = v.re; bdn[at + 1] = v.im; }
}

// B(a) = sum_p x[p,a] bu[p,a] + gamma sum_p x[p,a] bdn[p,a]; sigma = |B|^2 / (4k).  With bdn null,
// B(a) = sum_p x[p,a] bu[p,a]: the far field of one potential whose plane-wave moments bu holds.
__global__ void __launch_bounds__(256) farfield_kernel(int n, const double* __restrict__ x, int64_t ldx,
                                                       const double* __restrict__ bu, const double* __restrict__ bdn,
                                                       int64_t ldb, double k, double* __restrict__ amp,
                                                       double* __restrict__ sigma) {
    const int a = blockIdx.x;
    double s0r = 0, s0i = 0, s1r = 0, s1i = 0;
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const int64_t ix = 2 * ((int64_t)p + (int64_t)a * ldx), ib = 2 * ((int64_t)p + (int64_t)a * ldb);
        const double xr = x[ix], xi = x[ix + 1];
        s0r += xr * bu[ib] - xi * bu[ib + 1];
        s0i += xr * bu[ib + 1] + xi * bu[ib];
        if (bdn) {
            s1r += xr * bdn[ib] - xi * bdn[ib + 1];
            s1i += xr * bdn[ib + 1] + xi * bdn[ib];
        }
    }
    __shared__ double red[4][4];
    s0r = wave_sum(s0r); s0i = wave_sum(s0i); s1r = wave_sum(s1r); s1i = wave_sum(s1i);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[w][0] = s0r; red[w][1] = s0i; red[w][2] = s1r; red[w][3] = s1i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {0, 0, 0, 0};
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i)
            for (int j = 0; j < 4; ++j) t[j] += red[i][j];
        const double g = -1.0 / k;   // gamma = (0, g)
        const double br = t[0] - g * t[3], bi = t[1] + g * t[2];
        amp[2 * a] = br;
        amp[2 * a + 1] = bi;
        if (sigma) sigma[a] = 0.25 / k * (br * br + bi * bi);
    }
}

// ---------------------------------------------------------------------------
// Host launchers (called from api.cpp)
// ---------------------------------------------------------------------------
extern "C" hipError_t grim_upload_rules(const double* gl8t, const double* gl8w, const double* gl16t,
                                         const double* gl16w, const double* tx, const double* ty, const double* tw) {
    hipError_t err;
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_gl8_t), gl8t, 8 * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_gl8_w), gl8w, 8 * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_gl16_t), gl16t, 16 * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_gl16_w), gl16w, 16 * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_touch_x), tx, 800 * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_touch_y), ty, 800 * sizeof(double))) != hipSuccess) return err;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_touch_w), tw, 800 * sizeof(double));
}

// Far pass launcher: one wavefront per (row block I, column chunk j) with j in block J >= I.
template <int MODE>
static hipError_t launch_far(int q, int n, const double* p0, const double* p1, const int* next, const int* prev,
                             double k, double* ztm, int64_t ldtm, double* zte, int64_t ldte, const int* fix_rows,
                             int nfix, hipStream_t st) {
    // columns per wavefront: enough wavefronts (about 8-16 K) to fill 1024 SIMDs
    // with their tails overlapping; chunk boundaries share their column by atomics
    // onto entries far_zero cleared, so a chunk evaluates only its own columns.
    int chunk = grim_shape_value(GRIM_SHAPE_FAR_COLUMNS, n);
    if (chunk == 0) chunk = n < 1414 ? 1 : n < 5656 ? 2 : 8;   // measured on MI355X, n = 500-12000
    const int nc = (n + chunk - 1) / chunk, cpb = 64 / chunk, nb = (n + 63) / 64;
    long ntiles = 0;
    for (int I = 0; I < nb; ++I) ntiles += nc - I * cpb;
    if (nc > 1) {
        if (chunk <= 2 && n <= 65535)
            far_zero_all<<<dim3((n + 255) / 256, n), 256, 0, st>>>(n, ztm, ldtm, zte, ldte);
        else far_zero<<<(unsigned)(nc - 1), 256, 0, st>>>(n, chunk, prev, ztm, ldtm, zte, ldte);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    if (q == 8)
        far_kernel<8, MODE><<<(unsigned)ntiles, 64, 0, st>>>(n, p0, p1, next, prev, k, chunk, ztm, ldtm, zte, ldte);
    else
        far_kernel<16, MODE><<<(unsigned)ntiles, 64, 0, st>>>(n, p0, p1, next, prev, k, chunk, ztm, ldtm, zte, ldte);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || nfix == 0) return err;
    dim3 g2((n + 62) / 63, nfix);
    if (q == 8) far_fixup<8, MODE><<<g2, 64, 0, st>>>(n, p0, p1, next, prev, k, fix_rows, ztm, ldtm, zte, ldte);
    else far_fixup<16, MODE><<<g2, 64, 0, st>>>(n, p0, p1, next, prev, k, fix_rows, ztm, ldtm, zte, ldte);
    return hipGetLastError();
}

// mode 0: the combined-potential pair (TM, TE); 1: S and K'; 2: D and W (far_pair).
extern "C" hipError_t grim_launch_far(int q, int n, const double* p0, const double* p1, const int* next,
                                      const int* prev, double k, double* ztm, int64_t ldtm, double* zte,
                                      int64_t ldte, const int* fix_rows, int nfix, int mode, hipStream_t st) {
    if (mode == 1) return launch_far<1>(q, n, p0, p1, next, prev, k, ztm, ldtm, zte, ldte, fix_rows, nfix, st);
    if (mode == 2) return launch_far<2>(q, n, p0, p1, next, prev, k, ztm, ldtm, zte, ldte, fix_rows, nfix, st);
    return launch_far<0>(q, n, p0, p1, next, prev, k, ztm, ldtm, zte, ldte, fix_rows, nfix, st);
}

extern "C" hipError_t grim_launch_near(int n, const double* p0, const double* p1, const int* next, const int* prev,
                                       double k, int* counts, int* offsets, int capacity, int* pair_e, int* pair_f,
                                       double* blocks, int* kind, unsigned* status, double* ztm, int64_t ldtm,
                                       double* zte, int64_t ldte, int mode, double mass, const double* mass_coef,
                                       hipStream_t st) {
    near_list<<<n, 64, 0, st>>>(n, p0, p1, next, 0, counts, offsets, capacity, pair_e, pair_f, status);
    scan_counts<<<1, 1024, 0, st>>>(n, counts, offsets);
    near_list<<<n, 64, 0, st>>>(n, p0, p1, next, 1, counts, offsets, capacity, pair_e, pair_f, status);
    near_blocks<<<capacity, 64, 0, st>>>(n, p0, p1, next, offsets, pair_e, pair_f, capacity, k, blocks, kind,
                                         status);
    near_adaptive<<<capacity < 1024 ? capacity : 1024, 256, 0, st>>>(n, p0, p1, offsets, pair_e, pair_f, capacity, k,
                                                                     blocks, kind, status);
    near_gather<<<(n + 255) / 256, 256, 0, st>>>(n, p0, p1, next, prev, offsets, pair_f, capacity, blocks, k, mode,
                                                 mass,
                                                 mass_coef, ztm, ldtm, zte, ldte);
    return hipGetLastError();
}

// The near-pair lists and the linear S / K' blocks of every listed pair (no gather), for callers
// that form node entries from element-pair blocks themselves (sheet systems).
extern "C" hipError_t grim_launch_near_pairs(int n, const double* p0, const double* p1, const int* next, double k,
                                             int* counts, int* offsets, int capacity, int* pair_e, int* pair_f,
                                             double* blocks, int* kind, unsigned* status, hipStream_t st) {
    near_list<<<n, 64, 0, st>>>(n, p0, p1, next, 0, counts, offsets, capacity, pair_e, pair_f, status);
    scan_counts<<<1, 1024, 0, st>>>(n, counts, offsets);
    near_list<<<n, 64, 0, st>>>(n, p0, p1, next, 1, counts, offsets, capacity, pair_e, pair_f, status);
    near_blocks<<<capacity, 64, 0, st>>>(n, p0, p1, next, offsets, pair_e, pair_f, capacity, k, blocks, kind,
                                         status);
    near_adaptive<<<capacity < 1024 ? capacity : 1024, 256, 0, st>>>(n, p0, p1, offsets, pair_e, pair_f, capacity, k,
                                                                     blocks, kind, status);
    return hipGetLastError();
}

// The near-pair CSR lists alone (count, scan, list), for the polynomial near pass.
extern "C" hipError_t grim_launch_near_list(int n, const double* p0, const double* p1, const int* next, int* counts,
                                            int* offsets, int capacity, int* pair_e, int* pair_f, unsigned* status,
                                            hipStream_t st) {
    near_list<<<n, 64, 0, st>>>(n, p0, p1, next, 0, counts, offsets, capacity, pair_e, pair_f, status);
    scan_counts<<<1, 1024, 0, st>>>(n, counts, offsets);
    near_list<<<n, 64, 0, st>>>(n, p0, p1, next, 1, counts, offsets, capacity, pair_e, pair_f, status);
    return hipGetLastError();
}

// Impedance (Robin) rows of closed contours, in place over a column-major pair
// of rows x cols arrays: x <- x + ax[r] y on every row, y <- y on Dirichlet rows
// and x + ay[r] y elsewhere (both from the original x, y).  With x, y the TE and
// TM systems (or their loads) this forms the flux rows dn(u) + alpha u of the
// combined potential (twod/formulations/combined_regions.py:43-62,
// twod/formulations/robin.py:25-83) for a law that is constant on each contour.
__global__ void __launch_bounds__(256) robin_rows_kernel(int64_t rows, int64_t cols, double* __restrict__ x,
                                                         int64_t ldx, double* __restrict__ y, int64_t ldy,
                                                         const double* __restrict__ ax, const double* __restrict__ ay,
                                                         const int* __restrict__ dirichlet) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (r >= rows || c >= cols) return;
    double* px = x + 2 * (r + c * ldx);
    double* py = y + 2 * (r + c * ldy);
    const double xr = px[0], xi = px[1], yr = py[0], yi = py[1];
    const double axr = ax[2 * r], axi = ax[2 * r + 1];
    px[0] = xr + (axr * yr - axi * yi);
    px[1] = xi + (axr * yi + axi * yr);
    if (!dirichlet[r]) {
        const double ayr = ay[2 * r], ayi = ay[2 * r + 1];
        py[0] = xr + (ayr * yr - ayi * yi);
        py[1] = xi + (ayr * yi + ayi * yr);
    }
}

// ---------------------------------------------------------------------------
// Surface impedance that varies along a contour (zones, tapers, tables)
// ---------------------------------------------------------------------------
// The Robin row of node p weights S by each observation element's own coefficient:
//   alpha_e0 S0[p, :] + alpha_e1 S1[p, :] = alpha_p S[p, :] + delta_p (S0 - S1)[p, :],
// where e0 = p (basis 1 - t), e1 = prev p (basis t), alpha_p = (alpha_e0 + alpha_e1) / 2 and
// delta_p = (alpha_e0 - alpha_e1) / 2; its trace jump uses the element-weighted mass, which
// differs from alpha_p M by delta_p m_p, m_p = (L_e0/6 at next p, (L_e0 - L_e1)/3 at p,
// -L_e1/6 at prev p) (twod/formulations/regions.py:274-304, combined_regions.py:65-83).
// robin_rows forms Z_TE + alpha_p Z_TM; robin_split adds delta_p ((S0 - S1)[p, :] + gamma/2 m_p).

// S block of a far ordered pair (obs o, src s), the Q x Q rule of far_pair.
template <int Q>
__device__ __forceinline__ void far_pair_s(const Elem& o, const Elem& s, double k, cplx (&S)[2][2]) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) S[a][b] = {0.0, 0.0};
    for (int qi = 0; qi < Q; ++qi) {
        const double ti = Rule<Q>::t(qi), wi = Rule<Q>::w(qi);
        const double xo = fma(ti, o.sx, o.x0), yo = fma(ti, o.sy, o.y0);
        cplx ps0 = {0, 0}, ps1 = {0, 0};
        for (int qj = 0; qj < Q; ++qj) {
            const double tj = Rule<Q>::t(qj), wj = Rule<Q>::w(qj);
            const double dx = xo - fma(tj, s.sx, s.x0);
            const double dy = yo - fma(tj, s.sy, s.y0);
            const double r2 = fmax(fma(dx, dx, dy * dy), GRIM_EPS * GRIM_EPS);
            double inv = __builtin_amdgcn_rsq(r2);
            inv = inv * fma(-0.5 * r2, inv * inv, 1.5);
            inv = inv * fma(-0.5 * r2, inv * inv, 1.5);
            double r = r2 * inv;
            r = fma(0.5 * inv, fma(-r, r, r2), r);
            cplx g, h;
            green_kernels(k, r, g, h);
            const double w0 = wj * (1.0 - tj), w1 = wj * tj;
            ps0.re = fma(w0, g.re, ps0.re); ps0.im = fma(w0, g.im, ps0.im);
            ps1.re = fma(w1, g.re, ps1.re); ps1.im = fma(w1, g.im, ps1.im);
        }
        const double wo0 = wi * (1.0 - ti), wo1 = wi * ti;
        S[0][0] = cadd(S[0][0], cscale(ps0, wo0)); S[0][1] = cadd(S[0][1], cscale(ps1, wo0));
        S[1][0] = cadd(S[1][0], cscale(ps0, wo1)); S[1][1] = cadd(S[1][1], cscale(ps1, wo1));
    }
    const double scale = o.len * s.len;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) S[a][b] = cscale(S[a][b], scale);
}

// S block of the ordered pair (e, f): far rule, or the near block the near pass stored for it.
template <int Q>
__device__ __forceinline__ void pair_s(const Elem& o, int e, int ne, const double* __restrict__ p0,
                                       const double* __restrict__ p1, const int* __restrict__ next, int f, double k,
                                       const int* __restrict__ offsets, const int* __restrict__ pair_f,
                                       const double* __restrict__ blocks, cplx (&S)[2][2]) {
    const Elem s = load_elem(p0, p1, f);
    if (is_far(o, e, ne, s, f, next[f])) {
        far_pair_s<Q>(o, s, k, S);
        return;
    }
    const int idx = find_pair(offsets, pair_f, e, f);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int at = 16 * idx + 2 * (2 * a + b);
            S[a][b] = idx < 0 ? cplx{0.0, 0.0} : cplx{blocks[at], blocks[at + 1]};
        }
}

// One wavefront per (walk block, column chunk).  walk[64 b + l] is an element (-1: none): lane 0
// holds the predecessor of lane 1's row, lane l >= 1 a row p whose previous element is lane
// l - 1's, so (S1)[p, :] comes from lane l - 1 by a shuffle.  Along the chunk's columns q the
// block S(e, prev q) is carried from the previous column (re-evaluated at contour starts).
// Row p of each system gets c[p] ((S0 - S1)[p, q] + gamma/2 m_p(q)), the mass part only on rows
// whose flag bit is set (bit 0: TE system, bit 1: TM system).  Every entry has one writer.
template <int Q>
__global__ void __launch_bounds__(64) robin_split(int n, const double* __restrict__ p0, const double* __restrict__ p1,
                                                  const int* __restrict__ next, const int* __restrict__ prev,
                                                  double k, const int* __restrict__ walk, int chunk,
                                                  const int* __restrict__ offsets, const int* __restrict__ pair_f,
                                                  int capacity, const double* __restrict__ blocks,
                                                  const double* __restrict__ c_te,
                                                  const double* __restrict__ c_tm, const int* __restrict__ flags,
                                                  double* __restrict__ zte, int64_t ldte, double* __restrict__ ztm,
                                                  int64_t ldtm) {
    const int lane = threadIdx.x;
    if (offsets[n] > capacity) return;                      // overflow: the lists were not written
    const int w = walk[64 * (int64_t)blockIdx.x + lane];
    const bool active = w >= 0 && w < n;
    const int e = active ? w : 0;
    const Elem o = load_elem(p0, p1, e);
    const int ne = next[e];
    const bool row = active && lane > 0;
    cplx cte = {0.0, 0.0}, ctm = {0.0, 0.0};
    int flag = 0, pe = 0;
    double l0 = 0.0, l1 = 0.0;
    if (row) {
        cte = {c_te[2 * e], c_te[2 * e + 1]};
        ctm = {c_tm[2 * e], c_tm[2 * e + 1]};
        flag = flags[e];
        pe = prev[e];
        l0 = o.len;
        l1 = load_elem(p0, p1, pe).len;
    }
    const double half_g = -0.5 / k;   // gamma / 2 = (0, -1 / (2k))
    const int c0 = blockIdx.y * chunk;
    const int c1 = min(n, c0 + chunk);
    cplx carry[2][2], cur[2][2];
    for (int q = c0; q < c1; ++q) {
        const int pq = prev[q];
        if (q == c0 || pq != q - 1) {
            if (active) pair_s<Q>(o, e, ne, p0, p1, next, pq, k, offsets, pair_f, blocks, carry);
            else
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) carry[a][b] = {0.0, 0.0};
        }
        if (active) pair_s<Q>(o, e, ne, p0, p1, next, q, k, offsets, pair_f, blocks, cur);
        else
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) cur[a][b] = {0.0, 0.0};
        const cplx own = cadd(cur[0][0], carry[0][1]);
        const cplx up = cadd(cur[1][0], carry[1][1]);
        const double ur = __shfl_up(up.re, 1, 64), ui = __shfl_up(up.im, 1, 64);
        if (row) {
            const cplx d = {own.re - ur, own.im - ui};
            double m = 0.0;
            if (q == ne) m += l0 / 6.0;
            if (q == e) m += (l0 - l1) / 3.0;
            if (q == pe) m -= l1 / 6.0;
            const cplx dm = {d.re, d.im + half_g * m};
            const cplx vte = cmul(cte, (flag & 1) ? dm : d);
            const cplx vtm = cmul(ctm, (flag & 2) ? dm : d);
            double* t = zte + 2 * ((int64_t)e + (int64_t)q * ldte);
            t[0] += vte.re;
            t[1] += vte.im;
            double* u = ztm + 2 * ((int64_t)e + (int64_t)q * ldtm);
            u[0] += vtm.re;
            u[1] += vtm.im;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) carry[a][b] = cur[a][b];
    }
}

// Right-hand sides of one polarization's impedance rows: out[p, a] = -b_u[p, a] on Dirichlet
// rows, else -sum_e i_e(p) (j k d.n_e + alpha_e): the trace load weighted by each element's own
// Robin coefficient (twod/assembly/kernels.py:222-294, observation coefficients;
// twod/formulations/regions.py:557-590).
__global__ void __launch_bounds__(256) robin_rhs_kernel(int n, int na, const double* __restrict__ p0,
                                                        const double* __restrict__ p1, const int* __restrict__ prev,
                                                        double k, const double* __restrict__ angles_deg,
                                                        const double* __restrict__ alpha_e,
                                                        const int* __restrict__ dirichlet, double* __restrict__ out,
                                                        int64_t ld) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (p >= n) return;
    const double phi = angles_deg[a] * (GRIM_PI / 180.0);
    double dy, dx;
    sincos(phi, &dy, &dx);
    const int pp = prev[p];
    const Elem A = load_elem(p0, p1, p), B = load_elem(p0, p1, pp);
    cplx i0, i1, dnA, j0_, j1_, dnB;
    elem_moments(A, k, dx, dy, i0, i1, dnA);
    elem_moments(B, k, dx, dy, j0_, j1_, dnB);
    cplx v;
    if (dirichlet[p]) {
        v = cadd(i0, j1_);
    } else {
        const cplx wa = cadd(dnA, cplx{alpha_e[2 * p], alpha_e[2 * p + 1]});
        const cplx wb = cadd(dnB, cplx{alpha_e[2 * pp], alpha_e[2 * pp + 1]});
        v = cadd(cmul(i0, wa), cmul(j1_, wb));
    }
    const int64_t at = 2 * ((int64_t)p + (int64_t)a * ld);
    out[at] = -v.re;
    out[at + 1] = -v.im;
}

extern "C" hipError_t grim_launch_robin_split(int q, int n, const double* p0, const double* p1, const int* next,
                                              const int* prev, double k, int nblocks, const int* walk,
                                              const int* offsets, const int* pair_f, int capacity,
                                              const double* blocks, const double* c_te, const double* c_tm,
                                              const int* flags, double* zte, int64_t ldte, double* ztm, int64_t ldtm,
                                              hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    const int chunk = 32;
    dim3 grid((unsigned)nblocks, (unsigned)((n + chunk - 1) / chunk));
    if (q == 16)
        robin_split<16><<<grid, 64, 0, st>>>(n, p0, p1, next, prev, k, walk, chunk, offsets, pair_f, capacity, blocks,
                                             c_te, c_tm, flags, zte, ldte, ztm, ldtm);
    else
        robin_split<8><<<grid, 64, 0, st>>>(n, p0, p1, next, prev, k, walk, chunk, offsets, pair_f, capacity, blocks,
                                            c_te, c_tm, flags, zte, ldte, ztm, ldtm);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_robin_rhs(int n, int na, const double* p0, const double* p1, const int* prev,
                                            double k, const double* angles_deg, const double* alpha_e,
                                            const int* dirichlet, double* out, int64_t ld, hipStream_t st) {
    dim3 grid((n + 255) / 256, na);
    robin_rhs_kernel<<<grid, 256, 0, st>>>(n, na, p0, p1, prev, k, angles_deg, alpha_e, dirichlet, out, ld);
    return hipGetLastError();
}

// z(block i, block j) *= scale[i + 2 j] over the four n x n blocks of a 2n x 2n matrix.
__global__ void __launch_bounds__(256) scale_blocks_kernel(int64_t n, double* __restrict__ z, int64_t ld, double s00r,
                                                           double s00i, double s10r, double s10i, double s01r,
                                                           double s01i, double s11r, double s11i) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (r >= 2 * n || c >= 2 * n) return;
    const bool lo_r = r < n, lo_c = c < n;
    const double sr = lo_c ? (lo_r ? s00r : s10r) : (lo_r ? s01r : s11r);
    const double si = lo_c ? (lo_r ? s00i : s10i) : (lo_r ? s01i : s11i);
    double* t = z + 2 * (r + c * ld);
    const double xr = t[0], xi = t[1];
    t[0] = sr * xr - si * xi;
    t[1] = sr * xi + si * xr;
}

extern "C" hipError_t grim_launch_scale_blocks(int64_t n, double* z, int64_t ld, const double* s, hipStream_t st) {
    if (n == 0) return hipSuccess;
    dim3 grid((unsigned)((2 * n + 255) / 256), (unsigned)(2 * n));
    scale_blocks_kernel<<<grid, 256, 0, st>>>(n, z, ld, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]);
    return hipGetLastError();
}

// dst[rows_dst[i], cols_dst[j]] += w[i] src[rows_src[i], cols_src[j]] for i < m, j < n (column-major
// complex128; a negative destination index skips the entry).  Composes regional systems from
// assembled operator blocks (twod/assembly/scatter.py:65-116, SystemScatter).
__global__ void __launch_bounds__(256) route_kernel(int m, int n, const double* __restrict__ src, int64_t ld_src,
                                                    const int* __restrict__ rows_src, const int* __restrict__ cols_src,
                                                    double* __restrict__ dst, int64_t ld_dst,
                                                    const int* __restrict__ rows_dst, const int* __restrict__ cols_dst,
                                                    const double* __restrict__ w) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= m || j >= n) return;
    const int rd = rows_dst[i], cd = cols_dst[j];
    if (rd < 0 || cd < 0) return;
    const double* a = src + 2 * ((int64_t)rows_src[i] + (int64_t)cols_src[j] * ld_src);
    double* t = dst + 2 * ((int64_t)rd + (int64_t)cd * ld_dst);
    const double wr = w[2 * i], wi = w[2 * i + 1], ar = a[0], ai = a[1];
    t[0] += wr * ar - wi * ai;
    t[1] += wr * ai + wi * ar;
}

extern "C" hipError_t grim_launch_route(int m, int n, const double* src, int64_t ld_src, const int* rows_src,
                                       const int* cols_src, double* dst, int64_t ld_dst, const int* rows_dst,
                                       const int* cols_dst, const double* w, hipStream_t st) {
    if (m == 0 || n == 0) return hipSuccess;
    dim3 grid((unsigned)((m + 255) / 256), (unsigned)n);
    route_kernel<<<grid, 256, 0, st>>>(m, n, src, ld_src, rows_src, cols_src, dst, ld_dst, rows_dst, cols_dst, w);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_robin_rows(int64_t rows, int64_t cols, double* x, int64_t ldx, double* y,
                                             int64_t ldy, const double* ax, const double* ay, const int* dirichlet,
                                             hipStream_t st) {
    if (rows == 0 || cols == 0) return hipSuccess;
    dim3 grid((unsigned)((rows + 255) / 256), (unsigned)cols);
    robin_rows_kernel<<<grid, 256, 0, st>>>(rows, cols, x, ldx, y, ldy, ax, ay, dirichlet);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_loads(int n, int na, const double* p0, const double* p1, const int* prev, double k,
                                        const double* angles_deg, double* bu, double* bdn, int64_t ld,
                                        hipStream_t st) {
    dim3 grid((n + 255) / 256, na);
    loads_kernel<<<grid, 256, 0, st>>>(n, na, p0, p1, prev, k, angles_deg, bu, bdn, ld);
    return hipGetLastError();
}

extern "C" hipError_t grim_launch_farfield(int n, int na, const double* x, int64_t ldx, const double* bu,
                                           const double* bdn, int64_t ldb, double k, double* amp, double* sigma,
                                           hipStream_t st) {
    farfield_kernel<<<na, 256, 0, st>>>(n, x, ldx, bu, bdn, ldb, k, amp, sigma);
    return hipGetLastError();
}
