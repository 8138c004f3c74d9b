// Dense complex linear algebra on MFMA for the BEM systems: blocked LU with
// partial pivoting (complex128 or complex64), triangular solves for many
// right-hand sides, and the FP64 residual of iterative refinement.
//
// Replaces scipy.linalg.lu_factor / lu_solve as used by the reference's
// DenseFactor (tools/GHOST/ghost_backend/linalg/dense.py:114-155, 239-301) and
// RefinedLU (linalg/refined_lu.py:25-61).
//
// Storage: column-major, complex interleaved (re, im), element (r, c) at
// a[2 * (r + c * lda)].  Complex GEMM is four real MFMA products per k-step
// (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32).
//
// Factorization: right-looking, panels of GRIM_LU_NB = 64 columns, each panel
// factored recursively down to register-resident leaves.  Row interchanges
// follow LAPACK (whole rows, P A = L U): each leaf applies its own to the rest
// of the panel, and each panel's interchanges are composed into one list that
// is applied as a gather to the columns left of the panel and, fused into the
// U12 triangular solve, to the columns right of it; the same list updates the
// permutation vector used by the solves.  Triangular solves run lane-parallel
// (one row per lane, x_j broadcast with v_readlane): up to 64 rows per launch in
// the factorization, and one 256-row diagonal block per launch in the LU solve,
// whose four 64-row sub-blocks are chained inside the kernel; everything else is
// MFMA matrix products.
// All kernels take a batch of equally strided matrices (blockIdx.z or one
// workgroup per matrix), e.g. the TE and TM systems of one request.
#include "common.h"

template <typename T> struct cx { T re, im; };

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Output element v of lane l of a 16 x 16 tile sits at column l % 16 and row
// row(l / 16, v): f64 16x16x4 interleaves rows (l/16 + 4 v), f32 16x16x4 keeps
// them contiguous (4 (l/16) + v) -- measured on gfx950.
template <> struct Mfma<double> {
    typedef d4 acc;
    static __device__ __forceinline__ int row(int fk, int v) { return fk + 4 * v; }
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef f4 acc;
    static __device__ __forceinline__ int row(int fk, int v) { return 4 * fk + v; }
    static __device__ __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// ---------------------------------------------------------------------------
// C[m x n] (-)= A[m x k] * B[k x n]   (complex, column-major, interleaved)
// SET: C = A B, else C -= A B.  256 threads, (32 MI) x 64 output tile, 4 waves of
// (16 MI) x 32 (MI x 2 MFMA tiles, MI = 1 or 2); k staged through LDS GB_K<T> at a time (16 for
// complex128, 32 for complex64) with the next k-slab prefetched into
// registers while the current one is multiplied.  The MFMA computes the
// transposed tile (B^T A^T), so a lane's outputs run down a column of C and
// the read-modify-write of C is coalesced.  blockIdx.z selects the matrix of a
// strided batch (strides in complex elements).
// ---------------------------------------------------------------------------
#define GB_N 64
template <typename T> struct Cplx2;
template <> struct Cplx2<double> { typedef double2 type; };
template <> struct Cplx2<float> { typedef float2 type; };
template <typename T> constexpr int gb_k() { return sizeof(T) == 8 ? 16 : 32; }

// complex128 without the C read-ahead is held to 3 waves per SIMD (134 VGPRs instead of 210):
// the 2000 x 720 x 2000 refinement residual's 768 workgroups then run in one round
// (942 -> 845 us, measured); the other forms keep the compiler's choice.
// TA: A is used conjugate-transposed, C (-)= A^H B with A stored k x m (lda >= k); its k-slab is
// staged with the k index running along the workgroup's threads, so those reads stay coalesced.
template <typename T, bool SET, bool PRE = false, int MI = 2, bool TA = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && !PRE ? 3 : 1))) cgemm(int m, int n, int k, const T* __restrict__ A, int64_t lda, int64_t sa,
                                             const T* __restrict__ B, int64_t ldb, int64_t sb, T* __restrict__ C,
                                             int64_t ldc, int64_t sc) {
    typedef typename Mfma<T>::acc acc_t;
    typedef typename Cplx2<T>::type v2;
    constexpr int BK = gb_k<T>();
    constexpr int BM = 32 * MI;   // tile rows: each wave covers MI 16-row MFMA tiles
    constexpr int PA = BK * BM / 256;   // complex A elements each thread stages per k-slab
    constexpr int PB = BK / 4;   // complex B elements (BK x 64 / 256)
    __shared__ T a_re[BK][BM + 1], a_im[BK][BM + 1];
    __shared__ T b_re[BK][GB_N + 1], b_im[BK][GB_N + 1];
    A += 2 * sa * blockIdx.z;
    B += 2 * sb * blockIdx.z;
    C += 2 * sc * blockIdx.z;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * GB_N;
    const int wm = (wave & 1) * 16 * MI, wn = (wave >> 1) * 32;
    acc_t cr[MI][2], ci[MI][2];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cr[i][j] = acc_t{0, 0, 0, 0};
            ci[i][j] = acc_t{0, 0, 0, 0};
        }
    const int fr = lane & 15, fk = lane >> 4;
    // staging map: A rows tid % BM, columns tid / BM + (256 / BM) i;  B rows tid % BK, columns tid / BK + (256 / BK) i
    const int ar0 = tid % BM, ac0 = tid / BM;
    const int br0 = tid % BK, bc0 = tid / BK;
    v2 pa[PA], pb[PB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < PA; ++i) {
            if constexpr (TA) {
                const int gk = k0 + tid % BK, gm = m0 + tid / BK + (256 / BK) * i;
                v2 v = (gm < m && gk < k) ? *(const v2*)(A + 2 * ((int64_t)gk + (int64_t)gm * lda)) : v2{0, 0};
                v.y = -v.y;
                pa[i] = v;
            } else {
                const int gr = m0 + ar0, gc = k0 + ac0 + (256 / BM) * i;
                pa[i] = (gr < m && gc < k) ? *(const v2*)(A + 2 * ((int64_t)gr + (int64_t)gc * lda)) : v2{0, 0};
            }
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int gr = k0 + br0, gc = n0 + bc0 + (256 / BK) * i;
            pb[i] = (gr < k && gc < n) ? *(const v2*)(B + 2 * ((int64_t)gr + (int64_t)gc * ldb)) : v2{0, 0};
        }
    };
    // PRE: this thread's C values are read before the k-loop, so a short update (k <= 64, the
    // LU's rank-64 steps) overlaps their latency with the MFMA work instead of the epilogue
    v2 cpre[MI][2][4];
    if (PRE && !SET) {
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int gr = m0 + wm + 16 * i + fr;
                    const int gc = n0 + wn + 16 * j + Mfma<T>::row(fk, v);
                    cpre[i][j][v] = (gr < m && gc < n) ? *(const v2*)(C + 2 * ((int64_t)gr + (int64_t)gc * ldc))
                                                       : v2{0, 0};
                }
    }
    fetch(0);
    for (int k0 = 0; k0 < k; k0 += BK) {
#pragma unroll
        for (int i = 0; i < PA; ++i) {
            if constexpr (TA) {
                a_re[tid % BK][tid / BK + (256 / BK) * i] = pa[i].x;
                a_im[tid % BK][tid / BK + (256 / BK) * i] = pa[i].y;
            } else {
                a_re[ac0 + (256 / BM) * i][ar0] = pa[i].x;
                a_im[ac0 + (256 / BM) * i][ar0] = pa[i].y;
            }
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            b_re[br0][bc0 + (256 / BK) * i] = pb[i].x;
            b_im[br0][bc0 + (256 / BK) * i] = pb[i].y;
        }
        __syncthreads();
        if (k0 + BK < k) fetch(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T ar[MI], ai[MI], br[2], bi[2], nbi[2];
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                ar[i] = a_re[kk + fk][wm + 16 * i + fr];
                ai[i] = a_im[kk + fk][wm + 16 * i + fr];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                br[j] = b_re[kk + fk][wn + 16 * j + fr];
                bi[j] = b_im[kk + fk][wn + 16 * j + fr];
                nbi[j] = -bi[j];
            }
            // transposed tile: D[j-col][i-row] = sum_k B[k][j] A[i][k]
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    cr[i][j] = Mfma<T>::run(br[j], ar[i], cr[i][j]);
                    cr[i][j] = Mfma<T>::run(nbi[j], ai[i], cr[i][j]);
                    ci[i][j] = Mfma<T>::run(bi[j], ar[i], ci[i][j]);
                    ci[i][j] = Mfma<T>::run(br[j], ai[i], ci[i][j]);
                }
        }
        __syncthreads();
    }
    // lane holds D[row(lane / 16, v)][lane % 16] of each transposed 16 x 16 tile:
    // C row = lane % 16 (consecutive lanes, consecutive rows), C column = row(lane / 16, v)
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gr = m0 + wm + 16 * i + fr;
                const int gc = n0 + wn + 16 * j + Mfma<T>::row(fk, v);
                if (gr < m && gc < n) {
                    v2* p = (v2*)(C + 2 * ((int64_t)gr + (int64_t)gc * ldc));
                    if (SET) {
                        *p = v2{cr[i][j][v], ci[i][j][v]};
                    } else {
                        v2 c = PRE ? cpre[i][j][v] : *p;
                        c.x -= cr[i][j][v];
                        c.y -= ci[i][j][v];
                        *p = c;
                    }
                }
            }
}

// ---------------------------------------------------------------------------
// (value, index) argmax across lanes without LDS: larger value wins, ties go to
// the smaller index (LAPACK izamax picks the first maximum).  DPP quad
// permutations and row mirrors reduce within rows of 16 lanes, then gfx950's
// v_permlane16_swap / v_permlane32_swap combine the rows; every lane ends with
// the result.  The comparisons are branch-free selects.
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ void am_take(T& v, int& i, T ov, int oi) {
    const bool t = (ov > v) | ((ov == v) & (oi < i));
    v = t ? ov : v;
    i = t ? oi : i;
}
template <int CTRL> __device__ __forceinline__ int dpp_mov(int x) {
    return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}
template <int CTRL> __device__ __forceinline__ void am_dpp(double& v, int& i) {
    const int lo = dpp_mov<CTRL>(__double2loint(v)), hi = dpp_mov<CTRL>(__double2hiint(v));
    am_take(v, i, __hiloint2double(hi, lo), dpp_mov<CTRL>(i));
}
template <int CTRL> __device__ __forceinline__ void am_dpp(float& v, int& i) {
    am_take(v, i, __int_as_float(dpp_mov<CTRL>(__float_as_int(v))), dpp_mov<CTRL>(i));
}
template <typename T> __device__ __forceinline__ void argmax16(T& v, int& i) {
    am_dpp<0xB1>(v, i);    // quad_perm [1,0,3,2]
    am_dpp<0x4E>(v, i);    // quad_perm [2,3,0,1]
    am_dpp<0x141>(v, i);   // row_half_mirror
    am_dpp<0x140>(v, i);   // row_mirror
}
// v_permlane{16,32}_swap(x, x) returns (x of the even rows / low half, x of
// the odd rows / high half) in every lane, so one select combines partners.
#define GRIM_AM_SWAP(NAME, BUILTIN)                                                     \
    __device__ __forceinline__ void NAME(double& v, int& i) {                           \
        const auto lo = BUILTIN(__double2loint(v), __double2loint(v), false, false);    \
        const auto hi = BUILTIN(__double2hiint(v), __double2hiint(v), false, false);    \
        const auto ix = BUILTIN(i, i, false, false);                                    \
        v = __hiloint2double(hi[0], lo[0]);                                             \
        i = ix[0];                                                                      \
        am_take(v, i, __hiloint2double(hi[1], lo[1]), (int)ix[1]);                      \
    }                                                                                   \
    __device__ __forceinline__ void NAME(float& v, int& i) {                            \
        const auto x = BUILTIN(__float_as_int(v), __float_as_int(v), false, false);     \
        const auto ix = BUILTIN(i, i, false, false);                                    \
        v = __int_as_float(x[0]);                                                       \
        i = ix[0];                                                                      \
        am_take(v, i, __int_as_float(x[1]), (int)ix[1]);                                \
    }
GRIM_AM_SWAP(am_swap16, __builtin_amdgcn_permlane16_swap)
GRIM_AM_SWAP(am_swap32, __builtin_amdgcn_permlane32_swap)
#undef GRIM_AM_SWAP
__device__ __forceinline__ double first_lane(double v) {
    return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                            __builtin_amdgcn_readfirstlane(__double2loint(v)));
}
__device__ __forceinline__ float first_lane(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
template <typename T> __device__ __forceinline__ void argmax64(T& v, int& i) {
    argmax16(v, i);
    am_swap16(v, i);
    am_swap32(v, i);
}

// ---------------------------------------------------------------------------
// Composite of a sequence of row interchanges, built by one wave.
// The interchanges are (r0 + i) <-> piv[i] for i = 0 .. cnt-1 (cnt <= 64,
// piv[i] >= r0 + i, absolute rows, applied in order).  On return lane l holds
//   top  = the row whose content ends at position r0 + l   (l < cnt),
//   ldst / lsrc = a lower position (>= r0 + cnt) that changes and the row
//          whose content ends there (l < nlow).
// Every position whose content changes is a top slot or a lower entry, and
// every source is one of those positions, so applying the interchanges to a
// column is one gather: read all sources, then write all destinations.
// ---------------------------------------------------------------------------
struct Composite {
    int top, ldst, lsrc, nlow;
};

__device__ Composite wave_composite_of(int r0, int cnt, int pl) {   // lane i passes piv[i]
    const int lane = threadIdx.x & 63;
    // Every lane traces one final position back through the interchanges in
    // reverse order, so the cnt steps carry no cross-lane dependence: lane l
    // traces top slot r0 + l and lower position piv[l].
    int top = r0 + lane, src = pl;
    bool repeat = false;   // an earlier interchange already targets piv[l]
    for (int i = cnt - 1; i >= 0; --i) {
        const int p = __builtin_amdgcn_readlane(pl, i);
        const int t = r0 + i;
        top = top == t ? p : top == p ? t : top;
        src = src == t ? p : src == p ? t : src;
        repeat = repeat || (i < lane && p == pl);
    }
    Composite x;
    x.top = lane < cnt ? top : r0 + lane;
    x.ldst = -1;
    x.lsrc = -1;
    // lower entries in order of first occurrence: the k-th moves to lane k
    unsigned long long first = __ballot(lane < cnt && pl - r0 >= cnt && !repeat);
    x.nlow = __popcll(first);
    for (int k = 0; first; ++k) {
        const int i = __ffsll((long long)first) - 1;
        first &= first - 1;
        const int d = __builtin_amdgcn_readlane(pl, i), s = __builtin_amdgcn_readlane(src, i);
        if (lane == k) { x.ldst = d; x.lsrc = s; }
    }
    return x;
}

__device__ Composite wave_composite(int r0, int cnt, const int* piv) {
    const int lane = threadIdx.x & 63;
    return wave_composite_of(r0, cnt, lane < cnt ? piv[lane] : 0);
}

// Interchange list of one 64-column panel in the auxiliary storage:
// [0] = nlow, [1 + l] = top source of slot l, [65 + l] / [129 + l] = lower
// destination / source of entry l (absolute rows).
#define GRIM_LIST_INTS 256

// One wave: store the interchange list of the panel [pj, pj + pw) (lane i
// passes the panel's pivot i) and apply it to the running row permutation
// (perm[i] = original row now at position i).
__device__ void panel_list(int pj, int pw, int pl, int* __restrict__ perm, int* __restrict__ list) {
    const int lane = threadIdx.x & 63;
    const Composite x = wave_composite_of(pj, pw, pl);
    if (lane == 0) list[0] = x.nlow;
    list[1 + lane] = x.top;
    list[65 + lane] = x.ldst;
    list[129 + lane] = x.lsrc;
    const int vt = lane < pw ? perm[x.top] : 0;
    const int vl = lane < x.nlow ? perm[x.lsrc] : 0;
    if (lane < pw) perm[pj + lane] = vt;
    if (lane < x.nlow) perm[x.ldst] = vl;
}

// ---------------------------------------------------------------------------
// Panel leaf: factor leaf columns [lc0, lc0 + w) of the panel whose first
// column starts at `pan` (row r0 = first row of the leaf, panel width pw),
// rows r0 .. r0 + m - 1, with the leaf's block cached in registers.
// One workgroup of 1024 threads per matrix; thread t owns rows t, t + 1024, ...
// Pivot = first row maximising |re| + |im| (LAPACK izamax).  Per column: a
// register-only wave argmax, one LDS slot per wave and a barrier, a 16-lane
// argmax of the slots, then the pivot row and row c through double-buffered
// LDS and a second barrier.  Pivots stay in LDS until the end.  Finally the
// leaf's interchanges are applied, as one gather, to the other columns of the
// panel, so the panel recursion needs no separate row-swap passes.
// ---------------------------------------------------------------------------
template <typename T, int R, int W>
__global__ void __launch_bounds__(1024) panel_leaf(int m, int w, T* __restrict__ pan, int64_t lda, int64_t sa,
                                                   int pw, int lc0, int row0, int* __restrict__ piv, int64_t spiv,
                                                   unsigned* __restrict__ status, int* __restrict__ perm,
                                                   int64_t sperm, int* __restrict__ list) {
    __shared__ T cand_v[2][16];
    __shared__ int cand_i[2][16];
    __shared__ T prow_re[2][W], prow_im[2][W], crow_re[2][W], crow_im[2][W];
    __shared__ T pinv[2][2];   // 1 / pivot, formed once by the pivot row's thread
    __shared__ int lpiv[W];
    __shared__ int zero_pivot;
    __shared__ int list_dst[2 * W], list_src[2 * W];
    __shared__ int list_n;
    pan += 2 * sa * blockIdx.x;
    piv += spiv * blockIdx.x;
    T* a = pan + 2 * (int64_t)lc0 * lda;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) zero_pivot = 0;
    T vr[R][W], vi[R][W];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = tid + 1024 * r;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            T x = 0, y = 0;
            if (row < m && c < w) {
                const T* p = a + 2 * ((int64_t)row + (int64_t)c * lda);
                x = p[0];
                y = p[1];
            }
            vr[r][c] = x;
            vi[r][c] = y;
        }
    }
#pragma unroll
    for (int c = 0; c < W; ++c) {
        if (c < w) {
            const int buf = c & 1;
            T best = -1;
            int bi = 0x7fffffff;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = tid + 1024 * r;
                if (row >= c && row < m) {
                    T v = fabs(vr[r][c]) + fabs(vi[r][c]);
                    v = v == v ? v : T(-0.5);   // NaN ranks below every number: a valid row always wins
                    am_take(best, bi, v, row);
                }
            }
            argmax64(best, bi);
            if (lane == 0) { cand_v[buf][wave] = best; cand_i[buf][wave] = bi; }
            __syncthreads();
            T bv = T(-2);
            int p = 0x7fffffff;
            if (lane < 16) { bv = cand_v[buf][lane]; p = cand_i[buf][lane]; }
            argmax16(bv, p);
            p = __builtin_amdgcn_readfirstlane(p);
            if (tid == 0) {
                lpiv[c] = p;
                if (!(bv > 0)) zero_pivot = 1;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = tid + 1024 * r;
                if (row == p) {
#pragma unroll
                    for (int cc = 0; cc < W; ++cc) { prow_re[buf][cc] = vr[r][cc]; prow_im[buf][cc] = vi[r][cc]; }
                    const T dr = vr[r][c], di = vi[r][c];
                    const T den = dr * dr + di * di;
                    pinv[buf][0] = den > 0 ? dr / den : 0;
                    pinv[buf][1] = den > 0 ? -di / den : 0;
                }
                if (row == c) {
#pragma unroll
                    for (int cc = 0; cc < W; ++cc) { crow_re[buf][cc] = vr[r][cc]; crow_im[buf][cc] = vi[r][cc]; }
                }
            }
            __syncthreads();
            const T ir = pinv[buf][0], ii = pinv[buf][1];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = tid + 1024 * r;
                if (p != c) {
                    if (row == c) {
#pragma unroll
                        for (int cc = 0; cc < W; ++cc) { vr[r][cc] = prow_re[buf][cc]; vi[r][cc] = prow_im[buf][cc]; }
                    } else if (row == p) {
#pragma unroll
                        for (int cc = 0; cc < W; ++cc) { vr[r][cc] = crow_re[buf][cc]; vi[r][cc] = crow_im[buf][cc]; }
                    }
                }
                if (row > c && row < m) {
                    const T lr = vr[r][c] * ir - vi[r][c] * ii;
                    const T li = vr[r][c] * ii + vi[r][c] * ir;
                    vr[r][c] = lr;
                    vi[r][c] = li;
#pragma unroll
                    for (int cc = 0; cc < W; ++cc) {
                        if (cc > c) {
                            const T ur = prow_re[buf][cc], ui = prow_im[buf][cc];
                            vr[r][cc] -= lr * ur - li * ui;
                            vi[r][cc] -= lr * ui + li * ur;
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = tid + 1024 * r;
        if (row < m) {
#pragma unroll
            for (int c = 0; c < W; ++c) {
                if (c < w) {
                    T* p = a + 2 * ((int64_t)row + (int64_t)c * lda);
                    p[0] = vr[r][c];
                    p[1] = vi[r][c];
                }
            }
        }
    }
    __syncthreads();
    if (tid < w) piv[tid] = row0 + lpiv[tid];
    if (tid == 0 && zero_pivot) atomicOr(status, (unsigned)GRIM_DEV_ZERO_PIVOT);
    // The panel's last leaf (perm given) also builds the panel's interchange
    // list, on wave 1 while wave 0 builds the leaf's own.  Pivots of earlier
    // leaves come from memory (earlier launches), this leaf's from LDS.
    if (perm != nullptr && wave == 1) {
        const int pl = lane < lc0 ? piv[lane - lc0] : lane < lc0 + w ? row0 + lpiv[lane - lc0] : 0;
        panel_list(row0 - lc0, pw, pl, perm + sperm * blockIdx.x, list + sperm * blockIdx.x);
    }
    if (pw == w) return;
    // the leaf's interchanges on the panel's other columns (rows local to r0)
    if (wave == 0) {
        const Composite x = wave_composite(0, w, lpiv);
        if (lane < w) { list_dst[lane] = lane; list_src[lane] = x.top; }
        if (lane < x.nlow) { list_dst[w + lane] = x.ldst; list_src[w + lane] = x.lsrc; }
        if (lane == 0) list_n = w + x.nlow;
    }
    __syncthreads();
    const int cnt = list_n;
    const int others = pw - w;
    const int items = cnt * others;
    T sr[2], si[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int idx = tid + 1024 * k;
        sr[k] = 0;
        si[k] = 0;
        if (idx < items) {
            const int e = idx % cnt, oc = idx / cnt;
            const int col = oc < lc0 ? oc : oc + w;
            const T* p = pan + 2 * ((int64_t)list_src[e] + (int64_t)col * lda);
            sr[k] = p[0];
            si[k] = p[1];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int idx = tid + 1024 * k;
        if (idx < items) {
            const int e = idx % cnt, oc = idx / cnt;
            const int col = oc < lc0 ? oc : oc + w;
            T* p = pan + 2 * ((int64_t)list_dst[e] + (int64_t)col * lda);
            p[0] = sr[k];
            p[1] = si[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Tall-panel leaf: the same column steps as panel_leaf for panels taller than
// one workgroup's registers hold.  G workgroups (grid.x, G <= GRIM_TALL_G)
// each own a contiguous block of 1024 * R rows; per column every workgroup
// publishes its candidate (value, row and the row's TW values) and the owner
// of row c publishes row c in double-buffered device-memory slots, then all
// meet at a device-wide barrier: one atomic counter per matrix that only
// grows during a factorization (`base` = arrivals of earlier tall leaves), so
// it needs no reset between columns.  The wait is bounded: on expiry the
// kernel flags GRIM_DEV_BARRIER_TIMEOUT and carries on with clamped indices,
// so it always terminates and never leaves the matrix.  Matrices of a batch
// are grid.y.  TW = 2 leaf columns.
// ---------------------------------------------------------------------------
#define GRIM_TALL_G 8
#define GRIM_TALL_TW 2
template <typename T> struct TallScratch {
    unsigned counter;
    unsigned pad[63];
    T cand_v[2][GRIM_TALL_G];
    int cand_i[2][GRIM_TALL_G];
    T cand_re[2][GRIM_TALL_G][GRIM_TALL_TW], cand_im[2][GRIM_TALL_G][GRIM_TALL_TW];
    T crow_re[2][GRIM_TALL_TW], crow_im[2][GRIM_TALL_TW];
};

template <typename T> __device__ __forceinline__ void pub(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ T sub(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, int R>
__global__ void __launch_bounds__(1024) panel_leaf_tall(int m, int w, T* __restrict__ pan, int64_t lda, int64_t sa,
                                                        int pw, int lc0, int row0, int* __restrict__ piv,
                                                        int64_t spiv, unsigned* __restrict__ status, char* scratch,
                                                        int64_t sscr, unsigned base) {
    constexpr int W = GRIM_TALL_TW;
    __shared__ T wv[16];
    __shared__ int wi[16];
    __shared__ int lpiv[W];
    __shared__ int zero_pivot;
    const int G = gridDim.x, g = blockIdx.x;
    pan += 2 * sa * blockIdx.y;
    piv += spiv * blockIdx.y;
    TallScratch<T>* ts = (TallScratch<T>*)(scratch + sscr * blockIdx.y);
    T* a = pan + 2 * (int64_t)lc0 * lda;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rbase = g * 1024 * R;
    if (tid == 0) zero_pivot = 0;
    T vr[R][W], vi[R][W];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = rbase + tid + 1024 * r;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            T x = 0, y = 0;
            if (row < m && c < w) {
                const T* p = a + 2 * ((int64_t)row + (int64_t)c * lda);
                x = p[0];
                y = p[1];
            }
            vr[r][c] = x;
            vi[r][c] = y;
        }
    }
#pragma unroll
    for (int c = 0; c < W; ++c) {
        if (c < w) {
            const int buf = c & 1;
            T best = -1;
            int bi = 0x7fffffff;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = rbase + tid + 1024 * r;
                if (row >= c && row < m) {
                    T v = fabs(vr[r][c]) + fabs(vi[r][c]);
                    v = v == v ? v : T(-0.5);
                    am_take(best, bi, v, row);
                }
            }
            argmax64(best, bi);
            if (lane == 0) { wv[wave] = best; wi[wave] = bi; }
            __syncthreads();
            T gv = T(-2);
            int gp = 0x7fffffff;
            if (lane < 16) { gv = wv[lane]; gp = wi[lane]; }
            argmax16(gv, gp);   // lanes 0-15 hold the workgroup's result
            gp = __builtin_amdgcn_readfirstlane(gp);
            gv = first_lane(gv);
            // publish this workgroup's candidate and (if owned) row c
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = rbase + tid + 1024 * r;
                if (row == gp) {
                    pub(&ts->cand_v[buf][g], gv);
                    pub(&ts->cand_i[buf][g], gp);
#pragma unroll
                    for (int cc = 0; cc < W; ++cc) {
                        pub(&ts->cand_re[buf][g][cc], vr[r][cc]);
                        pub(&ts->cand_im[buf][g][cc], vi[r][cc]);
                    }
                }
                if (row == c) {
#pragma unroll
                    for (int cc = 0; cc < W; ++cc) {
                        pub(&ts->crow_re[buf][cc], vr[r][cc]);
                        pub(&ts->crow_im[buf][cc], vi[r][cc]);
                    }
                }
            }
            if (gp >= m) {   // no row >= c in this block: an empty candidate
                if (tid == 0) {
                    pub(&ts->cand_v[buf][g], T(-2));
                    pub(&ts->cand_i[buf][g], 0x7fffffff);
                }
            }
            __syncthreads();
            if (tid == 0) {
                __hip_atomic_fetch_add(&ts->counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned target = base + (unsigned)(c + 1) * (unsigned)G;
                unsigned spins = 0;
                while (__hip_atomic_load(&ts->counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                    if (++spins > (1u << 24)) {
                        atomicOr(status, (unsigned)GRIM_DEV_BARRIER_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __syncthreads();
            T bv = T(-2);
            int p = 0x7fffffff;
            if (lane < G) {
                int ci = sub(&ts->cand_i[buf][lane]);
                bv = sub(&ts->cand_v[buf][lane]);
                if (ci < 0 || ci >= m) {   // empty slot (or stale after an expired wait)
                    ci = m;
                    bv = T(-2);
                }
                p = ci * 16 + lane;
            }
            argmax16(bv, p);
            p = __builtin_amdgcn_readfirstlane(p);
            const int win = p & 15;
            p >>= 4;
            if (p < c || p >= m) p = c;   // only after an expired wait
            T ure[W], uim[W], cre[W], cim[W];
#pragma unroll
            for (int cc = 0; cc < W; ++cc) {
                ure[cc] = sub(&ts->cand_re[buf][win][cc]);
                uim[cc] = sub(&ts->cand_im[buf][win][cc]);
                cre[cc] = sub(&ts->crow_re[buf][cc]);
                cim[cc] = sub(&ts->crow_im[buf][cc]);
            }
            if (p == c) {
#pragma unroll
                for (int cc = 0; cc < W; ++cc) { ure[cc] = cre[cc]; uim[cc] = cim[cc]; }
            }
            if (tid == 0) {
                lpiv[c] = p;
                if (!(bv > 0)) zero_pivot = 1;
            }
            const T dr = ure[c], di = uim[c];
            const T den = dr * dr + di * di;
            const T ir = den > 0 ? dr / den : 0, ii = den > 0 ? -di / den : 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = rbase + tid + 1024 * r;
                if (p != c) {
                    if (row == c) {
#pragma unroll
                        for (int cc = 0; cc < W; ++cc) { vr[r][cc] = ure[cc]; vi[r][cc] = uim[cc]; }
                    } else if (row == p) {
#pragma unroll
                        for (int cc = 0; cc < W; ++cc) { vr[r][cc] = cre[cc]; vi[r][cc] = cim[cc]; }
                    }
                }
                if (row > c && row < m) {
                    const T lr = vr[r][c] * ir - vi[r][c] * ii;
                    const T li = vr[r][c] * ii + vi[r][c] * ir;
                    vr[r][c] = lr;
                    vi[r][c] = li;
#pragma unroll
                    for (int cc = 0; cc < W; ++cc) {
                        if (cc > c) {
                            vr[r][cc] -= lr * ure[cc] - li * uim[cc];
                            vi[r][cc] -= lr * uim[cc] + li * ure[cc];
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = rbase + tid + 1024 * r;
        if (row < m) {
#pragma unroll
            for (int c = 0; c < W; ++c) {
                if (c < w) {
                    T* p = a + 2 * ((int64_t)row + (int64_t)c * lda);
                    p[0] = vr[r][c];
                    p[1] = vi[r][c];
                }
            }
        }
    }
    __syncthreads();
    if (g != 0) return;
    if (tid < w) piv[tid] = row0 + lpiv[tid];
    if (tid == 0 && zero_pivot) atomicOr(status, (unsigned)GRIM_DEV_ZERO_PIVOT);
    if (pw == w) return;
    // the leaf's interchanges on the panel's other columns: at most 2 W rows,
    // applied by the first workgroup (rows of other columns are in memory only)
    __shared__ int list_dst[2 * W], list_src[2 * W];
    __shared__ int list_n;
    if (wave == 0) {
        const Composite x = wave_composite(0, w, lpiv);
        if (lane < w) { list_dst[lane] = lane; list_src[lane] = x.top; }
        if (lane < x.nlow) { list_dst[w + lane] = x.ldst; list_src[w + lane] = x.lsrc; }
        if (lane == 0) list_n = w + x.nlow;
    }
    __syncthreads();
    const int cnt = list_n;
    const int others = pw - w;
    const int items = cnt * others;
    T sr = 0, si = 0;
    int e = 0, col = 0;
    if (tid < items) {
        e = tid % cnt;
        const int oc = tid / cnt;
        col = oc < lc0 ? oc : oc + w;
        const T* p = pan + 2 * ((int64_t)list_src[e] + (int64_t)col * lda);
        sr = p[0];
        si = p[1];
    }
    __syncthreads();
    if (tid < items) {
        T* p = pan + 2 * ((int64_t)list_dst[e] + (int64_t)col * lda);
        p[0] = sr;
        p[1] = si;
    }
}

// Interchange list of the panel [pj, pj + pw) from its pivots; also applied to
// the running row permutation (perm[i] = original row now at position i).
// One wave per matrix.
__global__ void __launch_bounds__(64) panel_perm(int pj, int pw, const int* __restrict__ piv, int64_t spiv,
                                                 int* __restrict__ perm, int64_t sperm, int* __restrict__ list,
                                                 int64_t slist) {
    piv += spiv * blockIdx.x;
    const int lane = threadIdx.x;
    panel_list(pj, pw, lane < pw ? piv[pj + lane] : 0, perm + sperm * blockIdx.x, list + slist * blockIdx.x);
}

__global__ void __launch_bounds__(256) perm_init(int n, int* __restrict__ perm, int64_t sperm,
                                                 unsigned* __restrict__ counter) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) perm[sperm * blockIdx.z + i] = i;
    if (i == 0) counter[sperm * blockIdx.z] = 0;   // the tall leaves' barrier counter
}

// Apply a panel's interchange list to columns [0, ncols) of a (rows absolute).
// One wave per column, lanes over list entries; batch in blockIdx.z.
template <typename T>
__global__ void __launch_bounds__(256) apply_list(int ncols, int pj, int pw, T* __restrict__ a, int64_t lda,
                                                  int64_t sa, const int* __restrict__ list, int64_t slist) {
    const int lane = threadIdx.x & 63;
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= ncols) return;
    a += 2 * sa * blockIdx.z + 2 * (int64_t)col * lda;
    list += slist * blockIdx.z;
    const int nlow = list[0];
    const int ts = list[1 + lane], ld = list[65 + lane], ls = list[129 + lane];
    T tr = 0, ti = 0, lr = 0, li = 0;
    if (lane < pw) { tr = a[2 * (int64_t)ts]; ti = a[2 * (int64_t)ts + 1]; }
    if (lane < nlow) { lr = a[2 * (int64_t)ls]; li = a[2 * (int64_t)ls + 1]; }
    if (lane < pw) { a[2 * (int64_t)(pj + lane)] = tr; a[2 * (int64_t)(pj + lane) + 1] = ti; }
    if (lane < nlow) { a[2 * (int64_t)ld] = lr; a[2 * (int64_t)ld + 1] = li; }
}

#define GRIM_LU_NB 64
#ifndef GRIM_SOLVE_NB
#define GRIM_SOLVE_NB 256
#endif
static_assert(GRIM_SOLVE_NB % 128 == 0 && GRIM_SOLVE_NB <= 256, "block solve: 128 or 256 rows");

__device__ __forceinline__ double lane_value(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// X = L^{-1} B in place for a w x w unit lower triangle (w <= 64) and ncols
// columns (the factorization's U12 = L11^{-1} A12 steps).  Lane i of a wave holds
// row i of CPW columns; step j broadcasts x_j from lane j with v_readlane
// (wave-uniform, so it lands in SGPRs) and every lane subtracts T[i][j] x_j.
// The strict triangle is staged transposed in LDS (zeros elsewhere, so the
// update needs no lane mask).  256 threads = 4 waves; batch in blockIdx.z.
// With `list` (a panel's interchange list; rows minus list_base are relative
// to the top of B) each wave first applies the panel's row interchanges to its
// columns of B: it reads every source row, writes the lower destinations and
// solves on the permuted top rows, so no separate row-swap pass is needed.
template <typename T, int CPW>
__global__ void __launch_bounds__(256) trsm_lanes(int w, int ncols, const T* __restrict__ tri, int64_t ldt, int64_t st,
                                                  T* __restrict__ b, int64_t ldb, int64_t sb,
                                                  const int* __restrict__ list, int64_t slist, int list_base) {
    __shared__ T tr[GRIM_LU_NB][GRIM_LU_NB], ti[GRIM_LU_NB][GRIM_LU_NB];   // [col j][row i]
    tri += 2 * st * blockIdx.z;
    b += 2 * sb * blockIdx.z;
#pragma unroll 4
    for (int jm = 0; jm < GRIM_LU_NB; jm += 4) {   // lane: row i; wave: columns j = wave + 4 m
        const int i = threadIdx.x & 63, j = (threadIdx.x >> 6) + jm;
        T x = 0, y = 0;
        if (i < w && j < w && i > j) {
            const T* p = tri + 2 * ((int64_t)i + (int64_t)j * ldt);
            x = p[0];
            y = p[1];
        }
        tr[j][i] = x;
        ti[j][i] = y;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col0 = (blockIdx.x * 4 + wave) * CPW;
    T xr[CPW], xi[CPW];
    if (list) {
        list += slist * blockIdx.z;
        const int nlow = list[0];
        const int ts = list[1 + lane] - list_base, ld = list[65 + lane] - list_base,
                  ls = list[129 + lane] - list_base;
        T lr[CPW], li[CPW];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            xr[c] = 0;
            xi[c] = 0;
            lr[c] = 0;
            li[c] = 0;
            if (col0 + c < ncols) {
                const T* bc = b + 2 * (int64_t)(col0 + c) * ldb;
                if (lane < w) { xr[c] = bc[2 * (int64_t)ts]; xi[c] = bc[2 * (int64_t)ts + 1]; }
                if (lane < nlow) { lr[c] = bc[2 * (int64_t)ls]; li[c] = bc[2 * (int64_t)ls + 1]; }
            }
        }
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            if (col0 + c < ncols && lane < nlow) {
                T* bc = b + 2 * (int64_t)(col0 + c) * ldb;
                bc[2 * (int64_t)ld] = lr[c];
                bc[2 * (int64_t)ld + 1] = li[c];
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            xr[c] = 0;
            xi[c] = 0;
            if (lane < w && col0 + c < ncols) {
                const T* p = b + 2 * ((int64_t)lane + (int64_t)(col0 + c) * ldb);
                xr[c] = p[0];
                xi[c] = p[1];
            }
        }
    }
    __syncthreads();
    for (int j = 0; j < w - 1; ++j) {
        const T lr = tr[j][lane], li = ti[j][lane];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const T ur = lane_value(xr[c], j), ui = lane_value(xi[c], j);
            xr[c] -= lr * ur - li * ui;
            xi[c] -= lr * ui + li * ur;
        }
    }
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        if (lane < w && col0 + c < ncols) {
            T* p = b + 2 * ((int64_t)lane + (int64_t)(col0 + c) * ldb);
            p[0] = xr[c];
            p[1] = xi[c];
        }
    }
}

// One chunk of a step's coupling rows (CROWS = RB x 64 rows, CPT / RB x 4 columns), spread over
// the workgroup: thread entry k is row lane + 64 (k % RB) of the coupling rows (from row c_lo of
// the block) and column j0 + wave + 4 (k / RB) of the step's sub-block (from column r0), so every
// offset but the lane's is a compile-time constant.  Zero outside nrow x w.
template <int RB> __device__ __forceinline__ int chunk_row(int k) { return (int)(threadIdx.x & 63) + 64 * (k % RB); }
template <int RB> __device__ __forceinline__ int chunk_col(int k) { return (int)(threadIdx.x >> 6) + 4 * (k / RB); }
template <typename T, int CPT, int CROWS>
__device__ __forceinline__ void fetch_coupling(const T* __restrict__ a, int64_t lda, int c_lo, int nrow, int r0,
                                               int j0, int w, T (&pr)[CPT], T (&pi)[CPT]) {
    constexpr int RB = CROWS / 64;
    static_assert(CROWS % 64 == 0 && CPT % RB == 0, "chunk map: whole 64-row groups");
    typedef T T2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int row = chunk_row<RB>(k), col = j0 + chunk_col<RB>(k);
        T2 v = {0, 0};
        if (row < nrow && col < w)
            v = *reinterpret_cast<const T2*>(a + 2 * ((int64_t)(c_lo + row) + (int64_t)(r0 + col) * lda));
        pr[k] = v.x;
        pi[k] = v.y;
    }
}

// Complex updates of the block solve with one fixed rounding sequence (explicit FMAs, no
// compiler contraction), so that every launch shape gives bit-identical results.
template <typename T>
__device__ __forceinline__ void csub_mul(T& xr, T& xi, T ar, T ai, T ur, T ui) {   // x -= a u
#pragma clang fp contract(off)
    xr -= fma(ar, ur, -(ai * ui));
    xi -= fma(ar, ui, ai * ur);
}
template <typename T>
__device__ __forceinline__ void cmul_by(T& xr, T& xi, T ar, T ai) {   // x *= a
#pragma clang fp contract(off)
    const T br = xr, bi = xi;
    xr = fma(br, ar, -(bi * ai));
    xi = fma(br, ai, bi * ar);
}

// X = T^{-1} B in place for one diagonal block of up to GRIM_SOLVE_NB = 256 rows and ncols
// columns, in one launch (the in-block steps of lu_solve).  Each wave owns CPW columns and all
// rows of the block: lane l holds rows l, 64 + l, 128 + l and 192 + l.  Sub-block by sub-block
// (top-down for the unit lower triangle, bottom-up for the upper one) the 64 x 64 diagonal
// triangle is staged in LDS and shared by the four waves (the upper one as the unit triangle
// D^{-1} U, with B scaled by D^{-1} first); each wave solves its columns' chain as trsm_lanes
// does (x_j broadcast from lane j with v_readlane), then subtracts the coupling blocks of the
// sub-blocks still to come times the new rows, broadcast the same way; the coupling blocks
// pass through LDS in CH-column chunks loaded by the whole workgroup (16 columns; 4 for
// complex128 launches with more workgroups than CUs, so that two workgroups fit per CU).
template <typename T, bool LOWER, int CPW, int CH>
__global__ void __launch_bounds__(256, 2) solve_block(int nb, int ncols, const T* __restrict__ a, int64_t lda,
                                                   int64_t sa, T* __restrict__ b, int64_t ldb, int64_t sb) {
    constexpr int NS = GRIM_SOLVE_NB / GRIM_LU_NB;
    constexpr int CROWS = GRIM_SOLVE_NB - GRIM_LU_NB;              // coupling rows of one step, at most
    constexpr int CPT = CROWS * CH / 256;                          // chunk entries per thread
    __shared__ T tr[GRIM_LU_NB][GRIM_LU_NB], ti[GRIM_LU_NB][GRIM_LU_NB];   // [col j][row i]
    __shared__ T dinv_r[GRIM_LU_NB], dinv_i[GRIM_LU_NB];
    __shared__ T cp_r[CH][CROWS], cp_i[CH][CROWS];                 // coupling chunk [col][row]
    a += 2 * sa * blockIdx.z;
    b += 2 * sb * blockIdx.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col0 = (blockIdx.x * 4 + wave) * CPW;
    const int nsub = (nb + 63) / 64;
    T xr[NS][CPW], xi[NS][CPW];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        const int row = 64 * t + lane;
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            xr[t][c] = 0;
            xi[t][c] = 0;
            if (row < nb && col0 + c < ncols) {
                const T* p = b + 2 * ((int64_t)row + (int64_t)(col0 + c) * ldb);
                xr[t][c] = p[0];
                xi[t][c] = p[1];
            }
        }
    }
#pragma unroll
    for (int step = 0; step < NS; ++step) {
        const int s = LOWER ? step : NS - 1 - step;
        if (s >= nsub) continue;   // same for the whole workgroup
        const int r0 = 64 * s;
        const int w = min(64, nb - r0);
        const T* d = a + 2 * ((int64_t)r0 + (int64_t)r0 * lda);
        __syncthreads();   // every wave is done with the previous triangle
        if (!LOWER) {
            if (threadIdx.x < GRIM_LU_NB) {
                const int j = threadIdx.x;
                T ir = 0, ii = 0;
                if (j < w) {
                    const T* p = d + 2 * ((int64_t)j + (int64_t)j * lda);
                    const T dr = p[0], di = p[1];
                    const T den = fma(dr, dr, di * di);
                    ir = dr / den;
                    ii = -di / den;
                }
                dinv_r[j] = ir;
                dinv_i[j] = ii;
            }
            __syncthreads();
        }
#pragma unroll 4
        for (int jm = 0; jm < GRIM_LU_NB; jm += 4) {   // lane: row i; wave: columns j = wave + 4 m
            const int i = lane, j = wave + jm;
            T x = 0, y = 0;
            if (i < w && j < w && (LOWER ? i > j : i < j)) {
                const T* p = d + 2 * ((int64_t)i + (int64_t)j * lda);
                x = p[0];
                y = p[1];
                if (!LOWER) cmul_by(x, y, dinv_r[i], dinv_i[i]);   // row i of D^{-1} U
            }
            tr[j][i] = x;
            ti[j][i] = y;
        }
        __syncthreads();
        // coupling rows of this step: [r0 + 64, nb) below (lower) or [0, r0) above (upper); the
        // first column chunk of them is fetched while the chain runs
        const int c_lo = LOWER ? r0 + GRIM_LU_NB : 0;
        const int nrow = LOWER ? nb - c_lo : r0;
        const int nchunk = nrow > 0 ? (w + CH - 1) / CH : 0;
        T pr[CPT], pi[CPT];
        if (nchunk > 0) fetch_coupling<T, CPT, CROWS>(a, lda, c_lo, nrow, r0, 0, w, pr, pi);
        // chain over the 64 x 64 triangle (entries from row or column w on are zero)
        if (LOWER) {
#pragma unroll 8
            for (int j = 0; j < GRIM_LU_NB - 1; ++j) {
                const T lr = tr[j][lane], li = ti[j][lane];
#pragma unroll
                for (int c = 0; c < CPW; ++c)
                    csub_mul(xr[s][c], xi[s][c], lr, li, lane_value(xr[s][c], j), lane_value(xi[s][c], j));
            }
        } else {
            const T ir = dinv_r[lane], ii = dinv_i[lane];
#pragma unroll
            for (int c = 0; c < CPW; ++c) cmul_by(xr[s][c], xi[s][c], ir, ii);
#pragma unroll 8
            for (int j = GRIM_LU_NB - 1; j > 0; --j) {
                const T lr = tr[j][lane], li = ti[j][lane];
#pragma unroll
                for (int c = 0; c < CPW; ++c)
                    csub_mul(xr[s][c], xi[s][c], lr, li, lane_value(xr[s][c], j), lane_value(xi[s][c], j));
            }
        }
        // coupling: x_t -= A_ts x_s for the sub-blocks t still to be solved, one LDS-staged
        // chunk of columns at a time (the next chunk's loads in flight meanwhile)
#pragma unroll
        for (int q = 0; q < GRIM_LU_NB / CH; ++q) {
            if (q >= nchunk) break;   // same for the whole workgroup
            __syncthreads();          // the previous chunk's readers are done
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                cp_r[chunk_col<CROWS / 64>(k)][chunk_row<CROWS / 64>(k)] = pr[k];
                cp_i[chunk_col<CROWS / 64>(k)][chunk_row<CROWS / 64>(k)] = pi[k];
            }
            __syncthreads();
            if (q + 1 < nchunk)
                fetch_coupling<T, CPT, CROWS>(a, lda, c_lo, nrow, r0, (q + 1) * CH, w, pr, pi);
#pragma unroll 4
            for (int jj = 0; jj < CH; ++jj) {
                const int j = q * CH + jj;
                T ur[CPW], ui[CPW];
#pragma unroll
                for (int c = 0; c < CPW; ++c) {
                    ur[c] = lane_value(xr[s][c], j);
                    ui[c] = lane_value(xi[s][c], j);
                }
#pragma unroll
                for (int t = 0; t < NS; ++t) {
                    if (LOWER ? t <= s : t >= s) continue;
                    const int rl = GRIM_LU_NB * t + lane - c_lo;   // row within the coupling rows
                    const T cr = cp_r[jj][rl], ci = cp_i[jj][rl];
#pragma unroll
                    for (int c = 0; c < CPW; ++c) csub_mul(xr[t][c], xi[t][c], cr, ci, ur[c], ui[c]);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        const int row = 64 * t + lane;
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            if (row < nb && col0 + c < ncols) {
                T* p = b + 2 * ((int64_t)row + (int64_t)(col0 + c) * ldb);
                p[0] = xr[t][c];
                p[1] = xi[t][c];
            }
        }
    }
}

// y[i, col] = b[perm[i], col]
template <typename T>
__global__ void __launch_bounds__(256) gather_rows(int n, const int* __restrict__ perm, int64_t sperm,
                                                   const T* __restrict__ b, int64_t ldb, int64_t sb,
                                                   T* __restrict__ y, int64_t ldy, int64_t sy) {
    const int i = blockIdx.x * 256 + threadIdx.x, col = blockIdx.y;
    if (i >= n) return;
    perm += sperm * blockIdx.z;
    b += 2 * sb * blockIdx.z;
    y += 2 * sy * blockIdx.z;
    const T* s = b + 2 * ((int64_t)perm[i] + (int64_t)col * ldb);
    T* d = y + 2 * ((int64_t)i + (int64_t)col * ldy);
    d[0] = s[0];
    d[1] = s[1];
}

// dst = src (m x n complex of type T), batched; optional conversion to U.
template <typename T, typename U>
__global__ void __launch_bounds__(256) copy_cast(int m, const T* __restrict__ src, int64_t lds, int64_t ss,
                                                 U* __restrict__ dst, int64_t ldd, int64_t sd) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= m) return;
    src += 2 * ss * blockIdx.z;
    dst += 2 * sd * blockIdx.z;
    const T* s = src + 2 * ((int64_t)r + (int64_t)c * lds);
    U* d = dst + 2 * ((int64_t)r + (int64_t)c * ldd);
    d[0] = (U)s[0];
    d[1] = (U)s[1];
}

// x (c128) += d (c64), batched
__global__ void __launch_bounds__(256) add_c64(int m, const float* __restrict__ d, int64_t ldd, int64_t sd,
                                               double* __restrict__ x, int64_t ldx, int64_t sx) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= m) return;
    d += 2 * sd * blockIdx.z;
    x += 2 * sx * blockIdx.z;
    const float* s = d + 2 * ((int64_t)r + (int64_t)c * ldd);
    double* t = x + 2 * ((int64_t)r + (int64_t)c * ldx);
    t[0] += (double)s[0];
    t[1] += (double)s[1];
}

// out[z * n + c] = max_r |v[r, c]| (complex modulus), batched.
__global__ void __launch_bounds__(256) colmax_abs(int m, int ncols, const double* __restrict__ v, int64_t ld,
                                                  int64_t sv, double* __restrict__ out) {
    const int c = blockIdx.x;
    v += 2 * sv * blockIdx.z;
    double best = 0.0;
    for (int r = threadIdx.x; r < m; r += 256) {
        const double* p = v + 2 * ((int64_t)r + (int64_t)c * ld);
        best = fmax(best, sqrt(p[0] * p[0] + p[1] * p[1]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) best = fmax(best, __shfl_xor(best, off, 64));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) out[(int64_t)blockIdx.z * ncols + c] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Per column 2-norm (for the relative residual ||r|| / ||b||).
__global__ void __launch_bounds__(256) colnorm2(int m, const double* __restrict__ v, int64_t ld,
                                                double* __restrict__ out) {
    const int c = blockIdx.x;
    double s = 0.0;
    for (int r = threadIdx.x; r < m; r += 256) {
        const double* p = v + 2 * ((int64_t)r + (int64_t)c * ld);
        s += p[0] * p[0] + p[1] * p[1];
    }
    s = wave_sum(s);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[c] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// Host-side drivers
// ---------------------------------------------------------------------------
struct Batch {
    int count;
    hipStream_t st;
};

// 32-row output tiles for short updates (k <= 256: the LU's rank-64 steps and the solve's
// between-block products), twice the workgroups of 64-row tiles.  Measured at N = 2000, two
// systems, 720 columns: the seven between-block products 338 -> 311 us (complex64) and
// 529 -> 457 us (complex128), the complex128 factor 9.13 -> 8.70 ms.  Long products
// (k = 2000 residual: 444 -> 496 us complex64) keep 64-row tiles and their operand reuse.
static bool gemm_narrow_rows(int m, int k) { return k <= 256 && m <= 65535 * 32; }

template <typename T, bool SET = false>
static hipError_t gemm(const Batch& bt, int m, int n, int k, const T* A, int64_t lda, int64_t sa, const T* B,
                       int64_t ldb, int64_t sb, T* C, int64_t ldc, int64_t sc) {
    if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
    const bool narrow = gemm_narrow_rows(m, k);
    const int bm = narrow ? 32 : 64;
    dim3 grid((m + bm - 1) / bm, (n + GB_N - 1) / GB_N, bt.count);
    // complex64 updates with k <= 64 read C ahead (measured: 1000^2 x 64 19 -> 16 us); complex128
    // keeps the epilogue read (the prefetch registers cost it occupancy)
    const bool pre = sizeof(T) == 4 && !SET && k <= 64;
    if (pre && narrow)
        cgemm<T, SET, true, 1><<<grid, 256, 0, bt.st>>>(m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc);
    else if (pre)
        cgemm<T, SET, true><<<grid, 256, 0, bt.st>>>(m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc);
    else if (narrow)
        cgemm<T, SET, false, 1><<<grid, 256, 0, bt.st>>>(m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc);
    else
        cgemm<T, SET><<<grid, 256, 0, bt.st>>>(m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc);
    return hipGetLastError();
}

// C -= A^H B with A stored k x m (the off-diagonal updates of the conjugate-transposed solve and
// the residual b - Z^H x).
template <typename T>
static hipError_t gemm_ta(const Batch& bt, int m, int n, int k, const T* A, int64_t lda, int64_t sa, const T* B,
                          int64_t ldb, int64_t sb, T* C, int64_t ldc, int64_t sc) {
    if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
    const bool narrow = gemm_narrow_rows(m, k);
    const int bm = narrow ? 32 : 64;
    dim3 grid((m + bm - 1) / bm, (n + GB_N - 1) / GB_N, bt.count);
    if (narrow)
        cgemm<T, false, false, 1, true><<<grid, 256, 0, bt.st>>>(m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc);
    else
        cgemm<T, false, false, 2, true><<<grid, 256, 0, bt.st>>>(m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc);
    return hipGetLastError();
}

#define GRIM_LEAF_MAX_M 16384

template <typename T>
static hipError_t leaf(const Batch& bt, int m, int w, T* pan, int64_t lda, int64_t sa, int pw, int lc0, int row0,
                       int* piv, int64_t spiv, unsigned* status, char* tall, int64_t stall, unsigned& arrivals,
                       int* perm, int* list, bool& listed) {
    hipStream_t st = bt.st;
    const int b = bt.count;
    if (m > GRIM_LEAF_MAX_M) {
        // rows over G <= GRIM_TALL_G workgroups of 1024 * R rows each
        const int R = m <= 2 * GRIM_LEAF_MAX_M ? 4 : 8;
        const int G = (m + 1024 * R - 1) / (1024 * R);
        if (G > GRIM_TALL_G) return hipErrorInvalidValue;
        if (R == 4)
            panel_leaf_tall<T, 4><<<dim3(G, b), 1024, 0, st>>>(m, w, pan, lda, sa, pw, lc0, row0, piv, spiv, status,
                                                                tall, stall, arrivals);
        else
            panel_leaf_tall<T, 8><<<dim3(G, b), 1024, 0, st>>>(m, w, pan, lda, sa, pw, lc0, row0, piv, spiv, status,
                                                                tall, stall, arrivals);
        arrivals += (unsigned)(w * G);
        return hipGetLastError();
    }
    // the panel's last leaf also builds the panel's interchange list
    int* lperm = lc0 + w == pw ? perm : nullptr;
    if (lperm != nullptr) listed = true;
#define GRIM_LEAF(R, W) \
    panel_leaf<T, R, W><<<b, 1024, 0, st>>>(m, w, pan, lda, sa, pw, lc0, row0, piv, spiv, status, lperm, stall / 4, \
                                            list)
    if (m <= 1024) GRIM_LEAF(1, 16);
    else if (m <= 2048) {
        // complex64 holds twice the columns in the same registers (here and below): half the
        // leaves and in-panel updates (complex64 factor, measured: N = 2000 7.28 -> 6.88 ms,
        // N = 8000 87.4 -> 75.7 ms, N = 12 000 248 -> 203 ms)
        if constexpr (sizeof(T) == 4) GRIM_LEAF(2, 16);
        else GRIM_LEAF(2, 8);
    }
    else if (m <= 4096) {
        if constexpr (sizeof(T) == 4) GRIM_LEAF(4, 8);
        else GRIM_LEAF(4, 4);
    } else if (m <= 8192) {
        if constexpr (sizeof(T) == 4) GRIM_LEAF(8, 4);
        else GRIM_LEAF(8, 2);
    } else {
        if constexpr (sizeof(T) == 4) GRIM_LEAF(16, 2);
        else GRIM_LEAF(16, 1);
    }
#undef GRIM_LEAF
    return hipGetLastError();
}

template <typename T>
static int leaf_width(int m) {
    if (m <= 1024) return 16;
    if (m <= 2048) return sizeof(T) == 4 ? 16 : 8;
    if (m <= 4096) return sizeof(T) == 4 ? 8 : 4;
    if (m <= 8192) return sizeof(T) == 4 ? 4 : 2;
    if (m <= GRIM_LEAF_MAX_M) return sizeof(T) == 4 ? 2 : 1;
    return GRIM_TALL_TW;
}

template <typename T>
static hipError_t trsm(const Batch& bt, int w, int ncols, const T* tri, int64_t ldt, int64_t st, T* b, int64_t ldb,
                       int64_t sb, const int* list = nullptr, int64_t slist = 0, int list_base = 0) {
    if (w <= 0 || ncols <= 0) return hipSuccess;
    // columns per wave: enough waves to cover the chip first, then ILP
    const int64_t cols = (int64_t)ncols * bt.count;
#define GRIM_TRSM(CPW)                                                                                            \
    trsm_lanes<T, CPW><<<dim3((ncols + 4 * CPW - 1) / (4 * CPW), 1, bt.count), 256, 0, bt.st>>>(        \
        w, ncols, tri, ldt, st, b, ldb, sb, list, slist, list_base)
    int cpw = grim_shape_value(GRIM_SHAPE_TRSM_COLUMNS, cols);
    if (cpw == 0) cpw = cols >= 8192 ? 4 : cols >= 2048 ? 2 : 1;
    if (cpw == 4) GRIM_TRSM(4);
    else if (cpw == 2) GRIM_TRSM(2);
    else GRIM_TRSM(1);
#undef GRIM_TRSM
    return hipGetLastError();
}

// Recursive factorization of columns [j0, j0 + w) of the panel starting at
// column pj (width pw), rows j0 .. n-1.  Each leaf applies its interchanges to
// the whole panel, so the right-hand half sees permuted rows without a swap
// pass.
template <typename T>
static hipError_t panel_rec(const Batch& bt, int n, T* a, int64_t lda, int64_t sa, int pj, int pw, int j0, int w,
                            int* piv, int64_t spiv, unsigned* status, char* tall, int64_t stall,
                            unsigned& arrivals, int* perm, int* list, bool& listed) {
    const int m = n - j0;
    const int lw = leaf_width<T>(m);
    T* top = a + 2 * ((int64_t)j0 + (int64_t)j0 * lda);
    if (w <= lw)
        return leaf<T>(bt, m, w, a + 2 * ((int64_t)j0 + (int64_t)pj * lda), lda, sa, pw, j0 - pj, j0, piv + j0,
                       spiv, status, tall, stall, arrivals, perm, list, listed);
    int w1 = w / 2;
    if (w1 > lw) w1 = (w1 / lw) * lw;
    hipError_t err =
        panel_rec<T>(bt, n, a, lda, sa, pj, pw, j0, w1, piv, spiv, status, tall, stall, arrivals, perm, list, listed);
    if (err != hipSuccess) return err;
    const int w2 = w - w1;
    T* right = a + 2 * ((int64_t)(j0 + w1) * lda);
    if ((err = trsm<T>(bt, w1, w2, top, lda, sa, right + 2 * (int64_t)j0, lda, sa)) != hipSuccess) return err;
    if ((err = gemm<T>(bt, m - w1, w2, w1, top + 2 * (int64_t)w1, lda, sa, right + 2 * (int64_t)j0, lda, sa,
                       right + 2 * (int64_t)(j0 + w1), lda, sa)) != hipSuccess)
        return err;
    return panel_rec<T>(bt, n, a, lda, sa, pj, pw, j0 + w1, w2, piv, spiv, status, tall, stall, arrivals, perm, list,
                        listed);
}

// Factor-side auxiliary storage per matrix: the row permutation the pivots
// compose to (n int32, applied to right-hand sides in one gather) and the
// interchange list of the panel being factored.
static int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }
static int64_t aux_bytes(int n) {
    return align256(4 * (int64_t)n) + align256(4 * GRIM_LIST_INTS) + align256(sizeof(TallScratch<double>));
}

template <typename T>
static hipError_t lu_factor(const Batch& bt, int n, T* a, int64_t lda, int64_t sa, int* piv, int64_t spiv,
                            void* auxp, int64_t saux, unsigned* status) {
    hipError_t err;
    int* perm = (int*)auxp;
    int* list = (int*)((char*)auxp + align256(4 * (int64_t)n));
    char* tall = (char*)auxp + align256(4 * (int64_t)n) + align256(4 * GRIM_LIST_INTS);
    const int64_t sperm = saux / 4;
    perm_init<<<dim3((n + 255) / 256, 1, bt.count), 256, 0, bt.st>>>(n, perm, sperm, (unsigned*)tall);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    unsigned arrivals = 0;
    for (int j = 0; j < n; j += GRIM_LU_NB) {
        const int jb = min(GRIM_LU_NB, n - j);
        bool listed = false;
        if ((err = panel_rec<T>(bt, n, a, lda, sa, j, jb, j, jb, piv, spiv, status, tall, saux, arrivals, perm, list,
                                listed)) != hipSuccess)
            return err;
        if (!listed) {   // the last leaf was a tall one
            panel_perm<<<bt.count, 64, 0, bt.st>>>(j, jb, piv, spiv, perm, sperm, list, sperm);
            if ((err = hipGetLastError()) != hipSuccess) return err;
        }
        // the panel's interchanges on the columns left of it (LAPACK keeps L permuted)
        if (j > 0) {
            apply_list<T><<<dim3((j + 3) / 4, 1, bt.count), 256, 0, bt.st>>>(j, j, jb, a, lda, sa, list, sperm);
            if ((err = hipGetLastError()) != hipSuccess) return err;
        }
        const int rest = n - j - jb;
        if (rest > 0) {
            T* diag = a + 2 * ((int64_t)j + (int64_t)j * lda);
            T* a12 = a + 2 * ((int64_t)j + (int64_t)(j + jb) * lda);
            // interchanges on the right-hand columns fused with U12 = L11^{-1} A12, then A22 -= L21 U12
            if ((err = trsm<T>(bt, jb, rest, diag, lda, sa, a12, lda, sa, list, sperm, j)) != hipSuccess)
                return err;
            if ((err = gemm<T>(bt, rest, rest, jb, diag + 2 * (int64_t)jb, lda, sa, a12, lda, sa,
                               a12 + 2 * (int64_t)jb, lda, sa)) != hipSuccess)
                return err;
        }
    }
    return hipSuccess;
}

// Compute units of an MI355X (gfx950): the block-solve rule counts resident workgroups with it.
#define GRIM_CUS 256
template <typename T, bool LOWER, int CPW>
static void launch_solve_block(bool narrow, dim3 grid, hipStream_t st, int nb, int ncols, const T* a, int64_t lda,
                               int64_t sa, T* b, int64_t ldb, int64_t sb) {
    if constexpr (sizeof(T) == 8) {
        if (narrow) {
            solve_block<T, LOWER, CPW, 4><<<grid, 256, 0, st>>>(nb, ncols, a, lda, sa, b, ldb, sb);
            return;
        }
    }
    solve_block<T, LOWER, CPW, 16><<<grid, 256, 0, st>>>(nb, ncols, a, lda, sa, b, ldb, sb);
}
template <typename T, bool LOWER>
static hipError_t solve_diag_block(const Batch& bt, int nb, int ncols, const T* a, int64_t lda, int64_t sa, T* b,
                                   int64_t ldb, int64_t sb) {
    if (nb <= 0 || ncols <= 0) return hipSuccess;
    const int64_t cols = (int64_t)ncols * bt.count;
    int cpw = grim_shape_value(GRIM_SHAPE_SOLVE_COLUMNS, cols);
    // one column per wavefront while every workgroup fits at once (two per CU), else two
    if (cpw == 0) cpw = (cols + 3) / 4 <= 2 * GRIM_CUS ? 1 : 2;
    const dim3 grid((ncols + 4 * cpw - 1) / (4 * cpw), 1, bt.count);
    // complex128 with 16-column coupling chunks takes 114 KB of LDS, one workgroup per CU; past
    // one workgroup per CU, 4-column chunks (77 KB) let two share a CU
    const bool narrow = sizeof(T) == 8 && (int64_t)grid.x * grid.z > GRIM_CUS;
    if (cpw == 4) launch_solve_block<T, LOWER, 4>(narrow, grid, bt.st, nb, ncols, a, lda, sa, b, ldb, sb);
    else if (cpw == 2) launch_solve_block<T, LOWER, 2>(narrow, grid, bt.st, nb, ncols, a, lda, sa, b, ldb, sb);
    else launch_solve_block<T, LOWER, 1>(narrow, grid, bt.st, nb, ncols, a, lda, sa, b, ldb, sb);
    return hipGetLastError();
}

// Solve in place: b <- A^{-1} b with work (n x nrhs per matrix, ld n).
template <typename T>
static hipError_t lu_solve(const Batch& bt, int n, int nrhs, const T* a, int64_t lda, int64_t sa, const void* auxp,
                           int64_t saux, T* b, int64_t ldb, int64_t sb, T* w, int64_t sw) {
    const int64_t ldw = n;
    hipError_t err;
    gather_rows<T><<<dim3((n + 255) / 256, nrhs, bt.count), 256, 0, bt.st>>>(n, (const int*)auxp, saux / 4, b, ldb,
                                                                              sb, w, ldw, sw);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    // Blocks of GRIM_SOLVE_NB rows: inside a block, one solve_block launch (its 64-row
    // triangles and their coupling in LDS and lanes); between blocks, one
    // rank-GRIM_SOLVE_NB update, so most of the solve's flops run as deep-k GEMMs.
    // forward: w_J <- L_JJ^{-1} w_J (one solve_block launch) ; w_{J+} -= L_{J+, J} w_J
    for (int J = 0; J < n; J += GRIM_SOLVE_NB) {
        const int jend = min(n, J + GRIM_SOLVE_NB);
        if ((err = solve_diag_block<T, true>(bt, jend - J, nrhs, a + 2 * ((int64_t)J + (int64_t)J * lda), lda, sa,
                                             w + 2 * (int64_t)J, ldw, sw)) != hipSuccess)
            return err;
        if (jend < n)
            if ((err = gemm<T>(bt, n - jend, nrhs, jend - J, a + 2 * ((int64_t)jend + (int64_t)J * lda), lda, sa,
                               w + 2 * (int64_t)J, ldw, sw, w + 2 * (int64_t)jend, ldw, sw)) != hipSuccess)
                return err;
    }
    // backward: w_J <- U_JJ^{-1} w_J ; w_{<J} -= U_{<J, J} w_J
    const int lastJ = ((n - 1) / GRIM_SOLVE_NB) * GRIM_SOLVE_NB;
    for (int J = lastJ; J >= 0; J -= GRIM_SOLVE_NB) {
        const int jend = min(n, J + GRIM_SOLVE_NB);
        if ((err = solve_diag_block<T, false>(bt, jend - J, nrhs, a + 2 * ((int64_t)J + (int64_t)J * lda), lda, sa,
                                              w + 2 * (int64_t)J, ldw, sw)) != hipSuccess)
            return err;
        if (J > 0)
            if ((err = gemm<T>(bt, J, nrhs, jend - J, a + 2 * (int64_t)J * lda, lda, sa, w + 2 * (int64_t)J, ldw, sw,
                               w, ldw, sw)) != hipSuccess)
                return err;
    }
    copy_cast<T, T><<<dim3((n + 255) / 256, nrhs, bt.count), 256, 0, bt.st>>>(n, w, ldw, sw, b, ldb, sb);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Conjugate-transposed solve A^H x = b with the factors of P A = L U (scipy lu_solve with
// trans=2, LAPACK getrs 'C', as the reference's condition estimate and RefinedLU call it):
// A^H = U^H L^H P, so U^H z = b (lower, forward), L^H w = z (unit upper, backward), x = P^T w.
// Each 256-row diagonal block is staged conjugate-transposed into a 256 x 256 workspace and
// solved by solve_block: U_JJ^H as the unit lower triangle U_JJ^H D^{-1} (D = diag(conj U_jj),
// the rows then scaled by D^{-1}), L_JJ^H as a unit upper triangle (diagonal staged as 1).  The
// off-diagonal updates are conjugate-transposed GEMMs on the factors in place.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cdiv(T xr, T xi, T dr, T di, T& qr, T& qi) {   // q = x / d
    const T den = fma(dr, dr, di * di);
    qr = fma(xr, dr, xi * di) / den;
    qi = fma(xi, dr, -(xr * di)) / den;
}

// W[i, j] (i, j < jw, ld GRIM_SOLVE_NB) from the diagonal block D = a[j0.., j0..]:
// FROM_U: i > j: conj(D[j, i]) / conj(D[j, j]) (strict lower of U^H D^{-1}), else 0;
// else:   i < j: conj(D[j, i]) (strict upper of L^H), i == j: 1, else 0.
// 64 x 64 tiles through LDS: reads down columns of D, writes down columns of W.
template <typename T, bool FROM_U>
__global__ void __launch_bounds__(256) stage_trans_block(int jw, const T* __restrict__ d, int64_t lda, int64_t sa,
                                                         T* __restrict__ w, int64_t sw) {
    __shared__ T tr[64][65], ti[64][65];   // [column i of the tile of D][row j]
    __shared__ T gr[64], gi[64];            // diagonal of D for the tile's rows j
    d += 2 * sa * blockIdx.z;
    w += 2 * sw * blockIdx.z;
    const int i0 = 64 * blockIdx.x, j0 = 64 * blockIdx.y;   // W tile rows i0.., columns j0..
    const int t = threadIdx.x, l = t & 63, q = t >> 6;
    for (int c = q; c < 64; c += 4) {            // D rows j0 + l, columns i0 + c
        const int j = j0 + l, i = i0 + c;
        T x = 0, y = 0;
        if (j < jw && i < jw) {
            const T* p = d + 2 * ((int64_t)j + (int64_t)i * lda);
            x = p[0];
            y = p[1];
        }
        tr[c][l] = x;
        ti[c][l] = y;
    }
    if (FROM_U && t < 64) {
        const int j = j0 + t;
        T x = 1, y = 0;
        if (j < jw) {
            const T* p = d + 2 * ((int64_t)j + (int64_t)j * lda);
            x = p[0];
            y = p[1];
        }
        gr[t] = x;
        gi[t] = y;
    }
    __syncthreads();
    for (int c = q; c < 64; c += 4) {            // W rows i0 + l, columns j0 + c
        const int i = i0 + l, j = j0 + c;
        if (i >= jw || j >= jw) continue;
        const T x = tr[l][c], y = -ti[l][c];     // conj(D[j, i])
        T vr = 0, vi = 0;
        if (FROM_U) {
            if (i > j) cdiv(x, y, gr[c], -gi[c], vr, vi);
        } else {
            if (i < j) {
                vr = x;
                vi = y;
            } else if (i == j) {
                vr = 1;
            }
        }
        T* p = w + 2 * ((int64_t)i + (int64_t)j * GRIM_SOLVE_NB);
        p[0] = vr;
        p[1] = vi;
    }
}

// x[r, c] /= conj(D[r, r]) for the jw rows of a block (D: diagonal block of U).
template <typename T>
__global__ void __launch_bounds__(256) scale_inv_conj_diag(int jw, int ncols, const T* __restrict__ d, int64_t lda,
                                                           int64_t sa, T* __restrict__ x, int64_t ldx, int64_t sx) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= jw || c >= ncols) return;
    d += 2 * sa * blockIdx.z;
    x += 2 * sx * blockIdx.z;
    const T* g = d + 2 * ((int64_t)r + (int64_t)r * lda);
    T* p = x + 2 * ((int64_t)r + (int64_t)c * ldx);
    T qr, qi;
    cdiv(p[0], p[1], g[0], -g[1], qr, qi);
    p[0] = qr;
    p[1] = qi;
}

// b[perm[i], col] = y[i, col]
template <typename T>
__global__ void __launch_bounds__(256) scatter_rows(int n, const int* __restrict__ perm, int64_t sperm,
                                                    const T* __restrict__ y, int64_t ldy, int64_t sy,
                                                    T* __restrict__ b, int64_t ldb, int64_t sb) {
    const int i = blockIdx.x * 256 + threadIdx.x, col = blockIdx.y;
    if (i >= n) return;
    perm += sperm * blockIdx.z;
    b += 2 * sb * blockIdx.z;
    y += 2 * sy * blockIdx.z;
    const T* s = y + 2 * ((int64_t)i + (int64_t)col * ldy);
    T* d = b + 2 * ((int64_t)perm[i] + (int64_t)col * ldb);
    d[0] = s[0];
    d[1] = s[1];
}

// Solve A^H x = b in place.  w: n x nrhs per matrix (stride sw); blk: GRIM_SOLVE_NB^2 per matrix.
template <typename T>
static hipError_t lu_solve_trans(const Batch& bt, int n, int nrhs, const T* a, int64_t lda, int64_t sa,
                                 const void* auxp, int64_t saux, T* b, int64_t ldb, int64_t sb, T* w, int64_t sw,
                                 T* blk) {
    const int64_t ldw = n, sblk = (int64_t)GRIM_SOLVE_NB * GRIM_SOLVE_NB;
    const int tiles = GRIM_SOLVE_NB / 64;
    hipError_t err;
    copy_cast<T, T><<<dim3((n + 255) / 256, nrhs, bt.count), 256, 0, bt.st>>>(n, b, ldb, sb, w, ldw, sw);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    // forward: U^H z = b
    for (int J = 0; J < n; J += GRIM_SOLVE_NB) {
        const int jw = min(GRIM_SOLVE_NB, n - J);
        const T* djj = a + 2 * ((int64_t)J + (int64_t)J * lda);
        stage_trans_block<T, true><<<dim3(tiles, tiles, bt.count), 256, 0, bt.st>>>(jw, djj, lda, sa, blk, sblk);
        if ((err = hipGetLastError()) != hipSuccess) return err;
        if ((err = solve_diag_block<T, true>(bt, jw, nrhs, blk, GRIM_SOLVE_NB, sblk, w + 2 * (int64_t)J, ldw, sw)) !=
            hipSuccess)
            return err;
        scale_inv_conj_diag<T><<<dim3((jw + 255) / 256, nrhs, bt.count), 256, 0, bt.st>>>(
            jw, nrhs, djj, lda, sa, w + 2 * (int64_t)J, ldw, sw);
        if ((err = hipGetLastError()) != hipSuccess) return err;
        const int jend = J + jw;
        if (jend < n)   // z_{J+} -= (U_{J, J+})^H z_J
            if ((err = gemm_ta<T>(bt, n - jend, nrhs, jw, a + 2 * ((int64_t)J + (int64_t)jend * lda), lda, sa,
                                  w + 2 * (int64_t)J, ldw, sw, w + 2 * (int64_t)jend, ldw, sw)) != hipSuccess)
                return err;
    }
    // backward: L^H w = z
    const int lastJ = ((n - 1) / GRIM_SOLVE_NB) * GRIM_SOLVE_NB;
    for (int J = lastJ; J >= 0; J -= GRIM_SOLVE_NB) {
        const int jw = min(GRIM_SOLVE_NB, n - J);
        const T* djj = a + 2 * ((int64_t)J + (int64_t)J * lda);
        stage_trans_block<T, false><<<dim3(tiles, tiles, bt.count), 256, 0, bt.st>>>(jw, djj, lda, sa, blk, sblk);
        if ((err = hipGetLastError()) != hipSuccess) return err;
        if ((err = solve_diag_block<T, false>(bt, jw, nrhs, blk, GRIM_SOLVE_NB, sblk, w + 2 * (int64_t)J, ldw,
                                              sw)) != hipSuccess)
            return err;
        if (J > 0)      // w_{<J} -= (L_{J, <J})^H w_J
            if ((err = gemm_ta<T>(bt, J, nrhs, jw, a + 2 * (int64_t)J, lda, sa, w + 2 * (int64_t)J, ldw, sw, w, ldw,
                                  sw)) != hipSuccess)
                return err;
    }
    scatter_rows<T><<<dim3((n + 255) / 256, nrhs, bt.count), 256, 0, bt.st>>>(n, (const int*)auxp, saux / 4, w, ldw,
                                                                               sw, b, ldb, sb);
    return hipGetLastError();
}

extern "C" int64_t grim_dense_trans_block_elems(void) { return (int64_t)GRIM_SOLVE_NB * GRIM_SOLVE_NB; }

extern "C" hipError_t grim_dense_lu_solve_trans(int batch, int n, int nrhs, const void* a, int64_t lda, int64_t sa,
                                                int single, const void* aux, int64_t saux, void* b, int64_t ldb,
                                                int64_t sb, void* w, int64_t sw, void* blk, hipStream_t st) {
    Batch bt{batch, st};
    return single ? lu_solve_trans<float>(bt, n, nrhs, (const float*)a, lda, sa, aux, saux, (float*)b, ldb, sb,
                                          (float*)w, sw, (float*)blk)
                  : lu_solve_trans<double>(bt, n, nrhs, (const double*)a, lda, sa, aux, saux, (double*)b, ldb, sb,
                                           (double*)w, sw, (double*)blk);
}

extern "C" hipError_t grim_dense_gemm_minus_ta(int batch, int m, int n, int k, const void* a, int64_t lda,
                                               int64_t sa, const void* b, int64_t ldb, int64_t sb, void* c,
                                               int64_t ldc, int64_t sc, int single, hipStream_t st) {
    Batch bt{batch, st};
    return single ? gemm_ta<float>(bt, m, n, k, (const float*)a, lda, sa, (const float*)b, ldb, sb, (float*)c, ldc,
                                   sc)
                  : gemm_ta<double>(bt, m, n, k, (const double*)a, lda, sa, (const double*)b, ldb, sb, (double*)c,
                                    ldc, sc);
}

// ---------------------------------------------------------------------------
// Row/column equilibration of a complex128 matrix and the equilibrated 1-norm
// (twod/solver.py:472-525): r_i = max_j |a_ij| (1 if 0), c_j = max_i |a_ij| / r_i (1 if 0),
// norm = max_j sum_i |a_ij| / r_i / c_j.  The maxima are taken with atomicMax on the bit
// patterns of non-negative doubles (which order as the values), so they are exact and do not
// depend on the order of arrival; each column sum is a fixed tree over one workgroup.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fill_u64(int64_t count, unsigned long long* p) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0ull;
}

// rows i of one thread, 64 columns per workgroup column-chunk (grid.y), batch grid.z
__global__ void __launch_bounds__(256) row_absmax(int n, const double* __restrict__ a, int64_t lda, int64_t sa,
                                                  unsigned long long* __restrict__ rmax) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    a += 2 * sa * blockIdx.z;
    rmax += (int64_t)n * blockIdx.z;
    const int c0 = 64 * blockIdx.y, c1 = min(n, c0 + 64);
    double m = 0.0;
    for (int c = c0; c < c1; ++c) {
        const double* p = a + 2 * ((int64_t)i + (int64_t)c * lda);
        m = fmax(m, hypot(p[0], p[1]));
    }
    atomicMax(rmax + i, (unsigned long long)__double_as_longlong(m));
}

__global__ void __launch_bounds__(256) finish_row_scale(int n, int batch, double* __restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)n * batch) return;
    const double v = r[i];   // the bit pattern written by row_absmax
    r[i] = v > 0.0 ? v : 1.0;
}

__device__ __forceinline__ double block_reduce(double v, bool is_max, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] = is_max ? fmax(sh[t], sh[t + s]) : sh[t] + sh[t + s];
        __syncthreads();
    }
    const double out = sh[0];
    __syncthreads();
    return out;
}

// one workgroup per column j (grid.x), batch grid.z
__global__ void __launch_bounds__(256) col_equilibrate(int n, const double* __restrict__ a, int64_t lda, int64_t sa,
                                                       const double* __restrict__ rs, double* __restrict__ cs,
                                                       unsigned long long* __restrict__ norm) {
    __shared__ double sh[256];
    const int j = blockIdx.x;
    a += 2 * sa * blockIdx.z + 2 * (int64_t)j * lda;
    rs += (int64_t)n * blockIdx.z;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, hypot(a[2 * i], a[2 * i + 1]) / rs[i]);
    double c = block_reduce(m, true, sh);
    c = c > 0.0 ? c : 1.0;
    double sum = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) sum += hypot(a[2 * i], a[2 * i + 1]) / rs[i] / c;
    sum = block_reduce(sum, false, sh);
    if (threadIdx.x == 0) {
        cs[(int64_t)n * blockIdx.z + j] = c;
        atomicMax(norm + blockIdx.z, (unsigned long long)__double_as_longlong(sum));
    }
}

extern "C" hipError_t grim_dense_equilibration(int batch, int n, const double* a, int64_t lda, int64_t sa,
                                               double* row_scale, double* col_scale, double* norm, hipStream_t st) {
    unsigned long long* rbits = (unsigned long long*)row_scale;
    unsigned long long* nbits = (unsigned long long*)norm;
    const int64_t rcount = (int64_t)n * batch;
    fill_u64<<<(unsigned)((rcount + 255) / 256), 256, 0, st>>>(rcount, rbits);
    fill_u64<<<1, 256, 0, st>>>(batch, nbits);
    row_absmax<<<dim3((n + 255) / 256, (n + 63) / 64, batch), 256, 0, st>>>(n, a, lda, sa, rbits);
    finish_row_scale<<<(unsigned)((rcount + 255) / 256), 256, 0, st>>>(n, batch, row_scale);
    col_equilibrate<<<dim3(n, 1, batch), 256, 0, st>>>(n, a, lda, sa, row_scale, col_scale, nbits);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ||A||_inf = max_i sum_j |a_ij| of complex128 matrices (linalg/dense.py, the backward-error
// gate's matrix_inf_norm): partial sums of 64 columns per thread and row (coalesced down the
// columns), then each row adds its partials in chunk order, so the result is deterministic.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) row_abs_partial(int n, const double* __restrict__ a, int64_t lda, int64_t sa,
                                                       double* __restrict__ part) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    a += 2 * sa * blockIdx.z;
    const int chunks = (n + 63) / 64;
    part += (int64_t)n * chunks * blockIdx.z;
    const int c0 = 64 * blockIdx.y, c1 = min(n, c0 + 64);
    double s = 0.0;
    for (int c = c0; c < c1; ++c) {
        const double* p = a + 2 * ((int64_t)i + (int64_t)c * lda);
        s += hypot(p[0], p[1]);
    }
    part[(int64_t)blockIdx.y * n + i] = s;
}

__global__ void __launch_bounds__(256) row_abs_finish(int n, const double* __restrict__ part,
                                                      unsigned long long* __restrict__ out) {
    __shared__ double sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int chunks = (n + 63) / 64;
    part += (int64_t)n * chunks * blockIdx.z;
    double s = 0.0;
    if (i < n)
        for (int c = 0; c < chunks; ++c) s += part[(int64_t)c * n + i];
    const double m = block_reduce(s, true, sh);
    if (threadIdx.x == 0) atomicMax(out + blockIdx.z, (unsigned long long)__double_as_longlong(m));
}

extern "C" int64_t grim_dense_norm_inf_work(int n) { return (int64_t)n * ((n + 63) / 64) * 8; }

extern "C" hipError_t grim_dense_norm_inf(int batch, int n, const double* a, int64_t lda, int64_t sa, double* work,
                                          double* out, hipStream_t st) {
    fill_u64<<<1, 256, 0, st>>>(batch, (unsigned long long*)out);
    row_abs_partial<<<dim3((n + 255) / 256, (n + 63) / 64, batch), 256, 0, st>>>(n, a, lda, sa, work);
    row_abs_finish<<<dim3((n + 255) / 256, 1, batch), 256, 0, st>>>(n, work, (unsigned long long*)out);
    return hipGetLastError();
}

extern "C" int64_t grim_dense_aux_bytes(int n, int single) {
    (void)single;
    return aux_bytes(n);
}

extern "C" hipError_t grim_dense_lu_factor(int batch, int n, void* a, int64_t lda, int64_t sa, int single, int* piv,
                                           int64_t spiv, void* aux, int64_t saux, unsigned* status,
                                           hipStream_t st) {
    Batch bt{batch, st};
    return single ? lu_factor<float>(bt, n, (float*)a, lda, sa, piv, spiv, aux, saux, status)
                  : lu_factor<double>(bt, n, (double*)a, lda, sa, piv, spiv, aux, saux, status);
}

extern "C" hipError_t grim_dense_lu_solve(int batch, int n, int nrhs, const void* a, int64_t lda, int64_t sa,
                                          int single, const void* aux, int64_t saux, void* b, int64_t ldb,
                                          int64_t sb, void* w, int64_t sw, hipStream_t st) {
    Batch bt{batch, st};
    return single ? lu_solve<float>(bt, n, nrhs, (const float*)a, lda, sa, aux, saux, (float*)b, ldb, sb,
                                    (float*)w, sw)
                  : lu_solve<double>(bt, n, nrhs, (const double*)a, lda, sa, aux, saux, (double*)b, ldb, sb,
                                     (double*)w, sw);
}

extern "C" hipError_t grim_dense_gemm_minus(int batch, int m, int n, int k, const void* a, int64_t lda, int64_t sa,
                                            const void* b, int64_t ldb, int64_t sb, void* c, int64_t ldc,
                                            int64_t sc, int single, hipStream_t st) {
    Batch bt{batch, st};
    return single ? gemm<float>(bt, m, n, k, (const float*)a, lda, sa, (const float*)b, ldb, sb, (float*)c, ldc, sc)
                  : gemm<double>(bt, m, n, k, (const double*)a, lda, sa, (const double*)b, ldb, sb, (double*)c, ldc,
                                 sc);
}

extern "C" hipError_t grim_dense_to_c64(int batch, int m, int n, const double* src, int64_t lds, int64_t ss,
                                        float* dst, int64_t ldd, int64_t sd, hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    copy_cast<double, float><<<dim3((m + 255) / 256, n, batch), 256, 0, st>>>(m, src, lds, ss, dst, ldd, sd);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_copy(int batch, int m, int n, const double* src, int64_t lds, int64_t ss,
                                      double* dst, int64_t ldd, int64_t sd, hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    copy_cast<double, double><<<dim3((m + 255) / 256, n, batch), 256, 0, st>>>(m, src, lds, ss, dst, ldd, sd);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_add_c64(int batch, int m, int n, const float* d, int64_t ldd, int64_t sd, double* x,
                                        int64_t ldx, int64_t sx, hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    add_c64<<<dim3((m + 255) / 256, n, batch), 256, 0, st>>>(m, d, ldd, sd, x, ldx, sx);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_colmax(int batch, int m, int n, const double* v, int64_t ld, int64_t sv,
                                        double* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    colmax_abs<<<dim3(n, 1, batch), 256, 0, st>>>(m, n, v, ld, sv, out);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_colnorm(int m, int n, const double* v, int64_t ld, double* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    colnorm2<<<n, 256, 0, st>>>(m, v, ld, out);
    return hipGetLastError();
}
