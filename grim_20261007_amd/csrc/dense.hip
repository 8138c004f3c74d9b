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
// factored recursively down to register-resident leaves; row interchanges are
// applied to the whole row (LAPACK convention, P A = L U) and also folded into
// a permutation vector.  After each panel the inverses of its 64 x 64 unit
// lower and upper diagonal blocks are formed, so every triangular step of the
// factorization and of the many-column solve is an MFMA matrix product.
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
// SET: C = A B, else C -= A B.  256 threads, 64 x 64 output tile, 4 waves of
// 32 x 32 (2 x 2 MFMA tiles), k staged through LDS 16 at a time; blockIdx.z
// selects the matrix of a strided batch (strides in complex elements).
// ---------------------------------------------------------------------------
#define GB_M 64
#define GB_N 64
#define GB_K 16

template <typename T, bool SET>
__global__ void __launch_bounds__(256) cgemm(int m, int n, int k, const T* __restrict__ A, int64_t lda, int64_t sa,
                                             const T* __restrict__ B, int64_t ldb, int64_t sb, T* __restrict__ C,
                                             int64_t ldc, int64_t sc) {
    typedef typename Mfma<T>::acc acc_t;
    __shared__ T a_re[GB_K][GB_M + 1], a_im[GB_K][GB_M + 1];
    __shared__ T b_re[GB_K][GB_N + 1], b_im[GB_K][GB_N + 1];
    A += 2 * sa * blockIdx.z;
    B += 2 * sb * blockIdx.z;
    C += 2 * sc * blockIdx.z;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * GB_M, n0 = blockIdx.y * GB_N;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    acc_t cr[2][2], ci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cr[i][j] = acc_t{0, 0, 0, 0};
            ci[i][j] = acc_t{0, 0, 0, 0};
        }
    const int fr = lane & 15, fk = lane >> 4;
    for (int k0 = 0; k0 < k; k0 += GB_K) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = tid & 63, c = (tid >> 6) + 4 * i;
            const int gr = m0 + r, gc = k0 + c;
            T vr = 0, vi = 0;
            if (gr < m && gc < k) {
                const T* p = A + 2 * ((int64_t)gr + (int64_t)gc * lda);
                vr = p[0];
                vi = p[1];
            }
            a_re[c][r] = vr;
            a_im[c][r] = vi;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = tid & 15, c = (tid >> 4) + 16 * i;
            const int gr = k0 + r, gc = n0 + c;
            T vr = 0, vi = 0;
            if (gr < k && gc < n) {
                const T* p = B + 2 * ((int64_t)gr + (int64_t)gc * ldb);
                vr = p[0];
                vi = p[1];
            }
            b_re[r][c] = vr;
            b_im[r][c] = vi;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 4) {
            T ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ar[i] = a_re[kk + fk][wm + 16 * i + fr];
                ai[i] = a_im[kk + fk][wm + 16 * i + fr];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                br[j] = b_re[kk + fk][wn + 16 * j + fr];
                bi[j] = b_im[kk + fk][wn + 16 * j + fr];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    cr[i][j] = Mfma<T>::run(ar[i], br[j], cr[i][j]);
                    cr[i][j] = Mfma<T>::run(-ai[i], bi[j], cr[i][j]);
                    ci[i][j] = Mfma<T>::run(ar[i], bi[j], ci[i][j]);
                    ci[i][j] = Mfma<T>::run(ai[i], br[j], ci[i][j]);
                }
        }
        __syncthreads();
    }
    // lane holds D[row(lane / 16, v)][lane % 16] of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gr = m0 + wm + 16 * i + Mfma<T>::row(fk, v);
                const int gc = n0 + wn + 16 * j + fr;
                if (gr < m && gc < n) {
                    T* p = C + 2 * ((int64_t)gr + (int64_t)gc * ldc);
                    if (SET) {
                        p[0] = cr[i][j][v];
                        p[1] = ci[i][j][v];
                    } else {
                        p[0] -= cr[i][j][v];
                        p[1] -= ci[i][j][v];
                    }
                }
            }
}

// ---------------------------------------------------------------------------
// Panel leaf: factor columns [0, w) of an m x w block (rows cached in registers).
// One workgroup of 1024 threads per matrix; thread t owns rows t, t + 1024, ...
// Pivot = first row maximising |re| + |im| (LAPACK izamax).  Two barriers per
// column: wave candidates and the pivot / current rows go through
// double-buffered LDS slots, and every thread reduces the 16 wave candidates
// itself, so no thread has to publish the pivot index separately.
// ---------------------------------------------------------------------------
template <typename T, int R, int W>
__global__ void __launch_bounds__(1024) panel_leaf(int m, int w, T* __restrict__ a, int64_t lda, int64_t sa,
                                                   int row0, int* __restrict__ piv, int64_t spiv,
                                                   unsigned* __restrict__ status) {
    __shared__ T cand_v[2][16];
    __shared__ int cand_i[2][16];
    __shared__ T prow_re[2][W], prow_im[2][W], crow_re[2][W], crow_im[2][W];
    a += 2 * sa * blockIdx.x;
    piv += spiv * blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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
        if (c >= w) break;
        const int buf = c & 1;
        T best = -1;
        int bi = 0x7fffffff;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + 1024 * r;
            if (row >= c && row < m) {
                const T v = fabs(vr[r][c]) + fabs(vi[r][c]);
                if (v > best || (v == best && row < bi)) { best = v; bi = row; }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if (lane == 0) { cand_v[buf][wave] = best; cand_i[buf][wave] = bi; }
        __syncthreads();
        T bv = cand_v[buf][0];
        int p = cand_i[buf][0];
#pragma unroll
        for (int i = 1; i < 16; ++i) {
            const T v = cand_v[buf][i];
            const int q = cand_i[buf][i];
            if (v > bv || (v == bv && q < p)) { bv = v; p = q; }
        }
        if (tid == 0) {
            piv[c] = row0 + p;
            if (!(bv > 0)) atomicOr(status, (unsigned)GRIM_DEV_ZERO_PIVOT);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + 1024 * r;
            if (row == p) {
#pragma unroll
                for (int cc = 0; cc < W; ++cc) { prow_re[buf][cc] = vr[r][cc]; prow_im[buf][cc] = vi[r][cc]; }
            }
            if (row == c) {
#pragma unroll
                for (int cc = 0; cc < W; ++cc) { crow_re[buf][cc] = vr[r][cc]; crow_im[buf][cc] = vi[r][cc]; }
            }
        }
        __syncthreads();
        const T dr = prow_re[buf][c], di = prow_im[buf][c];
        const T den = dr * dr + di * di;
        const T ir = den > 0 ? dr / den : 0, ii = den > 0 ? -di / den : 0;
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
}

// Apply interchanges piv[c0..c1) (absolute rows) to columns [0, ncols) of a,
// skipping columns [s0, s1).  One thread per column; batch in blockIdx.z.
template <typename T>
__global__ void __launch_bounds__(256) laswp(int ncols, T* __restrict__ a, int64_t lda, int64_t sa,
                                             const int* __restrict__ piv, int64_t spiv, int c0, int c1, int s0,
                                             int s1) {
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= ncols || (col >= s0 && col < s1)) return;
    a += 2 * sa * blockIdx.z;
    piv += spiv * blockIdx.z;
    T* base = a + 2 * (int64_t)col * lda;
    for (int c = c0; c < c1; ++c) {
        const int p = piv[c];
        if (p != c) {
            T* x = base + 2 * (int64_t)c;
            T* y = base + 2 * (int64_t)p;
            const T xr = x[0], xi = x[1];
            x[0] = y[0];
            x[1] = y[1];
            y[0] = xr;
            y[1] = xi;
        }
    }
}

#define GRIM_LU_NB 64

__device__ __forceinline__ double lane_value(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// X = T^{-1} B in place for a w x w triangle (w <= 64) and ncols columns:
// LOWER: unit lower triangle; else upper with its diagonal.  Lane i of a wave
// holds row i of TRSM_CPW columns; step j broadcasts x_j from lane j with
// v_readlane (wave-uniform, so it lands in SGPRs) and every lane below (above)
// subtracts its T[i][j] x_j.  The triangle is staged transposed in LDS so lanes
// read consecutive words.  256 threads = 4 waves; batch in blockIdx.z.
#define TRSM_CPW 4
template <typename T, bool LOWER>
__global__ void __launch_bounds__(256) trsm_lanes(int w, int ncols, const T* __restrict__ tri, int64_t ldt, int64_t st,
                                                  T* __restrict__ b, int64_t ldb, int64_t sb) {
    __shared__ T tr[GRIM_LU_NB][GRIM_LU_NB], ti[GRIM_LU_NB][GRIM_LU_NB];   // [col j][row i]
    tri += 2 * st * blockIdx.z;
    b += 2 * sb * blockIdx.z;
    for (int idx = threadIdx.x; idx < GRIM_LU_NB * GRIM_LU_NB; idx += 256) {
        const int i = idx & 63, j = idx >> 6;
        T x = 0, y = 0;
        if (i < w && j < w) {
            const T* p = tri + 2 * ((int64_t)i + (int64_t)j * ldt);
            x = p[0];
            y = p[1];
        }
        tr[j][i] = x;
        ti[j][i] = y;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col0 = (blockIdx.x * 4 + wave) * TRSM_CPW;
    T xr[TRSM_CPW], xi[TRSM_CPW];
#pragma unroll
    for (int c = 0; c < TRSM_CPW; ++c) {
        xr[c] = 0;
        xi[c] = 0;
        if (lane < w && col0 + c < ncols) {
            const T* p = b + 2 * ((int64_t)lane + (int64_t)(col0 + c) * ldb);
            xr[c] = p[0];
            xi[c] = p[1];
        }
    }
    if (LOWER) {
        for (int j = 0; j < w - 1; ++j) {
            const T lr = tr[j][lane], li = ti[j][lane];
            const bool below = lane > j;
#pragma unroll
            for (int c = 0; c < TRSM_CPW; ++c) {
                const T ur = lane_value(xr[c], j), ui = lane_value(xi[c], j);
                if (below) {
                    xr[c] -= lr * ur - li * ui;
                    xi[c] -= lr * ui + li * ur;
                }
            }
        }
    } else {
        for (int j = w - 1; j >= 0; --j) {
            const T dr = tr[j][j], di = ti[j][j];
            const T den = dr * dr + di * di;
            const T ir = dr / den, ii = -di / den;
            const T lr = tr[j][lane], li = ti[j][lane];
            const bool above = lane < j;
#pragma unroll
            for (int c = 0; c < TRSM_CPW; ++c) {
                const T br = lane_value(xr[c], j), bi = lane_value(xi[c], j);
                const T ur = br * ir - bi * ii, ui = br * ii + bi * ir;
                if (lane == j) {
                    xr[c] = ur;
                    xi[c] = ui;
                } else if (above) {
                    xr[c] -= lr * ur - li * ui;
                    xi[c] -= lr * ui + li * ur;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < TRSM_CPW; ++c) {
        if (lane < w && col0 + c < ncols) {
            T* p = b + 2 * ((int64_t)lane + (int64_t)(col0 + c) * ldb);
            p[0] = xr[c];
            p[1] = xi[c];
        }
    }
}

// perm[i] = original row that ends at row i after the interchanges piv[0..n).
__global__ void build_perm(int n, const int* __restrict__ piv, int64_t spiv, int* __restrict__ perm,
                           int64_t sperm) {
    if (threadIdx.x != 0) return;
    piv += spiv * blockIdx.x;
    perm += sperm * blockIdx.x;
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int c = 0; c < n; ++c) {
        const int p = piv[c];
        const int t = perm[c];
        perm[c] = perm[p];
        perm[p] = t;
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

template <typename T, bool SET = false>
static hipError_t gemm(const Batch& bt, int m, int n, int k, const T* A, int64_t lda, int64_t sa, const T* B,
                       int64_t ldb, int64_t sb, T* C, int64_t ldc, int64_t sc) {
    if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
    dim3 grid((m + GB_M - 1) / GB_M, (n + GB_N - 1) / GB_N, bt.count);
    cgemm<T, SET><<<grid, 256, 0, bt.st>>>(m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc);
    return hipGetLastError();
}

template <typename T>
static hipError_t leaf(const Batch& bt, int m, int w, T* a, int64_t lda, int64_t sa, int row0, int* piv,
                       int64_t spiv, unsigned* status) {
    hipStream_t st = bt.st;
    const int b = bt.count;
    if (m <= 1024) panel_leaf<T, 1, 16><<<b, 1024, 0, st>>>(m, w, a, lda, sa, row0, piv, spiv, status);
    else if (m <= 2048) panel_leaf<T, 2, 8><<<b, 1024, 0, st>>>(m, w, a, lda, sa, row0, piv, spiv, status);
    else if (m <= 4096) panel_leaf<T, 4, 4><<<b, 1024, 0, st>>>(m, w, a, lda, sa, row0, piv, spiv, status);
    else if (m <= 8192) panel_leaf<T, 8, 2><<<b, 1024, 0, st>>>(m, w, a, lda, sa, row0, piv, spiv, status);
    else if (m <= 16384) panel_leaf<T, 16, 1><<<b, 1024, 0, st>>>(m, w, a, lda, sa, row0, piv, spiv, status);
    else panel_leaf<T, 32, 1><<<b, 1024, 0, st>>>(m, w, a, lda, sa, row0, piv, spiv, status);
    return hipGetLastError();
}

static int leaf_width(int m) {
    if (m <= 1024) return 16;
    if (m <= 2048) return 8;
    if (m <= 4096) return 4;
    if (m <= 8192) return 2;
    return 1;
}

template <typename T>
static hipError_t swap_cols(const Batch& bt, int ncols, T* a, int64_t lda, int64_t sa, const int* piv, int64_t spiv,
                            int c0, int c1, int s0, int s1) {
    if (ncols <= 0 || c1 <= c0) return hipSuccess;
    laswp<T><<<dim3((ncols + 255) / 256, 1, bt.count), 256, 0, bt.st>>>(ncols, a, lda, sa, piv, spiv, c0, c1, s0,
                                                                         s1);
    return hipGetLastError();
}

template <typename T, bool LOWER>
static hipError_t trsm(const Batch& bt, int w, int ncols, const T* tri, int64_t ldt, int64_t st, T* b, int64_t ldb,
                       int64_t sb) {
    if (w <= 0 || ncols <= 0) return hipSuccess;
    const int per_block = 4 * TRSM_CPW;
    trsm_lanes<T, LOWER><<<dim3((ncols + per_block - 1) / per_block, 1, bt.count), 256, 0, bt.st>>>(
        w, ncols, tri, ldt, st, b, ldb, sb);
    return hipGetLastError();
}

// Recursive factorization of columns [j0, j0 + w) (rows j0..n-1); interchanges
// are applied to all columns of this column range.
template <typename T>
static hipError_t panel_rec(const Batch& bt, int n, T* a, int64_t lda, int64_t sa, int j0, int w, int* piv,
                            int64_t spiv, unsigned* status) {
    const int m = n - j0;
    const int lw = leaf_width(m);
    T* top = a + 2 * ((int64_t)j0 + (int64_t)j0 * lda);
    if (w <= lw) return leaf<T>(bt, m, w, top, lda, sa, j0, piv + j0, spiv, status);
    int w1 = w / 2;
    if (w1 > lw) w1 = (w1 / lw) * lw;
    hipError_t err = panel_rec<T>(bt, n, a, lda, sa, j0, w1, piv, spiv, status);
    if (err != hipSuccess) return err;
    const int w2 = w - w1;
    T* right = a + 2 * ((int64_t)(j0 + w1) * lda);
    if ((err = swap_cols<T>(bt, w2, right, lda, sa, piv, spiv, j0, j0 + w1, -1, -1)) != hipSuccess) return err;
    if ((err = trsm<T, true>(bt, w1, w2, top, lda, sa, right + 2 * (int64_t)j0, lda, sa)) != hipSuccess) return err;
    if ((err = gemm<T>(bt, m - w1, w2, w1, top + 2 * (int64_t)w1, lda, sa, right + 2 * (int64_t)j0, lda, sa,
                       right + 2 * (int64_t)(j0 + w1), lda, sa)) != hipSuccess)
        return err;
    if ((err = panel_rec<T>(bt, n, a, lda, sa, j0 + w1, w2, piv, spiv, status)) != hipSuccess) return err;
    T* left = a + 2 * (int64_t)j0 * lda;
    return swap_cols<T>(bt, w1, left, lda, sa, piv, spiv, j0 + w1, j0 + w, -1, -1);
}

// Factor-side auxiliary storage per matrix: the row permutation (n int32)
// that the pivots compose to, applied to right-hand sides in one gather.
static int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }
static int64_t aux_bytes(int n) { return align256(4 * (int64_t)n); }

template <typename T>
static hipError_t lu_factor(const Batch& bt, int n, T* a, int64_t lda, int64_t sa, int* piv, int64_t spiv,
                            void* auxp, int64_t saux, unsigned* status) {
    hipError_t err;
    for (int j = 0; j < n; j += GRIM_LU_NB) {
        const int jb = min(GRIM_LU_NB, n - j);
        if ((err = panel_rec<T>(bt, n, a, lda, sa, j, jb, piv, spiv, status)) != hipSuccess) return err;
        // the panel's interchanges on every other column (LAPACK convention)
        if ((err = swap_cols<T>(bt, n, a, lda, sa, piv, spiv, j, j + jb, j, j + jb)) != hipSuccess) return err;
        const int rest = n - j - jb;
        if (rest > 0) {
            T* diag = a + 2 * ((int64_t)j + (int64_t)j * lda);
            T* a12 = a + 2 * ((int64_t)j + (int64_t)(j + jb) * lda);
            // U12 = L11^{-1} A12 in place, then A22 -= L21 U12
            if ((err = trsm<T, true>(bt, jb, rest, diag, lda, sa, a12, lda, sa)) != hipSuccess) return err;
            if ((err = gemm<T>(bt, rest, rest, jb, diag + 2 * (int64_t)jb, lda, sa, a12, lda, sa,
                               a12 + 2 * (int64_t)jb, lda, sa)) != hipSuccess)
                return err;
        }
    }
    build_perm<<<bt.count, 64, 0, bt.st>>>(n, piv, spiv, (int*)auxp, saux / 4);
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
    // forward: w_j <- L_jj^{-1} w_j ; w_{j+} -= L_{j+, j} w_j
    for (int j = 0; j < n; j += GRIM_LU_NB) {
        const int jb = min(GRIM_LU_NB, n - j);
        const T* djj = a + 2 * ((int64_t)j + (int64_t)j * lda);
        if ((err = trsm<T, true>(bt, jb, nrhs, djj, lda, sa, w + 2 * (int64_t)j, ldw, sw)) != hipSuccess) return err;
        if (j + jb < n)
            if ((err = gemm<T>(bt, n - j - jb, nrhs, jb, djj + 2 * (int64_t)jb, lda, sa, w + 2 * (int64_t)j, ldw, sw,
                               w + 2 * (int64_t)(j + jb), ldw, sw)) != hipSuccess)
                return err;
    }
    // backward: w_j <- U_jj^{-1} w_j ; w_{<j} -= U_{<j, j} w_j
    const int last = ((n - 1) / GRIM_LU_NB) * GRIM_LU_NB;
    for (int j = last; j >= 0; j -= GRIM_LU_NB) {
        const int jb = min(GRIM_LU_NB, n - j);
        const T* djj = a + 2 * ((int64_t)j + (int64_t)j * lda);
        if ((err = trsm<T, false>(bt, jb, nrhs, djj, lda, sa, w + 2 * (int64_t)j, ldw, sw)) != hipSuccess) return err;
        if (j > 0)
            if ((err = gemm<T>(bt, j, nrhs, jb, a + 2 * (int64_t)j * lda, lda, sa, w + 2 * (int64_t)j, ldw, sw, w,
                               ldw, sw)) != hipSuccess)
                return err;
    }
    copy_cast<T, T><<<dim3((n + 255) / 256, nrhs, bt.count), 256, 0, bt.st>>>(n, w, ldw, sw, b, ldb, sb);
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
