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
// Pivot convention ("panel-delayed"): the factorization runs in panels of
// GRIM_LU_NB columns; row interchanges of a panel are applied to every column
// of that panel and to all later columns, never to earlier panels.  piv[c] is
// the absolute row exchanged with row c.  The solve applies each panel's
// interchanges to the right-hand sides just before that panel's forward step,
// which reproduces P A = L U exactly.
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
// C[m x n] -= A[m x k] * B[k x n]   (complex, column-major, interleaved)
// 256 threads, 64 x 64 output tile, 4 waves of 32 x 32 (2 x 2 MFMA tiles).
// ---------------------------------------------------------------------------
#define GB_M 64
#define GB_N 64
#define GB_K 16

template <typename T>
__global__ void __launch_bounds__(256) cgemm_minus(int m, int n, int k, const T* __restrict__ A, int64_t lda,
                                                   const T* __restrict__ B, int64_t ldb, T* __restrict__ C,
                                                   int64_t ldc) {
    typedef typename Mfma<T>::acc acc_t;
    __shared__ T a_re[GB_K][GB_M + 1], a_im[GB_K][GB_M + 1];
    __shared__ T b_re[GB_K][GB_N + 1], b_im[GB_K][GB_N + 1];
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
        // A tile: 64 rows x 16 cols; thread -> row tid % 64, cols tid / 64 + 4 i
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
        // B tile: 16 rows x 64 cols; thread -> row tid % 16, col tid / 16 + 16 i
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
                    p[0] -= cr[i][j][v];
                    p[1] -= ci[i][j][v];
                }
            }
}

// ---------------------------------------------------------------------------
// Panel leaf: factor columns [0, w) of an m x w block (rows cached in registers).
// One workgroup of 1024 threads; thread t owns rows t, t + 1024, ...
// Pivot = first row maximising |re| + |im| (LAPACK izamax).
// ---------------------------------------------------------------------------
template <typename T, int R, int W>
__global__ void __launch_bounds__(1024) panel_leaf(int m, int w, T* __restrict__ a, int64_t lda, int row0,
                                                   int* __restrict__ piv, unsigned* __restrict__ status) {
    __shared__ T red_v[16];
    __shared__ int red_i[16];
    __shared__ T prow_re[W], prow_im[W];
    __shared__ T crow_re[W], crow_im[W];
    __shared__ int s_piv;
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
        // local argmax over own rows >= c
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
        if (lane == 0) { red_v[wave] = best; red_i[wave] = bi; }
        __syncthreads();
        if (tid == 0) {
            T bv = red_v[0];
            int bidx = red_i[0];
            for (int i = 1; i < 16; ++i)
                if (red_v[i] > bv || (red_v[i] == bv && red_i[i] < bidx)) { bv = red_v[i]; bidx = red_i[i]; }
            s_piv = bidx;
            piv[c] = row0 + bidx;
            if (!(bv > 0)) atomicOr(status, (unsigned)GRIM_DEV_ZERO_PIVOT);
        }
        __syncthreads();
        const int p = s_piv;
        // publish pivot row and row c (all W columns), then swap
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + 1024 * r;
            if (row == p) {
#pragma unroll
                for (int cc = 0; cc < W; ++cc) { prow_re[cc] = vr[r][cc]; prow_im[cc] = vi[r][cc]; }
            }
            if (row == c) {
#pragma unroll
                for (int cc = 0; cc < W; ++cc) { crow_re[cc] = vr[r][cc]; crow_im[cc] = vi[r][cc]; }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + 1024 * r;
            if (row == c && p != c) {
#pragma unroll
                for (int cc = 0; cc < W; ++cc) { vr[r][cc] = prow_re[cc]; vi[r][cc] = prow_im[cc]; }
            } else if (row == p && p != c) {
#pragma unroll
                for (int cc = 0; cc < W; ++cc) { vr[r][cc] = crow_re[cc]; vi[r][cc] = crow_im[cc]; }
            }
        }
        // scale and rank-1 update of rows > c using the pivot row (now row c)
        const T dr = prow_re[c], di = prow_im[c];
        const T den = dr * dr + di * di;
        const T ir = den > 0 ? dr / den : 0, ii = den > 0 ? -di / den : 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + 1024 * r;
            if (row > c && row < m) {
                const T lr = vr[r][c] * ir - vi[r][c] * ii;
                const T li = vr[r][c] * ii + vi[r][c] * ir;
                vr[r][c] = lr;
                vi[r][c] = li;
#pragma unroll
                for (int cc = 0; cc < W; ++cc) {
                    if (cc > c) {
                        const T ur = prow_re[cc], ui = prow_im[cc];
                        vr[r][cc] -= lr * ur - li * ui;
                        vi[r][cc] -= lr * ui + li * ur;
                    }
                }
            }
        }
        __syncthreads();
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

// Apply interchanges piv[c0..c1) (absolute rows) to columns [0, ncols) of a.
template <typename T>
__global__ void __launch_bounds__(256) laswp(int ncols, T* __restrict__ a, int64_t lda, const int* __restrict__ piv,
                                             int c0, int c1) {
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= ncols) return;
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

// Triangular solve with a w x w (w <= 32) triangle for many columns:
// lower = 1: X = L^{-1} B with unit diagonal; lower = 0: X = U^{-1} B.
template <typename T, int W>
__global__ void __launch_bounds__(64) trsm_small(int w, int ncols, const T* __restrict__ tri, int64_t ldt,
                                                 T* __restrict__ b, int64_t ldb, int lower) {
    __shared__ T t_re[W][W], t_im[W][W];
    for (int i = threadIdx.x; i < W * W; i += 64) {
        const int r = i % W, c = i / W;
        T x = 0, y = 0;
        if (r < w && c < w) {
            const T* p = tri + 2 * ((int64_t)r + (int64_t)c * ldt);
            x = p[0];
            y = p[1];
        }
        t_re[r][c] = x;
        t_im[r][c] = y;
    }
    __syncthreads();
    const int col = blockIdx.x * 64 + threadIdx.x;
    if (col >= ncols) return;
    T xr[W], xi[W];
    T* base = b + 2 * (int64_t)col * ldb;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        xr[i] = i < w ? base[2 * i] : 0;
        xi[i] = i < w ? base[2 * i + 1] : 0;
    }
    if (lower) {
#pragma unroll
        for (int i = 0; i < W; ++i) {
#pragma unroll
            for (int j = 0; j < W; ++j) {
                if (j < i) {
                    xr[i] -= t_re[i][j] * xr[j] - t_im[i][j] * xi[j];
                    xi[i] -= t_re[i][j] * xi[j] + t_im[i][j] * xr[j];
                }
            }
        }
    } else {
#pragma unroll
        for (int i = W - 1; i >= 0; --i) {
            if (i < w) {
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    if (j > i && j < w) {
                        xr[i] -= t_re[i][j] * xr[j] - t_im[i][j] * xi[j];
                        xi[i] -= t_re[i][j] * xi[j] + t_im[i][j] * xr[j];
                    }
                }
                const T dr = t_re[i][i], di = t_im[i][i];
                const T den = dr * dr + di * di;
                const T nr = (xr[i] * dr + xi[i] * di) / den;
                const T ni = (xi[i] * dr - xr[i] * di) / den;
                xr[i] = nr;
                xi[i] = ni;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < W; ++i) {
        if (i < w) {
            base[2 * i] = xr[i];
            base[2 * i + 1] = xi[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Elementwise helpers for mixed precision and refinement
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) to_c64(int m, int n, const double* __restrict__ src, int64_t lds,
                                              float* __restrict__ dst, int64_t ldd) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= m) return;
    const double* s = src + 2 * ((int64_t)r + (int64_t)c * lds);
    float* d = dst + 2 * ((int64_t)r + (int64_t)c * ldd);
    d[0] = (float)s[0];
    d[1] = (float)s[1];
}

__global__ void __launch_bounds__(256) copy_c128(int m, int n, const double* __restrict__ src, int64_t lds,
                                                 double* __restrict__ dst, int64_t ldd) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= m) return;
    const double* s = src + 2 * ((int64_t)r + (int64_t)c * lds);
    double* d = dst + 2 * ((int64_t)r + (int64_t)c * ldd);
    d[0] = s[0];
    d[1] = s[1];
}

// x (c128) += d (c64)
__global__ void __launch_bounds__(256) add_c64(int m, int n, const float* __restrict__ d, int64_t ldd,
                                               double* __restrict__ x, int64_t ldx) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= m) return;
    const float* s = d + 2 * ((int64_t)r + (int64_t)c * ldd);
    double* t = x + 2 * ((int64_t)r + (int64_t)c * ldx);
    t[0] += (double)s[0];
    t[1] += (double)s[1];
}

// Per column: out[c] = max_r |v[r, c]| (complex modulus).
__global__ void __launch_bounds__(256) colmax_abs(int m, const double* __restrict__ v, int64_t ld,
                                                  double* __restrict__ out) {
    const int c = blockIdx.x;
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
    if (threadIdx.x == 0) out[c] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
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
template <typename T>
static hipError_t gemm_minus(int m, int n, int k, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
                             int64_t ldc, hipStream_t st) {
    if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
    dim3 grid((m + GB_M - 1) / GB_M, (n + GB_N - 1) / GB_N);
    cgemm_minus<T><<<grid, 256, 0, st>>>(m, n, k, A, lda, B, ldb, C, ldc);
    return hipGetLastError();
}

template <typename T>
static hipError_t trsm(int w, int ncols, const T* tri, int64_t ldt, T* b, int64_t ldb, int lower, hipStream_t st) {
    // blocks of <= 32: X1 = T11^-1 B1, B2 -= T21 X1, ... (lower) or bottom-up (upper)
    if (w <= 0 || ncols <= 0) return hipSuccess;
    const int nb = 32;
    dim3 grid((ncols + 63) / 64);
    if (lower) {
        for (int j = 0; j < w; j += nb) {
            const int jb = min(nb, w - j);
            trsm_small<T, 32><<<grid, 64, 0, st>>>(jb, ncols, tri + 2 * (j + j * ldt), ldt, b + 2 * j, ldb, 1);
            if (j + jb < w)
                gemm_minus<T>(w - j - jb, ncols, jb, tri + 2 * ((j + jb) + j * ldt), ldt, b + 2 * j, ldb,
                              b + 2 * (j + jb), ldb, st);
        }
    } else {
        const int last = ((w - 1) / nb) * nb;
        for (int j = last; j >= 0; j -= nb) {
            const int jb = min(nb, w - j);
            trsm_small<T, 32><<<grid, 64, 0, st>>>(jb, ncols, tri + 2 * (j + j * ldt), ldt, b + 2 * j, ldb, 0);
            if (j > 0) gemm_minus<T>(j, ncols, jb, tri + 2 * ((int64_t)j * ldt), ldt, b + 2 * j, ldb, b, ldb, st);
        }
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t leaf(int m, int w, T* a, int64_t lda, int row0, int* piv, unsigned* status, hipStream_t st) {
    // rows cached per thread R and leaf width W with R * W <= 16 complex values
    if (m <= 1024) {
        panel_leaf<T, 1, 16><<<1, 1024, 0, st>>>(m, w, a, lda, row0, piv, status);
    } else if (m <= 2048) {
        panel_leaf<T, 2, 8><<<1, 1024, 0, st>>>(m, w, a, lda, row0, piv, status);
    } else if (m <= 4096) {
        panel_leaf<T, 4, 4><<<1, 1024, 0, st>>>(m, w, a, lda, row0, piv, status);
    } else if (m <= 8192) {
        panel_leaf<T, 8, 2><<<1, 1024, 0, st>>>(m, w, a, lda, row0, piv, status);
    } else if (m <= 16384) {
        panel_leaf<T, 16, 1><<<1, 1024, 0, st>>>(m, w, a, lda, row0, piv, status);
    } else if (m <= 32768) {
        panel_leaf<T, 32, 1><<<1, 1024, 0, st>>>(m, w, a, lda, row0, piv, status);
    } else {
        panel_leaf<T, 64, 1><<<1, 1024, 0, st>>>(m, w, a, lda, row0, piv, status);
    }
    return hipGetLastError();
}

static int leaf_width(int m) {
    if (m <= 1024) return 16;
    if (m <= 2048) return 8;
    if (m <= 4096) return 4;
    if (m <= 8192) return 2;
    return 1;
}

// Recursive factorization of the panel a[j0.., j0..j0+w) of an n x n matrix
// (rows j0..n-1).  Interchanges of the right half are applied to the left half.
template <typename T>
static hipError_t panel_rec(int n, T* a, int64_t lda, int j0, int w, int* piv, unsigned* status, hipStream_t st) {
    const int m = n - j0;
    const int lw = leaf_width(m);
    T* top = a + 2 * ((int64_t)j0 + (int64_t)j0 * lda);
    if (w <= lw) return leaf<T>(m, w, top, lda, j0, piv + j0, status, st);
    int w1 = w / 2;
    if (w1 > lw) w1 = (w1 / lw) * lw;
    hipError_t err = panel_rec<T>(n, a, lda, j0, w1, piv, status, st);
    if (err != hipSuccess) return err;
    const int w2 = w - w1;
    T* right = a + 2 * ((int64_t)j0 + w1) * lda;   // column j0 + w1, row 0
    laswp<T><<<(w2 + 255) / 256, 256, 0, st>>>(w2, right, lda, piv, j0, j0 + w1);
    // U12 = L11^-1 A12 ; A22 -= L21 U12
    err = trsm<T>(w1, w2, top, lda, right + 2 * (int64_t)j0, lda, 1, st);
    if (err != hipSuccess) return err;
    err = gemm_minus<T>(m - w1, w2, w1, top + 2 * (int64_t)w1, lda, right + 2 * (int64_t)j0, lda,
                        right + 2 * (int64_t)(j0 + w1), lda, st);
    if (err != hipSuccess) return err;
    err = panel_rec<T>(n, a, lda, j0 + w1, w2, piv, status, st);
    if (err != hipSuccess) return err;
    // right-half interchanges onto the left half of this panel
    T* left = a + 2 * (int64_t)j0 * lda;
    laswp<T><<<(w1 + 255) / 256, 256, 0, st>>>(w1, left, lda, piv, j0 + w1, j0 + w);
    return hipGetLastError();
}

#define GRIM_LU_NB 64

template <typename T>
static hipError_t lu_factor(int n, T* a, int64_t lda, int* piv, unsigned* status, hipStream_t st) {
    for (int j = 0; j < n; j += GRIM_LU_NB) {
        const int jb = min(GRIM_LU_NB, n - j);
        hipError_t err = panel_rec<T>(n, a, lda, j, jb, piv, status, st);
        if (err != hipSuccess) return err;
        const int rest = n - j - jb;
        if (rest > 0) {
            T* right = a + 2 * (int64_t)(j + jb) * lda;
            laswp<T><<<(rest + 255) / 256, 256, 0, st>>>(rest, right, lda, piv, j, j + jb);
            const T* diag = a + 2 * ((int64_t)j + (int64_t)j * lda);
            err = trsm<T>(jb, rest, diag, lda, right + 2 * (int64_t)j, lda, 1, st);
            if (err != hipSuccess) return err;
            err = gemm_minus<T>(rest, rest, jb, diag + 2 * (int64_t)jb, lda, right + 2 * (int64_t)j, lda,
                                right + 2 * (int64_t)(j + jb), lda, st);
            if (err != hipSuccess) return err;
        }
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t lu_solve(int n, int nrhs, const T* a, int64_t lda, const int* piv, T* b, int64_t ldb,
                           hipStream_t st) {
    for (int j = 0; j < n; j += GRIM_LU_NB) {
        const int jb = min(GRIM_LU_NB, n - j);
        laswp<T><<<(nrhs + 255) / 256, 256, 0, st>>>(nrhs, b, ldb, piv, j, j + jb);
        const T* diag = a + 2 * ((int64_t)j + (int64_t)j * lda);
        hipError_t err = trsm<T>(jb, nrhs, diag, lda, b + 2 * (int64_t)j, ldb, 1, st);
        if (err != hipSuccess) return err;
        if (j + jb < n) {
            err = gemm_minus<T>(n - j - jb, nrhs, jb, diag + 2 * (int64_t)jb, lda, b + 2 * (int64_t)j, ldb,
                                b + 2 * (int64_t)(j + jb), ldb, st);
            if (err != hipSuccess) return err;
        }
    }
    const int last = ((n - 1) / GRIM_LU_NB) * GRIM_LU_NB;
    for (int j = last; j >= 0; j -= GRIM_LU_NB) {
        const int jb = min(GRIM_LU_NB, n - j);
        const T* diag = a + 2 * ((int64_t)j + (int64_t)j * lda);
        hipError_t err = trsm<T>(jb, nrhs, diag, lda, b + 2 * (int64_t)j, ldb, 0, st);
        if (err != hipSuccess) return err;
        if (j > 0) {
            err = gemm_minus<T>(j, nrhs, jb, a + 2 * (int64_t)j * lda, lda, b + 2 * (int64_t)j, ldb, b, ldb, st);
            if (err != hipSuccess) return err;
        }
    }
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_lu_factor(int n, void* a, int64_t lda, int single, int* piv, unsigned* status,
                                           hipStream_t st) {
    return single ? lu_factor<float>(n, (float*)a, lda, piv, status, st)
                  : lu_factor<double>(n, (double*)a, lda, piv, status, st);
}

extern "C" hipError_t grim_dense_lu_solve(int n, int nrhs, const void* a, int64_t lda, int single, const int* piv,
                                          void* b, int64_t ldb, hipStream_t st) {
    return single ? lu_solve<float>(n, nrhs, (const float*)a, lda, piv, (float*)b, ldb, st)
                  : lu_solve<double>(n, nrhs, (const double*)a, lda, piv, (double*)b, ldb, st);
}

extern "C" hipError_t grim_dense_gemm_minus(int m, int n, int k, const void* a, int64_t lda, const void* b,
                                            int64_t ldb, void* c, int64_t ldc, int single, hipStream_t st) {
    return single ? gemm_minus<float>(m, n, k, (const float*)a, lda, (const float*)b, ldb, (float*)c, ldc, st)
                  : gemm_minus<double>(m, n, k, (const double*)a, lda, (const double*)b, ldb, (double*)c, ldc, st);
}

extern "C" hipError_t grim_dense_to_c64(int m, int n, const double* src, int64_t lds, float* dst, int64_t ldd,
                                        hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    to_c64<<<dim3((m + 255) / 256, n), 256, 0, st>>>(m, n, src, lds, dst, ldd);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_copy(int m, int n, const double* src, int64_t lds, double* dst, int64_t ldd,
                                      hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    copy_c128<<<dim3((m + 255) / 256, n), 256, 0, st>>>(m, n, src, lds, dst, ldd);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_add_c64(int m, int n, const float* d, int64_t ldd, double* x, int64_t ldx,
                                        hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    add_c64<<<dim3((m + 255) / 256, n), 256, 0, st>>>(m, n, d, ldd, x, ldx);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_colmax(int m, int n, const double* v, int64_t ld, double* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    colmax_abs<<<n, 256, 0, st>>>(m, v, ld, out);
    return hipGetLastError();
}

extern "C" hipError_t grim_dense_colnorm(int m, int n, const double* v, int64_t ld, double* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    colnorm2<<<n, 256, 0, st>>>(m, v, ld, out);
    return hipGetLastError();
}
