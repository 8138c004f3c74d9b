// C ABI of the GRIM-2D HIP library (declarations in include/grim2d.h).
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/grim2d.h"
#include "common.h"

extern "C" hipError_t grim_upload_rules(const double*, const double*, const double*, const double*, const double*,
                                        const double*, const double*);
extern "C" hipError_t grim_launch_far(int q, int n, const double* p0, const double* p1, const int* next,
                                      const int* prev, double k, double* ztm, int64_t ldtm, double* zte,
                                      int64_t ldte, const int* fix_rows, int nfix, hipStream_t st);
extern "C" hipError_t grim_launch_near(int n, const double* p0, const double* p1, const int* next, const int* prev,
                                       double k, int* counts, int* offsets, int capacity, int* pair_e, int* pair_f,
                                       double* blocks, int* kind, unsigned* status, double* ztm, int64_t ldtm,
                                       double* zte, int64_t ldte, hipStream_t st);
extern "C" hipError_t grim_launch_loads(int n, int na, const double* p0, const double* p1, const int* prev, double k,
                                        const double* angles_deg, double* bu, double* bdn, int64_t ld,
                                        hipStream_t st);
extern "C" hipError_t grim_launch_farfield(int n, int na, const double* x, int64_t ldx, const double* bu,
                                           const double* bdn, int64_t ldb, double k, double* amp, double* sigma,
                                           hipStream_t st);
extern "C" hipError_t grim_dense_lu_factor(int n, void* a, int64_t lda, int single, int* piv, unsigned* status,
                                           hipStream_t st);
extern "C" hipError_t grim_dense_lu_solve(int n, int nrhs, const void* a, int64_t lda, int single, const int* piv,
                                          void* b, int64_t ldb, hipStream_t st);
extern "C" hipError_t grim_dense_gemm_minus(int m, int n, int k, const void* a, int64_t lda, const void* b,
                                            int64_t ldb, void* c, int64_t ldc, int single, hipStream_t st);
extern "C" hipError_t grim_dense_to_c64(int m, int n, const double* src, int64_t lds, float* dst, int64_t ldd,
                                        hipStream_t st);
extern "C" hipError_t grim_dense_copy(int m, int n, const double* src, int64_t lds, double* dst, int64_t ldd,
                                      hipStream_t st);
extern "C" hipError_t grim_dense_add_c64(int m, int n, const float* d, int64_t ldd, double* x, int64_t ldx,
                                        hipStream_t st);
extern "C" hipError_t grim_dense_colmax(int m, int n, const double* v, int64_t ld, double* out, hipStream_t st);
extern "C" hipError_t grim_dense_colnorm(int m, int n, const double* v, int64_t ld, double* out, hipStream_t st);

#define CHECK(x)                                   \
    do {                                           \
        if ((x) != hipSuccess) return GRIM_ERR_HIP; \
    } while (0)

static const int64_t kAlign = 256;
static int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

__global__ void zero_words(unsigned* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void zero_c128(int m, int n, double* x, int64_t ld) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= m) return;
    double* t = x + 2 * ((int64_t)r + (int64_t)c * ld);
    t[0] = 0.0;
    t[1] = 0.0;
}

extern "C" int grim_version(void) { return 100; }

// Gauss-Legendre rules (numpy.polynomial.legendre.leggauss mapped to [0, 1],
// the reference's _quadrature_nodes, twod/operators.py:3319-3323) and the
// 20-point t^4 graded corner rule (twod/polynomial_quadrature.py:319-335).
extern "C" int grim_init(int device) {
    if (hipSetDevice(device) != hipSuccess) return GRIM_ERR_HIP;
    const int n = 20;
    std::vector<double> tx(800), ty(800), tw(800);
    for (int i = 0; i < n; ++i) {
        const double t = GRIM_GL20_T[i], w = GRIM_GL20_W[i];
        const double u = t * t * t * t;
        for (int j = 0; j < n; ++j) {
            const double v = GRIM_GL20_T[j];
            const double wt = 4.0 * (t * t * t * t * t * t * t) * w * GRIM_GL20_W[j];
            const int a = i * n + j;
            tx[a] = u;
            ty[a] = u * v;
            tw[a] = wt;
            tx[400 + a] = u * v;
            ty[400 + a] = u;
            tw[400 + a] = wt;
        }
    }
    CHECK(grim_upload_rules(GRIM_GL8_T, GRIM_GL8_W, GRIM_GL16_T, GRIM_GL16_W, tx.data(), ty.data(), tw.data()));
    CHECK(hipDeviceSynchronize());
    return GRIM_OK;
}

struct AssemblyWork {
    unsigned* status;
    int *counts, *offsets, *pair_e, *pair_f, *kind;
    double* blocks;
    int64_t bytes;
};

static AssemblyWork carve(void* base, int64_t n, int64_t cap) {
    AssemblyWork w;
    char* p = (char*)base;
    int64_t off = 0;
    w.status = (unsigned*)(p + off); off += kAlign;
    w.counts = (int*)(p + off); off += align_up(4 * n);
    w.offsets = (int*)(p + off); off += align_up(4 * (n + 1));
    w.pair_e = (int*)(p + off); off += align_up(4 * cap);
    w.pair_f = (int*)(p + off); off += align_up(4 * cap);
    w.kind = (int*)(p + off); off += align_up(4 * cap);
    w.blocks = (double*)(p + off); off += align_up(16 * 8 * cap);
    w.bytes = off;
    return w;
}

extern "C" int64_t grim_assemble_workspace_bytes(int64_t n, int64_t near_per_element) {
    if (n <= 0 || near_per_element <= 0) return -1;
    return carve(nullptr, n, n * near_per_element).bytes;
}


static int assemble_pass(int passes, int64_t n, const double* p0, const double* p1, const int32_t* next,
                         const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k, int far_order,
                         double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te, void* workspace,
                         int64_t workspace_bytes, int64_t near_per_element, void* stream) {
    if (n < 2 || n > (1 << 30) || !p0 || !p1 || !next || !prev || !z_tm || !z_te || !workspace ||
        ld_tm < n || ld_te < n || !(k > 0) || !isfinite(k) || near_per_element <= 0 || nfix < 0 ||
        (nfix > 0 && !fix_rows))
        return GRIM_ERR_ARGS;
    if (far_order != 8 && far_order != 16) return GRIM_ERR_ARGS;
    const int64_t cap = n * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    AssemblyWork w = carve(workspace, n, cap);
    if (w.bytes > workspace_bytes) return GRIM_ERR_RANGE;
    hipStream_t st = (hipStream_t)stream;
    if (passes & 1) {
        zero_words<<<1, 64, 0, st>>>(w.status, 1);
        CHECK(hipGetLastError());
        CHECK(grim_launch_far(far_order, (int)n, p0, p1, next, prev, k, z_tm, ld_tm, z_te, ld_te, fix_rows,
                              (int)nfix, st));
    }
    if (passes & 2)
        CHECK(grim_launch_near((int)n, p0, p1, next, prev, k, w.counts, w.offsets, (int)cap, w.pair_e, w.pair_f,
                               w.blocks, w.kind, w.status, z_tm, ld_tm, z_te, ld_te, st));
    return GRIM_OK;
}

extern "C" int grim_assemble(int64_t n, const double* p0, const double* p1, const int32_t* next,
                             const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k, int far_order,
                             double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te, void* workspace,
                             int64_t workspace_bytes, int64_t near_per_element, void* stream) {
    return assemble_pass(3, n, p0, p1, next, prev, fix_rows, nfix, k, far_order, z_tm, ld_tm, z_te, ld_te,
                         workspace, workspace_bytes, near_per_element, stream);
}

extern "C" int grim_assemble_far(int64_t n, const double* p0, const double* p1, const int32_t* next,
                                 const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k, int far_order,
                                 double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te, void* workspace,
                                 int64_t workspace_bytes, int64_t near_per_element, void* stream) {
    return assemble_pass(1, n, p0, p1, next, prev, fix_rows, nfix, k, far_order, z_tm, ld_tm, z_te, ld_te,
                         workspace, workspace_bytes, near_per_element, stream);
}

extern "C" int grim_assemble_near(int64_t n, const double* p0, const double* p1, const int32_t* next,
                                  const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k,
                                  int far_order, double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te,
                                  void* workspace, int64_t workspace_bytes, int64_t near_per_element, void* stream) {
    return assemble_pass(2, n, p0, p1, next, prev, fix_rows, nfix, k, far_order, z_tm, ld_tm, z_te, ld_te,
                         workspace, workspace_bytes, near_per_element, stream);
}

extern "C" int grim_workspace_status(void* workspace, void* stream) {
    if (!workspace) return GRIM_ERR_ARGS;
    unsigned h = 0;
    hipStream_t st = (hipStream_t)stream;
    CHECK(hipMemcpyAsync(&h, workspace, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    CHECK(hipStreamSynchronize(st));
    if (h & GRIM_DEV_NEAR_CAPACITY) return GRIM_ERR_CAPACITY;
    if (h & (GRIM_DEV_ADAPTIVE_DEPTH | GRIM_DEV_SELF_SERIES)) return GRIM_ERR_NOT_CONVERGED;
    if (h & GRIM_DEV_ZERO_PIVOT) return GRIM_ERR_SINGULAR;
    return GRIM_OK;
}

extern "C" int grim_loads(int64_t n, int64_t na, const double* p0, const double* p1, const int32_t* prev, double k,
                          const double* angles_deg, double* bu, double* bdn, int64_t ld, void* stream) {
    if (n < 1 || na < 0 || !p0 || !p1 || !prev || !angles_deg || (!bu && !bdn) || ld < n || !(k > 0) ||
        na > 65535)
        return GRIM_ERR_ARGS;
    if (na == 0) return GRIM_OK;
    CHECK(grim_launch_loads((int)n, (int)na, p0, p1, prev, k, angles_deg, bu, bdn, ld, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_farfield(int64_t n, int64_t na, const double* x, int64_t ldx, const double* bu,
                             const double* bdn, int64_t ldb, double k, double* amp, double* sigma, void* stream) {
    if (n < 1 || na < 0 || !x || !bu || !bdn || !amp || ldx < n || ldb < n || !(k > 0)) return GRIM_ERR_ARGS;
    if (na == 0) return GRIM_OK;
    CHECK(grim_launch_farfield((int)n, (int)na, x, ldx, bu, bdn, ldb, k, amp, sigma, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_lu_factor(int64_t n, void* a, int64_t lda, int precision, int32_t* piv, void* workspace,
                              void* stream) {
    if (n < 1 || n > 32768 || !a || !piv || lda < n || (precision != 64 && precision != 32)) return GRIM_ERR_ARGS;
    hipStream_t st = (hipStream_t)stream;
    unsigned* status = (unsigned*)workspace;
    if (status) {
        zero_words<<<1, 64, 0, st>>>(status, 1);
        CHECK(hipGetLastError());
    } else {
        return GRIM_ERR_ARGS;
    }
    CHECK(grim_dense_lu_factor((int)n, a, lda, precision == 32, piv, status, st));
    return GRIM_OK;
}

extern "C" int grim_lu_solve(int64_t n, int64_t nrhs, const void* lu, int64_t ldlu, int precision,
                             const int32_t* piv, void* b, int64_t ldb, void* stream) {
    if (n < 1 || nrhs < 0 || !lu || !piv || !b || ldlu < n || ldb < n || (precision != 64 && precision != 32))
        return GRIM_ERR_ARGS;
    if (nrhs == 0) return GRIM_OK;
    CHECK(grim_dense_lu_solve((int)n, (int)nrhs, lu, ldlu, precision == 32, piv, b, ldb, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_gemm_minus(int64_t m, int64_t n, int64_t k, const void* a, int64_t lda, const void* b,
                               int64_t ldb, void* c, int64_t ldc, int precision, void* stream) {
    if (m < 0 || n < 0 || k < 0 || !a || !b || !c || lda < m || ldb < k || ldc < m ||
        (precision != 64 && precision != 32) || n > 65535 * 64)
        return GRIM_ERR_ARGS;
    CHECK(grim_dense_gemm_minus((int)m, (int)n, (int)k, a, lda, b, ldb, c, ldc, precision == 32,
                                (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int64_t grim_refine_workspace_bytes(int64_t n, int64_t nrhs) {
    if (n < 1 || nrhs < 1) return -1;
    return align_up(16 * n * nrhs) + align_up(8 * n * nrhs) + 2 * align_up(8 * nrhs);
}

extern "C" int grim_refined_solve(int64_t n, int64_t nrhs, const double* z, int64_t ldz, const float* lu,
                                  int64_t ldlu, const int32_t* piv, const double* b, int64_t ldb, double* x,
                                  int64_t ldx, int max_steps, double tol, void* workspace, int64_t workspace_bytes,
                                  int* steps_out, double* error_out, void* stream) {
    if (n < 1 || nrhs < 1 || !z || !lu || !piv || !b || !x || ldz < n || ldlu < n || ldb < n || ldx < n ||
        max_steps < 0 || !(tol > 0) || !workspace || nrhs > 65535)
        return GRIM_ERR_ARGS;
    if (grim_refine_workspace_bytes(n, nrhs) > workspace_bytes) return GRIM_ERR_RANGE;
    hipStream_t st = (hipStream_t)stream;
    char* p = (char*)workspace;
    double* r = (double*)p; p += align_up(16 * n * nrhs);
    float* d = (float*)p; p += align_up(8 * n * nrhs);
    double* bmax = (double*)p; p += align_up(8 * nrhs);
    double* rmax = (double*)p;
    const int N = (int)n, R = (int)nrhs;
    std::vector<double> hb(R), hr(R);
    CHECK(grim_dense_colmax(N, R, b, ldb, bmax, st));
    CHECK(hipMemcpyAsync(hb.data(), bmax, 8 * R, hipMemcpyDeviceToHost, st));
    // x = LU32^-1 b
    CHECK(grim_dense_to_c64(N, R, b, ldb, d, n, st));
    CHECK(grim_dense_lu_solve(N, R, lu, ldlu, 1, piv, d, n, st));
    zero_c128<<<dim3((N + 255) / 256, R), 256, 0, st>>>(N, R, x, ldx);
    CHECK(hipGetLastError());
    CHECK(grim_dense_add_c64(N, R, d, n, x, ldx, st));
    CHECK(hipStreamSynchronize(st));
    for (int c = 0; c < R; ++c) hb[c] = hb[c] > 0 ? hb[c] : 1.0;
    double previous = INFINITY;
    for (int step = 0; step < max_steps; ++step) {
        // r = b - Z x (FP64)
        CHECK(grim_dense_copy(N, R, b, ldb, r, n, st));
        CHECK(grim_dense_gemm_minus(N, R, N, z, ldz, x, ldx, r, n, 0, st));
        CHECK(grim_dense_colmax(N, R, r, n, rmax, st));
        CHECK(hipMemcpyAsync(hr.data(), rmax, 8 * R, hipMemcpyDeviceToHost, st));
        CHECK(hipStreamSynchronize(st));
        double err = 0.0;
        for (int c = 0; c < R; ++c) err = fmax(err, hr[c] / hb[c]);
        if (error_out) *error_out = err;
        if (steps_out) *steps_out = step;
        if (isfinite(err) && err <= tol) return GRIM_OK;
        if (!isfinite(err) || err >= 0.98 * previous) return GRIM_ERR_NOT_CONVERGED;
        previous = err;
        CHECK(grim_dense_to_c64(N, R, r, n, d, n, st));
        CHECK(grim_dense_lu_solve(N, R, lu, ldlu, 1, piv, d, n, st));
        CHECK(grim_dense_add_c64(N, R, d, n, x, ldx, st));
    }
    return GRIM_ERR_NOT_CONVERGED;
}

extern "C" int grim_to_c64(int64_t m, int64_t n, const double* src, int64_t lds, float* dst, int64_t ldd,
                           void* stream) {
    if (m < 0 || n < 0 || !src || !dst || lds < m || ldd < m || n > 65535) return GRIM_ERR_ARGS;
    CHECK(grim_dense_to_c64((int)m, (int)n, src, lds, dst, ldd, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_copy_c128(int64_t m, int64_t n, const double* src, int64_t lds, double* dst, int64_t ldd,
                              void* stream) {
    if (m < 0 || n < 0 || !src || !dst || lds < m || ldd < m || n > 65535) return GRIM_ERR_ARGS;
    CHECK(grim_dense_copy((int)m, (int)n, src, lds, dst, ldd, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_column_norms(int64_t m, int64_t n, const double* v, int64_t ld, double* out, void* stream) {
    if (m < 1 || n < 0 || !v || !out || ld < m) return GRIM_ERR_ARGS;
    CHECK(grim_dense_colnorm((int)m, (int)n, v, ld, out, (hipStream_t)stream));
    return GRIM_OK;
}
