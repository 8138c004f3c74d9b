// C ABI of the GRIM-2D HIP library (declarations in include/grim2d.h).
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/grim2d.h"
#include "common.h"

extern "C" hipError_t grim_upload_rules(const double*, const double*, const double*, const double*, const double*,
                                        const double*, const double*);
extern "C" hipError_t grim_launch_far(int q, int n, const double* p0, const double* p1, const int* next,
                                      const int* prev, double k, double* ztm, int64_t ldtm, double* zte,
                                      int64_t ldte, const int* fix_rows, int nfix, int mode, hipStream_t st);
extern "C" hipError_t grim_launch_near(int n, const double* p0, const double* p1, const int* next, const int* prev,
                                       double k, int* counts, int* offsets, int capacity, int* pair_e, int* pair_f,
                                       double* blocks, int* kind, unsigned* status, double* ztm, int64_t ldtm,
                                       double* zte, int64_t ldte, int mode, double mass, const double* mass_coef,
                                       hipStream_t st);
extern "C" hipError_t grim_launch_near_list(int n, const double* p0, const double* p1, const int* next, int* counts,
                                            int* offsets, int capacity, int* pair_e, int* pair_f, unsigned* status,
                                            hipStream_t st);
extern "C" hipError_t grim_launch_poly(int degree, int n, int q, const double* p0, const double* p1, const int* next,
                                       const int* prev, double k, const int* offsets, int capacity,
                                       const int* pair_e, const int* pair_f, double* bs, double* bk, double* bte,
                                       unsigned* status, double* ztm, int64_t ldtm, double* zte, int64_t ldte,
                                       hipStream_t st);
extern "C" hipError_t grim_launch_poly_loads(int degree, int n, int na, const double* p0, const double* p1,
                                             const int* prev, double k, const double* angles_deg, int order,
                                             const double* qt, const double* qw, double* bu, double* bdn, int64_t ld,
                                             hipStream_t st);
extern "C" hipError_t grim_launch_loads(int n, int na, const double* p0, const double* p1, const int* prev, double k,
                                        const double* angles_deg, double* bu, double* bdn, int64_t ld,
                                        hipStream_t st);
extern "C" hipError_t grim_launch_farfield(int n, int na, const double* x, int64_t ldx, const double* bu,
                                           const double* bdn, int64_t ldb, double k, double* amp, double* sigma,
                                           hipStream_t st);
extern "C" hipError_t grim_launch_robin_rows(int64_t rows, int64_t cols, double* x, int64_t ldx, double* y,
                                             int64_t ldy, const double* ax, const double* ay, const int* dirichlet,
                                             hipStream_t st);
extern "C" hipError_t grim_launch_robin_split(int q, int n, const double* p0, const double* p1, const int* next,
                                              const int* prev, double k, int nblocks, const int* walk,
                                              const int* offsets, const int* pair_f, int capacity,
                                              const double* blocks, const double* c_te, const double* c_tm,
                                              const int* flags, double* zte, int64_t ldte, double* ztm, int64_t ldtm,
                                              hipStream_t st);
extern "C" hipError_t grim_launch_robin_rhs(int n, int na, const double* p0, const double* p1, const int* prev,
                                            double k, const double* angles_deg, const double* alpha_e,
                                            const int* dirichlet, double* out, int64_t ld, hipStream_t st);
extern "C" hipError_t grim_launch_route(int m, int n, const double* src, int64_t ld_src, const int* rows_src,
                                       const int* cols_src, double* dst, int64_t ld_dst, const int* rows_dst,
                                       const int* cols_dst, const double* w, hipStream_t st);
extern "C" hipError_t grim_launch_scale_blocks(int64_t n, double* z, int64_t ld, const double* s, hipStream_t st);
extern "C" hipError_t grim_launch_near_pairs(int n, const double* p0, const double* p1, const int* next, double k,
                                             int* counts, int* offsets, int capacity, int* pair_e, int* pair_f,
                                             double* blocks, int* kind, unsigned* status, hipStream_t st);
extern "C" hipError_t grim_launch_sheet_far(int q, int ne, const double* p0, const double* p1, const int* nid,
                                            double k, double* bs, hipStream_t st);
extern "C" hipError_t grim_launch_sheet_near_store(int ne, const int* offsets, int capacity, const int* pair_e,
                                                   const int* pair_f, const double* blocks, double* bs,
                                                   hipStream_t st);
extern "C" hipError_t grim_launch_sheet_gather(int hyper, int nn, int ne, const double* p0, const double* p1,
                                               const int* own_ptr, const int* own_el, const int* own_loc,
                                               const double* bs, double k, const double* coef, const int* pinned,
                                               double* z, int64_t ld, hipStream_t st);
extern "C" hipError_t grim_launch_sheet_loads(int nn, int na, const double* p0, const double* p1, const int* own_ptr,
                                              const int* own_el, const int* own_loc, double k,
                                              const double* angles_deg, double* bu, double* bdn, int64_t ld,
                                              hipStream_t st);
extern "C" hipError_t grim_launch_lossy(int q, int n, const double* p0, const double* p1, const int* next,
                                        const int* prev, double kr, double ki, const int* offsets, int capacity,
                                        const int* pair_e, const int* pair_f, double* blocks, int* kind,
                                        unsigned* status, double* ts, double* tk, double mass_k, double mass_d,
                                        const double* coef_k, const double* coef_d, double* s_out, double* k_out,
                                        double* d_out, double* w_out, int64_t ld, hipStream_t st);
extern "C" hipError_t grim_launch_hankel2_probe(int m, const double* z, double* out, hipStream_t st);
extern "C" int64_t grim_dense_aux_bytes(int n, int single);
extern "C" hipError_t grim_dense_lu_factor(int batch, int n, void* a, int64_t lda, int64_t sa, int single, int* piv,
                                           int64_t spiv, void* aux, int64_t saux, unsigned* status, hipStream_t st);
extern "C" hipError_t grim_dense_lu_solve(int batch, int n, int nrhs, const void* a, int64_t lda, int64_t sa,
                                          int single, const void* aux, int64_t saux, void* b, int64_t ldb,
                                          int64_t sb, void* w, int64_t sw, hipStream_t st);
extern "C" hipError_t grim_dense_gemm_minus(int batch, int m, int n, int k, const void* a, int64_t lda, int64_t sa,
                                            const void* b, int64_t ldb, int64_t sb, void* c, int64_t ldc,
                                            int64_t sc, int single, hipStream_t st);
extern "C" hipError_t grim_dense_to_c64(int batch, int m, int n, const double* src, int64_t lds, int64_t ss,
                                        float* dst, int64_t ldd, int64_t sd, hipStream_t st);
extern "C" hipError_t grim_dense_copy(int batch, int m, int n, const double* src, int64_t lds, int64_t ss,
                                      double* dst, int64_t ldd, int64_t sd, hipStream_t st);
extern "C" hipError_t grim_dense_add_c64(int batch, int m, int n, const float* d, int64_t ldd, int64_t sd, double* x,
                                        int64_t ldx, int64_t sx, hipStream_t st);
extern "C" hipError_t grim_dense_colmax(int batch, int m, int n, const double* v, int64_t ld, int64_t sv,
                                        double* out, hipStream_t st);
extern "C" hipError_t grim_dense_colnorm(int m, int n, const double* v, int64_t ld, double* out, hipStream_t st);
extern "C" int64_t grim_dense_trans_block_elems(void);
extern "C" hipError_t grim_dense_lu_solve_trans(int batch, int n, int nrhs, const void* a, int64_t lda, int64_t sa,
                                                int single, const void* aux, int64_t saux, void* b, int64_t ldb,
                                                int64_t sb, void* w, int64_t sw, void* blk, hipStream_t st);
extern "C" hipError_t grim_dense_gemm_minus_ta(int batch, int m, int n, int k, const void* a, int64_t lda,
                                               int64_t sa, const void* b, int64_t ldb, int64_t sb, void* c,
                                               int64_t ldc, int64_t sc, int single, hipStream_t st);
extern "C" hipError_t grim_dense_equilibration(int batch, int n, const double* a, int64_t lda, int64_t sa,
                                               double* row_scale, double* col_scale, double* norm, hipStream_t st);
extern "C" int64_t grim_dense_norm_inf_work(int n);
extern "C" hipError_t grim_dense_norm_inf(int batch, int n, const double* a, int64_t lda, int64_t sa, double* work,
                                          double* out, hipStream_t st);

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

__global__ void zero_c128(int m, int n, double* x, int64_t ld, int64_t stride) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r >= m) return;
    double* t = x + 2 * (stride * blockIdx.z + (int64_t)r + (int64_t)c * ld);
    t[0] = 0.0;
    t[1] = 0.0;
}

extern "C" int grim_version(void) { return 100; }

// Launch-shape table: per knob, entries (min_size, value) sorted by min_size.
namespace {
constexpr int kShapeEntries = 16;
struct ShapeTable {
    int count = 0;
    int64_t min_size[kShapeEntries];
    int value[kShapeEntries];
};
ShapeTable g_shapes[GRIM_SHAPE_KNOBS];
std::mutex g_shape_lock;

bool shape_allowed(int knob, int value) {
    if (knob == GRIM_SHAPE_FAR_COLUMNS) return value == 1 || value == 2 || value == 4 || value == 8 || value == 16 || value == 32;
    if (knob == GRIM_SHAPE_TRSM_COLUMNS || knob == GRIM_SHAPE_SOLVE_COLUMNS) return value == 1 || value == 2 || value == 4;
    return false;
}
}  // namespace

extern "C" int grim_set_launch_shape(int knob, int64_t min_size, int value) {
    if (knob < 0 || knob >= GRIM_SHAPE_KNOBS || min_size < 0) return GRIM_ERR_ARGS;
    if (value != 0 && !shape_allowed(knob, value)) return GRIM_ERR_ARGS;
    std::lock_guard<std::mutex> guard(g_shape_lock);
    ShapeTable& t = g_shapes[knob];
    int pos = 0;
    while (pos < t.count && t.min_size[pos] < min_size) ++pos;
    const bool exists = pos < t.count && t.min_size[pos] == min_size;
    if (value == 0) {   // remove the entry
        if (exists) {
            for (int i = pos; i + 1 < t.count; ++i) {
                t.min_size[i] = t.min_size[i + 1];
                t.value[i] = t.value[i + 1];
            }
            --t.count;
        }
        return GRIM_OK;
    }
    if (!exists) {
        if (t.count == kShapeEntries) return GRIM_ERR_RANGE;
        for (int i = t.count; i > pos; --i) {
            t.min_size[i] = t.min_size[i - 1];
            t.value[i] = t.value[i - 1];
        }
        ++t.count;
        t.min_size[pos] = min_size;
    }
    t.value[pos] = value;
    return GRIM_OK;
}

extern "C" int grim_clear_launch_shapes(void) {
    std::lock_guard<std::mutex> guard(g_shape_lock);
    for (auto& t : g_shapes) t.count = 0;
    return GRIM_OK;
}

extern "C" int grim_shape_value(int knob, int64_t size) {
    if (knob < 0 || knob >= GRIM_SHAPE_KNOBS) return 0;
    std::lock_guard<std::mutex> guard(g_shape_lock);
    const ShapeTable& t = g_shapes[knob];
    int v = 0;
    for (int i = 0; i < t.count && t.min_size[i] <= size; ++i) v = t.value[i];
    return v;
}

extern "C" int grim_launch_shape(int knob, int64_t size) { return grim_shape_value(knob, size); }

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

// ---------------------------------------------------------------------------
// Static scratch of grim_assemble_noalloc: the near-pair lists and blocks that grim_assemble
// takes from its workspace argument.  Each device the library runs on holds its own copy for the
// life of the process.  Host code never uses the array's name as an address: grim_init (or the
// entry point's first call on a device) asks the runtime for the device's address once, with
// that device current, and keeps it per device.
// ---------------------------------------------------------------------------
#define GRIM_NOALLOC_SCRATCH_BYTES (64ll << 20)
static __device__ __align__(256) unsigned char scratch[GRIM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices];
std::mutex g_scratch_lock;

// The current device's address of `scratch` (looked up once per device), or null.
void* device_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[device] = ptr;
    }
    return g_scratch[device];
}

// The near-pass arrays of carve() without the status word, which the caller owns here.
AssemblyWork carve_scratch(void* base, int64_t n, int64_t cap, unsigned* status) {
    AssemblyWork w;
    char* p = (char*)base;
    int64_t off = 0;
    w.status = status;
    w.counts = (int*)(p + off); off += align_up(4 * n);
    w.offsets = (int*)(p + off); off += align_up(4 * (n + 1));
    w.pair_e = (int*)(p + off); off += align_up(4 * cap);
    w.pair_f = (int*)(p + off); off += align_up(4 * cap);
    w.kind = (int*)(p + off); off += align_up(4 * cap);
    w.blocks = (double*)(p + off); off += align_up(16 * 8 * cap);
    w.bytes = off;
    return w;
}
}  // namespace

// Gauss-Legendre rules (numpy.polynomial.legendre.leggauss mapped to [0, 1],
// the reference's _quadrature_nodes, twod/operators.py:3319-3323) and the
// 20-point t^4 graded corner rule (twod/polynomial_quadrature.py:319-335).
extern "C" int grim_init(int device) {
    if (hipSetDevice(device) != hipSuccess) return GRIM_ERR_HIP;
    if (!device_scratch(device)) return GRIM_ERR_HIP;
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

extern "C" int64_t grim_assemble_workspace_bytes(int64_t n, int64_t near_per_element) {
    if (n <= 0 || near_per_element <= 0) return -1;
    return carve(nullptr, n, n * near_per_element).bytes;
}


static bool assemble_args_ok(int64_t n, const double* p0, const double* p1, const int32_t* next,
                             const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k, int far_order,
                             double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te, int64_t near_per_element) {
    return !(n < 2 || n > (1 << 30) || !p0 || !p1 || !next || !prev || !z_tm || !z_te || ld_tm < n || ld_te < n ||
             !(k > 0) || !isfinite(k) || near_per_element <= 0 || nfix < 0 || (nfix > 0 && !fix_rows) ||
             (far_order != 8 && far_order != 16));
}

// The far and near passes with the near-pair lists and blocks in `w` (status word included).
static int assemble_launch(int passes, int64_t n, const double* p0, const double* p1, const int32_t* next,
                           const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k, int far_order,
                           double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te, const AssemblyWork& w,
                           int64_t cap, hipStream_t st, int mode = 0, double mass = 0.0,
                           const double* mass_coef = nullptr) {
    if (passes & 1) {
        zero_words<<<1, 64, 0, st>>>(w.status, 1);
        CHECK(hipGetLastError());
        CHECK(grim_launch_far(far_order, (int)n, p0, p1, next, prev, k, z_tm, ld_tm, z_te, ld_te, fix_rows,
                              (int)nfix, mode, st));
    }
    if (passes & 2)
        CHECK(grim_launch_near((int)n, p0, p1, next, prev, k, w.counts, w.offsets, (int)cap, w.pair_e, w.pair_f,
                               w.blocks, w.kind, w.status, z_tm, ld_tm, z_te, ld_te, mode, mass, mass_coef, st));
    return GRIM_OK;
}

static int assemble_pass(int passes, int64_t n, const double* p0, const double* p1, const int32_t* next,
                         const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k, int far_order,
                         double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te, void* workspace,
                         int64_t workspace_bytes, int64_t near_per_element, void* stream) {
    if (!workspace || !assemble_args_ok(n, p0, p1, next, prev, fix_rows, nfix, k, far_order, z_tm, ld_tm, z_te,
                                        ld_te, near_per_element))
        return GRIM_ERR_ARGS;
    const int64_t cap = n * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    AssemblyWork w = carve(workspace, n, cap);
    if (w.bytes > workspace_bytes) return GRIM_ERR_RANGE;
    return assemble_launch(passes, n, p0, p1, next, prev, fix_rows, nfix, k, far_order, z_tm, ld_tm, z_te, ld_te, w,
                           cap, (hipStream_t)stream);
}

extern "C" int grim_assemble_operators(int64_t n, const double* p0, const double* p1, const int32_t* next,
                                       const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k,
                                       int far_order, int which, double* a, int64_t ld_a, double* b, int64_t ld_b,
                                       double mass, const double* mass_coef, void* workspace,
                                       int64_t workspace_bytes, int64_t near_per_element, void* stream) {
    if (!workspace || !isfinite(mass) || (which != 1 && which != 2) ||
        !assemble_args_ok(n, p0, p1, next, prev, fix_rows, nfix, k, far_order, a, ld_a, b, ld_b, near_per_element))
        return GRIM_ERR_ARGS;
    const int64_t cap = n * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    AssemblyWork w = carve(workspace, n, cap);
    if (w.bytes > workspace_bytes) return GRIM_ERR_RANGE;
    return assemble_launch(3, n, p0, p1, next, prev, fix_rows, nfix, k, far_order, a, ld_a, b, ld_b, w, cap,
                           (hipStream_t)stream, which, mass, mass_coef);
}

extern "C" int grim_route(int64_t m, int64_t n, const double* src, int64_t ld_src, const int32_t* rows_src,
                          const int32_t* cols_src, double* dst, int64_t ld_dst, const int32_t* rows_dst,
                          const int32_t* cols_dst, const double* weight, void* stream) {
    if (m < 0 || n < 0 || m > (1 << 30) || n > 65535 || ld_src < 1 || ld_dst < 1) return GRIM_ERR_ARGS;
    if (m == 0 || n == 0) return GRIM_OK;
    if (!src || !rows_src || !cols_src || !dst || !rows_dst || !cols_dst || !weight) return GRIM_ERR_ARGS;
    CHECK(grim_launch_route((int)m, (int)n, src, ld_src, rows_src, cols_src, dst, ld_dst, rows_dst, cols_dst, weight,
                            (hipStream_t)stream));
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

extern "C" int grim_assemble_noalloc(int64_t n, const double* p0, const double* p1, const int32_t* next,
                                     const int32_t* prev, const int32_t* fix_rows, int64_t nfix, double k,
                                     int far_order, double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te,
                                     int64_t near_per_element, void* status, void* stream) {
    if (!status || !assemble_args_ok(n, p0, p1, next, prev, fix_rows, nfix, k, far_order, z_tm, ld_tm, z_te, ld_te,
                                     near_per_element))
        return GRIM_ERR_ARGS;
    const int64_t cap = n * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    if (carve_scratch(nullptr, n, cap, nullptr).bytes > GRIM_NOALLOC_SCRATCH_BYTES) return GRIM_ERR_RANGE;
    int device = -1;
    CHECK(hipGetDevice(&device));
    void* base = device_scratch(device);
    if (!base) return GRIM_ERR_HIP;
    const AssemblyWork w = carve_scratch(base, n, cap, (unsigned*)status);
    return assemble_launch(3, n, p0, p1, next, prev, fix_rows, nfix, k, far_order, z_tm, ld_tm, z_te, ld_te, w, cap,
                           (hipStream_t)stream);
}

extern "C" int grim_workspace_status(void* workspace, void* stream) {
    if (!workspace) return GRIM_ERR_ARGS;
    unsigned h = 0;
    hipStream_t st = (hipStream_t)stream;
    CHECK(hipMemcpyAsync(&h, workspace, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    CHECK(hipStreamSynchronize(st));
    if (h & GRIM_DEV_NEAR_CAPACITY) return GRIM_ERR_CAPACITY;
    if (h & GRIM_DEV_POLY_INTERSECT) return GRIM_ERR_GEOMETRY;
    if (h & (GRIM_DEV_ADAPTIVE_DEPTH | GRIM_DEV_SELF_SERIES | GRIM_DEV_POLY_NEAR)) return GRIM_ERR_NOT_CONVERGED;
    if (h & GRIM_DEV_BARRIER_TIMEOUT) return GRIM_ERR_TIMEOUT;
    if (h & GRIM_DEV_ZERO_PIVOT) return GRIM_ERR_SINGULAR;
    return GRIM_OK;
}

// ---------------------------------------------------------------------------
// Degree-2 and degree-3 Galerkin bases (poly.hip)
// ---------------------------------------------------------------------------
struct PolyWork {
    unsigned* status;
    int *counts, *offsets, *pair_e, *pair_f;
    double *bs, *bk, *bte;
    int64_t bytes;
};

static PolyWork carve_poly(void* base, int64_t n, int64_t cap, int degree) {
    PolyWork w;
    char* p = (char*)base;
    int64_t off = 0;
    const int64_t table = align_up(16 * (int64_t)(degree + 1) * (degree + 1) * n * n);
    w.status = (unsigned*)(p + off); off += kAlign;
    w.counts = (int*)(p + off); off += align_up(4 * n);
    w.offsets = (int*)(p + off); off += align_up(4 * (n + 1));
    w.pair_e = (int*)(p + off); off += align_up(4 * cap);
    w.pair_f = (int*)(p + off); off += align_up(4 * cap);
    w.bs = (double*)(p + off); off += table;
    w.bk = (double*)(p + off); off += table;
    w.bte = (double*)(p + off); off += table;
    w.bytes = off;
    return w;
}

static bool poly_size_ok(int64_t n, int degree) {
    return (degree == 2 || degree == 3) && n >= 2 && n * degree <= 65535;
}

extern "C" int64_t grim_assemble_poly_workspace_bytes(int64_t n, int degree, int64_t near_per_element) {
    if (!poly_size_ok(n, degree) || near_per_element <= 0) return -1;
    return carve_poly(nullptr, n, n * near_per_element, degree).bytes;
}

extern "C" int grim_assemble_poly(int64_t n, int degree, const double* p0, const double* p1, const int32_t* next,
                                  const int32_t* prev, double k, int far_order, double* z_tm, int64_t ld_tm,
                                  double* z_te, int64_t ld_te, void* workspace, int64_t workspace_bytes,
                                  int64_t near_per_element, void* stream) {
    if (!poly_size_ok(n, degree) || !p0 || !p1 || !next || !prev || !z_tm || !z_te || !workspace ||
        ld_tm < n * degree || ld_te < n * degree || !(k > 0) || !isfinite(k) || near_per_element <= 0 ||
        (far_order != 9 && far_order != 16))
        return GRIM_ERR_ARGS;
    const int64_t cap = n * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    const PolyWork w = carve_poly(workspace, n, cap, degree);
    if (w.bytes > workspace_bytes) return GRIM_ERR_RANGE;
    hipStream_t st = (hipStream_t)stream;
    zero_words<<<1, 64, 0, st>>>(w.status, 1);
    CHECK(hipGetLastError());
    CHECK(grim_launch_near_list((int)n, p0, p1, next, w.counts, w.offsets, (int)cap, w.pair_e, w.pair_f, w.status,
                                st));
    CHECK(grim_launch_poly(degree, (int)n, far_order, p0, p1, next, prev, k, w.offsets, (int)cap, w.pair_e,
                           w.pair_f, w.bs, w.bk, w.bte, w.status, z_tm, ld_tm, z_te, ld_te, st));
    return GRIM_OK;
}

extern "C" int grim_loads_poly(int64_t n, int degree, int64_t na, const double* p0, const double* p1,
                               const int32_t* prev, double k, const double* angles_deg, int64_t order,
                               const double* rule_t, const double* rule_w, double* bu, double* bdn, int64_t ld,
                               void* stream) {
    if (!poly_size_ok(n, degree) || na < 0 || na > 65535 || !p0 || !p1 || !prev || !angles_deg || (!bu && !bdn) ||
        ld < n * degree || !(k > 0) || !isfinite(k) || order < 1 || order > 256 || !rule_t || !rule_w)
        return GRIM_ERR_ARGS;
    if (na == 0) return GRIM_OK;
    CHECK(grim_launch_poly_loads(degree, (int)n, (int)na, p0, p1, prev, k, angles_deg, (int)order, rule_t, rule_w,
                                 bu, bdn, ld, (hipStream_t)stream));
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

extern "C" int grim_farfield_potential(int64_t n, int64_t na, const double* x, int64_t ldx, const double* b,
                                       int64_t ldb, double k, double* amp, double* sigma, void* stream) {
    if (n < 1 || na < 0 || !x || !b || !amp || ldx < n || ldb < n || !(k > 0) || !isfinite(k)) return GRIM_ERR_ARGS;
    if (na == 0) return GRIM_OK;
    CHECK(grim_launch_farfield((int)n, (int)na, x, ldx, b, nullptr, ldb, k, amp, sigma, (hipStream_t)stream));
    return GRIM_OK;
}

// ---------------------------------------------------------------------------
// Sheet (TYPE 1) and sheet + PEC systems (sheet.hip)
// ---------------------------------------------------------------------------
struct SheetWork {
    AssemblyWork near;   // status word, near-pair lists and linear near blocks
    double* bs;          // element-pair S table, ne x ne blocks of 2 x 2 complex
    int64_t bytes;
};

static SheetWork carve_sheet(void* base, int64_t ne, int64_t cap) {
    SheetWork w;
    w.near = carve(base, ne, cap);
    w.bs = (double*)((char*)base + w.near.bytes);
    w.bytes = w.near.bytes + align_up(64 * ne * ne);
    return w;
}

static bool sheet_size_ok(int64_t ne, int64_t nn) {
    return ne >= 1 && ne <= 65535 && nn >= 2 && nn <= 65535;
}

extern "C" int64_t grim_sheet_workspace_bytes(int64_t ne, int64_t near_per_element) {
    if (ne < 1 || ne > 65535 || near_per_element <= 0) return -1;
    return carve_sheet(nullptr, ne, ne * near_per_element).bytes;
}

extern "C" int grim_assemble_sheet(int64_t ne, int64_t nn, const double* p0, const double* p1,
                                   const int32_t* next_ids, const int32_t* own_ptr, const int32_t* own_el,
                                   const int32_t* own_loc, double k, int far_order_tm, int far_order_te,
                                   const double* coef_tm, const double* coef_te, const int32_t* pinned,
                                   double* z_tm, int64_t ld_tm, double* z_te, int64_t ld_te, void* workspace,
                                   int64_t workspace_bytes, int64_t near_per_element, void* stream) {
    if (!sheet_size_ok(ne, nn) || !p0 || !p1 || !next_ids || !own_ptr || !own_el || !own_loc || !workspace ||
        (!z_tm && !z_te) || (z_tm && (!coef_tm || ld_tm < nn)) || (z_te && (!coef_te || ld_te < nn)) ||
        !(k > 0) || !isfinite(k) || near_per_element <= 0 ||
        (z_tm && far_order_tm != 8 && far_order_tm != 16) || (z_te && far_order_te != 8 && far_order_te != 16))
        return GRIM_ERR_ARGS;
    const int64_t cap = ne * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    const SheetWork w = carve_sheet(workspace, ne, cap);
    if (w.bytes > workspace_bytes) return GRIM_ERR_RANGE;
    hipStream_t st = (hipStream_t)stream;
    zero_words<<<1, 64, 0, st>>>(w.near.status, 1);
    CHECK(hipGetLastError());
    CHECK(grim_launch_near_pairs((int)ne, p0, p1, next_ids, k, w.near.counts, w.near.offsets, (int)cap,
                                 w.near.pair_e, w.near.pair_f, w.near.blocks, w.near.kind, w.near.status, st));
    const int q0 = z_tm ? far_order_tm : far_order_te;
    CHECK(grim_launch_sheet_far(q0, (int)ne, p0, p1, next_ids, k, w.bs, st));
    CHECK(grim_launch_sheet_near_store((int)ne, w.near.offsets, (int)cap, w.near.pair_e, w.near.pair_f,
                                       w.near.blocks, w.bs, st));
    if (z_tm)
        CHECK(grim_launch_sheet_gather(0, (int)nn, (int)ne, p0, p1, own_ptr, own_el, own_loc, w.bs, k, coef_tm,
                                       nullptr, z_tm, ld_tm, st));
    if (z_te) {
        // the far pass rewrites only far blocks: the near blocks stored above stay
        if (far_order_te != q0) CHECK(grim_launch_sheet_far(far_order_te, (int)ne, p0, p1, next_ids, k, w.bs, st));
        CHECK(grim_launch_sheet_gather(1, (int)nn, (int)ne, p0, p1, own_ptr, own_el, own_loc, w.bs, k, coef_te,
                                       pinned, z_te, ld_te, st));
    }
    return GRIM_OK;
}

extern "C" int grim_loads_sheet(int64_t ne, int64_t nn, int64_t na, const double* p0, const double* p1,
                                const int32_t* own_ptr, const int32_t* own_el, const int32_t* own_loc, double k,
                                const double* angles_deg, double* bu, double* bdn, int64_t ld, void* stream) {
    if (!sheet_size_ok(ne, nn) || na < 0 || na > 65535 || !p0 || !p1 || !own_ptr || !own_el || !own_loc ||
        !angles_deg || (!bu && !bdn) || ld < nn || !(k > 0) || !isfinite(k))
        return GRIM_ERR_ARGS;
    if (na == 0) return GRIM_OK;
    CHECK(grim_launch_sheet_loads((int)nn, (int)na, p0, p1, own_ptr, own_el, own_loc, k, angles_deg, bu, bdn, ld,
                                  (hipStream_t)stream));
    return GRIM_OK;
}

// ---------------------------------------------------------------------------
// Lossy media: operators at a complex wavenumber (lossy.hip)
// ---------------------------------------------------------------------------
struct LossyWork {
    AssemblyWork near;   // status word, near-pair lists and near blocks
    double *ts, *tk;     // element-pair S and K' tables
    int64_t bytes;
};

static LossyWork carve_lossy(void* base, int64_t n, int64_t cap) {
    LossyWork w;
    w.near = carve(base, n, cap);
    const int64_t table = align_up(64 * n * n);
    w.ts = (double*)((char*)base + w.near.bytes);
    w.tk = (double*)((char*)base + w.near.bytes + table);
    w.bytes = w.near.bytes + 2 * table;
    return w;
}

extern "C" int64_t grim_assemble_operators_lossy_workspace_bytes(int64_t n, int64_t near_per_element) {
    if (n < 2 || n > 65535 || near_per_element <= 0) return -1;
    return carve_lossy(nullptr, n, n * near_per_element).bytes;
}

extern "C" int grim_assemble_operators_lossy(int64_t n, const double* p0, const double* p1, const int32_t* next,
                                             const int32_t* prev, double k_re, double k_im, int far_order,
                                             double* s, double* kp, double* d, double* w, int64_t ld,
                                             double mass_k, double mass_d, const double* coef_k,
                                             const double* coef_d, void* workspace, int64_t workspace_bytes,
                                             int64_t near_per_element, void* stream) {
    if (n < 2 || n > 65535 || !p0 || !p1 || !next || !prev || !workspace || (!s && !kp && !d && !w) || ld < n ||
        !(k_re > 0) || !isfinite(k_re) || !isfinite(k_im) || k_im > 0 || near_per_element <= 0 ||
        !isfinite(mass_k) || !isfinite(mass_d) || (far_order != 8 && far_order != 16))
        return GRIM_ERR_ARGS;
    const int64_t cap = n * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    const LossyWork wk = carve_lossy(workspace, n, cap);
    if (wk.bytes > workspace_bytes) return GRIM_ERR_RANGE;
    hipStream_t st = (hipStream_t)stream;
    zero_words<<<1, 64, 0, st>>>(wk.near.status, 1);
    CHECK(hipGetLastError());
    CHECK(grim_launch_near_list((int)n, p0, p1, next, wk.near.counts, wk.near.offsets, (int)cap, wk.near.pair_e,
                                wk.near.pair_f, wk.near.status, st));
    CHECK(grim_launch_lossy(far_order, (int)n, p0, p1, next, prev, k_re, k_im, wk.near.offsets, (int)cap,
                            wk.near.pair_e, wk.near.pair_f, wk.near.blocks, wk.near.kind, wk.near.status, wk.ts,
                            wk.tk, mass_k, mass_d, coef_k, coef_d, s, kp, d, w, ld, st));
    return GRIM_OK;
}

extern "C" int grim_hankel2_probe(int64_t m, const double* z, double* out, void* stream) {
    if (m < 0 || m > (int64_t)1 << 30 || (m > 0 && (!z || !out))) return GRIM_ERR_ARGS;
    if (m == 0) return GRIM_OK;
    CHECK(grim_launch_hankel2_probe((int)m, z, out, (hipStream_t)stream));
    return GRIM_OK;
}

static bool batch_ok(int64_t batch) { return batch >= 1 && batch <= 65535; }

extern "C" int64_t grim_lu_aux_bytes(int64_t n, int precision) {
    if (n < 1 || n > 65536 || (precision != 64 && precision != 32)) return -1;
    return grim_dense_aux_bytes((int)n, precision == 32);
}

extern "C" int grim_lu_factor(int64_t batch, int64_t n, void* a, int64_t lda, int64_t stride_a, int precision,
                              int32_t* piv, int64_t stride_piv, void* aux, int64_t stride_aux, void* status,
                              void* stream) {
    if (!batch_ok(batch) || n < 1 || n > 65536 || !a || !piv || !aux || !status || lda < n ||
        (precision != 64 && precision != 32))
        return GRIM_ERR_ARGS;
    if (batch > 1 && (stride_a < lda * n || stride_piv < n || stride_aux < grim_lu_aux_bytes(n, precision) ||
                      stride_aux % 256 != 0))
        return GRIM_ERR_ARGS;
    hipStream_t st = (hipStream_t)stream;
    zero_words<<<1, 64, 0, st>>>((unsigned*)status, 1);
    CHECK(hipGetLastError());
    CHECK(grim_dense_lu_factor((int)batch, (int)n, a, lda, stride_a, precision == 32, piv, stride_piv, aux,
                               stride_aux, (unsigned*)status, st));
    return GRIM_OK;
}

extern "C" int64_t grim_lu_solve_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs, int precision) {
    if (!batch_ok(batch) || n < 1 || nrhs < 1) return -1;
    return batch * n * nrhs * (precision == 32 ? 8 : 16);
}

extern "C" int grim_lu_solve(int64_t batch, int64_t n, int64_t nrhs, const void* lu, int64_t ldlu, int64_t stride_lu,
                             int precision, const void* aux, int64_t stride_aux, void* b, int64_t ldb,
                             int64_t stride_b, void* work, int64_t work_bytes, void* stream) {
    if (!batch_ok(batch) || n < 1 || nrhs < 0 || !lu || !aux || !b || ldlu < n || ldb < n || nrhs > 65535 ||
        (precision != 64 && precision != 32))
        return GRIM_ERR_ARGS;
    if (nrhs == 0) return GRIM_OK;
    if (!work || work_bytes < grim_lu_solve_workspace_bytes(batch, n, nrhs, precision)) return GRIM_ERR_RANGE;
    CHECK(grim_dense_lu_solve((int)batch, (int)n, (int)nrhs, lu, ldlu, stride_lu, precision == 32, aux, stride_aux,
                              b, ldb, stride_b, work, n * nrhs, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_gemm_minus(int64_t batch, int64_t m, int64_t n, int64_t k, const void* a, int64_t lda,
                               int64_t stride_a, const void* b, int64_t ldb, int64_t stride_b, void* c, int64_t ldc,
                               int64_t stride_c, int precision, void* stream) {
    if (!batch_ok(batch) || m < 0 || n < 0 || k < 0 || !a || !b || !c || lda < m || ldb < k || ldc < m ||
        (precision != 64 && precision != 32) || n > 65535 * 64 || m > 65535 * 64)
        return GRIM_ERR_ARGS;
    CHECK(grim_dense_gemm_minus((int)batch, (int)m, (int)n, (int)k, a, lda, stride_a, b, ldb, stride_b, c, ldc,
                                stride_c, precision == 32, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int64_t grim_lu_solve_ex_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs, int precision,
                                                    int trans) {
    if (trans == 0) return grim_lu_solve_workspace_bytes(batch, n, nrhs, precision);
    if (trans != 2 || !batch_ok(batch) || n < 1 || nrhs < 1) return -1;
    const int64_t es = precision == 32 ? 8 : 16;
    return align_up(batch * n * nrhs * es) + batch * grim_dense_trans_block_elems() * es;
}

extern "C" int grim_lu_solve_ex(int64_t batch, int64_t n, int64_t nrhs, const void* lu, int64_t ldlu,
                                int64_t stride_lu, int precision, const void* aux, int64_t stride_aux, int trans,
                                void* b, int64_t ldb, int64_t stride_b, void* work, int64_t work_bytes, void* stream) {
    if (trans == 0)
        return grim_lu_solve(batch, n, nrhs, lu, ldlu, stride_lu, precision, aux, stride_aux, b, ldb, stride_b, work,
                             work_bytes, stream);
    if (trans != 2 || !batch_ok(batch) || n < 1 || nrhs < 0 || !lu || !aux || !b || ldlu < n || ldb < n ||
        nrhs > 65535 || (precision != 64 && precision != 32))
        return GRIM_ERR_ARGS;
    if (nrhs == 0) return GRIM_OK;
    if (!work || work_bytes < grim_lu_solve_ex_workspace_bytes(batch, n, nrhs, precision, trans)) return GRIM_ERR_RANGE;
    const int64_t es = precision == 32 ? 8 : 16;
    char* blk = (char*)work + align_up(batch * n * nrhs * es);
    CHECK(grim_dense_lu_solve_trans((int)batch, (int)n, (int)nrhs, lu, ldlu, stride_lu, precision == 32, aux,
                                    stride_aux, b, ldb, stride_b, work, n * nrhs, blk, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_gemm_minus_ex(int64_t batch, int64_t m, int64_t n, int64_t k, int trans_a, const void* a,
                                  int64_t lda, int64_t stride_a, const void* b, int64_t ldb, int64_t stride_b, void* c,
                                  int64_t ldc, int64_t stride_c, int precision, void* stream) {
    if (trans_a == 0)
        return grim_gemm_minus(batch, m, n, k, a, lda, stride_a, b, ldb, stride_b, c, ldc, stride_c, precision,
                               stream);
    if (trans_a != 2 || !batch_ok(batch) || m < 0 || n < 0 || k < 0 || !a || !b || !c || lda < k || ldb < k ||
        ldc < m || (precision != 64 && precision != 32) || n > 65535 * 64 || m > 65535 * 64)
        return GRIM_ERR_ARGS;
    CHECK(grim_dense_gemm_minus_ta((int)batch, (int)m, (int)n, (int)k, a, lda, stride_a, b, ldb, stride_b, c, ldc,
                                   stride_c, precision == 32, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_equilibration(int64_t batch, int64_t n, const double* z, int64_t ldz, int64_t stride_z,
                                  double* row_scale, double* col_scale, double* norm, void* stream) {
    if (!batch_ok(batch) || n < 1 || n > (1 << 30) || !z || !row_scale || !col_scale || !norm || ldz < n ||
        (batch > 1 && stride_z < ldz * n))
        return GRIM_ERR_ARGS;
    CHECK(grim_dense_equilibration((int)batch, (int)n, z, ldz, stride_z, row_scale, col_scale, norm,
                                   (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int64_t grim_norm_inf_workspace_bytes(int64_t batch, int64_t n) {
    if (!batch_ok(batch) || n < 1 || n > (1 << 30)) return -1;
    return batch * grim_dense_norm_inf_work((int)n);
}

extern "C" int grim_norm_inf(int64_t batch, int64_t n, const double* z, int64_t ldz, int64_t stride_z, void* work,
                             int64_t work_bytes, double* out, void* stream) {
    if (!batch_ok(batch) || n < 1 || n > (1 << 30) || !z || !out || !work || ldz < n ||
        (batch > 1 && stride_z < ldz * n))
        return GRIM_ERR_ARGS;
    if (work_bytes < grim_norm_inf_workspace_bytes(batch, n)) return GRIM_ERR_RANGE;
    CHECK(grim_dense_norm_inf((int)batch, (int)n, z, ldz, stride_z, (double*)work, out, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int64_t grim_refine_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs) {
    if (!batch_ok(batch) || n < 1 || nrhs < 1) return -1;
    return align_up(batch * 16 * n * nrhs) + 2 * align_up(batch * 8 * n * nrhs) + 2 * align_up(batch * 8 * nrhs);
}

extern "C" int64_t grim_refine_ex_workspace_bytes(int64_t batch, int64_t n, int64_t nrhs, int trans) {
    const int64_t base = grim_refine_workspace_bytes(batch, n, nrhs);
    if (base < 0 || (trans != 0 && trans != 2)) return -1;
    return trans == 0 ? base : base + align_up(batch * 8 * grim_dense_trans_block_elems());
}

// complex64 solve with the factors (trans 0) or their conjugate transpose (trans 2), and the
// FP64 residual b - op(Z) x
static hipError_t refine_lu_solve(int trans, int cnt, int N, int R, const float* lu, int64_t ldlu, int64_t stride_lu,
                                  const void* aux, int64_t stride_aux, float* d, int64_t snr, float* w, float* blk,
                                  hipStream_t st) {
    if (trans == 0)
        return grim_dense_lu_solve(cnt, N, R, lu, ldlu, stride_lu, 1, aux, stride_aux, d, N, snr, w, snr, st);
    return grim_dense_lu_solve_trans(cnt, N, R, lu, ldlu, stride_lu, 1, aux, stride_aux, d, N, snr, w, snr, blk, st);
}

static hipError_t refine_residual(int trans, int cnt, int N, int R, const double* z, int64_t ldz, int64_t stride_z,
                                  const double* x, int64_t ldx, int64_t stride_x, double* r, int64_t snr,
                                  hipStream_t st) {
    if (trans == 0) return grim_dense_gemm_minus(cnt, N, R, N, z, ldz, stride_z, x, ldx, stride_x, r, N, snr, 0, st);
    return grim_dense_gemm_minus_ta(cnt, N, R, N, z, ldz, stride_z, x, ldx, stride_x, r, N, snr, 0, st);
}

static int refined_solve_impl(int trans, int64_t batch, int64_t n, int64_t nrhs, const double* z, int64_t ldz,
                              int64_t stride_z, const float* lu, int64_t ldlu, int64_t stride_lu, const void* aux,
                              int64_t stride_aux, const double* b, int64_t ldb, int64_t stride_b, double* x,
                              int64_t ldx, int64_t stride_x, int max_steps, double tol, void* workspace,
                              int64_t workspace_bytes, int* steps_out, double* error_out, void* stream) {
    if (!batch_ok(batch) || n < 1 || nrhs < 1 || !z || !lu || !aux || !b || !x || ldz < n || ldlu < n || ldb < n ||
        ldx < n || max_steps < 0 || !(tol > 0) || !workspace || nrhs > 65535 || (trans != 0 && trans != 2))
        return GRIM_ERR_ARGS;
    if (grim_refine_ex_workspace_bytes(batch, n, nrhs, trans) > workspace_bytes) return GRIM_ERR_RANGE;
    hipStream_t st = (hipStream_t)stream;
    const int B = (int)batch, N = (int)n, R = (int)nrhs;
    const int64_t snr = n * nrhs;   // per-matrix stride of the n x nrhs work arrays
    const int64_t sblk = 2 * grim_dense_trans_block_elems();   // floats per matrix
    char* p = (char*)workspace;
    double* r = (double*)p; p += align_up(batch * 16 * n * nrhs);
    float* d = (float*)p; p += align_up(batch * 8 * n * nrhs);
    float* w = (float*)p; p += align_up(batch * 8 * n * nrhs);
    double* bmax = (double*)p; p += align_up(batch * 8 * nrhs);
    double* rmax = (double*)p; p += align_up(batch * 8 * nrhs);
    float* blk = (float*)p;
    std::vector<double> hb((size_t)B * R), hr((size_t)B * R);
    CHECK(grim_dense_colmax(B, N, R, b, ldb, stride_b, bmax, st));
    CHECK(hipMemcpyAsync(hb.data(), bmax, 8 * (size_t)B * R, hipMemcpyDeviceToHost, st));
    // x = LU32^-1 b
    CHECK(grim_dense_to_c64(B, N, R, b, ldb, stride_b, d, n, snr, st));
    CHECK(refine_lu_solve(trans, B, N, R, lu, ldlu, stride_lu, aux, stride_aux, d, snr, w, blk, st));
    zero_c128<<<dim3((N + 255) / 256, R, B), 256, 0, st>>>(N, R, x, ldx, stride_x);
    CHECK(hipGetLastError());
    CHECK(grim_dense_add_c64(B, N, R, d, n, snr, x, ldx, stride_x, st));
    CHECK(hipStreamSynchronize(st));
    for (auto& v : hb) v = v > 0 ? v : 1.0;
    // Each system refines on its own, as linalg/refined_lu.py:41-61 does per matrix: a system
    // whose error is within tol takes no further correction while the others continue, so a
    // batched call returns what separate calls would.  Work on a subset of systems runs per
    // system (batch 1, offset pointers).
    std::vector<double> previous((size_t)B, INFINITY);
    std::vector<char> done((size_t)B, 0);
    int active = B;
    double worst = 0.0;
    auto for_active = [&](auto&& op) -> hipError_t {
        if (active == B) return op(0, B);
        for (int i = 0; i < B; ++i)
            if (!done[(size_t)i]) {
                hipError_t e = op(i, 1);
                if (e != hipSuccess) return e;
            }
        return hipSuccess;
    };
    for (int step = 0; step < max_steps; ++step) {
        // r = b - Z x (FP64)
        CHECK(for_active([&](int i, int cnt) -> hipError_t {
            hipError_t e = grim_dense_copy(cnt, N, R, b + 2 * stride_b * i, ldb, stride_b, r + 2 * snr * i, n, snr,
                                           st);
            if (e == hipSuccess)
                e = refine_residual(trans, cnt, N, R, z + 2 * stride_z * i, ldz, stride_z, x + 2 * stride_x * i, ldx,
                                    stride_x, r + 2 * snr * i, snr, st);
            if (e == hipSuccess) e = grim_dense_colmax(cnt, N, R, r + 2 * snr * i, n, snr, rmax + (int64_t)R * i, st);
            return e;
        }));
        CHECK(hipMemcpyAsync(hr.data(), rmax, 8 * (size_t)B * R, hipMemcpyDeviceToHost, st));
        CHECK(hipStreamSynchronize(st));
        if (steps_out) *steps_out = step;
        for (int i = 0; i < B; ++i) {
            if (done[(size_t)i]) continue;
            double err = 0.0;
            for (int c = 0; c < R; ++c) err = fmax(err, hr[(size_t)i * R + c] / hb[(size_t)i * R + c]);
            if (isfinite(err) && err <= tol) {
                done[(size_t)i] = 1;
                --active;
                worst = fmax(worst, err);
                continue;
            }
            if (!isfinite(err) || err >= 0.98 * previous[(size_t)i]) {
                if (error_out) *error_out = err;
                return GRIM_ERR_NOT_CONVERGED;
            }
            previous[(size_t)i] = err;
        }
        if (active == 0) {
            if (error_out) *error_out = worst;
            return GRIM_OK;
        }
        if (error_out) {
            double cur = worst;
            for (int i = 0; i < B; ++i)
                if (!done[(size_t)i]) cur = fmax(cur, previous[(size_t)i]);
            *error_out = cur;
        }
        CHECK(for_active([&](int i, int cnt) -> hipError_t {
            hipError_t e = grim_dense_to_c64(cnt, N, R, r + 2 * snr * i, n, snr, d + 2 * snr * i, n, snr, st);
            if (e == hipSuccess)
                e = refine_lu_solve(trans, cnt, N, R, lu + 2 * stride_lu * i, ldlu, stride_lu,
                                    (const char*)aux + stride_aux * i, stride_aux, d + 2 * snr * i, snr,
                                    w + 2 * snr * i, blk + sblk * i, st);
            if (e == hipSuccess)
                e = grim_dense_add_c64(cnt, N, R, d + 2 * snr * i, n, snr, x + 2 * stride_x * i, ldx, stride_x, st);
            return e;
        }));
    }
    if (steps_out) *steps_out = max_steps;   // the correction limit, as distinct from a stall
    return GRIM_ERR_NOT_CONVERGED;
}

extern "C" int grim_refined_solve(int64_t batch, int64_t n, int64_t nrhs, const double* z, int64_t ldz,
                                  int64_t stride_z, const float* lu, int64_t ldlu, int64_t stride_lu, const void* aux,
                                  int64_t stride_aux, const double* b, int64_t ldb, int64_t stride_b, double* x,
                                  int64_t ldx, int64_t stride_x, int max_steps, double tol, void* workspace,
                                  int64_t workspace_bytes, int* steps_out, double* error_out, void* stream) {
    return refined_solve_impl(0, batch, n, nrhs, z, ldz, stride_z, lu, ldlu, stride_lu, aux, stride_aux, b, ldb,
                              stride_b, x, ldx, stride_x, max_steps, tol, workspace, workspace_bytes, steps_out,
                              error_out, stream);
}

extern "C" int grim_refined_solve_ex(int64_t batch, int64_t n, int64_t nrhs, const double* z, int64_t ldz,
                                     int64_t stride_z, const float* lu, int64_t ldlu, int64_t stride_lu,
                                     const void* aux, int64_t stride_aux, int trans, const double* b, int64_t ldb,
                                     int64_t stride_b, double* x, int64_t ldx, int64_t stride_x, int max_steps,
                                     double tol, void* workspace, int64_t workspace_bytes, int* steps_out,
                                     double* error_out, void* stream) {
    return refined_solve_impl(trans, batch, n, nrhs, z, ldz, stride_z, lu, ldlu, stride_lu, aux, stride_aux, b, ldb,
                              stride_b, x, ldx, stride_x, max_steps, tol, workspace, workspace_bytes, steps_out,
                              error_out, stream);
}

extern "C" int grim_to_c64(int64_t batch, int64_t m, int64_t n, const double* src, int64_t lds, int64_t stride_src,
                           float* dst, int64_t ldd, int64_t stride_dst, void* stream) {
    if (!batch_ok(batch) || m < 0 || n < 0 || !src || !dst || lds < m || ldd < m || n > 65535) return GRIM_ERR_ARGS;
    CHECK(grim_dense_to_c64((int)batch, (int)m, (int)n, src, lds, stride_src, dst, ldd, stride_dst,
                            (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_copy_c128(int64_t batch, int64_t m, int64_t n, const double* src, int64_t lds,
                              int64_t stride_src, double* dst, int64_t ldd, int64_t stride_dst, void* stream) {
    if (!batch_ok(batch) || m < 0 || n < 0 || !src || !dst || lds < m || ldd < m || n > 65535) return GRIM_ERR_ARGS;
    CHECK(grim_dense_copy((int)batch, (int)m, (int)n, src, lds, stride_src, dst, ldd, stride_dst,
                          (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_column_norms(int64_t m, int64_t n, const double* v, int64_t ld, double* out, void* stream) {
    if (m < 1 || n < 0 || !v || !out || ld < m) return GRIM_ERR_ARGS;
    CHECK(grim_dense_colnorm((int)m, (int)n, v, ld, out, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_robin_split(int64_t n, const double* p0, const double* p1, const int32_t* next,
                                const int32_t* prev, double k, int far_order, int64_t nblocks, const int32_t* walk,
                                const double* c_te, const double* c_tm, const int32_t* flags, const void* workspace,
                                int64_t workspace_bytes, int64_t near_per_element, double* z_te, int64_t ld_te,
                                double* z_tm, int64_t ld_tm, void* stream) {
    if (n < 2 || n > (1 << 30) || nblocks < 0 || nblocks > (int64_t)1 << 30 || !(k > 0) || !isfinite(k) ||
        (far_order != 8 && far_order != 16) || near_per_element <= 0 || ld_te < n || ld_tm < n ||
        (n + 31) / 32 > 65535)
        return GRIM_ERR_ARGS;
    if (nblocks == 0) return GRIM_OK;
    if (!p0 || !p1 || !next || !prev || !walk || !c_te || !c_tm || !flags || !workspace || !z_te || !z_tm)
        return GRIM_ERR_ARGS;
    const int64_t cap = n * near_per_element;
    if (cap > (int64_t)1 << 31) return GRIM_ERR_RANGE;
    const AssemblyWork w = carve(const_cast<void*>(workspace), n, cap);
    if (w.bytes > workspace_bytes) return GRIM_ERR_RANGE;
    CHECK(grim_launch_robin_split(far_order, (int)n, p0, p1, next, prev, k, (int)nblocks, walk, w.offsets, w.pair_f,
                                  (int)cap, w.blocks, c_te, c_tm, flags, z_te, ld_te, z_tm, ld_tm,
                                  (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_robin_rhs(int64_t n, int64_t na, const double* p0, const double* p1, const int32_t* prev,
                              double k, const double* angles_deg, const double* alpha_e, const int32_t* dirichlet,
                              double* out, int64_t ld, void* stream) {
    if (n < 1 || n > (1 << 30) || na < 0 || na > 65535 || ld < n || !(k > 0) || !isfinite(k)) return GRIM_ERR_ARGS;
    if (na == 0) return GRIM_OK;
    if (!p0 || !p1 || !prev || !angles_deg || !alpha_e || !dirichlet || !out) return GRIM_ERR_ARGS;
    CHECK(grim_launch_robin_rhs((int)n, (int)na, p0, p1, prev, k, angles_deg, alpha_e, dirichlet, out, ld,
                                (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_scale_blocks(int64_t n, double* z, int64_t ld, const double* scale, void* stream) {
    if (n < 0 || 2 * n > 65535 || ld < 2 * n || (n > 0 && (!z || !scale))) return GRIM_ERR_ARGS;
    for (int i = 0; i < 8; ++i)
        if (!isfinite(scale[i])) return GRIM_ERR_ARGS;
    CHECK(grim_launch_scale_blocks(n, z, ld, scale, (hipStream_t)stream));
    return GRIM_OK;
}

extern "C" int grim_robin_rows(int64_t rows, int64_t cols, double* x, int64_t ldx, double* y, int64_t ldy,
                               const double* alpha_x, const double* alpha_y, const int32_t* dirichlet, void* stream) {
    if (rows < 0 || cols < 0 || cols > 65535 || ldx < rows || ldy < rows) return GRIM_ERR_ARGS;
    if (rows > 0 && cols > 0 && (!x || !y || !alpha_x || !alpha_y || !dirichlet)) return GRIM_ERR_ARGS;
    CHECK(grim_launch_robin_rows(rows, cols, x, ldx, y, ldy, alpha_x, alpha_y, dirichlet, (hipStream_t)stream));
    return GRIM_OK;
}
