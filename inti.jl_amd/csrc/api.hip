// C-ABI entry points of libinti_amd (declared in include/inti_amd.h):
// context / stream / profiling, quadrature upload, operators, corrections.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>

#include "common.hpp"

namespace inti {

static thread_local std::string g_create_err;

// Partial sums of the allocation-free apply (inti_op_apply_device_noalloc):
// a static buffer that every GPU the library runs on holds from load until
// the process ends. Host code never uses this name as an address: it is a
// host-side stand-in; the device address is looked up once per context with
// hipGetSymbolAddress (Ctx::noalloc_scratch).
constexpr size_t kNoAllocBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

int set_error(Ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    else g_create_err = msg;
    return code;
}

int ctx_scratch(Ctx *ctx, int slot, size_t bytes, void **ptr) {
    if (bytes == 0) bytes = 256;
    if (ctx->scratch_bytes[slot] < bytes) {
        if (ctx->scratch[slot]) {
            INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
            INTI_HIP(ctx, hipFree(ctx->scratch[slot]));
            ctx->scratch[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
        }
        size_t want = bytes + bytes / 4;
        if (hipMalloc(&ctx->scratch[slot], want) != hipSuccess) {
            ctx->scratch[slot] = nullptr;
            return set_error(ctx, INTI_ERR_NOMEM, "scratch allocation failed");
        }
        ctx->scratch_bytes[slot] = want;
    }
    *ptr = ctx->scratch[slot];
    return INTI_OK;
}

int csr_alloc(Ctx *ctx, int64_t m, int64_t n, int64_t nnz, int cplx, Csr **out) {
    Csr *c = new inti_csr();
    c->ctx = ctx; c->m = m; c->n = n; c->nnz = nnz; c->is_complex = cplx;
    if (hipMalloc(&c->rowptr, sizeof(int32_t) * (m + 1)) != hipSuccess ||
        hipMalloc(&c->colind, sizeof(int32_t) * (nnz > 0 ? nnz : 1)) != hipSuccess ||
        hipMalloc(&c->val, sizeof(double) * (cplx ? 2 : 1) * (nnz > 0 ? nnz : 1)) != hipSuccess) {
        if (c->rowptr) (void)hipFree(c->rowptr);
        if (c->colind) (void)hipFree(c->colind);
        if (c->val) (void)hipFree(c->val);
        delete static_cast<inti_csr *>(c);
        return set_error(ctx, INTI_ERR_NOMEM, "CSR allocation failed");
    }
    *out = c;
    return INTI_OK;
}

void csr_free(Csr *c) {
    if (!c) return;
    if (c->rowptr) (void)hipFree(c->rowptr);
    if (c->colind) (void)hipFree(c->colind);
    if (c->val) (void)hipFree(c->val);
    delete static_cast<inti_csr *>(c);
}

void op_free_correction(Op *op) {
    csr_free(op->corr);
    op->corr = nullptr;
    int32_t *ps[] = {op->corr_tptr, op->corr_trow, op->corr_tperm};
    for (int32_t *p : ps)
        if (p) (void)hipFree(p);
    op->corr_tptr = op->corr_trow = op->corr_tperm = nullptr;
}

int bdim_build(Op *opS, Op *opD, int deriv, int nq, int64_t ne, const int64_t *qtags,
               const int64_t *near_ptr, const int64_t *near_idx, const double *mu,
               const double *params, Csr **outS, Csr **outD, double *diag_out);
int bdim_build_types(Op *opS, Op *opD, int deriv, int ntypes, const int *nq, const int64_t *ne,
                     const int64_t *qtags, const int64_t *near_ptr, const int64_t *near_idx, const double *mu,
                     const double *params, Csr **outS, Csr **outD, double *diag_out);
int lsq_build_types(Ctx *ctx, int cplx, int nrows, int ntypes, const int *ncols, const int64_t *ne, const double *A,
                    const int64_t *cols, int64_t m, int64_t n, const int64_t *near_ptr, const int64_t *near_idx,
                    const double *rhs, Csr **out, double *diag_out);
int lsq_build(Ctx *ctx, int cplx, int nrows, int ncols, int64_t ne, const double *A, const int64_t *cols,
              int64_t m, int64_t n, const int64_t *near_ptr, const int64_t *near_idx, const double *rhs,
              Csr **out, double *diag_out);

}  // namespace inti

using namespace inti;

#define LOCK(ctx) std::lock_guard<std::mutex> lock_((ctx)->mu)
#define REQUIRE(ctx, cond, msg) \
    do { if (!(cond)) return set_error((ctx), INTI_ERR_ARG, (msg)); } while (0)

extern "C" {

int inti_version(void) { return 100; }

int inti_ctx_create(int device, inti_ctx **out) {
    if (!out) return set_error(nullptr, INTI_ERR_ARG, "out is null");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return set_error(nullptr, INTI_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= count) return set_error(nullptr, INTI_ERR_ARG, "device out of range");
    if (hipSetDevice(device) != hipSuccess) return set_error(nullptr, INTI_ERR_HIP, "hipSetDevice failed");
    inti_ctx *c = new inti_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return set_error(nullptr, INTI_ERR_HIP, "stream/event creation failed");
    }
    // this device's copy of the static partial-sum buffer (the device is
    // current here)
    void *sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess || sp == nullptr) {
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return set_error(nullptr, INTI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
    }
    c->noalloc_scratch = (double *)sp;
    *out = c;
    return INTI_OK;
}

int inti_ctx_destroy(inti_ctx *ctx) {
    if (!ctx) return INTI_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int s = 0; s < 4; ++s)
        if (ctx->scratch[s]) (void)hipFree(ctx->scratch[s]);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return INTI_OK;
}

const char *inti_last_error(const inti_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int inti_microbench(inti_ctx *ctx, int op, double *rate) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, rate != nullptr, "microbench: rate is null");
    LOCK(ctx);
    return microbench(ctx, op, rate);
}

int inti_ctx_synchronize(inti_ctx *ctx) {
    if (!ctx) return INTI_ERR_ARG;
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

int inti_ctx_stream(inti_ctx *ctx, void **stream) {
    if (!ctx || !stream) return INTI_ERR_ARG;
    *stream = (void *)ctx->stream;
    return INTI_OK;
}

int inti_ctx_set_profiling(inti_ctx *ctx, int on) {
    if (!ctx) return INTI_ERR_ARG;
    ctx->profiling = on ? 1 : 0;
    return INTI_OK;
}

int inti_ctx_last_kernel_ms(inti_ctx *ctx, double *ms) {
    if (!ctx || !ms) return INTI_ERR_ARG;
    LOCK(ctx);
    INTI_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float t = 0.f;
    INTI_HIP(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t;
    return INTI_OK;
}

int inti_device_alloc(inti_ctx *ctx, size_t bytes, void **dptr) {
    if (!ctx || !dptr) return INTI_ERR_ARG;
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess)
        return set_error(ctx, INTI_ERR_NOMEM, "device allocation failed");
    return INTI_OK;
}

int inti_device_free(inti_ctx *ctx, void *dptr) {
    if (!ctx) return INTI_ERR_ARG;
    LOCK(ctx);
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    INTI_HIP(ctx, hipFree(dptr));
    return INTI_OK;
}

int inti_memcpy_h2d(inti_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return INTI_ERR_ARG;
    LOCK(ctx);
    INTI_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

int inti_memcpy_d2h(inti_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return INTI_ERR_ARG;
    LOCK(ctx);
    INTI_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

__global__ __launch_bounds__(256) void k_copy16(const double2 *__restrict__ src,
                                                 double2 *__restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

int inti_device_copy(inti_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, bytes % 16 == 0, "device_copy: bytes must be a multiple of 16");
    REQUIRE(ctx, bytes == 0 || (dst && src), "device_copy: null pointer");
    LOCK(ctx);
    const int64_t n = (int64_t)(bytes / 16);
    if (n == 0) return INTI_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8 * 1024) blocks = 8 * 1024;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       (const double2 *)src, (double2 *)dst, n);
    INTI_HIP(ctx, hipGetLastError());
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

// ------------------------------------------------------------- quadrature
// True if every two of the n nodes are farther apart than 2e-14 (twice the
// coincidence tolerance, so that rounding differences between this test and
// the kernels' own distances cannot matter). Nodes are binned in cubes of
// side h = 4e-14; a closer pair lies in the same or in neighbouring cubes.
// O(n log n). Coordinates beyond 1e5 in magnitude (cube indices near the
// int64 range) give false, which keeps the mask.
static bool nodes_distinct(int dim, int64_t n, const double *c) {
    const double h = 4.0e-14, tol2 = 4.0e-28;
    std::vector<std::array<int64_t, 3>> key(n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            if (d >= dim) { key[i][d] = 0; continue; }
            const double v = c[(int64_t)d * n + i];
            if (!(std::fabs(v) < 1.0e5)) return false;
            key[i][d] = (int64_t)std::floor(v / h);
        }
    }
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    std::vector<std::array<int64_t, 3>> sk(n);
    for (int64_t q = 0; q < n; ++q) sk[q] = key[idx[q]];
    const int zr = dim == 3 ? 1 : 0;
    for (int64_t q = 0; q < n; ++q) {
        const int64_t i = idx[q];
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -zr; dz <= zr; ++dz) {
                    const std::array<int64_t, 3> nb = {sk[q][0] + dx, sk[q][1] + dy, sk[q][2] + dz};
                    auto rg = std::equal_range(sk.begin(), sk.end(), nb);
                    for (auto it = rg.first; it != rg.second; ++it) {
                        const int64_t j = idx[it - sk.begin()];
                        if (j == i) continue;
                        double d2 = 0.0;
                        for (int d = 0; d < dim; ++d) {
                            const double t = c[(int64_t)d * n + i] - c[(int64_t)d * n + j];
                            d2 += t * t;
                        }
                        if (d2 <= tol2) return false;
                    }
                }
    }
    return true;
}

int inti_quad_upload(inti_ctx *ctx, int dim, int64_t n, const double *coords,
                     const double *normals, const double *weights, inti_quad **out) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, out != nullptr, "quad_upload: out is null");
    REQUIRE(ctx, dim == 2 || dim == 3, "quad_upload: dim must be 2 or 3");
    REQUIRE(ctx, n >= 0, "quad_upload: negative size");
    REQUIRE(ctx, n == 0 || (coords && weights), "quad_upload: coords and weights are required");
    LOCK(ctx);
    inti_quad *q = new inti_quad();
    q->ctx = ctx; q->dim = dim; q->n = n;
    const size_t cb = sizeof(double) * dim * (n > 0 ? n : 1);
    if (hipMalloc(&q->coords, cb) != hipSuccess ||
        hipMalloc(&q->weights, sizeof(double) * (n > 0 ? n : 1)) != hipSuccess ||
        (normals && hipMalloc(&q->normals, cb) != hipSuccess)) {
        if (q->coords) (void)hipFree(q->coords);
        if (q->weights) (void)hipFree(q->weights);
        if (q->normals) (void)hipFree(q->normals);
        delete q;
        return set_error(ctx, INTI_ERR_NOMEM, "quad_upload: allocation failed");
    }
    if (n > 0) {
        INTI_HIP(ctx, hipMemcpyAsync(q->coords, coords, sizeof(double) * dim * n, hipMemcpyHostToDevice, ctx->stream));
        INTI_HIP(ctx, hipMemcpyAsync(q->weights, weights, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        if (normals)
            INTI_HIP(ctx, hipMemcpyAsync(q->normals, normals, sizeof(double) * dim * n, hipMemcpyHostToDevice, ctx->stream));
        for (int d = 0; d < dim; ++d) {
            double lo = coords[(int64_t)d * n], hi = lo;
            for (int64_t i = 1; i < n; ++i) {
                lo = std::min(lo, coords[(int64_t)d * n + i]);
                hi = std::max(hi, coords[(int64_t)d * n + i]);
            }
            q->lo[d] = lo; q->hi[d] = hi;
        }
        q->distinct = nodes_distinct(dim, n, coords);
        INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    *out = q;
    return INTI_OK;
}

int inti_quad_destroy(inti_quad *q) {
    if (!q) return INTI_ERR_ARG;
    Ctx *ctx = q->ctx;
    {
        LOCK(ctx);
        (void)hipStreamSynchronize(ctx->stream);
        if (q->coords) (void)hipFree(q->coords);
        if (q->weights) (void)hipFree(q->weights);
        if (q->normals) (void)hipFree(q->normals);
    }
    delete q;
    return INTI_OK;
}

// --------------------------------------------------------------- operator
int inti_op_create(inti_ctx *ctx, int pde, int kind, int dim, double k, const double *coefs,
                   const inti_quad *trg, const inti_quad *src, int precision, inti_op **out) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, out && trg && src, "op_create: null argument");
    REQUIRE(ctx, pde == INTI_LAPLACE || pde == INTI_HELMHOLTZ, "op_create: unknown pde");
    REQUIRE(ctx, kind >= INTI_SL && kind <= INTI_FF_COMB, "op_create: unknown kernel kind");
    const bool farfield = kind >= INTI_FF_SL;
    if (farfield && pde != INTI_HELMHOLTZ)
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "op_create: far-field kinds are Helmholtz only");
    REQUIRE(ctx, dim == 2 || dim == 3, "op_create: dim must be 2 or 3");
    REQUIRE(ctx, precision == INTI_FP64 || precision == INTI_MIXED, "op_create: unknown precision");
    // IntegralOperator(k, X, Y) requires equal dimensions (src/nystrom.jl:79-85)
    if (trg->dim != dim || src->dim != dim)
        return set_error(ctx, INTI_ERR_SIZE, "op_create: quadrature dimension differs from the operator's");
    const bool needs_src_n = kind != INTI_SL && kind != INTI_ADL && kind != INTI_FF_SL;
    const bool needs_trg_n = kind == INTI_ADL || kind == INTI_HS || kind == INTI_COMB_AH;
    if (needs_src_n && !src->normals)
        return set_error(ctx, INTI_ERR_ARG, "op_create: this kernel needs source normals");
    if (needs_trg_n && !trg->normals)
        return set_error(ctx, INTI_ERR_ARG, "op_create: this kernel needs target normals");
    if (pde == INTI_HELMHOLTZ) REQUIRE(ctx, k > 0.0 && std::isfinite(k), "op_create: Helmholtz needs a real k > 0");
    const bool comb = kind == INTI_COMB_SD || kind == INTI_COMB_AH || kind == INTI_FF_COMB;
    if (comb) {
        REQUIRE(ctx, coefs != nullptr, "op_create: combined kinds need coefs");
        if (pde == INTI_LAPLACE && (coefs[1] != 0.0 || coefs[3] != 0.0))
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "op_create: Laplace combinations take real coefficients");
    }
    inti_op *op = new inti_op();
    op->ctx = ctx; op->pde = pde; op->kind = kind; op->dim = dim; op->k = k;
    op->precision = precision; op->trg = trg; op->src = src;
    if (comb) std::memcpy(op->coefs, coefs, sizeof(double) * 4);
    *out = op;
    return INTI_OK;
}

int inti_op_destroy(inti_op *op) {
    if (!op) return INTI_ERR_ARG;
    {
        LOCK(op->ctx);
        (void)hipStreamSynchronize(op->ctx->stream);
        op_free_correction(op);
    }
    delete op;
    return INTI_OK;
}

int inti_op_set_identity(inti_op *op, double re, double im) {
    if (!op) return INTI_ERR_ARG;
    if (op->trg->n != op->src->n)
        return set_error(op->ctx, INTI_ERR_SIZE, "op_set_identity: operator is not square");
    if (op->pde == INTI_LAPLACE && im != 0.0)
        return set_error(op->ctx, INTI_ERR_UNSUPPORTED, "op_set_identity: Laplace takes a real coefficient");
    op->cid[0] = re; op->cid[1] = im;
    return INTI_OK;
}

int inti_op_set_wavenumber_imag(inti_op *op, double kimag) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, std::isfinite(kimag), "op_set_wavenumber_imag: kimag must be finite");
    if (kimag != 0.0) {
        if (op->pde != INTI_HELMHOLTZ || op->dim != 3)
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "op_set_wavenumber_imag: complex wavenumbers are 3D Helmholtz only");
        if (op->kind >= INTI_FF_SL)
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "op_set_wavenumber_imag: no far-field kernels for complex wavenumbers");
        if (op->precision != INTI_FP64)
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "op_set_wavenumber_imag: complex wavenumbers run in FP64");
    }
    LOCK(ctx);
    op->kimag = kimag;
    return INTI_OK;
}

int inti_op_set_launch(inti_op *op, int block, int tpt, int split) {
    if (!op) return INTI_ERR_ARG;
    if (block != 256 || tpt != 1 || split < 0)
        return set_error(op->ctx, INTI_ERR_UNSUPPORTED, "op_set_launch: block 256, tpt 1, split >= 0");
    op->block = block; op->tpt = tpt; op->split = split;
    return INTI_OK;
}

int inti_op_set_symmetric(inti_op *op, int mode, int group) {
    if (!op) return INTI_ERR_ARG;
    if (mode < 0 || mode > 2) return set_error(op->ctx, INTI_ERR_ARG, "op_set_symmetric: mode 0, 1 or 2");
    if (group != 0 && group != 1 && group != 2 && group != 4)
        return set_error(op->ctx, INTI_ERR_ARG, "op_set_symmetric: group 0, 1, 2 or 4");
    op->sym = mode;
    op->sym_group = group;
    return INTI_OK;
}

static int apply_device_locked(inti_op *op, int nrhs, const double *x, int64_t ldx, double *y,
                               int64_t ldy, const double *alpha, const double *beta) {
    DirectLaunch L;
    L.op = op; L.nrhs = nrhs; L.x = x; L.ldx = ldx; L.y = y; L.ldy = ldy;
    L.alpha[0] = alpha ? alpha[0] : 1.0; L.alpha[1] = alpha ? alpha[1] : 0.0;
    L.beta[0] = beta ? beta[0] : 0.0; L.beta[1] = beta ? beta[1] : 0.0;
    if (op->pde == INTI_LAPLACE && (L.alpha[1] != 0.0 || L.beta[1] != 0.0))
        return set_error(op->ctx, INTI_ERR_UNSUPPORTED, "apply: Laplace operators take real alpha/beta");
    return launch_apply(L);
}

int inti_op_apply_device(inti_op *op, int nrhs, const double *x, int64_t ldx, double *y,
                         int64_t ldy, const double *alpha, const double *beta) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, nrhs >= 0, "apply: nrhs < 0");
    REQUIRE(ctx, nrhs == 0 || (x && y), "apply: null operand");
    REQUIRE(ctx, ldx >= op->src->n && ldy >= op->trg->n, "apply: leading dimension too small");
    LOCK(ctx);
    return apply_device_locked(op, nrhs, x, ldx, y, ldy, alpha, beta);
}

int inti_op_apply_device_noalloc(inti_op *op, int nrhs, const double *x, int64_t ldx, double *y,
                                 int64_t ldy, const double *alpha, const double *beta) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, nrhs >= 0, "apply_noalloc: nrhs < 0");
    REQUIRE(ctx, nrhs == 0 || (x && y), "apply_noalloc: null operand");
    REQUIRE(ctx, ldx >= op->src->n && ldy >= op->trg->n, "apply_noalloc: leading dimension too small");
    if (!ctx->noalloc_scratch) return set_error(ctx, INTI_ERR_HIP, "apply_noalloc: no scratch address");
    LOCK(ctx);
    DirectLaunch L;
    L.op = op; L.nrhs = nrhs; L.x = x; L.ldx = ldx; L.y = y; L.ldy = ldy;
    L.alpha[0] = alpha ? alpha[0] : 1.0; L.alpha[1] = alpha ? alpha[1] : 0.0;
    L.beta[0] = beta ? beta[0] : 0.0; L.beta[1] = beta ? beta[1] : 0.0;
    if (op->pde == INTI_LAPLACE && (L.alpha[1] != 0.0 || L.beta[1] != 0.0))
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "apply_noalloc: Laplace operators take real alpha/beta");
    // every partial slot the reduction reads is written by the sum of the
    // same call (each block writes its own chunk's partial of every row of
    // its tile): nothing needs clearing between calls
    L.part = ctx->noalloc_scratch;
    L.part_bytes = kNoAllocBytes;
    return launch_apply(L);
}

int inti_op_apply(inti_op *op, int nrhs, const double *x, double *y, const double *alpha,
                  const double *beta) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, nrhs >= 0, "apply: nrhs < 0");
    REQUIRE(ctx, nrhs == 0 || (x && y), "apply: null operand");
    if (nrhs == 0) return INTI_OK;
    LOCK(ctx);
    const int cw = op_is_complex(op) ? 2 : 1;
    const int64_t m = op->trg->n, n = op->src->n;
    void *dx = nullptr, *dy = nullptr;
    int st = ctx_scratch(ctx, kScrX, sizeof(double) * cw * n * nrhs, &dx);
    if (st) return st;
    st = ctx_scratch(ctx, kScrY, sizeof(double) * cw * m * nrhs, &dy);
    if (st) return st;
    INTI_HIP(ctx, hipMemcpyAsync(dx, x, sizeof(double) * cw * n * nrhs, hipMemcpyHostToDevice, ctx->stream));
    const bool beta_nz = beta && (beta[0] != 0.0 || beta[1] != 0.0);
    if (beta_nz)
        INTI_HIP(ctx, hipMemcpyAsync(dy, y, sizeof(double) * cw * m * nrhs, hipMemcpyHostToDevice, ctx->stream));
    st = apply_device_locked(op, nrhs, (const double *)dx, n, (double *)dy, m, alpha, beta);
    if (st) return st;
    INTI_HIP(ctx, hipMemcpyAsync(y, dy, sizeof(double) * cw * m * nrhs, hipMemcpyDeviceToHost, ctx->stream));
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

int inti_op_assemble_dense(inti_op *op, double *out) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, out != nullptr, "assemble: out is null");
    LOCK(ctx);
    const int cw = op_is_complex(op) ? 2 : 1;
    const size_t bytes = sizeof(double) * cw * (size_t)op->trg->n * op->src->n;
    if (bytes == 0) return INTI_OK;
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return set_error(ctx, INTI_ERR_NOMEM, "assemble: allocation failed");
    int st = launch_assemble(op, (double *)d);
    if (st == INTI_OK) {
        if (hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            st = set_error(ctx, INTI_ERR_HIP, "assemble: copy back failed");
    }
    (void)hipFree(d);
    return st;
}

int inti_op_assemble_dense_device(inti_op *op, double *out, int64_t ldout) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, out != nullptr, "assemble_device: out is null");
    REQUIRE(ctx, ldout == op->trg->n, "assemble_device: the leading dimension must equal the number of targets");
    LOCK(ctx);
    return launch_assemble(op, out);
}

// A[i + lda j] += coef * csr(i, j): the axpy!(true, dS, Smat) of src/api.jl:171-173
__global__ void k_dense_add_csr(int64_t m, const int32_t *rowptr, const int32_t *colind, const double *val,
                                int cplx, double cr, double ci, double *A, int64_t lda) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
        const int64_t o = i + lda * (int64_t)colind[q];
        if (cplx) {
            const double vr = val[2 * q], vi = val[2 * q + 1];
            A[2 * o] += cr * vr - ci * vi;
            A[2 * o + 1] += cr * vi + ci * vr;
        } else {
            A[o] += cr * val[q];
        }
    }
}

int inti_dense_add_csr(inti_ctx *ctx, const inti_csr *csr, double coef_re, double coef_im, double *A,
                       int64_t lda) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, csr && A, "dense_add_csr: null argument");
    REQUIRE(ctx, lda >= csr->m, "dense_add_csr: leading dimension below the row count");
    LOCK(ctx);
    if (csr->m > 0)
        hipLaunchKernelGGL(k_dense_add_csr, dim3((unsigned)((csr->m + 255) / 256)), dim3(256), 0, ctx->stream, csr->m,
                           csr->rowptr, csr->colind, csr->val, csr->is_complex, coef_re, coef_im, A, lda);
    INTI_HIP(ctx, hipGetLastError());
    return INTI_OK;
}

int inti_dense_gemm(inti_ctx *ctx, int is_complex, int64_t m, int64_t n, int nrhs, const double *A, int64_t lda,
                    const double *X, int64_t ldx, double *Y, int64_t ldy, const double *alpha, const double *beta) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, A && X && Y && alpha && beta, "dense_gemm: null argument");
    REQUIRE(ctx, m >= 0 && n >= 0 && nrhs >= 0, "dense_gemm: negative size");
    REQUIRE(ctx, lda >= m && ldx >= n && ldy >= m, "dense_gemm: leading dimension too small");
    LOCK(ctx);
    return dense_gemm(ctx, is_complex ? 1 : 0, m, n, nrhs, A, lda, X, ldx, Y, ldy, alpha, beta);
}

int inti_op_adjoint(inti_op *op, const double *ubar, const double *sigma, double *sigma_bar,
                    double *src_xyz_bar, double *src_normal_bar, double *src_weight_bar,
                    double *trg_xyz_bar, const double *dcoefs, double *k_bar) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, ubar != nullptr && sigma != nullptr, "adjoint: ubar and sigma are required");
    if (op->kimag != 0.0)
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "adjoint: complex wavenumbers have no pullback");
    LOCK(ctx);
    return op_adjoint(op, ubar, sigma, sigma_bar, src_xyz_bar, src_normal_bar, src_weight_bar, trg_xyz_bar,
                      dcoefs, k_bar);
}

// merged correction: c_new = c_old + coef * csr (same pattern), or a copy
__global__ void k_axpy_vals(int64_t nnz, int cplx, const double *src, double cr, double ci,
                            double *dst, int accumulate) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nnz) return;
    if (cplx) {
        const double vr = src[2 * q], vi = src[2 * q + 1];
        double yr = cr * vr - ci * vi, yi = cr * vi + ci * vr;
        if (accumulate) { yr += dst[2 * q]; yi += dst[2 * q + 1]; }
        dst[2 * q] = yr; dst[2 * q + 1] = yi;
    } else {
        double yr = cr * src[q];
        if (accumulate) yr += dst[q];
        dst[q] = yr;
    }
}

int inti_op_attach_correction(inti_op *op, const inti_csr *csr, double coef_re, double coef_im) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, csr != nullptr, "attach: csr is null");
    if (csr->m != op->trg->n || csr->n != op->src->n)
        return set_error(ctx, INTI_ERR_SIZE, "attach: correction shape differs from the operator's");
    if (csr->is_complex != (op_is_complex(op) ? 1 : 0))
        return set_error(ctx, INTI_ERR_ARG, "attach: correction value type differs from the operator's");
    if (!csr->is_complex && coef_im != 0.0)
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "attach: real correction takes a real coefficient");
    LOCK(ctx);
    const int cplx = csr->is_complex;
    int accumulate = 0;
    if (op->corr) {
        // same pattern required for an in-place merge
        bool same = op->corr->nnz == csr->nnz;
        if (same && csr->nnz > 0) {
            std::vector<int32_t> a(csr->m + 1 + csr->nnz), b(a.size());
            INTI_HIP(ctx, hipMemcpy(a.data(), op->corr->rowptr, sizeof(int32_t) * (csr->m + 1), hipMemcpyDeviceToHost));
            INTI_HIP(ctx, hipMemcpy(a.data() + csr->m + 1, op->corr->colind, sizeof(int32_t) * csr->nnz, hipMemcpyDeviceToHost));
            INTI_HIP(ctx, hipMemcpy(b.data(), csr->rowptr, sizeof(int32_t) * (csr->m + 1), hipMemcpyDeviceToHost));
            INTI_HIP(ctx, hipMemcpy(b.data() + csr->m + 1, csr->colind, sizeof(int32_t) * csr->nnz, hipMemcpyDeviceToHost));
            same = a == b;
        }
        if (!same)
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "attach: corrections with different patterns cannot be merged");
        accumulate = 1;
    } else {
        // transposed pattern for the adjoint's gather (counting sort by
        // column; rows stay increasing within a column)
        const int64_t m = csr->m, n = csr->n, nnz = csr->nnz;
        std::vector<int32_t> rp(m + 1), ci(nnz), tptr(n + 1, 0), trow(nnz), tperm(nnz);
        INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
        INTI_HIP(ctx, hipMemcpy(rp.data(), csr->rowptr, sizeof(int32_t) * (m + 1), hipMemcpyDeviceToHost));
        if (nnz > 0) INTI_HIP(ctx, hipMemcpy(ci.data(), csr->colind, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < nnz; ++q) ++tptr[ci[q] + 1];
        for (int64_t j = 0; j < n; ++j) tptr[j + 1] += tptr[j];
        std::vector<int32_t> fill(tptr.begin(), tptr.end() - 1);
        for (int64_t i = 0; i < m; ++i)
            for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
                const int32_t p = fill[ci[q]]++;
                trow[p] = (int32_t)i;
                tperm[p] = q;
            }
        int32_t *dt[3] = {nullptr, nullptr, nullptr};
        const size_t tb[3] = {sizeof(int32_t) * (size_t)(n + 1), sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1),
                              sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)};
        for (int t = 0; t < 3; ++t)
            if (hipMalloc((void **)&dt[t], tb[t]) != hipSuccess) {
                for (int u = 0; u < 3; ++u)
                    if (dt[u]) (void)hipFree(dt[u]);
                return set_error(ctx, INTI_ERR_NOMEM, "attach: transposed pattern allocation failed");
            }
        Csr *c = nullptr;
        int st = csr_alloc(ctx, csr->m, csr->n, csr->nnz, cplx, &c);
        if (st) {
            for (int u = 0; u < 3; ++u) (void)hipFree(dt[u]);
            return st;
        }
        op->corr = c;
        op->corr_tptr = dt[0]; op->corr_trow = dt[1]; op->corr_tperm = dt[2];
        INTI_HIP(ctx, hipMemcpyAsync(c->rowptr, csr->rowptr, sizeof(int32_t) * (csr->m + 1), hipMemcpyDeviceToDevice, ctx->stream));
        if (csr->nnz > 0)
            INTI_HIP(ctx, hipMemcpyAsync(c->colind, csr->colind, sizeof(int32_t) * csr->nnz, hipMemcpyDeviceToDevice, ctx->stream));
        INTI_HIP(ctx, hipMemcpyAsync(dt[0], tptr.data(), tb[0], hipMemcpyHostToDevice, ctx->stream));
        if (nnz > 0) {
            INTI_HIP(ctx, hipMemcpyAsync(dt[1], trow.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, ctx->stream));
            INTI_HIP(ctx, hipMemcpyAsync(dt[2], tperm.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, ctx->stream));
        }
        // the host vectors above must outlive the asynchronous uploads
        INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (csr->nnz > 0) {
        hipLaunchKernelGGL(k_axpy_vals, dim3((unsigned)((csr->nnz + 255) / 256)), dim3(256), 0, ctx->stream,
                           csr->nnz, cplx, csr->val, coef_re, coef_im, op->corr->val, accumulate);
        INTI_HIP(ctx, hipGetLastError());
    }
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

int inti_op_detach_correction(inti_op *op) {
    if (!op) return INTI_ERR_ARG;
    LOCK(op->ctx);
    (void)hipStreamSynchronize(op->ctx->stream);
    op_free_correction(op);
    return INTI_OK;
}

// ---------------------------------------------------------- DIM correction
int inti_bdim_build(inti_op *opS, inti_op *opD, int derivative, int nq, int64_t ne,
                    const int64_t *qtags, const int64_t *near_ptr, const int64_t *near_idx,
                    const double *mu, const double *params, inti_csr **dS, inti_csr **dD,
                    double *diag) {
    if (!opS) return INTI_ERR_ARG;
    Ctx *ctx = opS->ctx;
    REQUIRE(ctx, opD && qtags && mu && dS && dD, "bdim: null argument");
    REQUIRE(ctx, opD->ctx == ctx, "bdim: operators on different contexts");
    REQUIRE(ctx, opS->trg == opD->trg && opS->src == opD->src, "bdim: S and D must share target and source");
    REQUIRE(ctx, opS->pde == opD->pde && opS->dim == opD->dim && opS->k == opD->k && opS->kimag == opD->kimag,
            "bdim: S and D must be the same PDE");
    const int kS = derivative ? INTI_ADL : INTI_SL, kD = derivative ? INTI_HS : INTI_DL;
    REQUIRE(ctx, opS->kind == kS && opD->kind == kD,
            "bdim: expects (SL, DL) operators, or (ADL, HS) with derivative = 1");
    REQUIRE(ctx, nq > 0 && ne > 0, "bdim: empty element set");
    REQUIRE(ctx, (near_ptr == nullptr) == (near_idx == nullptr), "bdim: near_ptr and near_idx go together");
    REQUIRE(ctx, params == nullptr || (params[0] > 0 && params[1] > 0), "bdim: bad parameters");
    LOCK(ctx);
    Csr *s = nullptr, *d = nullptr;
    int st = bdim_build(opS, opD, derivative, nq, ne, qtags, near_ptr, near_idx, mu, params, &s, &d, diag);
    if (st) return st;
    *dS = static_cast<inti_csr *>(s);
    *dD = static_cast<inti_csr *>(d);
    return INTI_OK;
}

int inti_bdim_build_types(inti_op *opS, inti_op *opD, int derivative, int ntypes, const int *nq,
                          const int64_t *ne, const int64_t *qtags, const int64_t *near_ptr,
                          const int64_t *near_idx, const double *mu, const double *params, inti_csr **dS,
                          inti_csr **dD, double *diag) {
    if (!opS) return INTI_ERR_ARG;
    Ctx *ctx = opS->ctx;
    REQUIRE(ctx, opD && nq && ne && qtags && mu && dS && dD, "bdim: null argument");
    REQUIRE(ctx, opD->ctx == ctx, "bdim: operators on different contexts");
    REQUIRE(ctx, opS->trg == opD->trg && opS->src == opD->src, "bdim: S and D must share target and source");
    REQUIRE(ctx, opS->pde == opD->pde && opS->dim == opD->dim && opS->k == opD->k && opS->kimag == opD->kimag,
            "bdim: S and D must be the same PDE");
    const int kS = derivative ? INTI_ADL : INTI_SL, kD = derivative ? INTI_HS : INTI_DL;
    REQUIRE(ctx, opS->kind == kS && opD->kind == kD,
            "bdim: expects (SL, DL) operators, or (ADL, HS) with derivative = 1");
    REQUIRE(ctx, ntypes > 0, "bdim: no element type");
    int64_t total = 0;
    for (int t = 0; t < ntypes; ++t) {
        REQUIRE(ctx, nq[t] > 0 && ne[t] >= 0, "bdim: bad element type sizes");
        total += ne[t];
    }
    REQUIRE(ctx, total > 0, "bdim: empty element set");
    REQUIRE(ctx, (near_ptr == nullptr) == (near_idx == nullptr), "bdim: near_ptr and near_idx go together");
    REQUIRE(ctx, params == nullptr || (params[0] > 0 && params[1] > 0), "bdim: bad parameters");
    LOCK(ctx);
    Csr *s = nullptr, *d = nullptr;
    int st = bdim_build_types(opS, opD, derivative, ntypes, nq, ne, qtags, near_ptr, near_idx, mu, params, &s,
                              &d, diag);
    if (st) return st;
    *dS = static_cast<inti_csr *>(s);
    *dD = static_cast<inti_csr *>(d);
    return INTI_OK;
}

int inti_lsq_build(inti_ctx *ctx, int is_complex, int nrows, int ncols, int64_t ne, const double *A,
                   const int64_t *cols, int64_t m, int64_t n, const int64_t *near_ptr,
                   const int64_t *near_idx, const double *rhs, inti_csr **out, double *diag) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, A && cols && rhs && out, "lsq: null argument");
    REQUIRE(ctx, ne > 0 && m > 0 && n > 0, "lsq: empty problem");
    REQUIRE(ctx, (near_ptr == nullptr) == (near_idx == nullptr), "lsq: near_ptr and near_idx go together");
    LOCK(ctx);
    Csr *c = nullptr;
    int st = lsq_build(ctx, is_complex ? 1 : 0, nrows, ncols, ne, A, cols, m, n, near_ptr, near_idx, rhs, &c, diag);
    if (st) return st;
    *out = static_cast<inti_csr *>(c);
    return INTI_OK;
}

int inti_lsq_build_types(inti_ctx *ctx, int is_complex, int nrows, int ntypes, const int *ncols, const int64_t *ne,
                         const double *A, const int64_t *cols, int64_t m, int64_t n, const int64_t *near_ptr,
                         const int64_t *near_idx, const double *rhs, inti_csr **out, double *diag) {
    if (!ctx) return INTI_ERR_ARG;
    REQUIRE(ctx, ncols && ne && A && cols && rhs && out, "lsq: null argument");
    REQUIRE(ctx, ntypes > 0 && m > 0 && n > 0, "lsq: empty problem");
    int64_t total = 0;
    for (int t = 0; t < ntypes; ++t) {
        REQUIRE(ctx, ne[t] >= 0, "lsq: negative element count");
        total += ne[t];
    }
    REQUIRE(ctx, total > 0, "lsq: empty problem");
    REQUIRE(ctx, (near_ptr == nullptr) == (near_idx == nullptr), "lsq: near_ptr and near_idx go together");
    LOCK(ctx);
    Csr *c = nullptr;
    int st = lsq_build_types(ctx, is_complex ? 1 : 0, nrows, ntypes, ncols, ne, A, cols, m, n, near_ptr, near_idx,
                             rhs, &c, diag);
    if (st) return st;
    *out = static_cast<inti_csr *>(c);
    return INTI_OK;
}

// ----------------------------------------------------------------- solvers
// host pointers: stage b and x on the device for the solve
struct Staged {
    void *b = nullptr, *x = nullptr;
    ~Staged() {
        if (b) (void)hipFree(b);
        if (x) (void)hipFree(x);
    }
};

static int stage_in(Ctx *ctx, Staged &S, size_t bytes, const double *b, const double *x) {
    if (hipMalloc(&S.b, bytes) != hipSuccess || hipMalloc(&S.x, bytes) != hipSuccess)
        return set_error(ctx, INTI_ERR_NOMEM, "solve: allocation failed");
    INTI_HIP(ctx, hipMemcpyAsync(S.b, b, bytes, hipMemcpyHostToDevice, ctx->stream));
    INTI_HIP(ctx, hipMemcpyAsync(S.x, x, bytes, hipMemcpyHostToDevice, ctx->stream));
    return INTI_OK;
}

int inti_gmres_prec(inti_op *op, inti_op *prec, int on_device, const double *b, double *x, int restart,
                    int maxiter, double rtol, double atol, int use_graph, int *iters, double *relres) {
    if (!op) return INTI_ERR_ARG;
    Ctx *ctx = op->ctx;
    REQUIRE(ctx, b && x, "gmres: null vector");
    REQUIRE(ctx, rtol >= 0 && atol >= 0, "gmres: negative tolerance");
    LOCK(ctx);
    if (on_device) return gmres_solve(op, prec, b, x, restart, maxiter, rtol, atol, use_graph, iters, relres);
    const size_t bytes = sizeof(double) * (op_is_complex(op) ? 2 : 1) * op->src->n;
    Staged S;
    int st = stage_in(ctx, S, bytes, b, x);
    if (st) return st;
    st = gmres_solve(op, prec, (const double *)S.b, (double *)S.x, restart, maxiter, rtol, atol, use_graph, iters,
                     relres);
    if (st) return st;
    INTI_HIP(ctx, hipMemcpyAsync(x, S.x, bytes, hipMemcpyDeviceToHost, ctx->stream));
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

int inti_gmres(inti_op *op, int on_device, const double *b, double *x, int restart, int maxiter,
               double rtol, double atol, int use_graph, int *iters, double *relres) {
    return inti_gmres_prec(op, nullptr, on_device, b, x, restart, maxiter, rtol, atol, use_graph, iters, relres);
}

int inti_refine_solve_prec(inti_op *op_hi, inti_op *op_lo, inti_op *prec, int on_device, const double *b,
                           double *x, int restart, int maxiter_inner, double inner_rtol, int max_outer,
                           double rtol, int use_graph, int *outer, int *inner_total, double *relres) {
    if (!op_hi || !op_lo) return INTI_ERR_ARG;
    Ctx *ctx = op_hi->ctx;
    REQUIRE(ctx, op_lo->ctx == ctx, "refine: operators on different contexts");
    REQUIRE(ctx, b && x, "refine: null vector");
    REQUIRE(ctx, restart >= 1 && maxiter_inner >= 1 && max_outer >= 0, "refine: bad iteration limits");
    LOCK(ctx);
    if (on_device)
        return refine_solve(op_hi, op_lo, prec, b, x, restart, maxiter_inner, inner_rtol, max_outer, rtol, use_graph,
                            outer, inner_total, relres);
    const size_t bytes = sizeof(double) * (op_is_complex(op_hi) ? 2 : 1) * op_hi->src->n;
    Staged S;
    int st = stage_in(ctx, S, bytes, b, x);
    if (st) return st;
    st = refine_solve(op_hi, op_lo, prec, (const double *)S.b, (double *)S.x, restart, maxiter_inner, inner_rtol,
                      max_outer, rtol, use_graph, outer, inner_total, relres);
    if (st) return st;
    INTI_HIP(ctx, hipMemcpyAsync(x, S.x, bytes, hipMemcpyDeviceToHost, ctx->stream));
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

int inti_refine_solve(inti_op *op_hi, inti_op *op_lo, int on_device, const double *b, double *x,
                      int restart, int maxiter_inner, double inner_rtol, int max_outer, double rtol,
                      int use_graph, int *outer, int *inner_total, double *relres) {
    return inti_refine_solve_prec(op_hi, op_lo, nullptr, on_device, b, x, restart, maxiter_inner, inner_rtol,
                                  max_outer, rtol, use_graph, outer, inner_total, relres);
}

// --------------------------------------------------------------------- CSR
int inti_csr_info(const inti_csr *csr, int64_t *m, int64_t *n, int64_t *nnz, int *is_complex) {
    if (!csr) return INTI_ERR_ARG;
    if (m) *m = csr->m;
    if (n) *n = csr->n;
    if (nnz) *nnz = csr->nnz;
    if (is_complex) *is_complex = csr->is_complex;
    return INTI_OK;
}

int inti_csr_export(const inti_csr *csr, int64_t *rowptr, int64_t *colind, double *val) {
    if (!csr) return INTI_ERR_ARG;
    Ctx *ctx = csr->ctx;
    LOCK(ctx);
    std::vector<int32_t> rp(csr->m + 1), ci(csr->nnz);
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    INTI_HIP(ctx, hipMemcpy(rp.data(), csr->rowptr, sizeof(int32_t) * (csr->m + 1), hipMemcpyDeviceToHost));
    if (csr->nnz > 0) {
        INTI_HIP(ctx, hipMemcpy(ci.data(), csr->colind, sizeof(int32_t) * csr->nnz, hipMemcpyDeviceToHost));
        if (val)
            INTI_HIP(ctx, hipMemcpy(val, csr->val, sizeof(double) * (csr->is_complex ? 2 : 1) * csr->nnz,
                                    hipMemcpyDeviceToHost));
    }
    if (rowptr) for (int64_t i = 0; i <= csr->m; ++i) rowptr[i] = rp[i];
    if (colind) for (int64_t p = 0; p < csr->nnz; ++p) colind[p] = ci[p];
    return INTI_OK;
}

int inti_csr_destroy(inti_csr *csr) {
    if (!csr) return INTI_ERR_ARG;
    Ctx *ctx = csr->ctx;
    {
        LOCK(ctx);
        (void)hipStreamSynchronize(ctx->stream);
    }
    csr_free(csr);
    return INTI_OK;
}

}  // extern "C"
