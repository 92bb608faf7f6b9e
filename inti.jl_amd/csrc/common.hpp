// Shared host/device definitions of the inti_amd library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/inti_amd.h"

namespace inti {

// Two points closer than this are "the same point" and the kernel is zero
// (Inti.jl src/utils.jl:235, SAME_POINT_TOLERANCE = 1e-14, compared on the
// distance; here on its square).
constexpr double kSamePointTol2 = 1.0e-28;
constexpr double kPi = 3.14159265358979323846;
constexpr double kInv2Pi = 0.15915494309189533577;
constexpr double kInv4Pi = 0.079577471545947667884;

// ---------------------------------------------------------------- host side
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string err;
    int profiling = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_kernel_ms = 0.f;
    int num_cu = 256;
    // scratch slots (grown on demand, freed at destroy): 0 partial sums,
    // 1-2 staging of host operands, 3 correction-build temporaries
    void *scratch[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t scratch_bytes[4] = {0, 0, 0, 0};
    // device address of the library's static partial-sum buffer on this
    // context's device (the allocation-free apply)
    double *noalloc_scratch = nullptr;
};

struct Quad {
    Ctx *ctx = nullptr;
    int dim = 0;
    int64_t n = 0;
    double *coords = nullptr;   // dim*n SoA
    double *normals = nullptr;  // dim*n SoA or null
    double *weights = nullptr;  // n
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounding box
    // no two nodes within 2x the coincidence tolerance: pairs of different
    // nodes then never need the coincident-point mask
    bool distinct = false;
};

// CSR with 0-based int32 indices; values real (double) or complex
// (interleaved doubles).
struct Csr {
    Ctx *ctx = nullptr;
    int64_t m = 0, n = 0, nnz = 0;
    int is_complex = 0;
    int32_t *rowptr = nullptr;  // m+1
    int32_t *colind = nullptr;  // nnz
    double *val = nullptr;      // nnz or 2*nnz
};

struct Op {
    Ctx *ctx = nullptr;
    int pde = 0, dim = 0, kind = 0, precision = 0;
    double k = 0.0;
    double kimag = 0.0;  // imaginary part of k (3D Helmholtz, FP64, one-direction sums)
    double coefs[4] = {1, 0, 1, 0};  // c0, c1 (complex) of the combined kinds
    double cid[2] = {0, 0};          // identity coefficient (square ops)
    const Quad *trg = nullptr, *src = nullptr;
    Csr *corr = nullptr;  // owned merged correction (may be null)
    // the correction's transposed pattern (built at the first attach; the
    // pattern never changes while attached): column j's entries are
    // corr_trow[p] / corr->val[corr_tperm[p]] for p in [corr_tptr[j],
    // corr_tptr[j+1]), rows increasing. The adjoint gathers through it.
    int32_t *corr_tptr = nullptr, *corr_trow = nullptr, *corr_tperm = nullptr;
    // launch configuration
    int block = 256, tpt = 1, split = 0;  // split 0 = automatic
    // symmetric evaluation of square operators on one quadrature: 0 off,
    // 1 automatic (where the work fills the GPU), 2 wherever it applies
    int sym = 1;
    int sym_group = 0;  // tiles per work group; 0 = automatic
};

// error helpers: record message in ctx and return a status code
int set_error(Ctx *ctx, int code, const std::string &msg);

#define INTI_HIP(ctx, call)                                                     \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return ::inti::set_error((ctx), INTI_ERR_HIP,                       \
                                     std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// scratch memory of the context (grown on demand; never shrinks)
enum { kScrPartial = 0, kScrX = 1, kScrY = 2, kScrTmp = 3 };
int ctx_scratch(Ctx *ctx, int slot, size_t bytes, void **ptr);

// ------------------------------------------------------- kernel dispatchers
struct DirectLaunch {
    const Op *op;
    int nrhs;
    const double *x;  // device, n x nrhs (complex interleaved for Helmholtz)
    int64_t ldx;
    double *y;        // device, m x nrhs
    int64_t ldy;
    double alpha[2], beta[2];
    // caller-provided partial-sum buffer: when set, the one-direction sum
    // runs with its source split capped so that nr x split x m partials fit
    // part_bytes, and nothing is allocated
    double *part = nullptr;
    size_t part_bytes = 0;
};
int launch_apply(const DirectLaunch &L);
// allocate the scratch an apply of nrhs columns needs (before graph capture)
int apply_reserve(const Op *op, int nrhs);
int gmres_solve(Op *op, Op *prec, const double *b, double *x, int restart, int maxiter, double rtol,
                double atol, int use_graph, int *iters, double *relres);
int refine_solve(Op *op_hi, Op *op_lo, Op *prec, const double *b, double *x, int restart, int maxiter_inner,
                 double inner_rtol, int max_outer, double rtol, int use_graph, int *outer, int *inner_total,
                 double *relres);
int launch_assemble(const Op *op, double *out_dev);
void op_free_correction(Op *op);
int op_adjoint(Op *op, const double *ubar, const double *sigma, double *sigma_bar, double *ybar,
               double *nbar, double *wbar, double *xbar, const double *dcoefs, double *kbar);

bool op_is_complex(const Op *op);
// PDE code of the kernel templates: the complex-wavenumber 3D Helmholtz
// kernels (2) for a Helmholtz operator with kimag != 0, else op->pde
int op_kernel_pde(const Op *op);
int microbench(Ctx *ctx, int op, double *rate);
int dense_gemm(Ctx *ctx, int cplx, int64_t m, int64_t n, int nrhs, const double *A, int64_t lda, const double *X,
               int64_t ldx, double *Y, int64_t ldy, const double alpha[2], const double beta[2]);

}  // namespace inti

// The opaque handles of the C-ABI are these types.
struct inti_ctx : inti::Ctx {};
struct inti_quad : inti::Quad {};
struct inti_op : inti::Op {};
struct inti_csr : inti::Csr {};
