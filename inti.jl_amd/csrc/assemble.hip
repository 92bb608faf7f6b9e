// Dense assembly of an IntegralOperator (Inti.jl src/nystrom.jl:103-124):
// out[i + m*j] = K(x_i, y_j) * w_j, column-major, one thread per entry with
// i fastest so that the stores are coalesced.
#include "common.hpp"
#include "pair.hpp"

namespace inti {

struct AsmArgs {
    int64_t m, n;
    const double *tx, *tn, *sx, *sn, *sw;
    double *out;
    KParams p;
};

template <int PDE, int DIM, int KIND>
__global__ __launch_bounds__(256) void k_assemble(AsmArgs a) {
    constexpr bool CPLX = PDE != INTI_LAPLACE;
    const int64_t total = a.m * a.n;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int64_t i = e % a.m, j = e / a.m;
        double x[DIM], y[DIM], nx[DIM], ny[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            x[d] = a.tx[d * a.m + i];
            y[d] = a.sx[d * a.n + j];
            nx[d] = a.tn ? a.tn[d * a.m + i] : 0.0;
            ny[d] = a.sn ? a.sn[d * a.n + j] : 0.0;
        }
        double kr, ki;
        pair_raw_xy<PDE, DIM, KIND, false>(x, nx, y, ny, a.p, kr, ki);
        {
            const double tr = a.p.postr * kr - a.p.posti * ki;
            ki = a.p.postr * ki + a.p.posti * kr;
            kr = tr;
        }
        const double w = a.sw[j];
        if constexpr (CPLX) {
            a.out[2 * e] = kr * w;
            a.out[2 * e + 1] = ki * w;
        } else {
            a.out[e] = kr * w;
        }
    }
}

using AsmFn = void (*)(const AsmArgs &, dim3, hipStream_t);
template <int PDE, int DIM, int KIND>
static void asm_launcher(const AsmArgs &a, dim3 g, hipStream_t s) {
    hipLaunchKernelGGL((k_assemble<PDE, DIM, KIND>), g, dim3(256), 0, s, a);
}
template <int PDE, int DIM>
static AsmFn asm_kind(int kind) {
    switch (kind) {
        case INTI_SL: return asm_launcher<PDE, DIM, INTI_SL>;
        case INTI_DL: return asm_launcher<PDE, DIM, INTI_DL>;
        case INTI_ADL: return asm_launcher<PDE, DIM, INTI_ADL>;
        case INTI_HS: return asm_launcher<PDE, DIM, INTI_HS>;
        case INTI_COMB_SD: return asm_launcher<PDE, DIM, INTI_COMB_SD>;
        case INTI_COMB_AH: return asm_launcher<PDE, DIM, INTI_COMB_AH>;
    }
    if constexpr (PDE == INTI_HELMHOLTZ) {
        switch (kind) {
            case INTI_FF_SL: return asm_launcher<PDE, DIM, INTI_FF_SL>;
            case INTI_FF_DL: return asm_launcher<PDE, DIM, INTI_FF_DL>;
            case INTI_FF_COMB: return asm_launcher<PDE, DIM, INTI_FF_COMB>;
        }
    }
    return nullptr;
}

int launch_assemble(const Op *op, double *out_dev) {
    AsmFn fn = nullptr;
    int kind = op->kind;
    const KParams kp = op_params(op, &kind);
    if (op->pde == INTI_LAPLACE && op->dim == 2) fn = asm_kind<INTI_LAPLACE, 2>(kind);
    if (op->pde == INTI_LAPLACE && op->dim == 3) fn = asm_kind<INTI_LAPLACE, 3>(kind);
    if (op->pde == INTI_HELMHOLTZ && op->dim == 2) fn = asm_kind<INTI_HELMHOLTZ, 2>(kind);
    if (op->pde == INTI_HELMHOLTZ && op->dim == 3) fn = asm_kind<INTI_HELMHOLTZ, 3>(kind);
    if (op_kernel_pde(op) == kHelmCK) fn = op->dim == 3 ? asm_kind<kHelmCK, 3>(kind) : nullptr;
    if (!fn) return set_error(op->ctx, INTI_ERR_UNSUPPORTED, "no assembly kernel for this operator");
    AsmArgs a;
    a.m = op->trg->n; a.n = op->src->n;
    a.tx = op->trg->coords; a.tn = op->trg->normals;
    a.sx = op->src->coords; a.sn = op->src->normals; a.sw = op->src->weights;
    a.out = out_dev;
    a.p = kp;
    const int64_t total = a.m * a.n;
    if (total == 0) return INTI_OK;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    fn(a, dim3((unsigned)blocks), op->ctx->stream);
    INTI_HIP(op->ctx, hipGetLastError());
    return INTI_OK;
}

}  // namespace inti
