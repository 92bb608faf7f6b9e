// Issue-rate microbenchmarks of the instructions the direct sums are built
// from: FP64 and FP32 (non-packed) FMA and the transcendental unit's FP32
// rsq / sin / cos / log / rcp, plus FP64 rsq. Each lane runs 8 independent
// dependency chains of one instruction, so the rate is the unit's
// throughput, not its latency. The result is lane-operations per second
// over the whole GPU: the VALU and transcendental roofs of the roofline.
#include "common.hpp"

namespace inti {

constexpr int kMbChains = 8;
constexpr int kMbBS = 256;

template <int OP>
__global__ __launch_bounds__(kMbBS) void k_microbench(int iters, double *out) {
    const int t = blockIdx.x * kMbBS + threadIdx.x;
    if constexpr (OP == 0) {  // v_fma_f64
        double a[kMbChains];
        const double b = 1.0 - 1e-9 * (t & 7), c = 1e-12;
#pragma unroll
        for (int k = 0; k < kMbChains; ++k) a[k] = 1.0 + k * 1e-3;
#pragma unroll 4
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int k = 0; k < kMbChains; ++k) a[k] = __builtin_fma(a[k], b, c);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < kMbChains; ++k) s += a[k];
        out[t] = s;
    } else {
        float a[kMbChains];
#pragma unroll
        for (int k = 0; k < kMbChains; ++k) a[k] = 0.5f + 0.01f * k + 1e-4f * (t & 7);
        const float b = 1.0f - 1e-6f * (t & 3), c = 1e-7f;
#pragma unroll 4
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int k = 0; k < kMbChains; ++k) {
                if constexpr (OP == 1) a[k] = __builtin_fmaf(a[k], b, c);              // v_fma_f32
                else if constexpr (OP == 2) a[k] = __builtin_amdgcn_rsqf(a[k]);       // v_rsq_f32
                else if constexpr (OP == 3) a[k] = __builtin_amdgcn_sinf(a[k]);       // v_sin_f32
                else if constexpr (OP == 4) a[k] = __builtin_amdgcn_cosf(a[k]);       // v_cos_f32
                else if constexpr (OP == 5) a[k] = __builtin_amdgcn_logf(a[k] + 2.0f); // v_log_f32 (+ add)
                else if constexpr (OP == 6) a[k] = __builtin_amdgcn_rcpf(a[k]);       // v_rcp_f32
            }
        }
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < kMbChains; ++k) s += a[k];
        out[t] = (double)s;
    }
}

template <>
__global__ __launch_bounds__(kMbBS) void k_microbench<7>(int iters, double *out) {  // v_rsq_f64
    const int t = blockIdx.x * kMbBS + threadIdx.x;
    double a[kMbChains];
#pragma unroll
    for (int k = 0; k < kMbChains; ++k) a[k] = 0.5 + 0.01 * k + 1e-6 * (t & 7);
#pragma unroll 4
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < kMbChains; ++k) a[k] = __builtin_amdgcn_rsq(a[k]);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kMbChains; ++k) s += a[k];
    out[t] = s;
}

int microbench(Ctx *ctx, int op, double *rate) {
    using Fn = void (*)(int, double *);
    static const Fn fns[] = {k_microbench<0>, k_microbench<1>, k_microbench<2>, k_microbench<3>,
                             k_microbench<4>, k_microbench<5>, k_microbench<6>, k_microbench<7>};
    if (op < 0 || op >= 8) return set_error(ctx, INTI_ERR_ARG, "microbench: op out of range");
    const int blocks = 8 * ctx->num_cu;  // 32 waves per CU: 8 per SIMD
    const int64_t threads = (int64_t)blocks * kMbBS;
    void *out = nullptr;
    int st = ctx_scratch(ctx, kScrTmp, sizeof(double) * threads, &out);
    if (st) return st;
    hipStream_t s = ctx->stream;
    const int iters = op == 0 || op == 1 ? 65536 : 16384;  // a few ms per launch
    // warm-up, then the timed launch
    hipLaunchKernelGGL(fns[op], dim3(blocks), dim3(kMbBS), 0, s, iters / 8, (double *)out);
    INTI_HIP(ctx, hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(fns[op], dim3(blocks), dim3(kMbBS), 0, s, iters, (double *)out);
    INTI_HIP(ctx, hipEventRecord(ctx->ev1, s));
    INTI_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    INTI_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *rate = (double)threads * (double)iters * kMbChains / (ms * 1e-3);
    return INTI_OK;
}

}  // namespace inti
