// Dense multi-right-hand-side product Y = alpha A X + beta Y with an
// assembled matrix (the dense `S * sigma` of src/api.jl:171-173 applied to
// many densities at once, and the dense Theta = S B1 - D B0 of
// src/bdim.jl:116-118): a real GEMM on the FP64 matrix cores.
//
// A complex matrix stored interleaved (column-major, the memory of Julia's
// Matrix{ComplexF64}) is, read as reals, the 2m x n matrix At with rows
// (Re a_i, Im a_i) alternating. With B = [Re X | Im X] (n x 2R),
// Z = At B (2m x 2R) holds every product needed:
//   Re(AX)_i = Z[2i, r] - Z[2i+1, R+r],   Im(AX)_i = Z[2i, R+r] + Z[2i+1, r].
// A real matrix is its own At, B = X.
//
// Kernel: v_mfma_f64_16x16x4f64 (16 x 16 += 16 x 4 * 4 x 16 per wavefront).
// A block of 4 wavefronts owns 128 rows of At, 32 columns of B and one slice
// of the K range; each wavefront keeps a 32 x 32 tile of Z in four
// accumulators, streams its two 16 x 4 fragments of At straight from HBM
// (each element of At is read once per 32 columns of B) and reads the 4 x 16
// fragments of B from an LDS chunk of 128 rows shared by the block. The K
// range is split across blocks (partial products summed in the epilogue)
// until the grid holds several blocks per CU: streaming At at HBM speed
// needs many wavefronts in flight, and a narrow product (16 right-hand
// sides: 32 columns) alone gives one block per 128 rows.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace inti {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kGmRows = 128;  // rows of At per block (4 wavefronts x 32)

struct GemmArgs {
    int64_t M, K, N;  // At is M x K (lda), B is K x N (ldb), Z is M x N (ldz); N % (16 NC) == 0
    int64_t kslice;   // K range of one block (a multiple of the LDS chunk); Z slices of M x N per block.z
    const double *A;
    int64_t lda;
    const double *B;
    int64_t ldb;
    double *Z;
    int64_t ldz;
};

// NC 16-column fragments per wavefront (32 x 16 NC tile of Z). The K range
// is walked in chunks of KC = 32: while the matrix cores work on chunk t
// (its At fragments in registers, its B rows in one LDS buffer), the loads of
// chunk t + 1 are in flight (At fragments to registers, B rows to registers
// and then to the other LDS buffer), so HBM latency hides behind 8 k-steps
// of MFMA work.
template <int NC>
__global__ __launch_bounds__(256) void k_gemm_f64(GemmArgs g) {
    constexpr int COLS = 16 * NC;
    constexpr int KC = 32, KS = KC / 4;        // k-steps of 4 per chunk
    constexpr int BPT = KC * COLS / 256;       // B values staged per thread
    __shared__ double sB[2][KC][COLS + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;  // fragment row / k index of this lane
    const int64_t r0 = (int64_t)blockIdx.x * kGmRows + wave * 32;
    const int64_t c0 = (int64_t)blockIdx.y * COLS;
    v4d acc0[NC], acc1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        acc0[c] = v4d{0, 0, 0, 0};
        acc1[c] = v4d{0, 0, 0, 0};
    }
    const int64_t ra = r0 + li, rb = r0 + 16 + li;
    const bool oka = ra < g.M, okb = rb < g.M;
    const int64_t kbeg = (int64_t)blockIdx.z * g.kslice;
    const int64_t kend = kbeg + g.kslice < g.K ? kbeg + g.kslice : g.K;
    double a0[KS], a1[KS], n0[KS], n1[KS], bq[BPT];
    auto load_a = [&](int64_t k0, double (&x0)[KS], double (&x1)[KS]) {
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            const int64_t k = k0 + 4 * q + lk;
            const bool okk = k < kend;
            x0[q] = (oka && okk) ? g.A[ra + k * g.lda] : 0.0;
            x1[q] = (okb && okk) ? g.A[rb + k * g.lda] : 0.0;
        }
    };
    auto load_b = [&](int64_t k0) {  // B rows k0 .. k0 + KC (zero beyond the slice)
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int e = tid + 256 * q;
            const int kk = e % KC, cc = e / KC;
            const int64_t k = k0 + kk;
            bq[q] = k < kend ? g.B[k + (c0 + cc) * g.ldb] : 0.0;
        }
    };
    auto store_b = [&](int buf) {
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int e = tid + 256 * q;
            sB[buf][e % KC][e / KC] = bq[q];
        }
    };
    int buf = 0;
    if (kbeg < kend) {
        load_b(kbeg);
        load_a(kbeg, a0, a1);
        store_b(0);
    }
    __syncthreads();
    for (int64_t k0 = kbeg; k0 < kend; k0 += KC) {
        const bool more = k0 + KC < kend;
        if (more) {
            load_b(k0 + KC);
            load_a(k0 + KC, n0, n1);
        }
        // the slice tail: rows of B beyond it are zero in LDS and the
        // matching loads of At are masked, so the last chunk adds zeros
#pragma unroll
        for (int q = 0; q < KS; ++q) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const double b = sB[buf][4 * q + lk][16 * c + li];
                acc0[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[q], b, acc0[c], 0, 0, 0);
                acc1[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b, acc1[c], 0, 0, 0);
            }
        }
        if (more) {
            store_b(buf ^ 1);
#pragma unroll
            for (int q = 0; q < KS; ++q) {
                a0[q] = n0[q];
                a1[q] = n1[q];
            }
        }
        __syncthreads();
        buf ^= 1;
    }
    // accumulator layout (measured): element v of lane l is Z[4 v + l / 16][l % 16]
    // of the 16 x 16 tile
    double *Z = g.Z + (int64_t)blockIdx.z * g.ldz * g.N;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t ia = r0 + 4 * v + lk, ib = r0 + 16 + 4 * v + lk;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int64_t col = c0 + 16 * c + li;
            if (ia < g.M) Z[ia + col * g.ldz] = acc0[c][v];
            if (ib < g.M) Z[ib + col * g.ldz] = acc1[c][v];
        }
    }
}

// B (K x N, ld K) = [Re X | Im X | 0] or [X | 0]
__global__ __launch_bounds__(256) void k_gemm_pack(int64_t K, int R, int cplx, int64_t N, const double *X,
                                                   int64_t ldx, double *B) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= K * N) return;
    const int64_t k = e % K, c = e / K;
    double v = 0.0;
    if (cplx) {
        if (c < R) v = X[2 * (k + c * ldx)];
        else if (c < 2 * R) v = X[2 * (k + (c - R) * ldx) + 1];
    } else if (c < R) {
        v = X[k + c * ldx];
    }
    B[e] = v;
}

// Y = alpha (A X) + beta Y from the nsl slices of Z (see the file comment)
__global__ __launch_bounds__(256) void k_gemm_combine(int64_t m, int R, int cplx, const double *Z, int64_t ldz,
                                                      int64_t zslice, int nsl, double *Y, int64_t ldy, double ar,
                                                      double ai, double br, double bi, int beta_zero) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * R) return;
    const int64_t i = e % m;
    const int r = (int)(e / m);
    auto z = [&](int64_t row, int64_t col) {
        double t = 0.0;
        for (int q = 0; q < nsl; ++q) t += Z[q * zslice + row + col * ldz];
        return t;
    };
    if (cplx) {
        const double pr = z(2 * i, r) - z(2 * i + 1, R + r);
        const double pi = z(2 * i, R + r) + z(2 * i + 1, r);
        double yr = ar * pr - ai * pi, yi = ar * pi + ai * pr;
        double *y = Y + 2 * (i + r * ldy);
        if (!beta_zero) {
            yr += br * y[0] - bi * y[1];
            yi += br * y[1] + bi * y[0];
        }
        y[0] = yr;
        y[1] = yi;
    } else {
        double yv = ar * z(i, r);
        if (!beta_zero) yv += br * Y[i + r * ldy];
        Y[i + r * ldy] = yv;
    }
}

int dense_gemm(Ctx *ctx, int cplx, int64_t m, int64_t n, int nrhs, const double *A, int64_t lda, const double *X,
               int64_t ldx, double *Y, int64_t ldy, const double alpha[2], const double beta[2]) {
    if (m == 0 || nrhs == 0) return INTI_OK;
    const int cols = cplx ? 2 * nrhs : nrhs;
    // column tile: 32, 64 or 128 columns per block (one pass over At for up
    // to 64 complex right-hand sides)
    const int NC = cols <= 32 ? 2 : cols <= 64 ? 4 : 8;
    const int KC = 32;  // the kernel's chunk
    const int COLS = 16 * NC;
    const int64_t N = (cols + COLS - 1) / COLS * COLS;
    const int64_t M = cplx ? 2 * m : m;
    const int64_t K = n;
    // K slices of whole LDS chunks. Blocks run in rounds of `cap` resident
    // blocks (occupancy x CUs), so a grid just past a multiple of cap leaves
    // the last round nearly empty; every slice also writes and re-reads an
    // M x N partial product. The slice count minimises the modelled traffic
    //   A bytes x (rounds x cap / blocks) + 16 M N nsl.
    const int64_t base_blocks = (M + kGmRows - 1) / kGmRows * (N / COLS);
    const int64_t kchunks = std::max<int64_t>(1, (K + KC - 1) / KC);
    int per_cu = 0;
    const void *kfn = NC == 2 ? (const void *)k_gemm_f64<2> : NC == 4 ? (const void *)k_gemm_f64<4>
                                                                        : (const void *)k_gemm_f64<8>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    const double cap = (double)per_cu * ctx->num_cu;
    const double abytes = 8.0 * (double)M * (double)K;
    int64_t nsl = 1, kslice = kchunks * KC;
    double best = -1.0;
    for (int64_t t = 1; t <= std::min<int64_t>(kchunks, 64); ++t) {
        const int64_t ks = (kchunks + t - 1) / t * KC;
        const int64_t ns = std::max<int64_t>(1, (K + ks - 1) / ks);
        // the partials stay well below the matrix itself
        if (ns > 1 && 8.0 * (double)M * (double)N * (double)ns > std::max(abytes / 4.0, 64.0e6)) break;
        const double blocks = (double)(base_blocks * ns);
        const double rounds = std::ceil(blocks / cap);
        const double cost = abytes * rounds * cap / blocks + 16.0 * (double)M * (double)N * (double)ns;
        if (best < 0.0 || cost < best) {
            best = cost;
            nsl = ns;
            kslice = ks;
        }
    }
    const size_t bbytes = sizeof(double) * (size_t)std::max<int64_t>(K, 1) * N;
    const size_t zbytes = sizeof(double) * (size_t)M * N * nsl;
    void *scr = nullptr;
    int rc = ctx_scratch(ctx, kScrTmp, bbytes + zbytes, &scr);
    if (rc) return rc;
    double *B = (double *)scr;
    double *Z = B + (size_t)std::max<int64_t>(K, 1) * N;
    hipStream_t s = ctx->stream;
    if (K > 0) {
        hipLaunchKernelGGL(k_gemm_pack, dim3((unsigned)((K * N + 255) / 256)), dim3(256), 0, s, K, nrhs, cplx, N, X,
                           ldx, B);
    }
    GemmArgs g;
    g.M = M; g.K = K; g.N = N; g.kslice = kslice;
    g.A = A; g.lda = cplx ? 2 * lda : lda;
    g.B = B; g.ldb = std::max<int64_t>(K, 1);
    g.Z = Z; g.ldz = M;
    const dim3 grid((unsigned)((M + kGmRows - 1) / kGmRows), (unsigned)(N / COLS), (unsigned)(K > 0 ? nsl : 1));
    if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev0, s));
    if (NC == 2) hipLaunchKernelGGL((k_gemm_f64<2>), grid, dim3(256), 0, s, g);
    else if (NC == 4) hipLaunchKernelGGL((k_gemm_f64<4>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((k_gemm_f64<8>), grid, dim3(256), 0, s, g);
    if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev1, s));
    hipLaunchKernelGGL(k_gemm_combine, dim3((unsigned)((m * nrhs + 255) / 256)), dim3(256), 0, s, m, nrhs, cplx,
                       (const double *)Z, M, M * N, (int)(K > 0 ? nsl : 1), Y, ldy, alpha[0], alpha[1], beta[0],
                       beta[1], beta[0] == 0.0 && beta[1] == 0.0 ? 1 : 0);
    INTI_HIP(ctx, hipGetLastError());
    return INTI_OK;
}

}  // namespace inti
