// Boundary density-interpolation (DIM) correction on the GPU
// (Inti.jl src/bdim.jl:50-196):
//   1. monopole sources z_k on a circle / Fibonacci sphere around the source
//      bounding box (bdim.jl:79-92, utils.jl:142-171), computed on the host;
//   2. traces B0[j,k] = G(y_j, z_k), B1[j,k] = ADL(y_j, z_k) (bdim.jl:94-103);
//   3. Theta = mu*G(x, z) + S*B1 - D*B0 for the corrected targets, a
//      matrix-free multi-right-hand-side direct sum (bdim.jl:106-118);
//   4. one wavefront per element: Householder QR of M^T (ns x 2nq) with the
//      reflector convention of the generic unblocked QR that
//      qr!(transpose(M)) runs (bdim.jl:163), then the least-squares solve of
//      every near target (bdim.jl:166-170), one lane per target;
//   5. the (target, node) triplets are summed into CSR (sparse() semantics,
//      bdim.jl:193-194): dD = W[1:nq], dS = -W[nq+1:2nq] (bdim.jl:175-182).
#include <algorithm>
#include <cmath>

#include "lsq.hpp"
#include "pair.hpp"

namespace inti {


// ----------------------------------------------------------------- traces
struct TraceArgs {
    int64_t n;
    int ns, nsp;  // monopoles, padded count (columns of B)
    const double *sx, *sn, *zs;
    double2 *B0, *B1;  // n x nsp, column-major
    KParams p;
};

template <int PDE, int DIM>
__global__ __launch_bounds__(256) void k_traces(TraceArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= a.n * a.nsp) return;
    const int64_t j = e % a.n;
    const int kk = (int)(e / a.n);
    double g0r = 0, g0i = 0, g1r = 0, g1i = 0;
    if (kk < a.ns) {
        double r[DIM], ny[DIM], dummy[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            r[d] = a.sx[d * a.n + j] - a.zs[d * a.ns + kk];
            ny[d] = a.sn[d * a.n + j];
            dummy[d] = 0.0;
        }
        pair_value<PDE, DIM, INTI_SL>(r, ny, dummy, a.p, g0r, g0i);
        pair_value<PDE, DIM, INTI_ADL>(r, ny, dummy, a.p, g1r, g1i);
    }
    a.B0[e] = make_double2(g0r, g0i);
    a.B1[e] = make_double2(g1r, g1i);
}

// ------------------------------------------------------------------ Theta
// Theta[l, k] = mu_i G(x_i, z_k) + sum_j S(x_i, y_j) B1[j, k] - D(x_i, y_j) B0[j, k]
// for the corrected targets i = glob[l]: an ns-right-hand-side direct sum
// with two kernels per pair. Every kernel value feeds 2 ns complex
// multiply-adds (8 ns FP64 FMAs), so the sum is FMA-bound on the VALU; the
// right-hand sides are the same for all targets of a wavefront, so they are
// read as wave-uniform scalar loads (SGPR operands of the FMAs: no LDS, no
// VGPRs for them) from a packed per-source record:
//   rec[j] = { y_j (DIM), n_j (DIM), then per monopole k:
//              post_S w_j B1[j,k], -post_D w_j B0[j,k] (complex: 4 doubles,
//              real for Laplace: 2) }
// so that the pair loop accumulates raw_S * rec + raw_D * rec only.
// A block is 64 targets (one per lane) x 4 wavefronts, each wavefront a
// quarter of the block's sources; the quarters are added in LDS in wave
// order (deterministic). With one source chunk (split = 1) the block writes
// Theta itself, mu G included; otherwise partial sums go to `part` and
// k_theta_reduce adds them. (Measured at 63,480 nodes, 3D Helmholtz, ns = 18:
// 4 waves and 3 source chunks 32.3 ms, 24 chunks 29.6 ms; 3 waves and one
// chunk, all 992 blocks nominally resident, 46.6 ms.)
constexpr int kThetaWaves = 4;
constexpr int kThetaRed = 8;  // doubles per lane per LDS reduction round

typedef const __attribute__((address_space(4))) double uniform_double;

template <int PDE>
constexpr int theta_per_k() { return PDE != INTI_LAPLACE ? 4 : 2; }

struct ThetaPackArgs {
    int64_t n;
    int ns, nsp, nst, rl;  // monopoles, trace columns, record columns, record length
    const double *sx, *sn, *sw;
    const double2 *B0, *B1;  // n x nsp, column-major
    double *rec;             // n x rl
    double psr, psi, pdr, pdi;  // post factors of the S and D kernels
};

// column kk of node j's record (kk == nst: the node's coordinates and
// normal); node arrays and trace columns have leading dimension ld
template <int PDE, int DIM>
__device__ __forceinline__ void theta_pack_one(int64_t j, int kk, int64_t ld, int ns, int nst, double *R,
                                               const double *sx, const double *sn, const double *sw,
                                               const double2 *B0, const double2 *B1, double psr,
                                               double psi, double pdr, double pdi) {
    if (kk == nst) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            R[d] = sx[d * ld + j];
            R[DIM + d] = sn[d * ld + j];
        }
        return;
    }
    double2 b1 = make_double2(0.0, 0.0), b0 = b1;
    if (kk < ns) {
        b1 = B1[j + ld * kk];
        b0 = B0[j + ld * kk];
    }
    const double w = sw[j];
    double *q = R + 2 * DIM + theta_per_k<PDE>() * kk;
    if constexpr (PDE == INTI_LAPLACE) {
        q[0] = psr * (w * b1.x);
        q[1] = -pdr * (w * b0.x);
    } else {
        const double s1r = w * b1.x, s1i = w * b1.y, s0r = w * b0.x, s0i = w * b0.y;
        q[0] = psr * s1r - psi * s1i;
        q[1] = psr * s1i + psi * s1r;
        q[2] = -(pdr * s0r - pdi * s0i);
        q[3] = -(pdr * s0i + pdi * s0r);
    }
}

template <int PDE, int DIM>
__global__ __launch_bounds__(256) void k_theta_pack(ThetaPackArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= a.n * (a.nst + 1)) return;
    const int64_t j = e % a.n;
    const int kk = (int)(e / a.n);
    theta_pack_one<PDE, DIM>(j, kk, a.n, a.ns, a.nst, a.rec + j * a.rl, a.sx, a.sn, a.sw, a.B0, a.B1, a.psr,
                             a.psi, a.pdr, a.pdi);
}

struct ThetaArgs {
    int64_t m, n, nloc;
    int ns, rl, split;
    int64_t chunk;  // sources per block
    const int32_t *glob;  // local row -> global target
    const double *tx, *tn;
    const double *rec;
    const double *zs, *mu;  // monopoles (DIM x ns), green multiplier per target
    double2 *part;          // split x NS x nloc partial sums (split > 1)
    double2 *theta;         // nloc x ns, column-major (split == 1)
    KParams p;
};

// raw values of the S and D kernels of one pair, sharing the radial parts
template <int PDE, int DIM, int KS, int KD>
__device__ __forceinline__ void theta_pair(const double *r, const double *nx, const double *ny,
                                           const KParams &p, double &sr, double &si, double &dr,
                                           double &di) {
    if constexpr (PDE == INTI_HELMHOLTZ && DIM == 2) {
        const double d2 = dotv<2>(r, r);
        const bool ok = d2 > kSamePointTol2;
        const Helm2Radial R = helm2_radial<false>(fmax(d2, kClampD2), p);
        helm2_combine<KS, false>(r, ok, nx, ny, p, R, sr, si);
        helm2_combine<KD, false>(r, ok, nx, ny, p, R, dr, di);
    } else {
        pair_raw<PDE, DIM, KS, false>(r, nx, ny, p, sr, si);
        pair_raw<PDE, DIM, KD, false>(r, nx, ny, p, dr, di);
    }
}

template <int PDE>
constexpr int theta_accs(int ns) { return PDE != INTI_LAPLACE ? 2 * ns : ns; }  // accumulators per lane

// acc += the (S, D) sums of one target over the wavefront's share of the
// sources [b0, b1) of a record array (kThetaWaves shares, in wave order)
template <int PDE, int DIM, int DERIV, int NS>
__device__ __forceinline__ void theta_sources(const double *xt, const double *nt, const double *recp, int rl,
                                              int64_t b0, int64_t b1, int wave, const KParams &p,
                                              double *acc) {
    constexpr int KS = DERIV ? INTI_ADL : INTI_SL;
    constexpr int KD = DERIV ? INTI_HS : INTI_DL;
    constexpr int PK = theta_per_k<PDE>();
    const int64_t share = (b1 - b0 + kThetaWaves - 1) / kThetaWaves;
    const int64_t j0 = b0 + wave * share;
    const int64_t j1 = j0 + share < b1 ? j0 + share : b1;
    const uniform_double *rec = (const uniform_double *)recp;
    for (int64_t j = j0; j < j1; ++j) {
        const uniform_double *R = rec + j * rl;
        double r[DIM], ny[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            r[d] = xt[d] - R[d];
            ny[d] = R[DIM + d];
        }
        double sr, si, dr, di;
        theta_pair<PDE, DIM, KS, KD>(r, nt, ny, p, sr, si, dr, di);
        const uniform_double *Q = R + 2 * DIM;
        if constexpr (PDE == INTI_LAPLACE) {
#pragma unroll
            for (int k = 0; k < NS; ++k)
                acc[k] = __builtin_fma(sr, Q[PK * k], __builtin_fma(dr, Q[PK * k + 1], acc[k]));
        } else {
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const double q1r = Q[PK * k], q1i = Q[PK * k + 1], q0r = Q[PK * k + 2], q0i = Q[PK * k + 3];
                double ar = acc[2 * k], ai = acc[2 * k + 1];
                ar = __builtin_fma(sr, q1r, ar);
                ai = __builtin_fma(sr, q1i, ai);
                ar = __builtin_fma(-si, q1i, ar);
                ai = __builtin_fma(si, q1r, ai);
                ar = __builtin_fma(dr, q0r, ar);
                ai = __builtin_fma(dr, q0i, ai);
                ar = __builtin_fma(-di, q0i, ar);
                ai = __builtin_fma(di, q0r, ai);
                acc[2 * k] = ar;
                acc[2 * k + 1] = ai;
            }
        }
    }
}

// adds the wavefronts' sums into wave 0's, in wave order, kThetaRed
// accumulators per round
template <int NA>
__device__ __forceinline__ void theta_wave_sum(double *acc, double (*s_red)[kThetaRed][64], int wave, int lane) {
    constexpr int NR = (NA + kThetaRed - 1) / kThetaRed;
#pragma unroll
    for (int rd = 0; rd < NR; ++rd) {
        if (wave > 0) {
#pragma unroll
            for (int q = 0; q < kThetaRed; ++q)
                if (rd * kThetaRed + q < NA) s_red[wave - 1][q][lane] = acc[rd * kThetaRed + q];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int q = 0; q < kThetaRed; ++q)
                if (rd * kThetaRed + q < NA) {
                    double v = acc[rd * kThetaRed + q];
#pragma unroll
                    for (int w = 0; w < kThetaWaves - 1; ++w) v += s_red[w][q][lane];
                    acc[rd * kThetaRed + q] = v;
                }
        }
        __syncthreads();
    }
}

template <int PDE, int DIM, int DERIV, int NS>
__global__ __launch_bounds__(64 * kThetaWaves) void k_theta(ThetaArgs a) {
    constexpr int NA = theta_accs<PDE>(NS);
    __shared__ double s_red[kThetaWaves - 1][kThetaRed][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t chunk = blockIdx.x % a.split, tile = blockIdx.x / a.split;
    const int64_t l = tile * 64 + lane;
    const int64_t lc = l < a.nloc ? l : a.nloc - 1;
    const int64_t i = a.glob[lc];
    double xt[DIM], nt[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xt[d] = a.tx[d * a.m + i];
        nt[d] = a.tn ? a.tn[d * a.m + i] : 0.0;
    }
    double acc[NA];
#pragma unroll
    for (int q = 0; q < NA; ++q) acc[q] = 0.0;
    // this wavefront's sources: a quarter of the block's chunk
    const int64_t b0 = chunk * a.chunk < a.n ? chunk * a.chunk : a.n;
    const int64_t b1 = b0 + a.chunk < a.n ? b0 + a.chunk : a.n;
    theta_sources<PDE, DIM, DERIV, NS>(xt, nt, a.rec, a.rl, b0, b1, wave, a.p, acc);
    theta_wave_sum<NA>(acc, s_red, wave, lane);
    if (wave != 0 || l >= a.nloc) return;
    if (a.split == 1) {
        const double mu = a.mu[i];
        for (int k = 0; k < a.ns; ++k) {
            double r[DIM], dummy[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                r[d] = xt[d] - a.zs[d * a.ns + k];
                dummy[d] = 0.0;
            }
            double gr, gi;
            pair_value<PDE, DIM, DERIV ? INTI_ADL : INTI_SL>(r, nt, dummy, a.p, gr, gi);
            double vr = 0.0, vi = 0.0;
#pragma unroll
            for (int kk = 0; kk < NS; ++kk)
                if (kk == k) {
                    vr = PDE != INTI_LAPLACE ? acc[2 * kk] : acc[kk];
                    vi = PDE != INTI_LAPLACE ? acc[2 * kk + 1] : 0.0;
                }
            a.theta[l + a.nloc * k] = make_double2(mu * gr + vr, mu * gi + vi);
        }
    } else {
        double2 *pp = a.part + chunk * NS * a.nloc + l;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            pp[k * a.nloc] = make_double2(PDE != INTI_LAPLACE ? acc[2 * k] : acc[k],
                                          PDE != INTI_LAPLACE ? acc[2 * k + 1] : 0.0);
    }
}

struct ThetaRedArgs {
    int64_t m, nloc;
    int ns, nst, split;
    const int32_t *glob;
    const double *tx, *tn, *zs, *mu;
    const double2 *part;  // split x nst x nloc
    double2 *theta;       // nloc x ns, column-major
    KParams p;
};

template <int PDE, int DIM, int DERIV>
__global__ __launch_bounds__(256) void k_theta_reduce(ThetaRedArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= a.nloc * a.ns) return;
    const int64_t l = e % a.nloc;
    const int kk = (int)(e / a.nloc);
    const int64_t i = a.glob[l];
    double r[DIM], nx[DIM], dummy[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        r[d] = a.tx[d * a.m + i] - a.zs[d * a.ns + kk];
        nx[d] = a.tn ? a.tn[d * a.m + i] : 0.0;
        dummy[d] = 0.0;
    }
    double gr, gi;
    pair_value<PDE, DIM, DERIV ? INTI_ADL : INTI_SL>(r, nx, dummy, a.p, gr, gi);
    const double mu = a.mu[i];
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < a.split; ++c) {
        const double2 v = a.part[((int64_t)c * a.nst + kk) * a.nloc + l];
        sr += v.x; si += v.y;
    }
    a.theta[l + a.nloc * kk] = make_double2(mu * gr + sr, mu * gi + si);
}

// ------------------------------------------------------- batched QR / LS
__global__ __launch_bounds__(64) void k_ls(LsArgs a) {
    __shared__ double2 A[kMaxNs * kMaxNc];  // ns x nc column-major
    __shared__ double2 tau[kMaxNc];
    __shared__ double2 s_xi1;
    __shared__ double2 bb[64][kMaxNs + 1];
    const int lane = threadIdx.x;
    const int e = a.elems[blockIdx.x];
    const int nq = a.nq, ns = a.ns, nc = 2 * nq;
    const int32_t *jg = a.qtags + (int64_t)nq * e;
    // A = transpose(M), M = [B0[jg,:]; B1[jg,:]]
    for (int q = lane; q < ns * nc; q += 64) {
        const int kk = q % ns, r = q / ns;
        const double2 *B = r < nq ? a.B0 : a.B1;
        const int rr = r < nq ? r : r - nq;
        A[kk + ns * r] = B[jg[rr] + a.ldb * kk];
    }
    __syncthreads();
    wave_householder_qr(A, ns, nc, tau, &s_xi1);
    const double ratio = rdiag_ratio(A, ns, nc);
    // least-squares solves, one lane per near target
    const int p0 = a.pair_ptr[e], p1 = a.pair_ptr[e + 1];
    double wmax = 0.0, bmax = 0.0;
    for (int pb = p0; pb < p1; pb += 64) {
        const int p = pb + lane;
        if (p < p1) {
            const int64_t l = a.pair_loc[p];
            double2 *b = bb[lane];
            double bn = 0.0;
            for (int q = 0; q < ns; ++q) {
                b[q] = a.theta[l + a.nloc * q];
                bn += b[q].x * b[q].x + b[q].y * b[q].y;
            }
            bmax = fmax(bmax, sqrt(bn));
            lane_qr_solve(A, ns, nc, tau, b);
            double wn = 0.0;
            for (int q = 0; q < nc; ++q) wn += b[q].x * b[q].x + b[q].y * b[q].y;
            wmax = fmax(wmax, sqrt(wn));
            // dD = W[1:nq], dS = -W[nq+1:2nq] (the sign is applied by the gather)
            for (int r = 0; r < nq; ++r) {
                a.triD[(int64_t)p * nq + r] = b[r];
                a.triS[(int64_t)p * nq + r] = b[nq + r];
            }
        }
    }
    // wave-wide max of the per-lane diagnostics
    for (int off = 32; off > 0; off >>= 1) {
        wmax = fmax(wmax, __shfl_xor(wmax, off));
        bmax = fmax(bmax, __shfl_xor(bmax, off));
    }
    if (lane == 0) {
        a.diag[3 * (int64_t)blockIdx.x + 0] = ratio;
        a.diag[3 * (int64_t)blockIdx.x + 1] = wmax;
        a.diag[3 * (int64_t)blockIdx.x + 2] = bmax;
    }
}

// --------------------------------------------------------------- dispatch
// Theta kernels are instantiated for these monopole counts (multiples of 3:
// ns = ceil(3 nq) with the default oversampling); a build takes the smallest
// one >= ns, the extra record columns being zero.
constexpr int kThetaNs[] = {6, 9, 12, 15, 18, 21, 24, 27, 30, 32};

using ThetaFn = void (*)(const ThetaArgs &, dim3, hipStream_t);
template <int PDE, int DIM, int DERIV, int NS>
static void theta_launcher(const ThetaArgs &a, dim3 g, hipStream_t s) {
    hipLaunchKernelGGL((k_theta<PDE, DIM, DERIV, NS>), g, dim3(64 * kThetaWaves), 0, s, a);
}
template <int PDE, int DIM, int DERIV, int NS>
static int theta_occupancy() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_theta<PDE, DIM, DERIV, NS>, 64 * kThetaWaves, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}
template <int PDE, int DIM, int DERIV>
static ThetaFn theta_ns(int nst, int *occ) {
#define INTI_T(N) \
    case N: *occ = theta_occupancy<PDE, DIM, DERIV, N>(); return theta_launcher<PDE, DIM, DERIV, N>;
    switch (nst) {
        INTI_T(6) INTI_T(9) INTI_T(12) INTI_T(15) INTI_T(18) INTI_T(21) INTI_T(24) INTI_T(27) INTI_T(30)
        INTI_T(32)
    }
#undef INTI_T
    return nullptr;
}

struct BdimFns {
    void (*traces)(const TraceArgs &, dim3, hipStream_t);
    void (*pack)(const ThetaPackArgs &, dim3, hipStream_t);
    ThetaFn theta;
    int theta_occ;  // resident Theta blocks per CU
    void (*reduce)(const ThetaRedArgs &, dim3, hipStream_t);
};

template <int PDE, int DIM, int DERIV>
static BdimFns bdim_fns(int nst) {
    BdimFns f;
    f.traces = [](const TraceArgs &a, dim3 g, hipStream_t s) {
        hipLaunchKernelGGL((k_traces<PDE, DIM>), g, dim3(256), 0, s, a);
    };
    f.pack = [](const ThetaPackArgs &a, dim3 g, hipStream_t s) {
        hipLaunchKernelGGL((k_theta_pack<PDE, DIM>), g, dim3(256), 0, s, a);
    };
    f.theta = theta_ns<PDE, DIM, DERIV>(nst, &f.theta_occ);
    f.reduce = [](const ThetaRedArgs &a, dim3 g, hipStream_t s) {
        hipLaunchKernelGGL((k_theta_reduce<PDE, DIM, DERIV>), g, dim3(256), 0, s, a);
    };
    return f;
}

static bool pick_bdim(int pde, int dim, int deriv, int nst, BdimFns &f) {
#define INTI_B(P, D, R) \
    if (pde == P && dim == D && deriv == R) { f = bdim_fns<P, D, R>(nst); return f.theta != nullptr; }
    INTI_B(INTI_LAPLACE, 2, 0) INTI_B(INTI_LAPLACE, 2, 1)
    INTI_B(INTI_LAPLACE, 3, 0) INTI_B(INTI_LAPLACE, 3, 1)
    INTI_B(INTI_HELMHOLTZ, 2, 0) INTI_B(INTI_HELMHOLTZ, 2, 1)
    INTI_B(INTI_HELMHOLTZ, 3, 0) INTI_B(INTI_HELMHOLTZ, 3, 1)
    INTI_B(kHelmCK, 3, 0) INTI_B(kHelmCK, 3, 1)
#undef INTI_B
    return false;
}

// monopole centres: bdim.jl:79-92 with utils.jl:142-149 / 158-171
void monopoles_box(int dim, const double *lo, const double *hi, int ns, double radius_mult,
                   std::vector<double> &zs) {
    double xc[3], diag2 = 0.0;
    for (int d = 0; d < dim; ++d) {
        xc[d] = (lo[d] + hi[d]) / 2;
        diag2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    }
    const double R = radius_mult * std::sqrt(diag2) / 2;
    zs.assign((size_t)dim * ns, 0.0);
    if (dim == 2) {
        for (int i = 0; i < ns; ++i) {
            const double th = 2 * kPi * i / ns;
            zs[0 * ns + i] = R * std::cos(th) + xc[0];
            zs[1 * ns + i] = R * std::sin(th) + xc[1];
        }
    } else {
        const double phi = kPi * (3 - std::sqrt(5.0));
        for (int i = 1; i <= ns; ++i) {
            const double ytmp = 1 - ((double)(i - 1) / (ns - 1)) * 2;
            const double radius = std::sqrt(1 - ytmp * ytmp);
            const double theta = phi * i;
            zs[0 * ns + i - 1] = std::cos(theta) * radius * R + xc[0];
            zs[1 * ns + i - 1] = ytmp * R + xc[1];
            zs[2 * ns + i - 1] = std::sin(theta) * radius * R + xc[2];
        }
    }
}

static void monopoles(const Quad *src, int ns, double radius_mult, std::vector<double> &zs) {
    monopoles_box(src->dim, src->lo, src->hi, ns, radius_mult, zs);
}

int csr_alloc(Ctx *ctx, int64_t m, int64_t n, int64_t nnz, int cplx, Csr **out);
void csr_free(Csr *c);

// Element types: type t has nq[t] nodes per element and ne[t] elements; its
// node tags are the nq[t] x ne[t] column-major block of qtags at
// sum_{u<t} nq[u] ne[u], and (target != source) its near lists are the
// ne[t] + 1 entries of near_ptr at sum_{u<t} (ne[u] + 1) (each starting at 0)
// indexing the block of near_idx at sum_{u<t} near_ptr_u[ne[u]]. The monopole
// count uses qmax = sum_t nq[t] (bdim.jl:79-80); the least squares run type
// by type and the triplets are taken in that order (bdim.jl:138-185).
int bdim_build_types(Op *opS, Op *opD, int deriv, int ntypes, const int *nq, const int64_t *ne,
                     const int64_t *qtags, const int64_t *near_ptr, const int64_t *near_idx, const double *mu,
                     const double *params, Csr **outS, Csr **outD, double *diag_out) {
    Ctx *ctx = opS->ctx;
    const Quad *trg = opS->trg, *src = opS->src;
    const int dim = src->dim;
    const int64_t m = trg->n, n = src->n;
    const bool self = near_ptr == nullptr;
    const double oversample = params ? params[0] : 3.0;
    const double radius_mult = params ? params[1] : 1.5;
    if (ntypes < 1) return set_error(ctx, INTI_ERR_ARG, "bdim: no element type");
    int qmax = 0, nqmax = 0;
    int64_t ntags = 0;
    for (int t = 0; t < ntypes; ++t) {
        if (nq[t] < 1 || ne[t] < 0) return set_error(ctx, INTI_ERR_ARG, "bdim: bad element type sizes");
        qmax += nq[t];
        nqmax = std::max(nqmax, nq[t]);
        ntags += (int64_t)nq[t] * ne[t];
    }
    const int ns = (int)std::ceil(oversample * qmax);
    if (ns > kMaxNs || 2 * nqmax > kMaxNc || ns < 2 * nqmax)
        return set_error(ctx, INTI_ERR_UNSUPPORTED,
                         "bdim: need 2*nq <= ns <= 32 and 2*nq <= 24 (ns = ceil(oversample * sum of nq over types))");
    if (self && (trg != src))
        return set_error(ctx, INTI_ERR_ARG, "bdim: the self pattern needs target == source");
    if (!src->normals) return set_error(ctx, INTI_ERR_ARG, "bdim: source quadrature needs normals");
    if (deriv && !trg->normals)
        return set_error(ctx, INTI_ERR_ARG, "bdim: derivative correction needs target normals");
    if (n == 0 || ntags == 0) return set_error(ctx, INTI_ERR_ARG, "bdim: empty source");
    if (ntags != n && self)
        return set_error(ctx, INTI_ERR_SIZE, "bdim: the element types must cover every source node once");
    for (int64_t q = 0; q < ntags; ++q)
        if (qtags[q] < 0 || qtags[q] >= n) return set_error(ctx, INTI_ERR_ARG, "bdim: qtags out of range");
    const int nsp = ns <= 8 ? 8 : ns <= 12 ? 12 : ns <= 16 ? 16 : ns <= 20 ? 20 : ns <= 24 ? 24 : 32;
    int nst = 32;
    for (int c : kThetaNs)
        if (c >= ns) { nst = c; break; }
    BdimFns fns;
    if (!pick_bdim(op_kernel_pde(opS), dim, deriv, nst, fns))
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "bdim: no kernel for this operator");

    // ---- host-side lists per type (element -> targets), concatenated
    std::vector<int64_t> tag_off(ntypes + 1, 0), el_off(ntypes + 1, 0), pair_off(ntypes + 1, 0),
        tri_off(ntypes + 1, 0), elems_off(ntypes + 1, 0);
    std::vector<int32_t> pair_ptr, pair_tgt, elems;  // pair_ptr: ne[t] + 1 per type, local to the type
    int64_t nidx_off = 0, nptr_off = 0;
    for (int t = 0; t < ntypes; ++t) {
        tag_off[t + 1] = tag_off[t] + (int64_t)nq[t] * ne[t];
        const int64_t *qt = qtags + tag_off[t];
        const int64_t *np = self ? nullptr : near_ptr + nptr_off;
        const int64_t *ni = self ? nullptr : near_idx + nidx_off;
        const int64_t p0 = (int64_t)pair_tgt.size();
        pair_ptr.push_back(0);
        for (int64_t e = 0; e < ne[t]; ++e) {
            const int64_t cnt = self ? nq[t] : np[e + 1] - np[e];
            if (cnt < 0) return set_error(ctx, INTI_ERR_ARG, "bdim: near lists must be non-decreasing");
            for (int64_t q = 0; q < cnt; ++q) {
                const int64_t i = self ? qt[q + (int64_t)nq[t] * e] : ni[np[e] + q];
                if (i < 0 || i >= m) return set_error(ctx, INTI_ERR_ARG, "bdim: near target out of range");
                pair_tgt.push_back((int32_t)i);
            }
            pair_ptr.push_back((int32_t)((int64_t)pair_tgt.size() - p0));
            if (cnt > 0) elems.push_back((int32_t)e);
        }
        el_off[t + 1] = el_off[t] + ne[t] + 1;
        pair_off[t + 1] = (int64_t)pair_tgt.size();
        tri_off[t + 1] = tri_off[t] + (pair_off[t + 1] - pair_off[t]) * nq[t];
        elems_off[t + 1] = (int64_t)elems.size();
        if (!self) {
            nidx_off += near_ptr[nptr_off + ne[t]];
            nptr_off += ne[t] + 1;
        }
    }
    const int64_t npairs = (int64_t)pair_tgt.size();
    std::vector<int32_t> loc(m, -1), glob;
    for (int64_t p = 0; p < npairs; ++p) {
        const int32_t i = pair_tgt[p];
        if (loc[i] < 0) { loc[i] = (int32_t)glob.size(); glob.push_back(i); }
    }
    std::vector<int32_t> pair_loc(npairs);
    for (int64_t p = 0; p < npairs; ++p) pair_loc[p] = loc[pair_tgt[p]];
    const int64_t nloc = (int64_t)glob.size();
    // triplets (type, element, target, node) -> (row, col), CSR pattern
    const int64_t ntri = tri_off[ntypes];
    if (ntri >= (int64_t)INT32_MAX) return set_error(ctx, INTI_ERR_SIZE, "bdim: too many triplets");
    std::vector<int32_t> tri_row(ntri), tri_col(ntri);
    for (int t = 0; t < ntypes; ++t) {
        const int64_t *qt = qtags + tag_off[t];
        const int32_t *pp = pair_ptr.data() + el_off[t];
        for (int64_t e = 0; e < ne[t]; ++e)
            for (int32_t p = pp[e]; p < pp[e + 1]; ++p)
                for (int r = 0; r < nq[t]; ++r) {
                    const int64_t q = tri_off[t] + (int64_t)p * nq[t] + r;
                    tri_row[q] = pair_tgt[pair_off[t] + p];
                    tri_col[q] = (int32_t)qt[r + (int64_t)nq[t] * e];
                }
    }
    TripletPattern pat;
    build_pattern(m, tri_row, tri_col, pat);
    const std::vector<int32_t> &rowptr = pat.rowptr, &colind = pat.colind, &seg = pat.seg, &perm = pat.perm;
    const int64_t nnz = (int64_t)colind.size();

    // ---- device buffers
    std::vector<double> zs;
    monopoles(src, ns, radius_mult, zs);
    std::vector<int32_t> qt32((size_t)ntags);
    for (size_t q = 0; q < qt32.size(); ++q) qt32[q] = (int32_t)qtags[q];
    DevBuf dzs, dqt, dglob, dpp, dploc, delems, dB0, dB1, dpart, drec, dth, dtriS, dtriD, ddiag, dseg, dperm, dmu;
    const size_t cb = sizeof(double2);
    INTI_ALLOC(ctx, dzs, zs.size() * sizeof(double));
    INTI_ALLOC(ctx, dqt, qt32.size() * sizeof(int32_t));
    INTI_ALLOC(ctx, dglob, glob.size() * sizeof(int32_t));
    INTI_ALLOC(ctx, dpp, pair_ptr.size() * sizeof(int32_t));
    INTI_ALLOC(ctx, dploc, pair_loc.size() * sizeof(int32_t));
    INTI_ALLOC(ctx, delems, elems.size() * sizeof(int32_t));
    INTI_ALLOC(ctx, dB0, cb * n * nsp);
    INTI_ALLOC(ctx, dB1, cb * n * nsp);
    INTI_ALLOC(ctx, dth, cb * nloc * ns);
    INTI_ALLOC(ctx, dtriS, cb * ntri);
    INTI_ALLOC(ctx, dtriD, cb * ntri);
    INTI_ALLOC(ctx, ddiag, sizeof(double) * 3 * elems.size());
    INTI_ALLOC(ctx, dseg, sizeof(int32_t) * seg.size());
    INTI_ALLOC(ctx, dperm, sizeof(int32_t) * perm.size());
    INTI_ALLOC(ctx, dmu, sizeof(double) * m);
    hipStream_t st = ctx->stream;
    auto up = [&](void *d, const void *h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
    };
    INTI_HIP(ctx, up(dzs.p, zs.data(), zs.size() * sizeof(double)));
    INTI_HIP(ctx, up(dqt.p, qt32.data(), qt32.size() * sizeof(int32_t)));
    INTI_HIP(ctx, up(dglob.p, glob.data(), glob.size() * sizeof(int32_t)));
    INTI_HIP(ctx, up(dpp.p, pair_ptr.data(), pair_ptr.size() * sizeof(int32_t)));
    INTI_HIP(ctx, up(dploc.p, pair_loc.data(), pair_loc.size() * sizeof(int32_t)));
    INTI_HIP(ctx, up(delems.p, elems.data(), elems.size() * sizeof(int32_t)));
    INTI_HIP(ctx, up(dseg.p, seg.data(), seg.size() * sizeof(int32_t)));
    INTI_HIP(ctx, up(dperm.p, perm.data(), perm.size() * sizeof(int32_t)));
    INTI_HIP(ctx, up(dmu.p, mu, sizeof(double) * m));

    KParams kp = basic_params(opS->k);
    kp.kimag = opS->kimag;

    // traces
    {
        TraceArgs a;
        a.n = n; a.ns = ns; a.nsp = nsp;
        a.sx = src->coords; a.sn = src->normals; a.zs = (const double *)dzs.p;
        a.B0 = (double2 *)dB0.p; a.B1 = (double2 *)dB1.p;
        a.p = kp;
        fns.traces(a, dim3((unsigned)((n * nsp + 255) / 256)), st);
        INTI_HIP(ctx, hipGetLastError());
    }
    // Theta: pack the per-source records, then the sum (and the reduction of
    // its partial sums when the sources are split over several blocks)
    if (nloc > 0) {
        const int pk = opS->pde == INTI_HELMHOLTZ ? 4 : 2;
        const int rl = 2 * dim + pk * nst;
        INTI_ALLOC(ctx, drec, sizeof(double) * n * rl);
        ThetaPackArgs pa;
        pa.n = n; pa.ns = ns; pa.nsp = nsp; pa.nst = nst; pa.rl = rl;
        pa.sx = src->coords; pa.sn = src->normals; pa.sw = src->weights;
        pa.B0 = (const double2 *)dB0.p; pa.B1 = (const double2 *)dB1.p;
        pa.rec = (double *)drec.p;
        basic_post(opS->pde, dim, deriv ? INTI_ADL : INTI_SL, opS->k, pa.psr, pa.psi);
        basic_post(opS->pde, dim, deriv ? INTI_HS : INTI_DL, opS->k, pa.pdr, pa.pdi);
        fns.pack(pa, dim3((unsigned)((n * (nst + 1) + 255) / 256)), st);
        INTI_HIP(ctx, hipGetLastError());
        // source split: the fewest rounds of resident blocks per unit of work,
        // the smaller split on ties. Every split adds nloc x nst partial sums
        // written and read once: up to 3 when the tiles alone fill a round
        // (partial traffic <= 3x Theta), more only for small builds
        const int64_t ntiles = (nloc + 63) / 64;
        const int64_t slots = (int64_t)fns.theta_occ * ctx->num_cu;
        const int64_t want = std::max<int64_t>(3, (slots + ntiles - 1) / ntiles);
        const int64_t max_split = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(64, want),
                                                                         n / (64 * kThetaWaves)));
        int64_t best = 1;
        double best_cost = 1e300;
        for (int64_t sp = 1; sp <= max_split; ++sp) {
            const double rounds = (double)((ntiles * sp + slots - 1) / slots);
            const double cost = rounds / (double)sp;
            if (cost < best_cost * (1 - 1e-9)) { best_cost = cost; best = sp; }
        }
        const int split = (int)best;
        const int64_t chunk = (n + split - 1) / split;
        if (split > 1) INTI_ALLOC(ctx, dpart, sizeof(double2) * split * nst * nloc);
        ThetaArgs a;
        a.m = m; a.n = n; a.nloc = nloc;
        a.ns = ns; a.rl = rl; a.split = split; a.chunk = chunk;
        a.glob = (const int32_t *)dglob.p;
        a.tx = trg->coords; a.tn = trg->normals;
        a.rec = (const double *)drec.p;
        a.zs = (const double *)dzs.p; a.mu = (const double *)dmu.p;
        a.part = (double2 *)dpart.p;
        a.theta = (double2 *)dth.p;
        a.p = kp;
        fns.theta(a, dim3((unsigned)(ntiles * split)), st);
        INTI_HIP(ctx, hipGetLastError());
        if (split > 1) {
            ThetaRedArgs r;
            r.m = m; r.nloc = nloc; r.ns = ns; r.nst = nst; r.split = split;
            r.glob = (const int32_t *)dglob.p;
            r.tx = trg->coords; r.tn = trg->normals; r.zs = (const double *)dzs.p;
            r.mu = (const double *)dmu.p;
            r.part = (const double2 *)dpart.p;
            r.theta = (double2 *)dth.p;
            r.p = kp;
            fns.reduce(r, dim3((unsigned)((nloc * ns + 255) / 256)), st);
            INTI_HIP(ctx, hipGetLastError());
        }
    }
    // batched QR + LS, one launch per element type
    for (int t = 0; t < ntypes; ++t) {
        const int64_t nel = elems_off[t + 1] - elems_off[t];
        if (nel == 0) continue;
        LsArgs a;
        a.n = n; a.nloc = nloc; a.nq = nq[t]; a.ns = ns;
        a.elems = (const int32_t *)delems.p + elems_off[t];
        a.qtags = (const int32_t *)dqt.p + tag_off[t];
        a.pair_ptr = (const int32_t *)dpp.p + el_off[t];
        a.pair_loc = (const int32_t *)dploc.p + pair_off[t];
        a.B0 = (const double2 *)dB0.p; a.B1 = (const double2 *)dB1.p; a.ldb = n;
        a.theta = (const double2 *)dth.p;
        a.triS = (double2 *)dtriS.p + tri_off[t]; a.triD = (double2 *)dtriD.p + tri_off[t];
        a.diag = (double *)ddiag.p + 3 * elems_off[t];
        hipLaunchKernelGGL(k_ls, dim3((unsigned)nel), dim3(64), 0, st, a);
        INTI_HIP(ctx, hipGetLastError());
    }
    // CSR
    const int cplx = op_is_complex(opS) ? 1 : 0;
    Csr *cS = nullptr, *cD = nullptr;
    int rc = csr_alloc(ctx, m, n, nnz, cplx, &cS);
    if (rc) return rc;
    rc = csr_alloc(ctx, m, n, nnz, cplx, &cD);
    if (rc) { csr_free(cS); return rc; }
    auto fail = [&](int code) { csr_free(cS); csr_free(cD); return code; };
    if (up(cS->rowptr, rowptr.data(), sizeof(int32_t) * (m + 1)) != hipSuccess ||
        up(cS->colind, colind.data(), sizeof(int32_t) * nnz) != hipSuccess ||
        up(cD->rowptr, rowptr.data(), sizeof(int32_t) * (m + 1)) != hipSuccess ||
        up(cD->colind, colind.data(), sizeof(int32_t) * nnz) != hipSuccess)
        return fail(set_error(ctx, INTI_ERR_HIP, "bdim: CSR upload failed"));
    {
        Gather1Args g;
        g.nnz = nnz;
        g.seg = (const int32_t *)dseg.p; g.perm = (const int32_t *)dperm.p;
        g.cplx = cplx;
        g.tri = (const double2 *)dtriS.p; g.val = cS->val; g.sign = -1.0;
        launch_gather1(g, st);
        g.tri = (const double2 *)dtriD.p; g.val = cD->val; g.sign = 1.0;
        launch_gather1(g, st);
        if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, INTI_ERR_HIP, "bdim: gather launch failed"));
    }
    std::vector<double> dg(3 * elems.size());
    if (!dg.empty() &&
        hipMemcpyAsync(dg.data(), ddiag.p, sizeof(double) * dg.size(), hipMemcpyDeviceToHost, st) != hipSuccess)
        return fail(set_error(ctx, INTI_ERR_HIP, "bdim: diag copy failed"));
    if (hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(ctx, INTI_ERR_HIP, "bdim: kernel execution failed"));
    if (diag_out) {
        double d0 = 0, d2 = 0, d3 = 0;
        for (size_t q = 0; q < elems.size(); ++q) {
            d0 = std::max(d0, dg[3 * q]);
            d2 = std::max(d2, dg[3 * q + 1]);
            d3 = std::max(d3, dg[3 * q + 2]);
        }
        diag_out[0] = d0; diag_out[1] = 0.0; diag_out[2] = d2; diag_out[3] = d3;
    }
    *outS = cS;
    *outD = cD;
    return INTI_OK;
}

// --------------------------------------------- DIM corrections of a batch
template <int PDE, int DIM>
__global__ __launch_bounds__(256) void k_traces_batch(TraceBatchArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= a.ntot * a.ns) return;
    const int64_t j = e % a.ntot;
    const int kk = (int)(e / a.ntot);
    const int p = a.row2p[j];
    const double *z = a.zs + (int64_t)p * DIM * a.ns;
    double r[DIM], ny[DIM], dummy[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        r[d] = a.sx[d * a.ntot + j] - z[d * a.ns + kk];
        ny[d] = a.sn[d * a.ntot + j];
        dummy[d] = 0.0;
    }
    const KParams kp = a.kp[p];
    double g0r, g0i, g1r, g1i;
    pair_value<PDE, DIM, INTI_SL>(r, ny, dummy, kp, g0r, g0i);
    pair_value<PDE, DIM, INTI_ADL>(r, ny, dummy, kp, g1r, g1i);
    const double mu = a.mu[j];
    a.B0[e] = make_double2(g0r, g0i);
    a.B1[e] = make_double2(g1r, g1i);
    a.theta[e] = make_double2(mu * g0r, mu * g0i);
}

bool launch_dim_traces_batch(int pde, int dim, const TraceBatchArgs &a, hipStream_t s) {
    const dim3 g((unsigned)((a.ntot * a.ns + 255) / 256));
    if (pde == INTI_LAPLACE && dim == 2) hipLaunchKernelGGL((k_traces_batch<INTI_LAPLACE, 2>), g, dim3(256), 0, s, a);
    else if (pde == INTI_LAPLACE && dim == 3) hipLaunchKernelGGL((k_traces_batch<INTI_LAPLACE, 3>), g, dim3(256), 0, s, a);
    else if (pde == INTI_HELMHOLTZ && dim == 2) hipLaunchKernelGGL((k_traces_batch<INTI_HELMHOLTZ, 2>), g, dim3(256), 0, s, a);
    else if (pde == INTI_HELMHOLTZ && dim == 3) hipLaunchKernelGGL((k_traces_batch<INTI_HELMHOLTZ, 3>), g, dim3(256), 0, s, a);
    else return false;
    return true;
}

template <int PDE, int DIM>
__global__ __launch_bounds__(256) void k_theta_pack_batch(ThetaBatchArgs a, int nst, int rl) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= a.ntot * (nst + 1)) return;
    const int64_t j = e % a.ntot;
    const int kk = (int)(e / a.ntot);
    const double *pf = a.post + 4 * (int64_t)a.row2p[j];
    theta_pack_one<PDE, DIM>(j, kk, a.ntot, a.ns, nst, a.rec + j * rl, a.coords, a.normals, a.weights, a.B0,
                             a.B1, pf[0], pf[1], pf[2], pf[3]);
}

template <int PDE, int DIM, int NS>
__global__ __launch_bounds__(64 * kThetaWaves) void k_theta_batch(ThetaBatchArgs a, int rl) {
    constexpr int NA = theta_accs<PDE>(NS);
    __shared__ double s_red[kThetaWaves - 1][kThetaRed][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int p = a.items[2 * (int64_t)blockIdx.x], tile = a.items[2 * (int64_t)blockIdx.x + 1];
    const int64_t off = a.off[p], n = a.off[p + 1] - off;
    const int64_t l = (int64_t)tile * 64 + lane;
    const int64_t i = off + (l < n ? l : n - 1);
    double xt[DIM], nt[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xt[d] = a.coords[d * a.ntot + i];
        nt[d] = a.normals[d * a.ntot + i];
    }
    double acc[NA];
#pragma unroll
    for (int q = 0; q < NA; ++q) acc[q] = 0.0;
    theta_sources<PDE, DIM, 0, NS>(xt, nt, a.rec + off * rl, rl, 0, n, wave, a.kp[p], acc);
    theta_wave_sum<NA>(acc, s_red, wave, lane);
    if (wave != 0 || l >= n) return;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        if (k >= a.ns) break;
        double2 *t = a.theta + (off + l) + a.ntot * k;
        const double2 v = *t;
        *t = PDE != INTI_LAPLACE ? make_double2(v.x + acc[2 * k], v.y + acc[2 * k + 1])
                                 : make_double2(v.x + acc[k], v.y);
    }
}

template <int PDE, int DIM, int NS>
static void theta_batch_launch(const ThetaBatchArgs &a, int64_t nitems, hipStream_t s) {
    const int rl = 2 * DIM + theta_per_k<PDE>() * NS;
    hipLaunchKernelGGL((k_theta_pack_batch<PDE, DIM>), dim3((unsigned)((a.ntot * (NS + 1) + 255) / 256)),
                       dim3(256), 0, s, a, NS, rl);
    hipLaunchKernelGGL((k_theta_batch<PDE, DIM, NS>), dim3((unsigned)nitems), dim3(64 * kThetaWaves), 0, s, a,
                       rl);
}

static int theta_nst(int ns) {
    for (int c : kThetaNs)
        if (c >= ns) return c;
    return 0;
}

int dim_theta_batch_record(int pde, int dim, int ns) {
    const int nst = theta_nst(ns);
    if (!nst || (pde != INTI_LAPLACE && pde != INTI_HELMHOLTZ) || (dim != 2 && dim != 3)) return 0;
    return 2 * dim + (pde != INTI_LAPLACE ? 4 : 2) * nst;
}

template <int PDE, int DIM>
static bool theta_batch_ns(const ThetaBatchArgs &a, int64_t nitems, hipStream_t s) {
#define INTI_T(N) \
    case N: theta_batch_launch<PDE, DIM, N>(a, nitems, s); return true;
    switch (theta_nst(a.ns)) {
        INTI_T(6) INTI_T(9) INTI_T(12) INTI_T(15) INTI_T(18) INTI_T(21) INTI_T(24) INTI_T(27) INTI_T(30)
        INTI_T(32)
    }
#undef INTI_T
    return false;
}

bool launch_dim_theta_batch(int pde, int dim, const ThetaBatchArgs &a, int64_t nitems, hipStream_t s) {
    if (nitems <= 0) return true;
    if (pde == INTI_LAPLACE && dim == 2) return theta_batch_ns<INTI_LAPLACE, 2>(a, nitems, s);
    if (pde == INTI_LAPLACE && dim == 3) return theta_batch_ns<INTI_LAPLACE, 3>(a, nitems, s);
    if (pde == INTI_HELMHOLTZ && dim == 2) return theta_batch_ns<INTI_HELMHOLTZ, 2>(a, nitems, s);
    if (pde == INTI_HELMHOLTZ && dim == 3) return theta_batch_ns<INTI_HELMHOLTZ, 3>(a, nitems, s);
    return false;
}

void launch_dim_ls(const LsArgs &a, int64_t nel, hipStream_t s) {
    if (nel > 0) hipLaunchKernelGGL(k_ls, dim3((unsigned)nel), dim3(64), 0, s, a);
}

int bdim_build(Op *opS, Op *opD, int deriv, int nq, int64_t ne, const int64_t *qtags,
               const int64_t *near_ptr, const int64_t *near_idx, const double *mu,
               const double *params, Csr **outS, Csr **outD, double *diag_out) {
    return bdim_build_types(opS, opD, deriv, 1, &nq, &ne, qtags, near_ptr, near_idx, mu, params, outS, outD,
                            diag_out);
}

}  // namespace inti
