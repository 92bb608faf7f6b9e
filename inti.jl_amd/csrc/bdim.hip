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

constexpr int kThetaTT = 32;   // sources per LDS tile of the Theta kernel
constexpr int kThetaBS = 256;  // targets per block

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
struct ThetaArgs {
    int64_t m, n, nloc;
    int nsp, split;
    int64_t chunk;
    const int32_t *glob;  // local row -> global target
    const double *tx, *tn, *sx, *sn, *sw;
    const double2 *B0, *B1;
    double2 *part;  // split x nloc x nsp
    KParams p;
};

template <int PDE, int DIM, int DERIV, int NS>
__global__ __launch_bounds__(kThetaBS) void k_theta(ThetaArgs a) {
    constexpr int KS = DERIV ? INTI_ADL : INTI_SL;
    constexpr int KD = DERIV ? INTI_HS : INTI_DL;
    __shared__ double s_geo[kThetaTT][2 * DIM + 1];
    __shared__ double2 s_b0[kThetaTT][NS];
    __shared__ double2 s_b1[kThetaTT][NS];
    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x;
    const int64_t chunk = b % a.split, tile = b / a.split;
    const int64_t l = tile * kThetaBS + tid;
    const int64_t lc = l < a.nloc ? l : a.nloc - 1;
    const int64_t i = a.glob[lc];
    double xt[DIM], nt[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xt[d] = a.tx[d * a.m + i];
        nt[d] = a.tn ? a.tn[d * a.m + i] : 0.0;
    }
    double accr[NS], acci[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) { accr[k] = 0.0; acci[k] = 0.0; }
    const int64_t j0 = chunk * a.chunk;
    const int64_t j1 = j0 + a.chunk < a.n ? j0 + a.chunk : a.n;
    for (int64_t s0 = j0; s0 < j1; s0 += kThetaTT) {
        for (int q = tid; q < kThetaTT * (2 * DIM + 1); q += kThetaBS) {
            const int jj = q / (2 * DIM + 1), f = q % (2 * DIM + 1);
            const int64_t j = s0 + jj;
            double v = 0.0;
            if (j < j1) {
                if (f < DIM) v = a.sx[f * a.n + j];
                else if (f < 2 * DIM) v = a.sn[(f - DIM) * a.n + j];
                else v = a.sw[j];
            } else if (f < DIM) {
                v = 1.0e3;  // padding source far away, weight 0
            }
            s_geo[jj][f] = v;
        }
        for (int q = tid; q < kThetaTT * NS; q += kThetaBS) {
            const int jj = q % kThetaTT, k = q / kThetaTT;
            const int64_t j = s0 + jj;
            double2 v0 = make_double2(0.0, 0.0), v1 = v0;
            if (j < j1) {
                v0 = a.B0[j + a.n * k];
                v1 = a.B1[j + a.n * k];
            }
            s_b0[jj][k] = v0;
            s_b1[jj][k] = v1;
        }
        __syncthreads();
        for (int jj = 0; jj < kThetaTT; ++jj) {
            double r[DIM], ny[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                r[d] = xt[d] - s_geo[jj][d];
                ny[d] = s_geo[jj][DIM + d];
            }
            const double w = s_geo[jj][2 * DIM];
            double sr, si, dr, di;
            pair_value<PDE, DIM, KS>(r, nt, ny, a.p, sr, si);
            pair_value<PDE, DIM, KD>(r, nt, ny, a.p, dr, di);
            sr *= w; si *= w; dr *= w; di *= w;
            if constexpr (PDE == INTI_LAPLACE) {
                // real kernels and traces: two FMAs per monopole
#pragma unroll
                for (int k = 0; k < NS; ++k)
                    accr[k] = __builtin_fma(sr, s_b1[jj][k].x, __builtin_fma(-dr, s_b0[jj][k].x, accr[k]));
            } else {
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const double2 b1 = s_b1[jj][k], b0 = s_b0[jj][k];
                    accr[k] += sr * b1.x - si * b1.y - (dr * b0.x - di * b0.y);
                    acci[k] += sr * b1.y + si * b1.x - (dr * b0.y + di * b0.x);
                }
            }
        }
        __syncthreads();
    }
    if (l < a.nloc) {
        double2 *pp = a.part + (chunk * a.nloc + l) * a.nsp;
#pragma unroll
        for (int k = 0; k < NS; ++k) pp[k] = make_double2(accr[k], acci[k]);
    }
}

struct ThetaRedArgs {
    int64_t m, nloc;
    int ns, nsp, split;
    const int32_t *glob;
    const double *tx, *tn, *zs, *mu;
    const double2 *part;
    double2 *theta;  // nloc x ns, column-major
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
    double tr = mu * gr, ti = mu * gi;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < a.split; ++c) {
        const double2 v = a.part[((int64_t)c * a.nloc + l) * a.nsp + kk];
        sr += v.x; si += v.y;
    }
    a.theta[l + a.nloc * kk] = make_double2(tr + sr, ti + si);
}

// ------------------------------------------------------- batched QR / LS
struct LsArgs {
    int64_t n, nloc;
    int nq, ns;
    const int32_t *elems;     // elements with a non-empty near list
    const int32_t *qtags;     // nq x ne
    const int32_t *pair_ptr;  // ne+1 -> pairs of the element
    const int32_t *pair_loc;  // pair -> local target row
    const double2 *B0, *B1;   // n x nsp
    int64_t ldb;              // = n
    const double2 *theta;     // nloc x ns
    double2 *triS, *triD;     // pair*nq + r
    double *diag;             // per element: Rdiag ratio, max |W|, max |Theta|
};

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
using ThetaFn = void (*)(const ThetaArgs &, dim3, hipStream_t);
template <int PDE, int DIM, int DERIV, int NS>
static void theta_launcher(const ThetaArgs &a, dim3 g, hipStream_t s) {
    hipLaunchKernelGGL((k_theta<PDE, DIM, DERIV, NS>), g, dim3(kThetaBS), 0, s, a);
}
template <int PDE, int DIM, int DERIV>
static ThetaFn theta_ns(int nsp) {
    switch (nsp) {
        case 8: return theta_launcher<PDE, DIM, DERIV, 8>;
        case 12: return theta_launcher<PDE, DIM, DERIV, 12>;
        case 16: return theta_launcher<PDE, DIM, DERIV, 16>;
        case 20: return theta_launcher<PDE, DIM, DERIV, 20>;
        case 24: return theta_launcher<PDE, DIM, DERIV, 24>;
        case 32: return theta_launcher<PDE, DIM, DERIV, 32>;
    }
    return nullptr;
}

struct BdimFns {
    void (*traces)(const TraceArgs &, dim3, hipStream_t);
    ThetaFn theta;
    void (*reduce)(const ThetaRedArgs &, dim3, hipStream_t);
};

template <int PDE, int DIM, int DERIV>
static BdimFns bdim_fns(int nsp) {
    BdimFns f;
    f.traces = [](const TraceArgs &a, dim3 g, hipStream_t s) {
        hipLaunchKernelGGL((k_traces<PDE, DIM>), g, dim3(256), 0, s, a);
    };
    f.theta = theta_ns<PDE, DIM, DERIV>(nsp);
    f.reduce = [](const ThetaRedArgs &a, dim3 g, hipStream_t s) {
        hipLaunchKernelGGL((k_theta_reduce<PDE, DIM, DERIV>), g, dim3(256), 0, s, a);
    };
    return f;
}

static bool pick_bdim(int pde, int dim, int deriv, int nsp, BdimFns &f) {
#define INTI_B(P, D, R) \
    if (pde == P && dim == D && deriv == R) { f = bdim_fns<P, D, R>(nsp); return f.theta != nullptr; }
    INTI_B(INTI_LAPLACE, 2, 0) INTI_B(INTI_LAPLACE, 2, 1)
    INTI_B(INTI_LAPLACE, 3, 0) INTI_B(INTI_LAPLACE, 3, 1)
    INTI_B(INTI_HELMHOLTZ, 2, 0) INTI_B(INTI_HELMHOLTZ, 2, 1)
    INTI_B(INTI_HELMHOLTZ, 3, 0) INTI_B(INTI_HELMHOLTZ, 3, 1)
#undef INTI_B
    return false;
}

// monopole centres: bdim.jl:79-92 with utils.jl:142-149 / 158-171
static void monopoles(const Quad *src, int ns, double radius_mult, std::vector<double> &zs) {
    const int dim = src->dim;
    double xc[3], diag2 = 0.0;
    for (int d = 0; d < dim; ++d) {
        xc[d] = (src->lo[d] + src->hi[d]) / 2;
        diag2 += (src->hi[d] - src->lo[d]) * (src->hi[d] - src->lo[d]);
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

// device buffer owned for the duration of a build; the stream is drained
// before the memory is released so that no kernel still reads it
struct DevBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    ~DevBuf() {
        if (p) {
            if (s) (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
    }
};

#define INTI_ALLOC(ctx, buf, bytes)                                                    \
    do {                                                                               \
        (buf).s = (ctx)->stream;                                                       \
        if (hipMalloc(&(buf).p, (bytes) > 0 ? (bytes) : 1) != hipSuccess)              \
            return set_error((ctx), INTI_ERR_NOMEM, "device allocation failed");       \
    } while (0)

int csr_alloc(Ctx *ctx, int64_t m, int64_t n, int64_t nnz, int cplx, Csr **out);
void csr_free(Csr *c);

int bdim_build(Op *opS, Op *opD, int deriv, int nq, int64_t ne, const int64_t *qtags,
               const int64_t *near_ptr, const int64_t *near_idx, const double *mu,
               const double *params, Csr **outS, Csr **outD, double *diag_out) {
    Ctx *ctx = opS->ctx;
    const Quad *trg = opS->trg, *src = opS->src;
    const int dim = src->dim;
    const int64_t m = trg->n, n = src->n;
    const bool self = near_ptr == nullptr;
    const double oversample = params ? params[0] : 3.0;
    const double radius_mult = params ? params[1] : 1.5;
    const int ns = (int)std::ceil(oversample * nq);
    const int nc = 2 * nq;
    if (ns > kMaxNs || nc > kMaxNc || ns < nc)
        return set_error(ctx, INTI_ERR_UNSUPPORTED,
                         "bdim: need 2*nq <= ns <= 32 and 2*nq <= 24 (ns = ceil(oversample*nq))");
    if (self && (trg != src))
        return set_error(ctx, INTI_ERR_ARG, "bdim: the self pattern needs target == source");
    if (!src->normals) return set_error(ctx, INTI_ERR_ARG, "bdim: source quadrature needs normals");
    if (deriv && !trg->normals)
        return set_error(ctx, INTI_ERR_ARG, "bdim: derivative correction needs target normals");
    if (n == 0 || ne == 0) return set_error(ctx, INTI_ERR_ARG, "bdim: empty source");
    if ((int64_t)nq * ne != n && self)
        return set_error(ctx, INTI_ERR_SIZE, "bdim: nq*ne must equal the number of sources");
    for (int64_t q = 0; q < (int64_t)nq * ne; ++q)
        if (qtags[q] < 0 || qtags[q] >= n) return set_error(ctx, INTI_ERR_ARG, "bdim: qtags out of range");
    const int nsp = ns <= 8 ? 8 : ns <= 12 ? 12 : ns <= 16 ? 16 : ns <= 20 ? 20 : ns <= 24 ? 24 : 32;
    BdimFns fns;
    if (!pick_bdim(opS->pde, dim, deriv, nsp, fns))
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "bdim: no kernel for this operator");

    // ---- host-side lists (element -> targets) and the CSR pattern
    std::vector<int32_t> pair_ptr(ne + 1, 0), pair_tgt;
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t cnt = self ? nq : near_ptr[e + 1] - near_ptr[e];
        for (int64_t t = 0; t < cnt; ++t) {
            const int64_t i = self ? qtags[t + (int64_t)nq * e] : near_idx[near_ptr[e] + t];
            if (i < 0 || i >= m) return set_error(ctx, INTI_ERR_ARG, "bdim: near target out of range");
            pair_tgt.push_back((int32_t)i);
        }
        pair_ptr[e + 1] = (int32_t)pair_tgt.size();
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
    std::vector<int32_t> elems;
    for (int64_t e = 0; e < ne; ++e)
        if (pair_ptr[e + 1] > pair_ptr[e]) elems.push_back((int32_t)e);
    // triplets (pair*nq + r) -> (row, col), CSR pattern
    const int64_t ntri = npairs * nq;
    if (ntri >= (int64_t)INT32_MAX) return set_error(ctx, INTI_ERR_SIZE, "bdim: too many triplets");
    std::vector<int32_t> tri_row(ntri), tri_col(ntri);
    for (int64_t e = 0; e < ne; ++e)
        for (int32_t p = pair_ptr[e]; p < pair_ptr[e + 1]; ++p)
            for (int r = 0; r < nq; ++r) {
                tri_row[(int64_t)p * nq + r] = pair_tgt[p];
                tri_col[(int64_t)p * nq + r] = (int32_t)qtags[r + (int64_t)nq * e];
            }
    TripletPattern pat;
    build_pattern(m, tri_row, tri_col, pat);
    const std::vector<int32_t> &rowptr = pat.rowptr, &colind = pat.colind, &seg = pat.seg, &perm = pat.perm;
    const int64_t nnz = (int64_t)colind.size();

    // ---- device buffers
    std::vector<double> zs;
    monopoles(src, ns, radius_mult, zs);
    std::vector<int32_t> qt32((size_t)nq * ne);
    for (size_t q = 0; q < qt32.size(); ++q) qt32[q] = (int32_t)qtags[q];
    DevBuf dzs, dqt, dglob, dpp, dploc, delems, dB0, dB1, dpart, dth, dtriS, dtriD, ddiag, dseg, dperm, dmu;
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
    // Theta partial sums
    int split = 1;
    int64_t chunk = n;
    if (nloc > 0) {
        const int64_t ntiles = (nloc + kThetaBS - 1) / kThetaBS;
        const int64_t stiles = (n + kThetaTT - 1) / kThetaTT;
        int64_t s = (8LL * ctx->num_cu + ntiles - 1) / ntiles;
        if (s >= 8) s = (s + 7) / 8 * 8;
        if (s > stiles) s = stiles;
        if (s < 1) s = 1;
        chunk = (stiles + s - 1) / s * kThetaTT;
        split = (int)((n + chunk - 1) / chunk);
        INTI_ALLOC(ctx, dpart, cb * split * nloc * nsp);
        ThetaArgs a;
        a.m = m; a.n = n; a.nloc = nloc; a.nsp = nsp; a.split = split; a.chunk = chunk;
        a.glob = (const int32_t *)dglob.p;
        a.tx = trg->coords; a.tn = trg->normals;
        a.sx = src->coords; a.sn = src->normals; a.sw = src->weights;
        a.B0 = (const double2 *)dB0.p; a.B1 = (const double2 *)dB1.p;
        a.part = (double2 *)dpart.p;
        a.p = kp;
        fns.theta(a, dim3((unsigned)(ntiles * split)), st);
        INTI_HIP(ctx, hipGetLastError());
        ThetaRedArgs r;
        r.m = m; r.nloc = nloc; r.ns = ns; r.nsp = nsp; r.split = split;
        r.glob = (const int32_t *)dglob.p;
        r.tx = trg->coords; r.tn = trg->normals; r.zs = (const double *)dzs.p;
        r.mu = (const double *)dmu.p;
        r.part = (const double2 *)dpart.p;
        r.theta = (double2 *)dth.p;
        r.p = kp;
        fns.reduce(r, dim3((unsigned)((nloc * ns + 255) / 256)), st);
        INTI_HIP(ctx, hipGetLastError());
    }
    // batched QR + LS
    if (!elems.empty()) {
        LsArgs a;
        a.n = n; a.nloc = nloc; a.nq = nq; a.ns = ns;
        a.elems = (const int32_t *)delems.p;
        a.qtags = (const int32_t *)dqt.p;
        a.pair_ptr = (const int32_t *)dpp.p;
        a.pair_loc = (const int32_t *)dploc.p;
        a.B0 = (const double2 *)dB0.p; a.B1 = (const double2 *)dB1.p; a.ldb = n;
        a.theta = (const double2 *)dth.p;
        a.triS = (double2 *)dtriS.p; a.triD = (double2 *)dtriD.p;
        a.diag = (double *)ddiag.p;
        hipLaunchKernelGGL(k_ls, dim3((unsigned)elems.size()), dim3(64), 0, st, a);
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

}  // namespace inti
