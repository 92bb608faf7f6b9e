// Batches of independent problems (many boundaries of uneven size, each with
// its own wavenumber) as one block-diagonal operator, applied in one
// persistent launch (k_apply_batch, apply_kernel.hpp) and solved by one
// batched restarted GMRES whose Arnoldi step, for every (problem,
// right-hand side) system at once, is captured in a hipGraph.
//
// Layout: the nodes of problem p are rows [off_p, off_p + n_p) of
// concatenated SoA arrays (stride ntot); a block vector is ntot x R complex,
// column-major (leading dimension ntot); system s = p * R + r is column r
// restricted to the rows of problem p. The reference counterpart is a loop
// over problems calling `gmres(L, rhs)` on each (IterativeSolvers on Inti's
// operators); nothing in the reference batches problems.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <vector>

#include "apply_kernel.hpp"
#include "common.hpp"
#include "lsq.hpp"

namespace inti {

int csr_alloc(Ctx *ctx, int64_t m, int64_t n, int64_t nnz, int cplx, Csr **out);
void csr_free(Csr *c);

struct Batch {
    Ctx *ctx = nullptr;
    int pde = 0, kind = 0, dim = 0, precision = 0;
    int nprob = 0;
    int64_t ntot = 0;
    std::vector<int64_t> off;   // nprob + 1
    std::vector<double> coefs;  // 4 per problem
    std::vector<double> cid;    // 2 per problem
    std::vector<KParams> kp;
    double *coords = nullptr, *normals = nullptr, *weights = nullptr;
    int32_t *row2p = nullptr;
    double *cid_d = nullptr;
    // launch plan
    int64_t chunk_len = 2048;  // source chunk length of a work item
    int grid = 0;              // persistent grid; 0 = 4 blocks per CU
    bool planned = false;
    BatchItem *items = nullptr;
    int64_t nitems = 0;
    BatchProb *probs = nullptr;
    std::vector<BatchProb> hprobs;
    double *part = nullptr;
    size_t part_bytes = 0;
    // corrections (host copies per problem, merged into one CSR on use)
    std::vector<std::vector<int32_t>> c_rowptr, c_colind;
    std::vector<std::vector<double>> c_val;
    Csr *corr = nullptr;
    bool corr_dirty = false;
    ~Batch() {
        void *ps[] = {coords, normals, weights, row2p, cid_d, items, probs, part};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        csr_free(corr);
    }
};

}  // namespace inti

struct inti_batch : inti::Batch {};

namespace inti {

static BatchFn pick_batch(int pde, int dim, int kind, int mixed, int nr) {
    if (pde == INTI_LAPLACE && dim == 2) return batch_pick_lap2(kind, mixed, nr);
    if (pde == INTI_LAPLACE && dim == 3) return batch_pick_lap3(kind, mixed, nr);
    if (pde == INTI_HELMHOLTZ && dim == 2) return batch_pick_helm2(kind, mixed, nr);
    if (pde == INTI_HELMHOLTZ && dim == 3) return batch_pick_helm3(kind, mixed, nr);
    return nullptr;
}

// ------------------------------------------------------------------ plan
static int batch_plan(Batch *B) {
    if (B->planned) return INTI_OK;
    Ctx *ctx = B->ctx;
    const int64_t L = std::max<int64_t>(kApplyBS, (B->chunk_len + kApplyBS - 1) / kApplyBS * kApplyBS);
    B->hprobs.assign(B->nprob, BatchProb{});
    struct It { BatchItem w; int64_t cost; };
    std::vector<It> its;
    int64_t poff = 0;
    for (int p = 0; p < B->nprob; ++p) {
        const int64_t n = B->off[p + 1] - B->off[p];
        BatchProb &P = B->hprobs[p];
        P.off = B->off[p];
        P.n = n;
        P.kp = B->kp[p];
        const int64_t stiles = (n + kApplyBS - 1) / kApplyBS;
        int64_t split = (n + L - 1) / L;
        split = std::max<int64_t>(1, std::min(split, stiles));
        const int64_t chunk = (stiles + split - 1) / split * kApplyBS;
        split = (n + chunk - 1) / chunk;
        P.chunk = chunk;
        P.split = (int32_t)std::max<int64_t>(split, 1);
        P.poff = poff;
        poff += (int64_t)16 * P.split * n;  // room for 16 right-hand sides
        for (int64_t t = 0; t < stiles; ++t)
            for (int64_t c = 0; c < P.split; ++c) {
                const int64_t rows = std::min<int64_t>(kApplyBS, n - t * kApplyBS);
                const int64_t cols = std::min<int64_t>(chunk, n - c * chunk);
                its.push_back({BatchItem{p, (int32_t)t, (int32_t)c, 0}, rows * cols});
            }
    }
    std::stable_sort(its.begin(), its.end(), [](const It &a, const It &b) { return a.cost > b.cost; });
    std::vector<BatchItem> hit(its.size());
    for (size_t q = 0; q < its.size(); ++q) hit[q] = its[q].w;
    if (B->items) (void)hipFree(B->items);
    if (B->probs) (void)hipFree(B->probs);
    B->items = nullptr;
    B->probs = nullptr;
    B->nitems = (int64_t)hit.size();
    const size_t cw = B->pde == INTI_HELMHOLTZ ? 2 : 1;
    const size_t pbytes = sizeof(double) * cw * (size_t)std::max<int64_t>(poff, 1);
    if (hipMalloc((void **)&B->items, sizeof(BatchItem) * std::max<size_t>(hit.size(), 1)) != hipSuccess ||
        hipMalloc((void **)&B->probs, sizeof(BatchProb) * std::max(B->nprob, 1)) != hipSuccess)
        return set_error(ctx, INTI_ERR_NOMEM, "batch: plan allocation failed");
    if (pbytes > B->part_bytes) {
        if (B->part) (void)hipFree(B->part);
        B->part = nullptr;
        if (hipMalloc((void **)&B->part, pbytes) != hipSuccess)
            return set_error(ctx, INTI_ERR_NOMEM, "batch: partial-sum allocation failed");
        B->part_bytes = pbytes;
    }
    hipStream_t s = ctx->stream;
    if (!hit.empty())
        INTI_HIP(ctx, hipMemcpyAsync(B->items, hit.data(), sizeof(BatchItem) * hit.size(), hipMemcpyHostToDevice, s));
    INTI_HIP(ctx, hipMemcpyAsync(B->probs, B->hprobs.data(), sizeof(BatchProb) * B->nprob, hipMemcpyHostToDevice, s));
    INTI_HIP(ctx, hipStreamSynchronize(s));
    B->planned = true;
    return INTI_OK;
}

// merged block-diagonal correction, global indices
static int batch_corr(Batch *B) {
    if (!B->corr_dirty) return INTI_OK;
    Ctx *ctx = B->ctx;
    const int cw = B->pde == INTI_HELMHOLTZ ? 2 : 1;
    std::vector<int32_t> rp(B->ntot + 1, 0), ci;
    std::vector<double> va;
    for (int p = 0; p < B->nprob; ++p) {
        const int64_t n = B->off[p + 1] - B->off[p], o = B->off[p];
        const bool has = !B->c_rowptr[p].empty();
        for (int64_t i = 0; i < n; ++i) {
            const int32_t a = has ? B->c_rowptr[p][i] : 0, b = has ? B->c_rowptr[p][i + 1] : 0;
            for (int32_t q = a; q < b; ++q) {
                ci.push_back((int32_t)(B->c_colind[p][q] + o));
                for (int c = 0; c < cw; ++c) va.push_back(B->c_val[p][(size_t)cw * q + c]);
            }
            rp[o + i + 1] = (int32_t)ci.size();
        }
    }
    csr_free(B->corr);
    B->corr = nullptr;
    Csr *c = nullptr;
    int rc = csr_alloc(ctx, B->ntot, B->ntot, (int64_t)ci.size(), cw == 2, &c);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    if (hipMemcpyAsync(c->rowptr, rp.data(), sizeof(int32_t) * rp.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        (!ci.empty() && (hipMemcpyAsync(c->colind, ci.data(), sizeof(int32_t) * ci.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
                         hipMemcpyAsync(c->val, va.data(), sizeof(double) * va.size(), hipMemcpyHostToDevice, s) != hipSuccess)) ||
        hipStreamSynchronize(s) != hipSuccess) {
        csr_free(c);
        return set_error(ctx, INTI_ERR_HIP, "batch: correction upload failed");
    }
    B->corr = c;
    B->corr_dirty = false;
    return INTI_OK;
}

// ----------------------------------------------------------------- apply
struct BReduceArgs {
    int64_t ntot;
    const int32_t *row2p;
    const BatchProb *probs;
    const double *part;
    int nr;
    const int32_t *rowptr, *colind;
    const double *cval;
    const double *cid;  // 2 per problem
    const double *x;
    int64_t ldx;
    double *y;
    int64_t ldy;
    double ar, ai, br, bi;
    int beta_zero;
};

// y = alpha (post_p * sum of partials + correction x + cid_p x) + beta y
template <bool CPLX>
__global__ __launch_bounds__(256) void k_reduce_batch(BReduceArgs a) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= a.ntot) return;
    const int r = blockIdx.y;
    const int p = a.row2p[g];
    const BatchProb P = a.probs[p];
    const int64_t i = g - P.off;
    const double *x = a.x + (CPLX ? 2 : 1) * (int64_t)r * a.ldx;
    double *y = a.y + (CPLX ? 2 : 1) * (int64_t)r * a.ldy;
    const int64_t base = P.poff + (int64_t)r * P.split * P.n;
    double sr = 0.0, si = 0.0;
    if constexpr (CPLX) {
        const double2 *pp = reinterpret_cast<const double2 *>(a.part) + base;
        for (int c = 0; c < P.split; ++c) {
            const double2 v = pp[(int64_t)c * P.n + i];
            sr += v.x; si += v.y;
        }
    } else {
        for (int c = 0; c < P.split; ++c) sr += a.part[base + (int64_t)c * P.n + i];
    }
    {
        const double tr = P.kp.postr * sr - P.kp.posti * si, ti = P.kp.postr * si + P.kp.posti * sr;
        sr = tr; si = ti;
    }
    if (a.rowptr) {
        for (int32_t q = a.rowptr[g]; q < a.rowptr[g + 1]; ++q) {
            const int32_t j = a.colind[q];
            if constexpr (CPLX) {
                const double vr = a.cval[2 * q], vi = a.cval[2 * q + 1];
                const double xr = x[2 * j], xi = x[2 * j + 1];
                sr += vr * xr - vi * xi;
                si += vr * xi + vi * xr;
            } else {
                sr += a.cval[q] * x[j];
            }
        }
    }
    const double cr = a.cid[2 * p], ci = a.cid[2 * p + 1];
    if constexpr (CPLX) {
        const double xr = x[2 * g], xi = x[2 * g + 1];
        sr += cr * xr - ci * xi;
        si += cr * xi + ci * xr;
        double yr = a.ar * sr - a.ai * si, yi = a.ar * si + a.ai * sr;
        if (!a.beta_zero) {
            const double o_r = y[2 * g], o_i = y[2 * g + 1];
            yr += a.br * o_r - a.bi * o_i;
            yi += a.br * o_i + a.bi * o_r;
        }
        y[2 * g] = yr;
        y[2 * g + 1] = yi;
    } else {
        sr += cr * x[g];
        double yr = a.ar * sr;
        if (!a.beta_zero) yr += a.br * y[g];
        y[g] = yr;
    }
}

// everything an apply needs allocated (before graph capture)
static int batch_reserve(Batch *B) {
    int rc = batch_plan(B);
    if (rc) return rc;
    return batch_corr(B);
}

static int batch_apply_dev(Batch *B, int nrhs, const double *x, int64_t ldx, double *y, int64_t ldy,
                           const double alpha[2], const double beta[2]) {
    Ctx *ctx = B->ctx;
    int rc = batch_reserve(B);
    if (rc) return rc;
    const bool cplx = B->pde == INTI_HELMHOLTZ;
    const int cw = cplx ? 2 : 1;
    hipStream_t s = ctx->stream;
    const int G = B->grid > 0 ? B->grid : 4 * ctx->num_cu;
    const bool prof = ctx->profiling != 0;
    for (int r0 = 0; r0 < nrhs;) {
        const int left = nrhs - r0;
        const int nr = left >= 16 ? 16 : left >= 4 ? 4 : 1;
        BatchFn fn = pick_batch(B->pde, B->dim, B->kind, B->precision == INTI_MIXED, nr);
        if (!fn) return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch: no kernel for this operator");
        BatchArgs a;
        a.items = B->items; a.nitems = B->nitems; a.probs = B->probs; a.ntot = B->ntot;
        a.coords = B->coords; a.normals = B->normals; a.weights = B->weights;
        a.x = x + (int64_t)r0 * ldx * cw; a.ldx = ldx; a.part = B->part;
        if (prof && r0 == 0) INTI_HIP(ctx, hipEventRecord(ctx->ev0, s));
        if (B->nitems > 0) fn(a, dim3((unsigned)std::min<int64_t>(G, B->nitems)), s);
        INTI_HIP(ctx, hipGetLastError());
        BReduceArgs ra;
        ra.ntot = B->ntot; ra.row2p = B->row2p; ra.probs = B->probs; ra.part = B->part; ra.nr = nr;
        ra.rowptr = B->corr ? B->corr->rowptr : nullptr;
        ra.colind = B->corr ? B->corr->colind : nullptr;
        ra.cval = B->corr ? B->corr->val : nullptr;
        ra.cid = B->cid_d;
        ra.x = a.x; ra.ldx = ldx; ra.y = y + (int64_t)r0 * ldy * cw; ra.ldy = ldy;
        ra.ar = alpha[0]; ra.ai = alpha[1]; ra.br = beta[0]; ra.bi = beta[1];
        ra.beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
        const dim3 rg((unsigned)((B->ntot + 255) / 256), (unsigned)nr);
        if (cplx) hipLaunchKernelGGL(k_reduce_batch<true>, rg, dim3(256), 0, s, ra);
        else hipLaunchKernelGGL(k_reduce_batch<false>, rg, dim3(256), 0, s, ra);
        INTI_HIP(ctx, hipGetLastError());
        if (prof && r0 + nr >= nrhs) INTI_HIP(ctx, hipEventRecord(ctx->ev1, s));
        r0 += nr;
    }
    return INTI_OK;
}

// --------------------------------------------------------- batched GMRES
// Scalar helpers (complex as double2, or real)
template <bool CPLX>
struct BS;
template <>
struct BS<true> {
    using T = double2;
    __device__ static T zero() { return make_double2(0.0, 0.0); }
    __device__ static T conjmul(T a, T b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
    __device__ static T mul(T a, T b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
    __device__ static T add(T a, T b) { return make_double2(a.x + b.x, a.y + b.y); }
    __device__ static T sub(T a, T b) { return make_double2(a.x - b.x, a.y - b.y); }
    __device__ static T scale(T a, double s) { return make_double2(a.x * s, a.y * s); }
    __device__ static double abs2(T a) { return a.x * a.x + a.y * a.y; }
};
template <>
struct BS<false> {
    using T = double;
    __device__ static T zero() { return 0.0; }
    __device__ static T conjmul(T a, T b) { return a * b; }
    __device__ static T mul(T a, T b) { return a * b; }
    __device__ static T add(T a, T b) { return a + b; }
    __device__ static T sub(T a, T b) { return a - b; }
    __device__ static T scale(T a, double s) { return a * s; }
    __device__ static double abs2(T a) { return a * a; }
};

constexpr int kBG = 256;        // threads of the vector kernels
constexpr int kSegRows = 4096;  // rows of one reduction segment

struct BGState {
    int k;  // current Arnoldi column (shared by all systems)
    int pad;
};

struct BGDims {
    int64_t ntot;
    int R;         // right-hand sides (a divisor of 256)
    int m;         // restart length
    int nsys;      // nprob * R
    int nseg;
    const int32_t *row2p;
    const int32_t *segp;     // segment -> problem
    const int64_t *segr;     // segment -> [r0, r1) rows (2 per segment)
    const int32_t *segptr;   // problem -> segments [segptr[p], segptr[p+1])
};

// vk = V[:, k] (ntot x R block)
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bcopy(const typename BS<CPLX>::T *V, const BGState *st,
                                               typename BS<CPLX>::T *vk, int64_t len) {
    const int64_t e = (int64_t)blockIdx.x * kBG + threadIdx.x;
    if (e < len) vk[e] = V[(int64_t)st->k * len + e];
}

// partial[(seg * (m+1) + j) * R + r] = sum over the segment's rows of conj(V_j) w
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bdots(BGDims d, const typename BS<CPLX>::T *V,
                                               const BGState *st, const typename BS<CPLX>::T *w,
                                               typename BS<CPLX>::T *partial) {
    using S = BS<CPLX>;
    __shared__ typename S::T red[kBG];
    const int seg = blockIdx.x, j = blockIdx.y;
    if (j > st->k) return;
    const int64_t r0 = d.segr[2 * seg], r1 = d.segr[2 * seg + 1];
    const int R = d.R;
    const int t = threadIdx.x;
    const int r = t % R;
    const int64_t len = d.ntot * R;
    const typename S::T *vj = V + (int64_t)j * len;
    typename S::T acc = S::zero();
    const int64_t cnt = (r1 - r0) * R;
    for (int64_t e = t; e < cnt; e += kBG) {
        const int64_t row = r0 + e / R;
        const int64_t o = row + d.ntot * (int64_t)(e % R);
        acc = S::add(acc, S::conjmul(vj[o], w[o]));
    }
    red[t] = acc;
    __syncthreads();
    if (t < R) {
        typename S::T s = S::zero();
        for (int q = t; q < kBG; q += R) s = S::add(s, red[q]);
        partial[((int64_t)seg * (d.m + 1) + j) * R + r] = s;
    }
}

// h[sys][j] = sum over the problem's segments; H[sys][j][k] += h   (sys, j <= k)
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bsum(BGDims d, const typename BS<CPLX>::T *partial,
                                              const BGState *st, typename BS<CPLX>::T *h,
                                              typename BS<CPLX>::T *H) {
    using S = BS<CPLX>;
    const int64_t e = (int64_t)blockIdx.x * kBG + threadIdx.x;
    const int k = st->k;
    const int64_t total = (int64_t)d.nsys * (k + 1);
    if (e >= total) return;
    const int sys = (int)(e / (k + 1)), j = (int)(e % (k + 1));
    const int p = sys / d.R, r = sys % d.R;
    typename S::T s = S::zero();
    for (int g = d.segptr[p]; g < d.segptr[p + 1]; ++g)
        s = S::add(s, partial[((int64_t)g * (d.m + 1) + j) * d.R + r]);
    h[(int64_t)sys * (d.m + 1) + j] = s;
    typename S::T *Hc = H + (int64_t)sys * (d.m + 1) * d.m + (int64_t)(d.m + 1) * k;
    Hc[j] = S::add(Hc[j], s);
}

// w -= sum_{j<=k} V_j h[sys][j]
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bupdate(BGDims d, const typename BS<CPLX>::T *V,
                                                 const BGState *st, const typename BS<CPLX>::T *h,
                                                 typename BS<CPLX>::T *w) {
    using S = BS<CPLX>;
    const int64_t len = d.ntot * d.R;
    const int64_t e = (int64_t)blockIdx.x * kBG + threadIdx.x;
    if (e >= len) return;
    const int64_t row = e % d.ntot;
    const int r = (int)(e / d.ntot);
    const int sys = d.row2p[row] * d.R + r;
    const int k = st->k;
    typename S::T s = w[e];
    for (int j = 0; j <= k; ++j) s = S::sub(s, S::mul(V[(int64_t)j * len + e], h[(int64_t)sys * (d.m + 1) + j]));
    w[e] = s;
}

// npart[seg * R + r] = sum over the segment of |w|^2
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bnorm(BGDims d, const typename BS<CPLX>::T *w, double *npart) {
    using S = BS<CPLX>;
    __shared__ double red[kBG];
    const int seg = blockIdx.x;
    const int64_t r0 = d.segr[2 * seg], r1 = d.segr[2 * seg + 1];
    const int R = d.R, t = threadIdx.x;
    double acc = 0.0;
    const int64_t cnt = (r1 - r0) * R;
    for (int64_t e = t; e < cnt; e += kBG) {
        const int64_t row = r0 + e / R;
        acc += S::abs2(w[row + d.ntot * (int64_t)(e % R)]);
    }
    red[t] = acc;
    __syncthreads();
    if (t < R) {
        double s = 0.0;
        for (int q = t; q < kBG; q += R) s += red[q];
        npart[(int64_t)seg * R + t] = s;
    }
}

// nrm[sys] = sqrt(sum over the problem's segments); zero for inactive systems
// when `mask` (Arnoldi), plain norms otherwise (residuals)
__global__ __launch_bounds__(kBG) void k_bnormsum(BGDims d, const double *npart, const int *active,
                                                  double *nrm) {
    const int sys = blockIdx.x * kBG + threadIdx.x;
    if (sys >= d.nsys) return;
    const int p = sys / d.R, r = sys % d.R;
    double s = 0.0;
    for (int g = d.segptr[p]; g < d.segptr[p + 1]; ++g) s += npart[(int64_t)g * d.R + r];
    nrm[sys] = (active && !active[sys]) ? 0.0 : sqrt(s);
}

// V[:, k+1] = w / nrm[sys] (0 where nrm is 0)
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bscale(BGDims d, typename BS<CPLX>::T *V, const BGState *st,
                                                const typename BS<CPLX>::T *w, const double *nrm) {
    using S = BS<CPLX>;
    const int64_t len = d.ntot * d.R;
    const int64_t e = (int64_t)blockIdx.x * kBG + threadIdx.x;
    if (e >= len) return;
    const int sys = d.row2p[e % d.ntot] * d.R + (int)(e / d.ntot);
    const double nv = nrm[sys];
    V[(int64_t)(st->k + 1) * len + e] = S::scale(w[e], nv > 0.0 ? 1.0 / nv : 0.0);
}

// v = src / nrm[sys] into V[:, 0]
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bstart(BGDims d, typename BS<CPLX>::T *V,
                                                const typename BS<CPLX>::T *src, const double *nrm,
                                                const int *active) {
    using S = BS<CPLX>;
    const int64_t len = d.ntot * d.R;
    const int64_t e = (int64_t)blockIdx.x * kBG + threadIdx.x;
    if (e >= len) return;
    const int sys = d.row2p[e % d.ntot] * d.R + (int)(e / d.ntot);
    const double nv = nrm[sys];
    V[e] = S::scale(src[e], (active[sys] && nv > 0.0) ? 1.0 / nv : 0.0);
}

// per system: H = 0, g = (rn, 0, ...), kk = 0; k = 0
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bsetup(BGDims d, typename BS<CPLX>::T *H, typename BS<CPLX>::T *g,
                                                const double *rn, int *kk, BGState *st) {
    using S = BS<CPLX>;
    const int sys = blockIdx.x * kBG + threadIdx.x;
    if (sys == 0) st->k = 0;
    if (sys >= d.nsys) return;
    typename S::T *Hs = H + (int64_t)sys * (d.m + 1) * d.m;
    for (int q = 0; q < (d.m + 1) * d.m; ++q) Hs[q] = S::zero();
    typename S::T *gs = g + (int64_t)sys * (d.m + 1);
    for (int q = 0; q <= d.m; ++q) gs[q] = S::zero();
    if constexpr (CPLX) gs[0] = make_double2(rn[sys], 0.0);
    else gs[0] = rn[sys];
    kk[sys] = 0;
}

// Givens update of Hessenberg column k of every active system; a system
// whose residual estimate reaches its tolerance stops (kk = k + 1).
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bgivens(BGDims d, typename BS<CPLX>::T *H, double *cs,
                                                 typename BS<CPLX>::T *sn, typename BS<CPLX>::T *g,
                                                 const double *wnorm, const double *tol, int *active,
                                                 int *kk, double *res, const BGState *st) {
    const int sys = blockIdx.x * kBG + threadIdx.x;
    if (sys >= d.nsys || !active[sys]) return;
    const int k = st->k, m = d.m;
    if constexpr (CPLX) {
        double2 *h = H + (int64_t)sys * (m + 1) * m + (int64_t)(m + 1) * k;
        double *c_ = cs + (int64_t)sys * m;
        double2 *s_ = sn + (int64_t)sys * m;
        double2 *gg = g + (int64_t)sys * (m + 1);
        h[k + 1] = make_double2(wnorm[sys], 0.0);
        for (int i = 0; i < k; ++i) {
            const double c = c_[i];
            const double2 s = s_[i], a = h[i], b = h[i + 1];
            h[i] = make_double2(c * a.x + s.x * b.x - s.y * b.y, c * a.y + s.x * b.y + s.y * b.x);
            h[i + 1] = make_double2(-(s.x * a.x + s.y * a.y) + c * b.x, -(s.x * a.y - s.y * a.x) + c * b.y);
        }
        const double2 a = h[k];
        const double bn = h[k + 1].x;
        const double an = hypot(a.x, a.y);
        double c;
        double2 s;
        if (an == 0.0) {
            c = 0.0;
            s = make_double2(1.0, 0.0);
            h[k] = make_double2(bn, 0.0);
        } else {
            const double t = hypot(an, bn);
            c = an / t;
            s = make_double2(a.x / an * bn / t, a.y / an * bn / t);
            h[k] = make_double2(a.x / an * t, a.y / an * t);
        }
        h[k + 1] = make_double2(0.0, 0.0);
        c_[k] = c;
        s_[k] = s;
        const double2 gk = gg[k];
        gg[k] = make_double2(c * gk.x, c * gk.y);
        gg[k + 1] = make_double2(-(s.x * gk.x + s.y * gk.y), -(s.x * gk.y - s.y * gk.x));
        res[(int64_t)sys * m + k] = hypot(gg[k + 1].x, gg[k + 1].y);
    } else {
        double *h = H + (int64_t)sys * (m + 1) * m + (int64_t)(m + 1) * k;
        double *c_ = cs + (int64_t)sys * m;
        double *s_ = sn + (int64_t)sys * m;
        double *gg = g + (int64_t)sys * (m + 1);
        h[k + 1] = wnorm[sys];
        for (int i = 0; i < k; ++i) {
            const double c = c_[i], s = s_[i], a = h[i], b = h[i + 1];
            h[i] = c * a + s * b;
            h[i + 1] = -s * a + c * b;
        }
        const double a = h[k], b = h[k + 1];
        double c, s;
        if (b == 0.0) { c = 1.0; s = 0.0; }
        else {
            const double t = hypot(a, b);
            c = a / t; s = b / t;
            h[k] = t;
        }
        h[k + 1] = 0.0;
        c_[k] = c; s_[k] = s;
        const double gk = gg[k];
        gg[k] = c * gk;
        gg[k + 1] = -s * gk;
        res[(int64_t)sys * m + k] = fabs(gg[k + 1]);
    }
    if (res[(int64_t)sys * m + k] <= tol[sys]) {
        active[sys] = 0;
        kk[sys] = k + 1;
    }
}

__global__ void k_bbump(BGState *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->k += 1;
}

// systems still active after `steps` columns use all of them
__global__ __launch_bounds__(kBG) void k_bfinal(int nsys, const int *active, int *kk, int steps) {
    const int sys = blockIdx.x * kBG + threadIdx.x;
    if (sys < nsys && active[sys]) kk[sys] = steps;
}

// y = H^{-1} g (upper triangular, kk x kk) per system
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_btrisolve(BGDims d, const typename BS<CPLX>::T *H,
                                                   const typename BS<CPLX>::T *g, const int *kk,
                                                   typename BS<CPLX>::T *y) {
    using S = BS<CPLX>;
    const int sys = blockIdx.x * kBG + threadIdx.x;
    if (sys >= d.nsys) return;
    const int m = d.m, n = kk[sys];
    const typename S::T *Hs = H + (int64_t)sys * (m + 1) * m;
    const typename S::T *gs = g + (int64_t)sys * (m + 1);
    typename S::T *ys = y + (int64_t)sys * (m + 1);
    for (int i = n - 1; i >= 0; --i) {
        typename S::T t = gs[i];
        for (int j = i + 1; j < n; ++j) t = S::sub(t, S::mul(Hs[i + (int64_t)(m + 1) * j], ys[j]));
        const typename S::T dd = Hs[i + (int64_t)(m + 1) * i];
        if constexpr (CPLX) {
            const double den = dd.x * dd.x + dd.y * dd.y;
            ys[i] = den > 0.0 ? make_double2((t.x * dd.x + t.y * dd.y) / den, (t.y * dd.x - t.x * dd.y) / den)
                              : make_double2(0.0, 0.0);
        } else {
            ys[i] = dd != 0.0 ? t / dd : 0.0;
        }
    }
}

// x += sum_{j < kk[sys]} V_j y[sys][j]
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bxupdate(BGDims d, const typename BS<CPLX>::T *V, const int *kk,
                                                  const typename BS<CPLX>::T *y, typename BS<CPLX>::T *x,
                                                  int64_t ldx) {
    using S = BS<CPLX>;
    const int64_t len = d.ntot * d.R;
    const int64_t e = (int64_t)blockIdx.x * kBG + threadIdx.x;
    if (e >= len) return;
    const int64_t row = e % d.ntot;
    const int r = (int)(e / d.ntot);
    const int sys = d.row2p[row] * d.R + r;
    typename S::T *xp = x + row + ldx * r;
    typename S::T s = *xp;
    for (int j = 0; j < kk[sys]; ++j) s = S::add(s, S::mul(V[(int64_t)j * len + e], y[(int64_t)sys * (d.m + 1) + j]));
    *xp = s;
}

// dst (ntot x R, ld ntot) = src (ld lds)
template <bool CPLX>
__global__ __launch_bounds__(kBG) void k_bgather(int64_t ntot, int R, const typename BS<CPLX>::T *src,
                                                 int64_t lds, typename BS<CPLX>::T *dst) {
    const int64_t e = (int64_t)blockIdx.x * kBG + threadIdx.x;
    if (e >= ntot * R) return;
    dst[e] = src[(e % ntot) + lds * (e / ntot)];
}

struct BWork {
    int m = 0, R = 0, nsys = 0, nseg = 0;
    void *V = nullptr, *w = nullptr, *vk = nullptr, *r = nullptr, *partial = nullptr, *h = nullptr;
    void *H = nullptr, *sn = nullptr, *g = nullptr, *y = nullptr;
    double *cs = nullptr, *res = nullptr, *npart = nullptr, *nrm = nullptr, *tol = nullptr;
    int *active = nullptr, *kk = nullptr;
    int32_t *segp = nullptr, *segptr = nullptr;
    int64_t *segr = nullptr;
    BGState *st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    // Teardown order: drain the stream (an error return may leave launches
    // queued), destroy the executable graph and the graph whose kernel nodes
    // hold pointers into the buffers, and only then release the buffers.
    ~BWork() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        void *ps[] = {V, w, vk, r, partial, h, H, sn, g, y, cs, res, npart, nrm, tol, active, kk, segp, segptr, segr, st};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
};

template <bool CPLX>
static int barnoldi(Batch *B, BWork &W, const BGDims &d, hipStream_t s) {
    using T = typename BS<CPLX>::T;
    const int64_t len = B->ntot * W.R;
    const dim3 gv((unsigned)((len + kBG - 1) / kBG));
    const dim3 gs((unsigned)((W.nsys + kBG - 1) / kBG));
    hipLaunchKernelGGL(k_bcopy<CPLX>, gv, dim3(kBG), 0, s, (const T *)W.V, W.st, (T *)W.vk, len);
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    int rc = batch_apply_dev(B, W.R, (const double *)W.vk, B->ntot, (double *)W.w, B->ntot, one, zero);
    if (rc) return rc;
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_bdots<CPLX>, dim3((unsigned)W.nseg, (unsigned)(W.m + 1)), dim3(kBG), 0, s, d,
                           (const T *)W.V, W.st, (const T *)W.w, (T *)W.partial);
        const int64_t ns = (int64_t)W.nsys * (W.m + 1);
        hipLaunchKernelGGL(k_bsum<CPLX>, dim3((unsigned)((ns + kBG - 1) / kBG)), dim3(kBG), 0, s, d,
                           (const T *)W.partial, W.st, (T *)W.h, (T *)W.H);
        hipLaunchKernelGGL(k_bupdate<CPLX>, gv, dim3(kBG), 0, s, d, (const T *)W.V, W.st, (const T *)W.h, (T *)W.w);
    }
    hipLaunchKernelGGL(k_bnorm<CPLX>, dim3((unsigned)W.nseg), dim3(kBG), 0, s, d, (const T *)W.w, W.npart);
    hipLaunchKernelGGL(k_bnormsum, gs, dim3(kBG), 0, s, d, (const double *)W.npart, (const int *)W.active, W.nrm);
    hipLaunchKernelGGL(k_bscale<CPLX>, gv, dim3(kBG), 0, s, d, (T *)W.V, W.st, (const T *)W.w, (const double *)W.nrm);
    hipLaunchKernelGGL(k_bgivens<CPLX>, gs, dim3(kBG), 0, s, d, (T *)W.H, W.cs, (T *)W.sn, (T *)W.g,
                       (const double *)W.nrm, (const double *)W.tol, W.active, W.kk, W.res, (const BGState *)W.st);
    hipLaunchKernelGGL(k_bbump, dim3(1), dim3(64), 0, s, W.st);
    return hipGetLastError() == hipSuccess ? INTI_OK : set_error(B->ctx, INTI_ERR_HIP, "batch gmres: launch failed");
}

// r = b - A x; nrm[sys] = ||r_sys||
template <bool CPLX>
static int bresidual(Batch *B, BWork &W, const BGDims &d, const double *b, const double *x, int64_t ldx) {
    using T = typename BS<CPLX>::T;
    Ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    const int64_t len = B->ntot * W.R;
    hipLaunchKernelGGL(k_bgather<CPLX>, dim3((unsigned)((len + kBG - 1) / kBG)), dim3(kBG), 0, s, B->ntot, W.R,
                       (const T *)b, B->ntot, (T *)W.r);
    const double m1[2] = {-1.0, 0.0}, one[2] = {1.0, 0.0};
    int rc = batch_apply_dev(B, W.R, x, ldx, (double *)W.r, B->ntot, m1, one);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bnorm<CPLX>, dim3((unsigned)W.nseg), dim3(kBG), 0, s, d, (const T *)W.r, W.npart);
    hipLaunchKernelGGL(k_bnormsum, dim3((unsigned)((W.nsys + kBG - 1) / kBG)), dim3(kBG), 0, s, d,
                       (const double *)W.npart, (const int *)nullptr, W.nrm);
    return hipGetLastError() == hipSuccess ? INTI_OK : set_error(ctx, INTI_ERR_HIP, "batch gmres: launch failed");
}

template <bool CPLX>
static int batch_gmres_t(Batch *B, int R, const double *b, double *x, int restart, int maxiter, double rtol,
                         double atol, int use_graph, int *iters, double *relres) {
    using T = typename BS<CPLX>::T;
    Ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    const size_t es = CPLX ? 16 : 8;
    const int m = restart;
    BWork W;
    W.stream = s;
    W.m = m; W.R = R; W.nsys = B->nprob * R;
    // reduction segments: rows of one problem, at most kSegRows each
    std::vector<int32_t> segp, segptr(B->nprob + 1, 0);
    std::vector<int64_t> segr;
    for (int p = 0; p < B->nprob; ++p) {
        for (int64_t r0 = B->off[p]; r0 < B->off[p + 1]; r0 += kSegRows) {
            segp.push_back(p);
            segr.push_back(r0);
            segr.push_back(std::min<int64_t>(r0 + kSegRows, B->off[p + 1]));
        }
        segptr[p + 1] = (int32_t)segp.size();
    }
    W.nseg = (int)segp.size();
    const int64_t len = B->ntot * R;
    const size_t hm = (size_t)(m + 1) * m;
    bool ok = hipMalloc(&W.V, es * (size_t)len * (m + 1)) == hipSuccess && hipMalloc(&W.w, es * len) == hipSuccess &&
              hipMalloc(&W.vk, es * len) == hipSuccess && hipMalloc(&W.r, es * len) == hipSuccess &&
              hipMalloc(&W.partial, es * (size_t)std::max(W.nseg, 1) * (m + 1) * R) == hipSuccess &&
              hipMalloc(&W.h, es * (size_t)W.nsys * (m + 1)) == hipSuccess &&
              hipMalloc(&W.H, es * (size_t)W.nsys * hm) == hipSuccess &&
              hipMalloc(&W.sn, es * (size_t)W.nsys * m) == hipSuccess &&
              hipMalloc(&W.g, es * (size_t)W.nsys * (m + 1)) == hipSuccess &&
              hipMalloc(&W.y, es * (size_t)W.nsys * (m + 1)) == hipSuccess &&
              hipMalloc((void **)&W.cs, sizeof(double) * W.nsys * m) == hipSuccess &&
              hipMalloc((void **)&W.res, sizeof(double) * W.nsys * m) == hipSuccess &&
              hipMalloc((void **)&W.npart, sizeof(double) * std::max(W.nseg, 1) * R) == hipSuccess &&
              hipMalloc((void **)&W.nrm, sizeof(double) * W.nsys) == hipSuccess &&
              hipMalloc((void **)&W.tol, sizeof(double) * W.nsys) == hipSuccess &&
              hipMalloc((void **)&W.active, sizeof(int) * W.nsys) == hipSuccess &&
              hipMalloc((void **)&W.kk, sizeof(int) * W.nsys) == hipSuccess &&
              hipMalloc((void **)&W.segp, sizeof(int32_t) * std::max(W.nseg, 1)) == hipSuccess &&
              hipMalloc((void **)&W.segptr, sizeof(int32_t) * (B->nprob + 1)) == hipSuccess &&
              hipMalloc((void **)&W.segr, sizeof(int64_t) * std::max<size_t>(segr.size(), 2)) == hipSuccess &&
              hipMalloc((void **)&W.st, sizeof(BGState)) == hipSuccess;
    if (!ok) return set_error(ctx, INTI_ERR_NOMEM, "batch gmres: workspace allocation failed");
    INTI_HIP(ctx, hipMemcpyAsync(W.segp, segp.data(), sizeof(int32_t) * segp.size(), hipMemcpyHostToDevice, s));
    INTI_HIP(ctx, hipMemcpyAsync(W.segptr, segptr.data(), sizeof(int32_t) * segptr.size(), hipMemcpyHostToDevice, s));
    INTI_HIP(ctx, hipMemcpyAsync(W.segr, segr.data(), sizeof(int64_t) * segr.size(), hipMemcpyHostToDevice, s));
    BGDims d;
    d.ntot = B->ntot; d.R = R; d.m = m; d.nsys = W.nsys; d.nseg = W.nseg;
    d.row2p = B->row2p; d.segp = W.segp; d.segr = W.segr; d.segptr = W.segptr;
    int rc = batch_reserve(B);
    if (rc) return rc;
    const dim3 gv((unsigned)((len + kBG - 1) / kBG)), gs((unsigned)((W.nsys + kBG - 1) / kBG));
    // tolerances from ||b_sys||
    hipLaunchKernelGGL(k_bnorm<CPLX>, dim3((unsigned)W.nseg), dim3(kBG), 0, s, d, (const T *)b, W.npart);
    hipLaunchKernelGGL(k_bnormsum, gs, dim3(kBG), 0, s, d, (const double *)W.npart, (const int *)nullptr, W.nrm);
    std::vector<double> bn(W.nsys), rn(W.nsys), tol(W.nsys);
    INTI_HIP(ctx, hipMemcpyAsync(bn.data(), W.nrm, sizeof(double) * W.nsys, hipMemcpyDeviceToHost, s));
    INTI_HIP(ctx, hipStreamSynchronize(s));
    for (int q = 0; q < W.nsys; ++q) tol[q] = std::max(rtol * bn[q], atol);
    INTI_HIP(ctx, hipMemcpyAsync(W.tol, tol.data(), sizeof(double) * W.nsys, hipMemcpyHostToDevice, s));
    const int saved_prof = ctx->profiling;
    ctx->profiling = 0;
    struct Restore { Ctx *c; int v; ~Restore() { c->profiling = v; } } restore{ctx, saved_prof};
    rc = bresidual<CPLX>(B, W, d, b, x, B->ntot);
    if (rc) return rc;
    INTI_HIP(ctx, hipMemcpyAsync(rn.data(), W.nrm, sizeof(double) * W.nsys, hipMemcpyDeviceToHost, s));
    INTI_HIP(ctx, hipStreamSynchronize(s));
    auto n_open = [&]() {
        int c = 0;
        for (int q = 0; q < W.nsys; ++q) c += rn[q] > tol[q];
        return c;
    };
    int it = 0;
    std::vector<int> act(W.nsys);
    while (n_open() > 0 && it < maxiter) {
        for (int q = 0; q < W.nsys; ++q) act[q] = rn[q] > tol[q];
        INTI_HIP(ctx, hipMemcpyAsync(W.active, act.data(), sizeof(int) * W.nsys, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_bsetup<CPLX>, gs, dim3(kBG), 0, s, d, (T *)W.H, (T *)W.g, (const double *)W.nrm, W.kk, W.st);
        hipLaunchKernelGGL(k_bstart<CPLX>, gv, dim3(kBG), 0, s, d, (T *)W.V, (const T *)W.r, (const double *)W.nrm,
                           (const int *)W.active);
        INTI_HIP(ctx, hipGetLastError());
        int k = 0;
        while (k < m && it + k < maxiter) {
            const int steps = std::min({8, m - k, maxiter - it - k});
            for (int q = 0; q < steps; ++q) {
                if (use_graph) {
                    if (!W.exec) {
                        INTI_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                        rc = barnoldi<CPLX>(B, W, d, s);
                        hipGraph_t gph = nullptr;
                        hipError_t e = hipStreamEndCapture(s, &gph);
                        if (rc) { if (gph) (void)hipGraphDestroy(gph); return rc; }
                        if (e != hipSuccess) return set_error(ctx, INTI_ERR_HIP, "batch gmres: graph capture failed");
                        W.graph = gph;
                        INTI_HIP(ctx, hipGraphInstantiate(&W.exec, W.graph, nullptr, nullptr, 0));
                    }
                    INTI_HIP(ctx, hipGraphLaunch(W.exec, s));
                } else {
                    rc = barnoldi<CPLX>(B, W, d, s);
                    if (rc) return rc;
                }
            }
            k += steps;
            INTI_HIP(ctx, hipMemcpyAsync(act.data(), W.active, sizeof(int) * W.nsys, hipMemcpyDeviceToHost, s));
            INTI_HIP(ctx, hipStreamSynchronize(s));
            int any = 0;
            for (int q = 0; q < W.nsys; ++q) any |= act[q];
            if (!any) break;
        }
        hipLaunchKernelGGL(k_bfinal, gs, dim3(kBG), 0, s, W.nsys, (const int *)W.active, W.kk, k);
        hipLaunchKernelGGL(k_btrisolve<CPLX>, gs, dim3(kBG), 0, s, d, (const T *)W.H, (const T *)W.g,
                           (const int *)W.kk, (T *)W.y);
        hipLaunchKernelGGL(k_bxupdate<CPLX>, gv, dim3(kBG), 0, s, d, (const T *)W.V, (const int *)W.kk,
                           (const T *)W.y, (T *)x, B->ntot);
        INTI_HIP(ctx, hipGetLastError());
        it += k;
        rc = bresidual<CPLX>(B, W, d, b, x, B->ntot);
        if (rc) return rc;
        INTI_HIP(ctx, hipMemcpyAsync(rn.data(), W.nrm, sizeof(double) * W.nsys, hipMemcpyDeviceToHost, s));
        INTI_HIP(ctx, hipStreamSynchronize(s));
    }
    if (iters) *iters = it;
    if (relres)
        for (int q = 0; q < W.nsys; ++q) relres[q] = bn[q] > 0 ? rn[q] / bn[q] : rn[q];
    return INTI_OK;
}

}  // namespace inti

// ================================================================== C-ABI
using namespace inti;

#define BLOCK(ctx) std::lock_guard<std::mutex> lock_((ctx)->mu)
#define BREQ(ctx, cond, msg) \
    do { if (!(cond)) return set_error((ctx), INTI_ERR_ARG, (msg)); } while (0)

extern "C" {

int inti_batch_create(inti_ctx *ctx, int pde, int kind, int dim, int nprob, const inti_quad *const *quads,
                      const double *k, const double *coefs, int precision, inti_batch **out) {
    if (!ctx) return INTI_ERR_ARG;
    BREQ(ctx, out && quads && nprob > 0, "batch_create: null argument or empty batch");
    BREQ(ctx, pde == INTI_LAPLACE || pde == INTI_HELMHOLTZ, "batch_create: unknown pde");
    BREQ(ctx, dim == 2 || dim == 3, "batch_create: dim must be 2 or 3");
    if (!(kind == INTI_SL || kind == INTI_DL || kind == INTI_COMB_SD))
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_create: batched kinds are SL, DL and c0 SL + c1 DL");
    BREQ(ctx, precision == INTI_FP64 || precision == INTI_MIXED, "batch_create: unknown precision");
    if (kind == INTI_COMB_SD) BREQ(ctx, coefs != nullptr, "batch_create: combined kinds need coefs");
    if (pde == INTI_HELMHOLTZ) BREQ(ctx, k != nullptr, "batch_create: Helmholtz needs one k per problem");
    for (int p = 0; p < nprob; ++p) {
        BREQ(ctx, quads[p] != nullptr, "batch_create: null quadrature");
        if (quads[p]->n < 1) return set_error(ctx, INTI_ERR_SIZE, "batch_create: empty quadrature");
        if (quads[p]->dim != dim) return set_error(ctx, INTI_ERR_SIZE, "batch_create: quadrature dimension differs");
        if (kind != INTI_SL && !quads[p]->normals)
            return set_error(ctx, INTI_ERR_ARG, "batch_create: this kernel needs normals");
        if (pde == INTI_HELMHOLTZ)
            BREQ(ctx, k[p] > 0.0 && std::isfinite(k[p]), "batch_create: Helmholtz needs real k > 0");
        if (pde == INTI_LAPLACE && kind == INTI_COMB_SD && (coefs[4 * p + 1] != 0.0 || coefs[4 * p + 3] != 0.0))
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_create: Laplace combinations take real coefficients");
    }
    BLOCK(ctx);
    inti_batch *B = new inti_batch();
    B->ctx = ctx; B->pde = pde; B->kind = kind; B->dim = dim; B->precision = precision; B->nprob = nprob;
    B->off.assign(nprob + 1, 0);
    for (int p = 0; p < nprob; ++p) B->off[p + 1] = B->off[p] + quads[p]->n;
    B->ntot = B->off[nprob];
    if (B->ntot >= INT32_MAX) { delete B; return set_error(ctx, INTI_ERR_SIZE, "batch_create: too many nodes"); }
    B->coefs.assign(4 * (size_t)nprob, 0.0);
    B->cid.assign(2 * (size_t)nprob, 0.0);
    B->kp.resize(nprob);
    for (int p = 0; p < nprob; ++p) {
        Op tmp;
        tmp.ctx = ctx; tmp.pde = pde; tmp.kind = kind; tmp.dim = dim; tmp.precision = precision;
        tmp.k = pde == INTI_HELMHOLTZ ? k[p] : 0.0;
        if (kind == INTI_COMB_SD) std::memcpy(tmp.coefs, coefs + 4 * p, sizeof(double) * 4);
        std::memcpy(&B->coefs[4 * p], tmp.coefs, sizeof(double) * 4);
        int ek = kind;
        B->kp[p] = op_params(&tmp, &ek);
        if (ek != kind) { delete B; return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_create: c1 = 0 combinations are not batched; use SL"); }
    }
    hipStream_t s = ctx->stream;
    const int64_t nt = B->ntot;
    bool ok = hipMalloc((void **)&B->coords, sizeof(double) * dim * nt) == hipSuccess &&
              hipMalloc((void **)&B->weights, sizeof(double) * nt) == hipSuccess &&
              hipMalloc((void **)&B->row2p, sizeof(int32_t) * nt) == hipSuccess &&
              hipMalloc((void **)&B->cid_d, sizeof(double) * 2 * nprob) == hipSuccess &&
              (kind == INTI_SL || hipMalloc((void **)&B->normals, sizeof(double) * dim * nt) == hipSuccess);
    if (!ok) { delete B; return set_error(ctx, INTI_ERR_NOMEM, "batch_create: allocation failed"); }
    std::vector<int32_t> r2p(nt);
    int rc = INTI_OK;
    for (int p = 0; p < nprob && rc == INTI_OK; ++p) {
        const Quad *q = quads[p];
        const int64_t n = q->n, o = B->off[p];
        for (int d = 0; d < dim; ++d) {
            if (hipMemcpyAsync(B->coords + d * nt + o, q->coords + d * n, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                (B->normals && hipMemcpyAsync(B->normals + d * nt + o, q->normals + d * n, sizeof(double) * n,
                                              hipMemcpyDeviceToDevice, s) != hipSuccess))
                rc = set_error(ctx, INTI_ERR_HIP, "batch_create: geometry copy failed");
        }
        if (hipMemcpyAsync(B->weights + o, q->weights, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
            rc = set_error(ctx, INTI_ERR_HIP, "batch_create: weight copy failed");
        for (int64_t i = 0; i < n; ++i) r2p[o + i] = p;
    }
    if (rc == INTI_OK &&
        (hipMemcpyAsync(B->row2p, r2p.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice, s) != hipSuccess ||
         hipMemcpyAsync(B->cid_d, B->cid.data(), sizeof(double) * 2 * nprob, hipMemcpyHostToDevice, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        rc = set_error(ctx, INTI_ERR_HIP, "batch_create: upload failed");
    if (rc) { delete B; return rc; }
    B->c_rowptr.resize(nprob);
    B->c_colind.resize(nprob);
    B->c_val.resize(nprob);
    *out = B;
    return INTI_OK;
}

int inti_batch_destroy(inti_batch *B) {
    if (!B) return INTI_ERR_ARG;
    {
        BLOCK(B->ctx);
        (void)hipStreamSynchronize(B->ctx->stream);
    }
    delete B;
    return INTI_OK;
}

int inti_batch_info(const inti_batch *B, int64_t *ntot, int64_t *offsets) {
    if (!B) return INTI_ERR_ARG;
    if (ntot) *ntot = B->ntot;
    if (offsets) std::copy(B->off.begin(), B->off.end(), offsets);
    return INTI_OK;
}

int inti_batch_set_identity(inti_batch *B, const double *cid) {
    if (!B) return INTI_ERR_ARG;
    Ctx *ctx = B->ctx;
    BREQ(ctx, cid != nullptr, "batch_set_identity: null coefficients");
    BLOCK(ctx);
    std::memcpy(B->cid.data(), cid, sizeof(double) * 2 * B->nprob);
    INTI_HIP(ctx, hipMemcpyAsync(B->cid_d, B->cid.data(), sizeof(double) * 2 * B->nprob, hipMemcpyHostToDevice, ctx->stream));
    INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return INTI_OK;
}

namespace inti {

// The DIM corrections (dS, dD) of every problem of B (bdim.jl:50-196 per
// problem: quadrature with itself, the self pattern, one element type) in
// one pass over the batch: the traces of all problems in one launch (each
// against its own monopoles and wavenumber), Theta = mu B0 + S B1 - D B0 in
// one launch over (problem, target tile) work items (the S and D values of a
// pair share its radial part, all ns columns at once), the least squares of
// all elements in one launch, and the triplets summed into
// one block-diagonal CSR that is then cut into one CSR per problem.
static int batch_bdim_build(Batch *B, int nq, const int64_t *ne, const int64_t *qtags, const double *mu,
                            const double *params, Csr **outS, Csr **outD) {
    Ctx *ctx = B->ctx;
    const int P = B->nprob, dim = B->dim;
    const int64_t ntot = B->ntot;
    const double oversample = params ? params[0] : 3.0, radius_mult = params ? params[1] : 1.5;
    const int ns = (int)std::ceil(oversample * nq);
    if (ns > kMaxNs || 2 * nq > kMaxNc || ns < 2 * nq)
        return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_bdim: need 2*nq <= ns <= 32 and 2*nq <= 24");
    if (!B->normals) return set_error(ctx, INTI_ERR_ARG, "batch_bdim: the batch's quadratures need normals");
    // element tags, global; the self pattern covers every node once
    int64_t ne_tot = 0;
    for (int p = 0; p < P; ++p) {
        if (ne[p] < 1) return set_error(ctx, INTI_ERR_ARG, "batch_bdim: every problem needs elements");
        if ((int64_t)nq * ne[p] != B->off[p + 1] - B->off[p])
            return set_error(ctx, INTI_ERR_SIZE, "batch_bdim: the elements of a problem must cover its nodes once");
        ne_tot += ne[p];
    }
    std::vector<int32_t> qt((size_t)nq * ne_tot);
    {
        int64_t q = 0;
        for (int p = 0; p < P; ++p) {
            const int64_t n = B->off[p + 1] - B->off[p];
            for (int64_t t = 0; t < (int64_t)nq * ne[p]; ++t, ++q) {
                if (qtags[q] < 0 || qtags[q] >= n) return set_error(ctx, INTI_ERR_ARG, "batch_bdim: qtags out of range");
                qt[q] = (int32_t)(B->off[p] + qtags[q]);
            }
        }
    }
    hipStream_t st = ctx->stream;
    // monopoles per problem around its bounding box
    std::vector<double> hx((size_t)dim * ntot);
    INTI_HIP(ctx, hipMemcpyAsync(hx.data(), B->coords, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, st));
    INTI_HIP(ctx, hipStreamSynchronize(st));
    std::vector<double> zs_all((size_t)P * dim * ns), zs;
    for (int p = 0; p < P; ++p) {
        double lo[3], hi[3];
        for (int d = 0; d < dim; ++d) {
            lo[d] = INFINITY; hi[d] = -INFINITY;
            for (int64_t j = B->off[p]; j < B->off[p + 1]; ++j) {
                lo[d] = std::min(lo[d], hx[(size_t)d * ntot + j]);
                hi[d] = std::max(hi[d], hx[(size_t)d * ntot + j]);
            }
        }
        monopoles_box(dim, lo, hi, ns, radius_mult, zs);
        std::copy(zs.begin(), zs.end(), zs_all.begin() + (size_t)p * dim * ns);
    }
    // device buffers
    const bool cplx = B->pde == INTI_HELMHOLTZ;
    const size_t nb = (size_t)ntot * ns;
    DevBuf dzs, dkp, dmu, dB0, dB1, dth, dqt, dpp, del, dtriS, dtriD, ddiag, dseg, dperm, dvS, dvD, drec, ditems,
        doff, dtkp, dpost;
    INTI_ALLOC(ctx, dzs, sizeof(double) * zs_all.size());
    INTI_ALLOC(ctx, dkp, sizeof(KParams) * P);
    INTI_ALLOC(ctx, dmu, sizeof(double) * ntot);
    INTI_ALLOC(ctx, dB0, sizeof(double2) * nb);
    INTI_ALLOC(ctx, dB1, sizeof(double2) * nb);
    INTI_ALLOC(ctx, dth, sizeof(double2) * nb);
    INTI_HIP(ctx, hipMemcpyAsync(dzs.p, zs_all.data(), sizeof(double) * zs_all.size(), hipMemcpyHostToDevice, st));
    INTI_HIP(ctx, hipMemcpyAsync(dkp.p, B->kp.data(), sizeof(KParams) * P, hipMemcpyHostToDevice, st));
    INTI_HIP(ctx, hipMemcpyAsync(dmu.p, mu, sizeof(double) * ntot, hipMemcpyHostToDevice, st));
    {
        TraceBatchArgs a;
        a.ntot = ntot; a.ns = ns;
        a.sx = B->coords; a.sn = B->normals; a.zs = (const double *)dzs.p;
        a.row2p = B->row2p; a.kp = (const KParams *)dkp.p; a.mu = (const double *)dmu.p;
        a.B0 = (double2 *)dB0.p; a.B1 = (double2 *)dB1.p; a.theta = (double2 *)dth.p;
        if (!launch_dim_traces_batch(B->pde, dim, a, st))
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_bdim: no trace kernel for this operator");
        INTI_HIP(ctx, hipGetLastError());
    }
    // Theta += S B1 - D B0: one block per (problem, 64-target tile), largest
    // problems first, each summing S and D over its own problem's nodes with
    // one evaluation of the pair's radial part for both
    const int64_t npairs = (int64_t)nq * ne_tot, ntri = npairs * nq;
    if (ntri >= (int64_t)INT32_MAX) return set_error(ctx, INTI_ERR_SIZE, "batch_bdim: too many triplets");
    TripletPattern pat;
    {
        const int rl = dim_theta_batch_record(B->pde, dim, ns);
        if (!rl) return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_bdim: no Theta kernel for this operator");
        std::vector<int> order(P);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return B->off[a + 1] - B->off[a] > B->off[b + 1] - B->off[b];
        });
        std::vector<int32_t> items;
        for (int p : order)
            for (int64_t t = 0; t < (B->off[p + 1] - B->off[p] + 63) / 64; ++t) {
                items.push_back(p);
                items.push_back((int32_t)t);
            }
        const int64_t nitems = (int64_t)items.size() / 2;
        std::vector<KParams> tkp(P);
        std::vector<double> post(4 * (size_t)P);
        for (int p = 0; p < P; ++p) {
            tkp[p] = basic_params(B->kp[p].k);
            basic_post(B->pde, dim, INTI_SL, B->kp[p].k, post[4 * p], post[4 * p + 1]);
            basic_post(B->pde, dim, INTI_DL, B->kp[p].k, post[4 * p + 2], post[4 * p + 3]);
        }
        INTI_ALLOC(ctx, drec, sizeof(double) * (size_t)ntot * rl);
        INTI_ALLOC(ctx, ditems, sizeof(int32_t) * items.size());
        INTI_ALLOC(ctx, doff, sizeof(int64_t) * (P + 1));
        INTI_ALLOC(ctx, dtkp, sizeof(KParams) * P);
        INTI_ALLOC(ctx, dpost, sizeof(double) * post.size());
        INTI_HIP(ctx, hipMemcpyAsync(ditems.p, items.data(), sizeof(int32_t) * items.size(), hipMemcpyHostToDevice, st));
        INTI_HIP(ctx, hipMemcpyAsync(doff.p, B->off.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, st));
        INTI_HIP(ctx, hipMemcpyAsync(dtkp.p, tkp.data(), sizeof(KParams) * P, hipMemcpyHostToDevice, st));
        INTI_HIP(ctx, hipMemcpyAsync(dpost.p, post.data(), sizeof(double) * post.size(), hipMemcpyHostToDevice, st));
        ThetaBatchArgs a;
        a.ntot = ntot; a.ns = ns;
        a.items = (const int32_t *)ditems.p; a.off = (const int64_t *)doff.p; a.row2p = B->row2p;
        a.kp = (const KParams *)dtkp.p; a.post = (const double *)dpost.p;
        a.coords = B->coords; a.normals = B->normals; a.weights = B->weights;
        a.B0 = (const double2 *)dB0.p; a.B1 = (const double2 *)dB1.p;
        a.rec = (double *)drec.p; a.theta = (double2 *)dth.p;
        if (!launch_dim_theta_batch(B->pde, dim, a, nitems, st))
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_bdim: no Theta kernel for this operator");
        INTI_HIP(ctx, hipGetLastError());
        // the triplets (target qt[e*nq + t], node qt[e*nq + r]) and their
        // block-diagonal CSR pattern, on the host while the device sums
        {
            std::vector<int32_t> tri_row(ntri), tri_col(ntri);
            for (int64_t e = 0; e < ne_tot; ++e)
                for (int t = 0; t < nq; ++t)
                    for (int r = 0; r < nq; ++r) {
                        const int64_t q = ((e * nq) + t) * nq + r;
                        tri_row[q] = qt[e * nq + t];
                        tri_col[q] = qt[e * nq + r];
                    }
            build_pattern(ntot, tri_row, tri_col, pat);
        }
    }
    // least squares of every element: the element's own nodes are its targets
    std::vector<int32_t> pptr(ne_tot + 1), elems(ne_tot);
    for (int64_t e = 0; e <= ne_tot; ++e) pptr[e] = (int32_t)(e * nq);
    std::iota(elems.begin(), elems.end(), 0);
    INTI_ALLOC(ctx, dqt, sizeof(int32_t) * qt.size());
    INTI_ALLOC(ctx, dpp, sizeof(int32_t) * pptr.size());
    INTI_ALLOC(ctx, del, sizeof(int32_t) * elems.size());
    INTI_ALLOC(ctx, dtriS, sizeof(double2) * ntri);
    INTI_ALLOC(ctx, dtriD, sizeof(double2) * ntri);
    INTI_ALLOC(ctx, ddiag, sizeof(double) * 3 * ne_tot);
    INTI_HIP(ctx, hipMemcpyAsync(dqt.p, qt.data(), sizeof(int32_t) * qt.size(), hipMemcpyHostToDevice, st));
    INTI_HIP(ctx, hipMemcpyAsync(dpp.p, pptr.data(), sizeof(int32_t) * pptr.size(), hipMemcpyHostToDevice, st));
    INTI_HIP(ctx, hipMemcpyAsync(del.p, elems.data(), sizeof(int32_t) * elems.size(), hipMemcpyHostToDevice, st));
    {
        LsArgs a;
        a.n = ntot; a.nloc = ntot; a.nq = nq; a.ns = ns;
        a.elems = (const int32_t *)del.p; a.qtags = (const int32_t *)dqt.p;
        a.pair_ptr = (const int32_t *)dpp.p; a.pair_loc = (const int32_t *)dqt.p;  // pair e*nq + t -> node qt[e*nq + t]
        a.B0 = (const double2 *)dB0.p; a.B1 = (const double2 *)dB1.p; a.ldb = ntot;
        a.theta = (const double2 *)dth.p;
        a.triS = (double2 *)dtriS.p; a.triD = (double2 *)dtriD.p;
        a.diag = (double *)ddiag.p;
        launch_dim_ls(a, ne_tot, st);
        INTI_HIP(ctx, hipGetLastError());
    }
    const int64_t nnz = (int64_t)pat.colind.size();
    const int cw = cplx ? 2 : 1;
    INTI_ALLOC(ctx, dseg, sizeof(int32_t) * pat.seg.size());
    INTI_ALLOC(ctx, dperm, sizeof(int32_t) * pat.perm.size());
    INTI_ALLOC(ctx, dvS, sizeof(double) * cw * nnz);
    INTI_ALLOC(ctx, dvD, sizeof(double) * cw * nnz);
    INTI_HIP(ctx, hipMemcpyAsync(dseg.p, pat.seg.data(), sizeof(int32_t) * pat.seg.size(), hipMemcpyHostToDevice, st));
    INTI_HIP(ctx, hipMemcpyAsync(dperm.p, pat.perm.data(), sizeof(int32_t) * pat.perm.size(), hipMemcpyHostToDevice, st));
    {
        Gather1Args g;
        g.nnz = nnz; g.seg = (const int32_t *)dseg.p; g.perm = (const int32_t *)dperm.p; g.cplx = cplx ? 1 : 0;
        g.tri = (const double2 *)dtriS.p; g.val = (double *)dvS.p; g.sign = -1.0;
        launch_gather1(g, st);
        g.tri = (const double2 *)dtriD.p; g.val = (double *)dvD.p; g.sign = 1.0;
        launch_gather1(g, st);
        INTI_HIP(ctx, hipGetLastError());
    }
    // one CSR per problem: rows [off_p, off_p+1), columns shifted by off_p
    std::vector<Csr *> made;
    auto fail = [&](int code) {
        for (Csr *c : made) csr_free(c);
        return code;
    };
    // the host rows and columns of every problem stay alive until the one
    // synchronisation after the last asynchronous upload
    std::vector<std::vector<int32_t>> rps(P), cis(P);
    for (int p = 0; p < P; ++p) {
        const int64_t r0 = B->off[p], n = B->off[p + 1] - r0;
        const int64_t q0 = pat.rowptr[r0], q1 = pat.rowptr[r0 + n], nzp = q1 - q0;
        std::vector<int32_t> &rp = rps[p], &ci = cis[p];
        rp.resize(n + 1);
        ci.resize(nzp);
        for (int64_t i = 0; i <= n; ++i) rp[i] = (int32_t)(pat.rowptr[r0 + i] - q0);
        for (int64_t q = 0; q < nzp; ++q) ci[q] = (int32_t)(pat.colind[q0 + q] - r0);
        Csr *cS = nullptr, *cD = nullptr;
        int rc = csr_alloc(ctx, n, n, nzp, cplx ? 1 : 0, &cS);
        if (rc) return fail(rc);
        made.push_back(cS);
        rc = csr_alloc(ctx, n, n, nzp, cplx ? 1 : 0, &cD);
        if (rc) return fail(rc);
        made.push_back(cD);
        for (Csr *c : {cS, cD})
            if (hipMemcpyAsync(c->rowptr, rp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
                (nzp && hipMemcpyAsync(c->colind, ci.data(), sizeof(int32_t) * nzp, hipMemcpyHostToDevice, st) != hipSuccess))
                return fail(set_error(ctx, INTI_ERR_HIP, "batch_bdim: CSR upload failed"));
        if (nzp && (hipMemcpyAsync(cS->val, (const double *)dvS.p + cw * q0, sizeof(double) * cw * nzp,
                                   hipMemcpyDeviceToDevice, st) != hipSuccess ||
                    hipMemcpyAsync(cD->val, (const double *)dvD.p + cw * q0, sizeof(double) * cw * nzp,
                                   hipMemcpyDeviceToDevice, st) != hipSuccess))
            return fail(set_error(ctx, INTI_ERR_HIP, "batch_bdim: CSR copy failed"));
        outS[p] = cS;
        outD[p] = cD;
    }
    if (hipStreamSynchronize(st) != hipSuccess) return fail(set_error(ctx, INTI_ERR_HIP, "batch_bdim: upload failed"));
    return INTI_OK;
}

}  // namespace inti

int inti_batch_bdim_build(inti_batch *B, int nq, const int64_t *ne, const int64_t *qtags, const double *mu,
                          const double *params, inti_csr **dS, inti_csr **dD) {
    if (!B) return INTI_ERR_ARG;
    Ctx *ctx = B->ctx;
    BREQ(ctx, ne && qtags && mu && dS && dD && nq > 0, "batch_bdim: null argument");
    BREQ(ctx, params == nullptr || (params[0] > 0 && params[1] > 0), "batch_bdim: bad parameters");
    BLOCK(ctx);
    std::vector<Csr *> s(B->nprob, nullptr), d(B->nprob, nullptr);
    const int rc = batch_bdim_build(B, nq, ne, qtags, mu, params, s.data(), d.data());
    if (rc) return rc;
    for (int p = 0; p < B->nprob; ++p) {
        dS[p] = static_cast<inti_csr *>(s[p]);
        dD[p] = static_cast<inti_csr *>(d[p]);
    }
    return INTI_OK;
}

int inti_batch_attach_correction(inti_batch *B, int p, const inti_csr *csr, double coef_re, double coef_im) {
    if (!B) return INTI_ERR_ARG;
    Ctx *ctx = B->ctx;
    BREQ(ctx, csr != nullptr && p >= 0 && p < B->nprob, "batch_attach: bad problem index or null CSR");
    const int64_t n = B->off[p + 1] - B->off[p];
    if (csr->m != n || csr->n != n) return set_error(ctx, INTI_ERR_SIZE, "batch_attach: CSR shape differs from the problem's");
    const int cw = B->pde == INTI_HELMHOLTZ ? 2 : 1;
    if (csr->is_complex != (cw == 2)) return set_error(ctx, INTI_ERR_ARG, "batch_attach: value type differs");
    BLOCK(ctx);
    std::vector<int32_t> rp(n + 1), ci(csr->nnz);
    std::vector<double> va((size_t)cw * csr->nnz);
    hipStream_t s = ctx->stream;
    INTI_HIP(ctx, hipMemcpyAsync(rp.data(), csr->rowptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, s));
    if (csr->nnz) {
        INTI_HIP(ctx, hipMemcpyAsync(ci.data(), csr->colind, sizeof(int32_t) * csr->nnz, hipMemcpyDeviceToHost, s));
        INTI_HIP(ctx, hipMemcpyAsync(va.data(), csr->val, sizeof(double) * va.size(), hipMemcpyDeviceToHost, s));
    }
    INTI_HIP(ctx, hipStreamSynchronize(s));
    for (int64_t q = 0; q < csr->nnz; ++q) {
        if (cw == 2) {
            const double vr = va[2 * q], vi = va[2 * q + 1];
            va[2 * q] = coef_re * vr - coef_im * vi;
            va[2 * q + 1] = coef_re * vi + coef_im * vr;
        } else {
            va[q] *= coef_re;
        }
    }
    if (B->c_rowptr[p].empty()) {
        B->c_rowptr[p] = std::move(rp);
        B->c_colind[p] = std::move(ci);
        B->c_val[p] = std::move(va);
    } else {
        if (B->c_rowptr[p] != rp || B->c_colind[p] != ci)
            return set_error(ctx, INTI_ERR_UNSUPPORTED, "batch_attach: corrections of one problem must share a pattern");
        for (size_t q = 0; q < va.size(); ++q) B->c_val[p][q] += va[q];
    }
    B->corr_dirty = true;
    return INTI_OK;
}

int inti_batch_set_launch(inti_batch *B, int64_t chunk_len, int grid) {
    if (!B) return INTI_ERR_ARG;
    BREQ(B->ctx, chunk_len >= 0 && grid >= 0, "batch_set_launch: negative value");
    BLOCK(B->ctx);
    const int64_t c = chunk_len > 0 ? chunk_len : 2048;
    if (c != B->chunk_len) B->planned = false;
    B->chunk_len = c;
    B->grid = grid;
    return INTI_OK;
}

int inti_batch_apply_device(inti_batch *B, int nrhs, const double *x, int64_t ldx, double *y, int64_t ldy,
                            const double *alpha, const double *beta) {
    if (!B) return INTI_ERR_ARG;
    Ctx *ctx = B->ctx;
    BREQ(ctx, x && y && alpha && beta && nrhs >= 1, "batch_apply: null argument or nrhs < 1");
    BREQ(ctx, ldx >= B->ntot && ldy >= B->ntot, "batch_apply: leading dimension below ntot");
    BLOCK(ctx);
    return batch_apply_dev(B, nrhs, x, ldx, y, ldy, alpha, beta);
}

int inti_batch_apply(inti_batch *B, int nrhs, const double *x, double *y, const double *alpha, const double *beta) {
    if (!B) return INTI_ERR_ARG;
    Ctx *ctx = B->ctx;
    BREQ(ctx, x && y && alpha && beta && nrhs >= 1, "batch_apply: null argument or nrhs < 1");
    BLOCK(ctx);
    const size_t bytes = sizeof(double) * (B->pde == INTI_HELMHOLTZ ? 2 : 1) * (size_t)B->ntot * nrhs;
    void *dx = nullptr, *dy = nullptr;
    int rc = ctx_scratch(ctx, kScrX, bytes, &dx);
    if (!rc) rc = ctx_scratch(ctx, kScrY, bytes, &dy);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    INTI_HIP(ctx, hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, s));
    if (beta[0] != 0.0 || beta[1] != 0.0) INTI_HIP(ctx, hipMemcpyAsync(dy, y, bytes, hipMemcpyHostToDevice, s));
    rc = batch_apply_dev(B, nrhs, (const double *)dx, B->ntot, (double *)dy, B->ntot, alpha, beta);
    if (rc) return rc;
    INTI_HIP(ctx, hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, s));
    INTI_HIP(ctx, hipStreamSynchronize(s));
    return INTI_OK;
}

int inti_batch_autotune(inti_batch *B, int nrhs, int reps, int64_t *chunk_out, int *grid_out, double *ms_out) {
    if (!B) return INTI_ERR_ARG;
    Ctx *ctx = B->ctx;
    BREQ(ctx, nrhs >= 1 && reps >= 1, "batch_autotune: nrhs and reps must be >= 1");
    BLOCK(ctx);
    const bool cplx = B->pde == INTI_HELMHOLTZ;
    const size_t bytes = sizeof(double) * (cplx ? 2 : 1) * (size_t)B->ntot * nrhs;
    void *dx = nullptr, *dy = nullptr;
    int rc = ctx_scratch(ctx, kScrX, bytes, &dx);
    if (!rc) rc = ctx_scratch(ctx, kScrY, bytes, &dy);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    INTI_HIP(ctx, hipMemsetAsync(dx, 0, bytes, s));
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    const int64_t chunks[] = {1024, 2048, 4096, 8192};
    const int grids[] = {2, 4, 8};
    double best = 1e300;
    int64_t bc = B->chunk_len;
    int bg = B->grid;
    const int saved = ctx->profiling;
    ctx->profiling = 0;
    for (int64_t c : chunks)
        for (int gm : grids) {
            if (c != B->chunk_len) B->planned = false;
            B->chunk_len = c;
            B->grid = gm * ctx->num_cu;
            rc = batch_apply_dev(B, nrhs, (const double *)dx, B->ntot, (double *)dy, B->ntot, one, zero);  // warm-up + plan
            if (rc) { ctx->profiling = saved; return rc; }
            INTI_HIP(ctx, hipEventRecord(ctx->ev0, s));
            for (int q = 0; q < reps; ++q) {
                rc = batch_apply_dev(B, nrhs, (const double *)dx, B->ntot, (double *)dy, B->ntot, one, zero);
                if (rc) { ctx->profiling = saved; return rc; }
            }
            INTI_HIP(ctx, hipEventRecord(ctx->ev1, s));
            INTI_HIP(ctx, hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            INTI_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            const double t = ms / reps;
            if (t < best) { best = t; bc = c; bg = B->grid; }
        }
    ctx->profiling = saved;
    if (bc != B->chunk_len) B->planned = false;
    B->chunk_len = bc;
    B->grid = bg;
    if (chunk_out) *chunk_out = bc;
    if (grid_out) *grid_out = bg;
    if (ms_out) *ms_out = best;
    return INTI_OK;
}

int inti_batch_gmres(inti_batch *B, int on_device, int nrhs, const double *b, double *x, int restart, int maxiter,
                     double rtol, double atol, int use_graph, int *iters, double *relres) {
    if (!B) return INTI_ERR_ARG;
    Ctx *ctx = B->ctx;
    BREQ(ctx, b && x, "batch_gmres: null vector");
    BREQ(ctx, nrhs >= 1 && nrhs <= 64 && 256 % nrhs == 0, "batch_gmres: nrhs must divide 256 (1, 2, 4, ..., 64)");
    BREQ(ctx, restart >= 1 && maxiter >= 0, "batch_gmres: restart >= 1, maxiter >= 0");
    BLOCK(ctx);
    const bool cplx = B->pde == INTI_HELMHOLTZ;
    const size_t bytes = sizeof(double) * (cplx ? 2 : 1) * (size_t)B->ntot * nrhs;
    const double *db = b;
    double *dxp = x;
    void *tb = nullptr, *tx = nullptr;
    if (!on_device) {
        if (hipMalloc(&tb, bytes) != hipSuccess || hipMalloc(&tx, bytes) != hipSuccess) {
            if (tb) (void)hipFree(tb);
            return set_error(ctx, INTI_ERR_NOMEM, "batch_gmres: staging allocation failed");
        }
        INTI_HIP(ctx, hipMemcpyAsync(tb, b, bytes, hipMemcpyHostToDevice, ctx->stream));
        INTI_HIP(ctx, hipMemcpyAsync(tx, x, bytes, hipMemcpyHostToDevice, ctx->stream));
        db = (const double *)tb;
        dxp = (double *)tx;
    }
    struct F { void *a, *b; hipStream_t s; ~F() { if (a || b) (void)hipStreamSynchronize(s); if (a) (void)hipFree(a); if (b) (void)hipFree(b); } } fr{tb, tx, ctx->stream};
    int rc = cplx ? batch_gmres_t<true>(B, nrhs, db, dxp, restart, maxiter, rtol, atol, use_graph, iters, relres)
                  : batch_gmres_t<false>(B, nrhs, db, dxp, restart, maxiter, rtol, atol, use_graph, iters, relres);
    if (rc) return rc;
    if (!on_device) {
        INTI_HIP(ctx, hipMemcpyAsync(x, dxp, bytes, hipMemcpyDeviceToHost, ctx->stream));
        INTI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return INTI_OK;
}

}  // extern "C"
