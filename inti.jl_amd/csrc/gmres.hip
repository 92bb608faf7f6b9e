// Restarted GMRES on a (corrected) operator, device resident, and mixed
// precision iterative refinement around it.
//
// The user-side solve of Inti.jl's examples is IterativeSolvers.gmres on
// L = I/2 + D - i k S (docs/src/examples/helmholtz_scattering.jl:230,494-495;
// docs/src/examples/poisson.jl:167-168). Here one Arnoldi step is a fixed
// sequence of kernels that read the current column index k from a device
// state block: copy V[:,k], apply the operator, two passes of classical
// Gram-Schmidt (dots, sums, update), norm, normalisation, and a one-thread
// Givens update of the Hessenberg column and of the residual estimate. The
// same sequence therefore serves every step and is captured once in a
// hipGraph and replayed; the host only looks at the residual estimates every
// few steps. No memset is recorded in the graph: clearing is done by kernels.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <complex>
#include <vector>

#include "common.hpp"

namespace inti {

struct GState {
    int k;       // current Arnoldi column
    int pad;
    double wnorm;
};

template <bool CPLX>
struct Sc;
template <>
struct Sc<true> {
    using T = double2;
    __device__ static T zero() { return make_double2(0.0, 0.0); }
    __device__ static T conjmul(T a, T b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
    __device__ static T mul(T a, T b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
    __device__ static T add(T a, T b) { return make_double2(a.x + b.x, a.y + b.y); }
    __device__ static T sub(T a, T b) { return make_double2(a.x - b.x, a.y - b.y); }
    __device__ static T scale(T a, double s) { return make_double2(a.x * s, a.y * s); }
    __device__ static double abs2(T a) { return a.x * a.x + a.y * a.y; }
    __device__ static T shfl(T v, int off) { return make_double2(__shfl_xor(v.x, off), __shfl_xor(v.y, off)); }
};
template <>
struct Sc<false> {
    using T = double;
    __device__ static T zero() { return 0.0; }
    __device__ static T conjmul(T a, T b) { return a * b; }
    __device__ static T mul(T a, T b) { return a * b; }
    __device__ static T add(T a, T b) { return a + b; }
    __device__ static T sub(T a, T b) { return a - b; }
    __device__ static T scale(T a, double s) { return a * s; }
    __device__ static double abs2(T a) { return a * a; }
    __device__ static T shfl(T v, int off) { return __shfl_xor(v, off); }
};

constexpr int kGB = 256;     // threads per block of the vector kernels
constexpr int kGRows = 4096; // rows per block of the dot kernels

template <bool CPLX>
__device__ typename Sc<CPLX>::T block_sum(typename Sc<CPLX>::T v) {
    using S = Sc<CPLX>;
    __shared__ typename S::T red[kGB / 64];
    for (int off = 32; off > 0; off >>= 1) v = S::add(v, S::shfl(v, off));
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) red[w] = v;
    __syncthreads();
    typename S::T s = S::zero();
    if (threadIdx.x == 0)
        for (int q = 0; q < kGB / 64; ++q) s = S::add(s, red[q]);
    return s;  // valid in thread 0
}

// vk = V[:, k]
template <bool CPLX>
__global__ __launch_bounds__(kGB) void k_gcopy(const typename Sc<CPLX>::T *V, int64_t ldv,
                                               const GState *st, typename Sc<CPLX>::T *vk,
                                               int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x;
    if (i < n) vk[i] = V[i + ldv * st->k];
}

// partial[j * nb + b] = sum over the rows of block b of conj(V[i,j]) * w[i], j <= k
template <bool CPLX>
__global__ __launch_bounds__(kGB) void k_gdots(const typename Sc<CPLX>::T *V, int64_t ldv,
                                               const GState *st, const typename Sc<CPLX>::T *w,
                                               int64_t n, typename Sc<CPLX>::T *partial) {
    using S = Sc<CPLX>;
    const int j = blockIdx.y;
    if (j > st->k) return;
    const int64_t r0 = (int64_t)blockIdx.x * kGRows;
    const int64_t r1 = r0 + kGRows < n ? r0 + kGRows : n;
    typename S::T acc = S::zero();
    const typename S::T *vj = V + ldv * j;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kGB) acc = S::add(acc, S::conjmul(vj[i], w[i]));
    acc = block_sum<CPLX>(acc);
    if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = acc;
}

// h[j] = sum_b partial[j, b]; H[j, k] += h[j]   (one block per column)
template <bool CPLX>
__global__ __launch_bounds__(64) void k_gsum(const typename Sc<CPLX>::T *partial, int nb,
                                             const GState *st, typename Sc<CPLX>::T *h,
                                             typename Sc<CPLX>::T *H, int ldh) {
    using S = Sc<CPLX>;
    const int j = blockIdx.x;
    if (j > st->k) return;
    if (threadIdx.x == 0) {
        typename S::T s = S::zero();
        for (int b = 0; b < nb; ++b) s = S::add(s, partial[(int64_t)j * nb + b]);
        h[j] = s;
        H[j + (int64_t)ldh * st->k] = S::add(H[j + (int64_t)ldh * st->k], s);
    }
}

// w -= V[:, 0..k] h
template <bool CPLX>
__global__ __launch_bounds__(kGB) void k_gupdate(const typename Sc<CPLX>::T *V, int64_t ldv,
                                                 const GState *st, const typename Sc<CPLX>::T *h,
                                                 typename Sc<CPLX>::T *w, int64_t n) {
    using S = Sc<CPLX>;
    const int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x;
    if (i >= n) return;
    const int k = st->k;
    typename S::T s = w[i];
    for (int j = 0; j <= k; ++j) s = S::sub(s, S::mul(V[i + ldv * j], h[j]));
    w[i] = s;
}

// partial[b] = sum over block b of |w|^2
template <bool CPLX>
__global__ __launch_bounds__(kGB) void k_gnorm(const typename Sc<CPLX>::T *w, int64_t n,
                                               double *partial) {
    using S = Sc<CPLX>;
    const int64_t r0 = (int64_t)blockIdx.x * kGRows;
    const int64_t r1 = r0 + kGRows < n ? r0 + kGRows : n;
    double acc = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kGB) acc += S::abs2(w[i]);
    acc = block_sum<false>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// V[:, k+1] = w / ||w|| (the norm is summed redundantly by every block)
template <bool CPLX>
__global__ __launch_bounds__(kGB) void k_gscale(typename Sc<CPLX>::T *V, int64_t ldv, GState *st,
                                                const typename Sc<CPLX>::T *w, int64_t n,
                                                const double *partial, int nb) {
    using S = Sc<CPLX>;
    __shared__ double s_inv;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partial[b];
        const double nrm = sqrt(s);
        s_inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
        if (blockIdx.x == 0) st->wnorm = nrm;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x;
    if (i < n) V[i + ldv * (st->k + 1)] = S::scale(w[i], s_inv);
}

// Givens update of Hessenberg column k, residual estimate, k += 1; also
// clears column k+1 of H for the next step.
__global__ void k_ggivens_c(double2 *H, int ldh, double *cs, double2 *sn, double2 *g, double *res,
                            GState *st, int m) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int k = st->k;
    double2 *h = H + (int64_t)ldh * k;
    h[k + 1] = make_double2(st->wnorm, 0.0);
    for (int i = 0; i < k; ++i) {
        const double c = cs[i];
        const double2 s = sn[i], a = h[i], b = h[i + 1];
        // [a; b] <- [c a + s b; -conj(s) a + c b]
        h[i] = make_double2(c * a.x + s.x * b.x - s.y * b.y, c * a.y + s.x * b.y + s.y * b.x);
        h[i + 1] = make_double2(-(s.x * a.x + s.y * a.y) + c * b.x, -(s.x * a.y - s.y * a.x) + c * b.y);
    }
    const double2 a = h[k];
    const double bn = h[k + 1].x;  // real, >= 0
    const double an = hypot(a.x, a.y);
    double c;
    double2 s;
    if (an == 0.0) {
        c = 0.0;
        s = make_double2(1.0, 0.0);
    } else {
        const double t = hypot(an, bn);
        c = an / t;
        s = make_double2(a.x / an * bn / t, a.y / an * bn / t);
        h[k] = make_double2(a.x / an * t, a.y / an * t);
    }
    if (an == 0.0) h[k] = make_double2(bn, 0.0);
    h[k + 1] = make_double2(0.0, 0.0);
    cs[k] = c;
    sn[k] = s;
    const double2 gk = g[k];
    g[k] = make_double2(c * gk.x, c * gk.y);
    g[k + 1] = make_double2(-(s.x * gk.x + s.y * gk.y), -(s.x * gk.y - s.y * gk.x));
    res[k] = hypot(g[k + 1].x, g[k + 1].y);
    if (k + 1 < m)
        for (int q = 0; q <= m; ++q) H[q + (int64_t)ldh * (k + 1)] = make_double2(0.0, 0.0);
    st->k = k + 1;
}

__global__ void k_ggivens_r(double *H, int ldh, double *cs, double *sn, double *g, double *res,
                            GState *st, int m) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int k = st->k;
    double *h = H + (int64_t)ldh * k;
    h[k + 1] = st->wnorm;
    for (int i = 0; i < k; ++i) {
        const double c = cs[i], s = sn[i], a = h[i], b = h[i + 1];
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
    cs[k] = c; sn[k] = s;
    const double gk = g[k];
    g[k] = c * gk;
    g[k + 1] = -s * gk;
    res[k] = fabs(g[k + 1]);
    if (k + 1 < m)
        for (int q = 0; q <= m; ++q) H[q + (int64_t)ldh * (k + 1)] = 0.0;
    st->k = k + 1;
}

// x += V[:, 0..kk-1] y
template <bool CPLX>
__global__ __launch_bounds__(kGB) void k_gxupdate(const typename Sc<CPLX>::T *V, int64_t ldv, int kk,
                                                  const typename Sc<CPLX>::T *y,
                                                  typename Sc<CPLX>::T *x, int64_t n) {
    using S = Sc<CPLX>;
    const int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x;
    if (i >= n) return;
    typename S::T s = x[i];
    for (int j = 0; j < kk; ++j) s = S::add(s, S::mul(V[i + ldv * j], y[j]));
    x[i] = s;
}

// generic vector helpers: y = a x + b y (real scalars), fill
template <bool CPLX>
__global__ __launch_bounds__(kGB) void k_gaxpby(int64_t n, double a, const typename Sc<CPLX>::T *x,
                                                double b, typename Sc<CPLX>::T *y) {
    using S = Sc<CPLX>;
    const int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x;
    if (i >= n) return;
    // b == 0 does not read y (fresh workspace may hold any bit pattern)
    y[i] = b == 0.0 ? S::scale(x[i], a) : S::add(S::scale(x[i], a), S::scale(y[i], b));
}

template <typename T>
__global__ __launch_bounds__(kGB) void k_gfill(int64_t n, T *p, T v) {
    const int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x;
    if (i < n) p[i] = v;
}

// ----------------------------------------------------------------- host side
struct GmresWork {
    int m;
    int64_t n;
    int cplx;
    void *V = nullptr, *w = nullptr, *vk = nullptr, *r = nullptr, *t = nullptr;  // t: P v (preconditioned)
    void *partial = nullptr, *h = nullptr, *H = nullptr, *sn = nullptr, *g = nullptr, *y = nullptr;
    double *cs = nullptr, *res = nullptr, *npart = nullptr;
    GState *st = nullptr;
    int nb = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    // Teardown order: drain the stream (an error return may leave launches
    // queued), destroy the executable graph and the graph whose kernel nodes
    // hold pointers into the buffers, and only then release the buffers.
    ~GmresWork() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        void *ps[] = {V, w, vk, r, t, partial, h, H, sn, g, y, cs, res, npart, st};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
};

static int gwork_alloc(Ctx *ctx, GmresWork &W, int m, int64_t n, int cplx) {
    W.m = m; W.n = n; W.cplx = cplx;
    W.stream = ctx->stream;
    const size_t es = cplx ? 16 : 8;
    W.nb = (int)((n + kGRows - 1) / kGRows);
    const size_t hm = (size_t)(m + 1) * m;
    bool ok = hipMalloc(&W.V, es * (size_t)n * (m + 1)) == hipSuccess &&
              hipMalloc(&W.w, es * n) == hipSuccess && hipMalloc(&W.vk, es * n) == hipSuccess &&
              hipMalloc(&W.r, es * n) == hipSuccess && hipMalloc(&W.t, es * n) == hipSuccess &&
              hipMalloc(&W.partial, es * (size_t)W.nb * (m + 1)) == hipSuccess &&
              hipMalloc(&W.h, es * (m + 1)) == hipSuccess && hipMalloc(&W.H, es * hm) == hipSuccess &&
              hipMalloc(&W.sn, es * m) == hipSuccess && hipMalloc(&W.g, es * (m + 1)) == hipSuccess &&
              hipMalloc(&W.y, es * (m + 1)) == hipSuccess &&
              hipMalloc((void **)&W.cs, sizeof(double) * m) == hipSuccess &&
              hipMalloc((void **)&W.res, sizeof(double) * m) == hipSuccess &&
              hipMalloc((void **)&W.npart, sizeof(double) * W.nb) == hipSuccess &&
              hipMalloc((void **)&W.st, sizeof(GState)) == hipSuccess;
    if (!ok) return set_error(ctx, INTI_ERR_NOMEM, "gmres: workspace allocation failed");
    return INTI_OK;
}

// y = alpha A x + beta y (real alpha, beta) on the context's stream
static int apply_vec(Op *op, const void *x, void *y, int64_t n, double alpha, double beta) {
    DirectLaunch L;
    L.op = op; L.nrhs = 1; L.x = (const double *)x; L.ldx = n; L.y = (double *)y; L.ldy = n;
    L.alpha[0] = alpha; L.alpha[1] = 0.0; L.beta[0] = beta; L.beta[1] = 0.0;
    return launch_apply(L);
}

// one Arnoldi step on the stream (also what the graph records); with a
// right preconditioner P the step applies A P
template <bool CPLX>
static int arnoldi_step(Op *op, Op *prec, GmresWork &W, hipStream_t s) {
    using T = typename Sc<CPLX>::T;
    const int64_t n = W.n;
    const dim3 gv((unsigned)((n + kGB - 1) / kGB));
    hipLaunchKernelGGL(k_gcopy<CPLX>, gv, dim3(kGB), 0, s, (const T *)W.V, n, W.st, (T *)W.vk, n);
    int rc = prec ? apply_vec(prec, W.vk, W.t, n, 1.0, 0.0) : INTI_OK;
    if (rc) return rc;
    rc = apply_vec(op, prec ? W.t : W.vk, W.w, n, 1.0, 0.0);
    if (rc) return rc;
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_gdots<CPLX>, dim3((unsigned)W.nb, (unsigned)(W.m + 1)), dim3(kGB), 0, s,
                           (const T *)W.V, n, W.st, (const T *)W.w, n, (T *)W.partial);
        hipLaunchKernelGGL(k_gsum<CPLX>, dim3((unsigned)(W.m + 1)), dim3(64), 0, s, (const T *)W.partial,
                           W.nb, W.st, (T *)W.h, (T *)W.H, W.m + 1);
        hipLaunchKernelGGL(k_gupdate<CPLX>, gv, dim3(kGB), 0, s, (const T *)W.V, n, W.st,
                           (const T *)W.h, (T *)W.w, n);
    }
    hipLaunchKernelGGL(k_gnorm<CPLX>, dim3((unsigned)W.nb), dim3(kGB), 0, s, (const T *)W.w, n, W.npart);
    hipLaunchKernelGGL(k_gscale<CPLX>, gv, dim3(kGB), 0, s, (T *)W.V, n, W.st, (const T *)W.w, n,
                       (const double *)W.npart, W.nb);
    if (CPLX)
        hipLaunchKernelGGL(k_ggivens_c, dim3(1), dim3(1), 0, s, (double2 *)W.H, W.m + 1, W.cs,
                           (double2 *)W.sn, (double2 *)W.g, W.res, W.st, W.m);
    else
        hipLaunchKernelGGL(k_ggivens_r, dim3(1), dim3(1), 0, s, (double *)W.H, W.m + 1, W.cs,
                           (double *)W.sn, (double *)W.g, W.res, W.st, W.m);
    return hipGetLastError() == hipSuccess ? INTI_OK
                                           : set_error(op->ctx, INTI_ERR_HIP, "gmres: kernel launch failed");
}

// r = b - A x (device vectors), returns ||r||
template <bool CPLX>
static int residual(Op *op, GmresWork &W, const double *b, const double *x, double *r, double &nrm) {
    using T = typename Sc<CPLX>::T;
    Ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = W.n;
    const dim3 gv((unsigned)((n + kGB - 1) / kGB));
    INTI_HIP(ctx, hipMemcpyAsync(r, b, (CPLX ? 16 : 8) * n, hipMemcpyDeviceToDevice, s));
    DirectLaunch L;
    L.op = op; L.nrhs = 1; L.x = x; L.ldx = n; L.y = r; L.ldy = n;
    L.alpha[0] = -1.0; L.alpha[1] = 0.0; L.beta[0] = 1.0; L.beta[1] = 0.0;
    int rc = launch_apply(L);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gnorm<CPLX>, dim3((unsigned)W.nb), dim3(kGB), 0, s, (const T *)r, n, W.npart);
    std::vector<double> np(W.nb);
    INTI_HIP(ctx, hipMemcpyAsync(np.data(), W.npart, sizeof(double) * W.nb, hipMemcpyDeviceToHost, s));
    INTI_HIP(ctx, hipStreamSynchronize(s));
    double t = 0.0;
    for (double v : np) t += v;
    nrm = std::sqrt(t);
    (void)gv;
    return INTI_OK;
}

// GMRES(m) solving A x = b to ||b - A x|| <= tol_abs; x in/out (device).
// prec (or null): a right preconditioner P, GMRES on A P y = b - A x0 with
// x = x0 + P y (the residual it minimises is the true one).
template <bool CPLX>
static int gmres_core(Op *op, Op *prec, GmresWork &W, const double *b, double *x, int maxiter, double tol_abs,
                      int use_graph, int check_every, int *iters, double *resnorm) {
    using T = typename Sc<CPLX>::T;
    using cd = std::complex<double>;
    Ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = W.n;
    const int m = W.m;
    const dim3 gv((unsigned)((n + kGB - 1) / kGB));
    const size_t es = CPLX ? 16 : 8;
    int it = 0;
    double rn = 0.0;
    int rc = residual<CPLX>(op, W, b, x, (double *)W.r, rn);
    if (rc) return rc;
    const int saved_prof = ctx->profiling;
    ctx->profiling = 0;  // no event records inside the solver (or its graph)
    // INTI_GMRES_LOG=N: a progress line on stderr every N Arnoldi steps
    const char *log_env = std::getenv("INTI_GMRES_LOG");
    const int log_every = log_env ? std::max(0, std::atoi(log_env)) : 0;
    while (rn > tol_abs && it < maxiter) {
        // V[:,0] = r / rn; g = (rn, 0, ...); H = 0; k = 0
        hipLaunchKernelGGL(k_gaxpby<CPLX>, gv, dim3(kGB), 0, s, n, 1.0 / rn, (const T *)W.r, 0.0, (T *)W.V);
        hipLaunchKernelGGL(k_gfill<double>, dim3((unsigned)(((m + 1) * (CPLX ? 2 : 1) + kGB - 1) / kGB)),
                           dim3(kGB), 0, s, (int64_t)(m + 1) * (CPLX ? 2 : 1), (double *)W.g, 0.0);
        hipLaunchKernelGGL(k_gfill<double>, dim3((unsigned)(((m + 1) * (CPLX ? 2 : 1) + kGB - 1) / kGB)),
                           dim3(kGB), 0, s, (int64_t)(m + 1) * (CPLX ? 2 : 1), (double *)W.H, 0.0);
        hipLaunchKernelGGL(k_gfill<double>, dim3(1), dim3(kGB), 0, s, (int64_t)1, (double *)W.g, rn);
        hipLaunchKernelGGL(k_gfill<int>, dim3(1), dim3(kGB), 0, s, (int64_t)2, (int *)W.st, 0);
        INTI_HIP(ctx, hipGetLastError());
        int k = 0, kk = 0;
        std::vector<double> res(m);
        while (k < m && it + k < maxiter) {
            const int steps = std::min({check_every, m - k, maxiter - it - k});
            for (int q = 0; q < steps; ++q) {
                if (use_graph) {
                    if (!W.exec) {
                        // the apply's scratch must exist before capture: no
                        // allocation may happen while the stream records
                        rc = apply_reserve(op, 1);
                        if (!rc && prec) rc = apply_reserve(prec, 1);
                        if (rc) { ctx->profiling = saved_prof; return rc; }
                        INTI_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                        rc = arnoldi_step<CPLX>(op, prec, W, s);
                        hipGraph_t gph = nullptr;
                        hipError_t e = hipStreamEndCapture(s, &gph);
                        if (rc) { if (gph) (void)hipGraphDestroy(gph); ctx->profiling = saved_prof; return rc; }
                        if (e != hipSuccess) { ctx->profiling = saved_prof; return set_error(ctx, INTI_ERR_HIP, "gmres: graph capture failed"); }
                        W.graph = gph;
                        INTI_HIP(ctx, hipGraphInstantiate(&W.exec, W.graph, nullptr, nullptr, 0));
                    }
                    INTI_HIP(ctx, hipGraphLaunch(W.exec, s));
                } else {
                    rc = arnoldi_step<CPLX>(op, prec, W, s);
                    if (rc) { ctx->profiling = saved_prof; return rc; }
                }
            }
            k += steps;
            INTI_HIP(ctx, hipMemcpyAsync(res.data(), W.res, sizeof(double) * k, hipMemcpyDeviceToHost, s));
            INTI_HIP(ctx, hipStreamSynchronize(s));
            kk = k;
            if (log_every > 0 && (it + k) / log_every != (it + k - steps) / log_every)
                std::fprintf(stderr, "gmres: step %d, residual estimate %.3e (target %.3e)\n", it + k, res[k - 1],
                             tol_abs);
            for (int q = k - steps; q < k; ++q)
                if (res[q] <= tol_abs) { kk = q + 1; break; }
            if (kk < k || res[k - 1] <= tol_abs) break;
        }
        // solve the kk x kk triangular system H y = g on the host
        std::vector<double> Hh((size_t)(m + 1) * m * (CPLX ? 2 : 1)), gh((m + 1) * (CPLX ? 2 : 1));
        INTI_HIP(ctx, hipMemcpyAsync(Hh.data(), W.H, es * (size_t)(m + 1) * m, hipMemcpyDeviceToHost, s));
        INTI_HIP(ctx, hipMemcpyAsync(gh.data(), W.g, es * (m + 1), hipMemcpyDeviceToHost, s));
        INTI_HIP(ctx, hipStreamSynchronize(s));
        std::vector<cd> y(kk);
        auto Hv = [&](int i, int j) {
            const size_t o = (size_t)i + (size_t)(m + 1) * j;
            return CPLX ? cd(Hh[2 * o], Hh[2 * o + 1]) : cd(Hh[o], 0.0);
        };
        for (int i = kk - 1; i >= 0; --i) {
            cd t = CPLX ? cd(gh[2 * i], gh[2 * i + 1]) : cd(gh[i], 0.0);
            for (int j = i + 1; j < kk; ++j) t -= Hv(i, j) * y[j];
            y[i] = t / Hv(i, i);
        }
        std::vector<double> yd(kk * (CPLX ? 2 : 1));
        for (int i = 0; i < kk; ++i) {
            if (CPLX) { yd[2 * i] = y[i].real(); yd[2 * i + 1] = y[i].imag(); }
            else yd[i] = y[i].real();
        }
        INTI_HIP(ctx, hipMemcpyAsync(W.y, yd.data(), es * kk, hipMemcpyHostToDevice, s));
        if (prec) {
            // x += P (V y): V y into t, then the preconditioner's apply adds onto x
            hipLaunchKernelGGL(k_gfill<double>, dim3((unsigned)((n * (CPLX ? 2 : 1) + kGB - 1) / kGB)), dim3(kGB),
                               0, s, n * (CPLX ? 2 : 1), (double *)W.t, 0.0);
            hipLaunchKernelGGL(k_gxupdate<CPLX>, gv, dim3(kGB), 0, s, (const T *)W.V, n, kk, (const T *)W.y,
                               (T *)W.t, n);
            INTI_HIP(ctx, hipGetLastError());
            rc = apply_vec(prec, W.t, x, n, 1.0, 1.0);
            if (rc) { ctx->profiling = saved_prof; return rc; }
        } else {
            hipLaunchKernelGGL(k_gxupdate<CPLX>, gv, dim3(kGB), 0, s, (const T *)W.V, n, kk, (const T *)W.y,
                               (T *)x, n);
        }
        INTI_HIP(ctx, hipGetLastError());
        it += kk;
        rc = residual<CPLX>(op, W, b, x, (double *)W.r, rn);
        if (rc) { ctx->profiling = saved_prof; return rc; }
    }
    ctx->profiling = saved_prof;
    if (iters) *iters = it;
    if (resnorm) *resnorm = rn;
    return INTI_OK;
}

static double dev_norm(Ctx *ctx, GmresWork &W, const double *v, bool cplx) {
    if (cplx)
        hipLaunchKernelGGL(k_gnorm<true>, dim3((unsigned)W.nb), dim3(kGB), 0, ctx->stream, (const double2 *)v, W.n, W.npart);
    else
        hipLaunchKernelGGL(k_gnorm<false>, dim3((unsigned)W.nb), dim3(kGB), 0, ctx->stream, v, W.n, W.npart);
    std::vector<double> np(W.nb);
    (void)hipMemcpyAsync(np.data(), W.npart, sizeof(double) * W.nb, hipMemcpyDeviceToHost, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    double t = 0.0;
    for (double x : np) t += x;
    return std::sqrt(t);
}

static int check_prec(Op *op, Op *prec) {
    if (!prec) return INTI_OK;
    if (prec->ctx != op->ctx) return set_error(op->ctx, INTI_ERR_ARG, "gmres: preconditioner on another context");
    if (prec->trg->n != op->src->n || prec->src->n != op->src->n)
        return set_error(op->ctx, INTI_ERR_SIZE, "gmres: preconditioner must be square of the operator's size");
    if (op_is_complex(prec) != op_is_complex(op))
        return set_error(op->ctx, INTI_ERR_ARG, "gmres: preconditioner and operator value types differ");
    return INTI_OK;
}

int gmres_solve(Op *op, Op *prec, const double *b, double *x, int restart, int maxiter, double rtol, double atol,
                int use_graph, int *iters, double *relres) {
    Ctx *ctx = op->ctx;
    const int64_t n = op->src->n;
    if (op->trg->n != n) return set_error(ctx, INTI_ERR_SIZE, "gmres: operator is not square");
    if (int rc = check_prec(op, prec)) return rc;
    if (restart < 1 || maxiter < 0) return set_error(ctx, INTI_ERR_ARG, "gmres: restart >= 1, maxiter >= 0");
    const bool cplx = op_is_complex(op);
    GmresWork W;
    int rc = gwork_alloc(ctx, W, restart, n, cplx);
    if (rc) return rc;
    const double bn = dev_norm(ctx, W, b, cplx);
    const double tol = std::max(rtol * bn, atol);
    double rn = 0.0;
    rc = cplx ? gmres_core<true>(op, prec, W, b, x, maxiter, tol, use_graph, 8, iters, &rn)
              : gmres_core<false>(op, prec, W, b, x, maxiter, tol, use_graph, 8, iters, &rn);
    if (rc) return rc;
    if (relres) *relres = bn > 0 ? rn / bn : rn;
    return INTI_OK;
}

int refine_solve(Op *op_hi, Op *op_lo, Op *prec, const double *b, double *x, int restart, int maxiter_inner,
                 double inner_rtol, int max_outer, double rtol, int use_graph, int *outer,
                 int *inner_total, double *relres) {
    Ctx *ctx = op_hi->ctx;
    const int64_t n = op_hi->src->n;
    if (op_lo->src != op_hi->src || op_lo->trg != op_hi->trg)
        return set_error(ctx, INTI_ERR_ARG, "refine: operators must share target and source");
    if (int rc = check_prec(op_hi, prec)) return rc;
    if (op_hi->trg->n != n) return set_error(ctx, INTI_ERR_SIZE, "refine: operator is not square");
    const bool cplx = op_is_complex(op_hi);
    if (cplx != op_is_complex(op_lo)) return set_error(ctx, INTI_ERR_ARG, "refine: value types differ");
    GmresWork W;
    int rc = gwork_alloc(ctx, W, restart, n, cplx);
    if (rc) return rc;
    const size_t es = cplx ? 16 : 8;
    void *r = nullptr, *d = nullptr;
    if (hipMalloc(&r, es * n) != hipSuccess || hipMalloc(&d, es * n) != hipSuccess) {
        if (r) (void)hipFree(r);
        return set_error(ctx, INTI_ERR_NOMEM, "refine: allocation failed");
    }
    struct Free { void *a, *b; ~Free() { (void)hipFree(a); (void)hipFree(b); } } fr{r, d};
    const double bn = dev_norm(ctx, W, b, cplx);
    double rn = 0.0;
    rc = cplx ? residual<true>(op_hi, W, b, x, (double *)r, rn) : residual<false>(op_hi, W, b, x, (double *)r, rn);
    if (rc) return rc;
    int o = 0, tot = 0;
    const dim3 gv((unsigned)((n + kGB - 1) / kGB));
    while (rn > rtol * bn && o < max_outer) {
        // d = 0; solve A_lo d = r to inner_rtol; x += d; r = b - A_hi x
        hipLaunchKernelGGL(k_gfill<double>, dim3((unsigned)((n * (cplx ? 2 : 1) + kGB - 1) / kGB)), dim3(kGB), 0,
                           ctx->stream, n * (cplx ? 2 : 1), (double *)d, 0.0);
        int it = 0;
        double irn = 0.0;
        const double itol = inner_rtol * rn;
        rc = cplx ? gmres_core<true>(op_lo, prec, W, (const double *)r, (double *)d, maxiter_inner, itol, use_graph, 8,
                                     &it, &irn)
                  : gmres_core<false>(op_lo, prec, W, (const double *)r, (double *)d, maxiter_inner, itol, use_graph, 8,
                                      &it, &irn);
        if (rc) return rc;
        tot += it;
        if (cplx)
            hipLaunchKernelGGL(k_gaxpby<true>, gv, dim3(kGB), 0, ctx->stream, n, 1.0, (const double2 *)d, 1.0, (double2 *)x);
        else
            hipLaunchKernelGGL(k_gaxpby<false>, gv, dim3(kGB), 0, ctx->stream, n, 1.0, (const double *)d, 1.0, x);
        rc = cplx ? residual<true>(op_hi, W, b, x, (double *)r, rn) : residual<false>(op_hi, W, b, x, (double *)r, rn);
        if (rc) return rc;
        ++o;
    }
    if (outer) *outer = o;
    if (inner_total) *inner_total = tot;
    if (relres) *relres = bn > 0 ? rn / bn : rn;
    return INTI_OK;
}

}  // namespace inti
