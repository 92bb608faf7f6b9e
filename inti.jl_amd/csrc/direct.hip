// Direct-sum driver: operator parameters (post factors), the choice of
// source chunks, the multi-right-hand-side dispatch of k_apply
// (apply_kernel.hpp), and the epilogue pass (split reduction + post factor +
// fused CSR correction + identity + alpha/beta).
#include <complex>

#include "apply_kernel.hpp"
#include "apply_sym.hpp"

namespace inti {

struct ReduceArgs {
    int64_t m;
    int split;
    const double *part;  // nr x split x m
    int64_t ldx, ldy;    // leading dimensions of x and y (elements)
    const int32_t *rowptr;
    const int32_t *colind;
    const double *cval;
    const double *x;  // density (for the correction and identity terms)
    double *y;
    double ar, ai, br, bi, cr, ci;
    double pr, pi;  // post factor of the direct sum
    int beta_zero;
    int has_id;
};

// y_i = alpha * (post * sum_c part[c][i] + (dA x)_i + cid x_i) + beta * y_i,
// one right-hand side per blockIdx.y
template <bool CPLX>
__global__ __launch_bounds__(256) void k_reduce(ReduceArgs a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.m) return;
    const int r = blockIdx.y;
    const double *x = a.x + (CPLX ? 2 : 1) * (int64_t)r * a.ldx;
    double *y = a.y + (CPLX ? 2 : 1) * (int64_t)r * a.ldy;
    const int64_t pbase = (int64_t)r * a.split * a.m;
    double sr = 0.0, si = 0.0;
    // many partials per row (one per tile group of the symmetric sum): the
    // unrolled loop keeps several loads in flight per thread (the order of
    // the additions is unchanged)
    if constexpr (CPLX) {
        const double2 *pp = reinterpret_cast<const double2 *>(a.part) + pbase;
#pragma unroll 8
        for (int c = 0; c < a.split; ++c) {
            double2 v = pp[(int64_t)c * a.m + i];
            sr += v.x; si += v.y;
        }
    } else {
#pragma unroll 8
        for (int c = 0; c < a.split; ++c) sr += a.part[pbase + (int64_t)c * a.m + i];
    }
    {
        const double tr = a.pr * sr - a.pi * si, ti = a.pr * si + a.pi * sr;
        sr = tr;
        si = ti;
    }
    if (a.rowptr) {
        const int32_t p0 = a.rowptr[i], p1 = a.rowptr[i + 1];
        for (int32_t p = p0; p < p1; ++p) {
            const int32_t j = a.colind[p];
            if constexpr (CPLX) {
                const double vr = a.cval[2 * p], vi = a.cval[2 * p + 1];
                const double xr = x[2 * j], xi = x[2 * j + 1];
                sr += vr * xr - vi * xi;
                si += vr * xi + vi * xr;
            } else {
                sr += a.cval[p] * x[j];
            }
        }
    }
    if (a.has_id) {
        if constexpr (CPLX) {
            const double xr = x[2 * i], xi = x[2 * i + 1];
            sr += a.cr * xr - a.ci * xi;
            si += a.cr * xi + a.ci * xr;
        } else {
            sr += a.cr * x[i];
        }
    }
    if constexpr (CPLX) {
        double yr = a.ar * sr - a.ai * si, yi = a.ar * si + a.ai * sr;
        if (!a.beta_zero) {
            const double or_ = y[2 * i], oi = y[2 * i + 1];
            yr += a.br * or_ - a.bi * oi;
            yi += a.br * oi + a.bi * or_;
        }
        y[2 * i] = yr;
        y[2 * i + 1] = yi;
    } else {
        double yr = a.ar * sr;
        if (!a.beta_zero) yr += a.br * y[i];
        y[i] = yr;
    }
}

// Long correction rows (a stencil of many elements per target: hundreds of
// entries per row): W lanes share a row, so the value and index loads of a
// wave are contiguous, and y_i += alpha (dA x)_i after k_reduce (which then
// skips the rows). The W partial sums of a row combine by a fixed butterfly,
// so the result does not depend on the launch. One row per W lanes, 256
// threads per block, one right-hand side per blockIdx.y.
struct CsrAddArgs {
    int64_t m;
    int64_t ldx, ldy;
    const int32_t *rowptr;
    const int32_t *colind;
    const double *cval;
    const double *x;
    double *y;
    double ar, ai;
};

template <bool CPLX, int W>
__global__ __launch_bounds__(256) void k_csr_add(CsrAddArgs a) {
    const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) / W;
    const int lane = threadIdx.x % W;
    if (i >= a.m) return;  // whole groups of W lanes leave together
    const int r = blockIdx.y;
    const double *x = a.x + (CPLX ? 2 : 1) * (int64_t)r * a.ldx;
    double *y = a.y + (CPLX ? 2 : 1) * (int64_t)r * a.ldy;
    const int32_t p0 = a.rowptr[i], p1 = a.rowptr[i + 1];
    double sr = 0.0, si = 0.0;
    if constexpr (CPLX) {
        const double2 *v = reinterpret_cast<const double2 *>(a.cval);
        const double2 *xv = reinterpret_cast<const double2 *>(x);
#pragma unroll 4
        for (int32_t p = p0 + lane; p < p1; p += W) {
            const double2 c = v[p], xx = xv[a.colind[p]];
            sr = fma(c.x, xx.x, fma(-c.y, xx.y, sr));
            si = fma(c.x, xx.y, fma(c.y, xx.x, si));
        }
    } else {
#pragma unroll 4
        for (int32_t p = p0 + lane; p < p1; p += W) sr = fma(a.cval[p], x[a.colind[p]], sr);
    }
#pragma unroll
    for (int o = W / 2; o > 0; o /= 2) {
        sr += __shfl_xor(sr, o, W);
        if constexpr (CPLX) si += __shfl_xor(si, o, W);
    }
    if (lane == 0) {
        if constexpr (CPLX) {
            y[2 * i] += a.ar * sr - a.ai * si;
            y[2 * i + 1] += a.ar * si + a.ai * sr;
        } else {
            y[i] += a.ar * sr;
        }
    }
}

// lanes per row of the correction: the fused per-thread walk of k_reduce
// for short rows (the element's own nodes: a few entries), a group of 16 or
// 64 lanes for long ones; 0 = fused
static int csr_lanes(const Csr *c) {
    if (!c || c->m == 0) return 0;
    const int64_t avg = c->nnz / c->m;
    return avg >= 96 ? 64 : avg >= 16 ? 16 : 0;
}

// ------------------------------------------------------------- dispatch
static ApplyFn pick_apply(int pde, int dim, int kind, int mixed, int nr) {
    if (pde == INTI_LAPLACE && dim == 2) return apply_pick_lap2(kind, mixed, nr);
    if (pde == INTI_LAPLACE && dim == 3) return apply_pick_lap3(kind, mixed, nr);
    if (pde == INTI_HELMHOLTZ && dim == 2) return apply_pick_helm2(kind, mixed, nr);
    if (pde == INTI_HELMHOLTZ && dim == 3) return apply_pick_helm3(kind, mixed, nr);
    if (pde == kHelmCK && dim == 3 && !mixed) return apply_pick_helm3c(kind, nr);
    return nullptr;
}

static SymFn pick_sym(int pde, int dim, int kind, int mixed, int G) {
    if (pde == INTI_LAPLACE && dim == 2) return sym_pick_lap2(kind, mixed, G);
    if (pde == INTI_LAPLACE && dim == 3) return sym_pick_lap3(kind, mixed, G);
    if (pde == INTI_HELMHOLTZ && dim == 2) return sym_pick_helm2(kind, mixed, G);
    if (pde == INTI_HELMHOLTZ && dim == 3) return sym_pick_helm3(kind, mixed, G);
    return nullptr;
}

bool op_is_complex(const Op *op) { return op->pde == INTI_HELMHOLTZ; }
int op_kernel_pde(const Op *op) { return op->pde == INTI_HELMHOLTZ && op->kimag != 0.0 ? kHelmCK : op->pde; }

// partial sums of the symmetric evaluation are capped: beyond this the
// plain direct sum runs
constexpr size_t kSymMaxBytes = size_t(4) << 30;

// Symmetric plan for one right-hand side of op: tiles of 256 in groups of G
// (the smallest of 1, 2, 4 whose S x n partials fit kSymMaxBytes). False if
// the operator does not qualify.
static bool plan_sym(const Op *op, int kind, int nr, int &G, int64_t &T, int64_t &S) {
    if (op->sym == 0 || nr != 1 || op->trg != op->src || kind >= INTI_FF_SL || op->kimag != 0.0) return false;
    const int64_t n = op->src->n;
    T = (n + kSymBS - 1) / kSymBS;
    if (T < 2) return false;
    // automatic mode: only where the triangle of tile pairs fills the GPU
    // (two blocks per CU); below that the plain sum, whose source range is
    // split across blocks, is faster
    if (op->sym == 1 && T * (T + 1) / 2 < 2LL * op->ctx->num_cu) return false;
    const size_t cw = op_is_complex(op) ? 2 : 1;
    for (int g : {1, 2, 4}) {
        if (op->sym_group && g != op->sym_group) continue;
        S = (T + g - 1) / g;
        if ((size_t)S * (size_t)n * cw * sizeof(double) <= kSymMaxBytes) {
            G = g;
            return true;
        }
    }
    return false;
}

KParams op_params(const Op *op, int *eff_kind) {
    using cd = std::complex<double>;
    KParams p = basic_params(op->k);
    p.kimag = op->kimag;
    int kind = op->kind;
    const cd c0(op->coefs[0], op->coefs[1]), c1(op->coefs[2], op->coefs[3]);
    cd post, ch(0.0, 0.0);
    double pr, pi;
    if (kind == INTI_COMB_SD || kind == INTI_COMB_AH || kind == INTI_FF_COMB) {
        const int ka = kind == INTI_COMB_SD ? INTI_SL : kind == INTI_COMB_AH ? INTI_ADL : INTI_FF_SL;
        const int kb = kind == INTI_COMB_SD ? INTI_DL : kind == INTI_COMB_AH ? INTI_HS : INTI_FF_DL;
        double ar, ai, br, bi;
        basic_post(op->pde, op->dim, ka, op->k, ar, ai);
        basic_post(op->pde, op->dim, kb, op->k, br, bi);
        if (c1 == cd(0.0, 0.0)) {  // only the first component remains
            kind = ka;
            post = c0 * cd(ar, ai);
        } else {
            // K = c1 postB (rawB + ch rawA), ch = c0 postA / (c1 postB); the
            // 3D/2D Helmholtz raw values of the combined kinds fold the
            // remaining constants as documented with helm3_raw / helm2_raw
            post = c1 * cd(br, bi);
            ch = c0 * cd(ar, ai) / post;
            if (op->pde == INTI_HELMHOLTZ && op->dim == 3 && kind == INTI_COMB_AH)
                ch = -ch;  // the ADL raw part inside COMB_AH carries -rdnx
            if (op->pde == INTI_HELMHOLTZ && op->dim == 2 && kind == INTI_COMB_AH)
                ch = -ch;
        }
    } else {
        basic_post(op->pde, op->dim, kind, op->k, pr, pi);
        post = cd(pr, pi);
    }
    p.postr = post.real(); p.posti = post.imag();
    p.chr = ch.real(); p.chi = ch.imag();
    if (eff_kind) *eff_kind = kind;
    return p;
}

// number of source chunks: enough workgroups to fill 256 CUs several
// times over, a multiple of the 8 XCDs when more than 8.
// Source chunks of the one-direction sum: block b works on target tile
// b / split and chunk b % split. The base split gives about 8 blocks per CU
// in whole 256-source tiles, a multiple of 8 from 8 on (all blocks of one XCD
// then share a chunk in their L2); an explicit split (inti_op_set_launch)
// replaces it. Without one, the plan then looks at splits from half to twice
// the base (multiples of 8 from 8 on) with chunks of whole LDS source tiles of
// the kernel (256, 128 or 64 sources for 1, 4 or 16 right-hand sides) and
// takes the one with the least rounds x chunk, where blocks run in rounds of
// `cap` resident blocks (occupancy x CUs): a grid just past a multiple of cap
// leaves its last round nearly empty.
static void plan_split(const Op *op, int64_t m, int64_t n, ApplyFn fn, int nr, int &split, int64_t &chunk) {
    const int64_t ntiles = (m + kApplyBS - 1) / kApplyBS;
    const int64_t stiles = (n + kApplyBS - 1) / kApplyBS;
    int64_t s0 = op->split;
    if (s0 <= 0) {
        s0 = (8LL * op->ctx->num_cu + ntiles - 1) / ntiles;
        if (s0 >= 8) s0 = (s0 + 7) / 8 * 8;
    }
    s0 = std::max<int64_t>(1, std::min(s0, stiles));
    chunk = (stiles + s0 - 1) / s0 * kApplyBS;
    if (op->split <= 0 && fn && n > 0) {
        const int64_t gran = nr == 16 ? kMfmaTT : nr == 4 ? ApplyTile<4>::TT : ApplyTile<1>::TT;
        const int64_t gtiles = (n + gran - 1) / gran;
        const double cap = (double)fn(nullptr, dim3(0), nullptr) * op->ctx->num_cu;
        auto cost_of = [&](int64_t c) { return std::ceil((double)(ntiles * ((n + c - 1) / c)) / cap) * (double)c; };
        double best = cost_of(chunk);  // the base plan stays unless a candidate beats it
        for (int64_t s = std::max<int64_t>(1, s0 / 2); s <= std::min(2 * s0, gtiles); s += s < 8 ? 1 : 8) {
            const int64_t c = (gtiles + s - 1) / s * gran;
            const double cost = cost_of(c);
            if (cost < best) {
                best = cost;
                chunk = c;
            }
        }
    }
    split = (int)((n + chunk - 1) / chunk);
    if (split < 1) split = 1;
}

int apply_reserve(const Op *op, int nrhs) {
    const int cw = op_is_complex(op) ? 2 : 1;
    const int64_t m = op->trg->n, n = op->src->n;
    int kind = op->kind, G;
    int64_t T, S;
    (void)op_params(op, &kind);
    // the largest partial buffer of the column blocks (16, 4, 1) a call uses
    size_t bytes = 0;
    for (int nr : {16, 4, 1}) {
        if (nrhs < nr && nr != 1) continue;
        int split;
        int64_t chunk;
        plan_split(op, m, n, pick_apply(op_kernel_pde(op), op->dim, kind, op->precision == INTI_MIXED, nr), nr, split, chunk);
        bytes = std::max(bytes, sizeof(double) * cw * (size_t)m * split * nr);
    }
    if (nrhs >= 1 && plan_sym(op, kind, 1, G, T, S))
        bytes = std::max(bytes, sizeof(double) * cw * (size_t)S * (size_t)n);
    void *scr = nullptr;
    return ctx_scratch(op->ctx, kScrPartial, bytes, &scr);
}

int launch_apply(const DirectLaunch &L) {
    const Op *op = L.op;
    Ctx *ctx = op->ctx;
    const bool cplx = op_is_complex(op);
    const int cw = cplx ? 2 : 1;
    const int64_t m = op->trg->n, n = op->src->n;
    if (m == 0 || L.nrhs == 0) return INTI_OK;
    int kind = op->kind;
    const KParams kp = op_params(op, &kind);
    const bool given = L.part != nullptr;
    void *scr = nullptr;
    double *part = L.part;
    if (!given) {
        int st = apply_reserve(op, L.nrhs);
        if (st) return st;
        st = ctx_scratch(ctx, kScrPartial, 0, &scr);
        if (st) return st;
        part = (double *)scr;
    }
    int st = INTI_OK;
    const int64_t ntiles = (m + kApplyBS - 1) / kApplyBS;
    int symG = 0;
    int64_t symT = 0, symS = 0;
    const bool sym = !given && plan_sym(op, kind, 1, symG, symT, symS);
    if (sym) {
        const size_t bytes = sizeof(double) * cw * (size_t)symS * (size_t)n;
        st = ctx_scratch(ctx, kScrPartial, bytes, &scr);
        if (st) return st;
        part = (double *)scr;
    }
    for (int r0 = 0; r0 < L.nrhs;) {
        const int left = L.nrhs - r0;
        const int nr = left >= 16 ? 16 : left >= 4 ? 4 : 1;
        const double *xr = L.x + (int64_t)r0 * L.ldx * cw;
        double *yr = L.y + (int64_t)r0 * L.ldy * cw;
        ApplyFn fn = sym && nr == 1 ? nullptr : pick_apply(op_kernel_pde(op), op->dim, kind, op->precision == INTI_MIXED, nr);
        int split = 1;
        int64_t chunk = n;
        if (fn) plan_split(op, m, n, fn, nr, split, chunk);
        if (given && n > 0) {
            // the largest split whose partials fit the caller's buffer, in
            // chunks of whole 256-source tiles
            const size_t per = sizeof(double) * cw * (size_t)m * nr;
            const int64_t cap = (int64_t)(L.part_bytes / per);
            if (cap < 1)
                return set_error(ctx, INTI_ERR_SIZE, "apply: the partial sums of this shape exceed the buffer");
            if (split > cap) {
                const int64_t stiles = (n + kApplyBS - 1) / kApplyBS;
                chunk = (stiles + cap - 1) / cap * kApplyBS;
                split = (int)((n + chunk - 1) / chunk);
            }
        }
        int nsplit = split;
        if (n > 0 && nr == 1 && sym) {
            SymFn sfn = pick_sym(op->pde, op->dim, kind, op->precision == INTI_MIXED, symG);
            if (!sfn) return set_error(ctx, INTI_ERR_UNSUPPORTED, "no symmetric kernel for this operator");
            SymArgs a;
            a.n = n;
            a.xy = op->src->coords; a.nrm = op->src->normals; a.w = op->src->weights;
            a.x = xr; a.part = part;
            a.ntiles = symT; a.ngroups = symS;
            a.distinct = op->src->distinct ? 1 : 0;
            // c0 SL + c1 DL with c0 = -i k c1 (the combined field of the
            // reference's examples): ch q needs no LDS slot of its own
            a.ch_ik = kind == INTI_COMB_SD && op->pde == INTI_HELMHOLTZ && op->dim == 3 &&
                              std::fabs(kp.chr) <= 1e-15 * kp.k && std::fabs(kp.chi + kp.k) <= 1e-15 * kp.k
                          ? 1
                          : 0;
            a.p = kp;
            if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
            sfn(a, symS * (symS + 1) / 2, ctx->stream);
            INTI_HIP(ctx, hipGetLastError());
            if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
            nsplit = (int)symS;
        } else if (n > 0) {
            if (!fn) return set_error(ctx, INTI_ERR_UNSUPPORTED, "no direct-sum kernel for this operator");
            ApplyArgs a;
            a.m = m; a.n = n;
            a.tstride = m; a.sstride = n;
            a.tx = op->trg->coords; a.tn = op->trg->normals;
            a.sx = op->src->coords; a.sn = op->src->normals; a.sw = op->src->weights;
            a.x = xr; a.ldx = L.ldx; a.part = part; a.split = split; a.chunk = chunk;
            a.p = kp;
            dim3 grid((unsigned)(ntiles * split));
            if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
            fn(&a, grid, ctx->stream);
            INTI_HIP(ctx, hipGetLastError());
            if (ctx->profiling) INTI_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        }
        ReduceArgs ra;
        ra.m = m; ra.split = n > 0 ? nsplit : 0; ra.part = part;
        ra.ldx = L.ldx; ra.ldy = L.ldy;
        const int cl = csr_lanes(op->corr);
        ra.rowptr = op->corr && !cl ? op->corr->rowptr : nullptr;
        ra.colind = op->corr && !cl ? op->corr->colind : nullptr;
        ra.cval = op->corr && !cl ? op->corr->val : nullptr;
        ra.x = xr; ra.y = yr;
        ra.ar = L.alpha[0]; ra.ai = L.alpha[1];
        ra.br = L.beta[0]; ra.bi = L.beta[1];
        ra.beta_zero = (L.beta[0] == 0.0 && L.beta[1] == 0.0);
        ra.cr = op->cid[0]; ra.ci = op->cid[1];
        ra.pr = kp.postr; ra.pi = kp.posti;
        ra.has_id = (op->cid[0] != 0.0 || op->cid[1] != 0.0);
        dim3 rg((unsigned)((m + 255) / 256), (unsigned)nr);
        if (cplx) hipLaunchKernelGGL(k_reduce<true>, rg, dim3(256), 0, ctx->stream, ra);
        else hipLaunchKernelGGL(k_reduce<false>, rg, dim3(256), 0, ctx->stream, ra);
        INTI_HIP(ctx, hipGetLastError());
        if (cl) {
            CsrAddArgs ca;
            ca.m = m; ca.ldx = L.ldx; ca.ldy = L.ldy;
            ca.rowptr = op->corr->rowptr; ca.colind = op->corr->colind; ca.cval = op->corr->val;
            ca.x = xr; ca.y = yr;
            ca.ar = L.alpha[0]; ca.ai = L.alpha[1];
            dim3 cg((unsigned)((m * cl + 255) / 256), (unsigned)nr);
            if (cplx) {
                if (cl == 64) hipLaunchKernelGGL((k_csr_add<true, 64>), cg, dim3(256), 0, ctx->stream, ca);
                else hipLaunchKernelGGL((k_csr_add<true, 16>), cg, dim3(256), 0, ctx->stream, ca);
            } else {
                if (cl == 64) hipLaunchKernelGGL((k_csr_add<false, 64>), cg, dim3(256), 0, ctx->stream, ca);
                else hipLaunchKernelGGL((k_csr_add<false, 16>), cg, dim3(256), 0, ctx->stream, ca);
            }
            INTI_HIP(ctx, hipGetLastError());
        }
        r0 += nr;
    }
    return INTI_OK;
}

}  // namespace inti
