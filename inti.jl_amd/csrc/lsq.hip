// Generic batched per-element least-squares correction builder (the vdim
// step, src/vdim.jl:84-127): element e has an explicit nrows x ncols matrix
// A_e (the Vandermonde of the interpolation basis at its nodes) and every
// target i listed for e has a right-hand side row Theta[i, :]; the solution
// w of min ||A_e w - Theta_i|| gives the CSR entries (i, cols[c, e]) = w[c].
// Also the CSR pattern helper shared with the boundary correction.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "lsq.hpp"

namespace inti {

void build_pattern(int64_t m, const std::vector<int32_t> &tri_row, const std::vector<int32_t> &tri_col,
                   TripletPattern &P) {
    // triplets in (row, col) order, equal pairs in their original order: two
    // stable counting sorts (by column, then by row), linear in the number
    // of triplets (tens of millions for a stencil of many elements per
    // target)
    const int64_t ntri = (int64_t)tri_row.size();
    int32_t ncol = 0;
    for (int64_t q = 0; q < ntri; ++q) ncol = std::max(ncol, tri_col[q] + 1);
    std::vector<int32_t> bycol(ntri), order(ntri);
    {
        std::vector<int64_t> at((size_t)ncol + 1, 0);
        for (int64_t q = 0; q < ntri; ++q) ++at[tri_col[q] + 1];
        for (int32_t j = 0; j < ncol; ++j) at[j + 1] += at[j];
        for (int64_t q = 0; q < ntri; ++q) bycol[at[tri_col[q]]++] = (int32_t)q;
    }
    {
        std::vector<int64_t> at((size_t)m + 1, 0);
        for (int64_t q = 0; q < ntri; ++q) ++at[tri_row[q] + 1];
        for (int64_t i = 0; i < m; ++i) at[i + 1] += at[i];
        for (int64_t q = 0; q < ntri; ++q) {
            const int32_t u = bycol[q];
            order[at[tri_row[u]]++] = u;
        }
    }
    P.seg.clear(); P.colind.clear();
    P.seg.reserve(ntri + 1); P.colind.reserve(ntri);
    P.perm.assign(ntri, 0);
    P.rowptr.assign(m + 1, 0);
    for (int64_t q = 0; q < ntri; ++q) {
        const int32_t u = order[q];
        P.perm[q] = u;
        if (q == 0 || tri_row[u] != tri_row[order[q - 1]] || tri_col[u] != tri_col[order[q - 1]]) {
            P.seg.push_back((int32_t)q);
            P.colind.push_back(tri_col[u]);
            P.rowptr[tri_row[u] + 1]++;
        }
    }
    P.seg.push_back((int32_t)ntri);
    for (int64_t i = 0; i < m; ++i) P.rowptr[i + 1] += P.rowptr[i];
}

static __global__ __launch_bounds__(256) void k_gather1(Gather1Args a) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= a.nnz) return;
    double sr = 0, si = 0;
    for (int32_t t = a.seg[q]; t < a.seg[q + 1]; ++t) {
        const double2 v = a.tri[a.perm[t]];
        sr += v.x; si += v.y;
    }
    if (a.cplx) {
        a.val[2 * q] = a.sign * sr;
        a.val[2 * q + 1] = a.sign * si;
    } else {
        a.val[q] = a.sign * sr;
    }
}

void launch_gather1(const Gather1Args &g, hipStream_t s) {
    if (g.nnz > 0) hipLaunchKernelGGL(k_gather1, dim3((unsigned)((g.nnz + 255) / 256)), dim3(256), 0, s, g);
}

struct LsqArgs {
    int nrows, ncols;
    int64_t m;
    const double *A;          // ne x (nrows x ncols) real, column-major per element
    const int32_t *elems;     // elements with targets
    const int32_t *pair_ptr;  // ne+1
    const int32_t *pair_tgt;  // pair -> target row of rhs
    const double2 *rhs;       // m x nrows (complex)
    double2 *tri;             // pair * ncols + c
    double *diag;             // per element: rdiag ratio, max |w|
};

// A wide system (nrows < ncols: fewer basis functions than nodes) takes the
// minimum-norm solution, which is what the pseudo-inverse of the SVD gives for
// a full-rank matrix.
static __global__ __launch_bounds__(64) void k_lsq(LsqArgs a) {
    __shared__ double2 A[kMaxNs * kMaxNc];
    __shared__ double2 tau[kMaxNc];
    __shared__ double2 s_xi1;
    __shared__ double2 bb[64][kMaxNs + 1];
    const int lane = threadIdx.x;
    const int e = a.elems[blockIdx.x];
    const bool wide = a.nrows < a.ncols;
    const int ns = wide ? a.ncols : a.nrows, nc = wide ? a.nrows : a.ncols;
    const double *Ae = a.A + (int64_t)e * a.nrows * a.ncols;
    for (int q = lane; q < ns * nc; q += 64) {
        const int r = q % ns, c = q / ns;
        A[q] = make_double2(wide ? Ae[c + a.nrows * r] : Ae[q], 0.0);
    }
    __syncthreads();
    wave_householder_qr(A, ns, nc, tau, &s_xi1);
    const double ratio = rdiag_ratio(A, ns, nc);
    const int p0 = a.pair_ptr[e], p1 = a.pair_ptr[e + 1];
    double wmax = 0.0;
    for (int pb = p0; pb < p1; pb += 64) {
        const int p = pb + lane;
        if (p < p1) {
            const int64_t i = a.pair_tgt[p];
            double2 *b = bb[lane];
            for (int q = 0; q < a.nrows; ++q) b[q] = a.rhs[i + a.m * q];
            if (wide)
                lane_qr_minnorm(A, ns, nc, tau, b);
            else
                lane_qr_solve(A, ns, nc, tau, b);
            double wn = 0.0;
            for (int c = 0; c < a.ncols; ++c) {
                wn += b[c].x * b[c].x + b[c].y * b[c].y;
                a.tri[(int64_t)p * a.ncols + c] = b[c];
            }
            wmax = fmax(wmax, sqrt(wn));
        }
    }
    for (int off = 32; off > 0; off >>= 1) wmax = fmax(wmax, __shfl_xor(wmax, off));
    if (lane == 0) {
        a.diag[2 * (int64_t)blockIdx.x] = ratio;
        a.diag[2 * (int64_t)blockIdx.x + 1] = wmax;
    }
}

int csr_alloc(Ctx *ctx, int64_t m, int64_t n, int64_t nnz, int cplx, Csr **out);
void csr_free(Csr *c);

struct HipBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    ~HipBuf() {
        if (p) {
            if (s) (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
    }
    int alloc(Ctx *ctx, size_t bytes) {
        s = ctx->stream;
        return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? INTI_OK
                                                              : set_error(ctx, INTI_ERR_NOMEM, "lsq: allocation failed");
    }
};

// Element types: type t has ncols[t] columns (nodes) and ne[t] elements; A
// holds the nrows x ncols[t] matrices of its elements after the earlier
// types', cols the ncols[t] x ne[t] column-major block after theirs, and
// (near_ptr non-NULL) near_ptr ne[t] + 1 entries per type (each block from
// 0) indexing the block of near_idx after the earlier types' targets. One
// k_lsq launch per type (vdim.jl:76-121 loops over the element types).
int lsq_build_types(Ctx *ctx, int cplx, int nrows, int ntypes, const int *ncols, const int64_t *ne, const double *A,
                    const int64_t *cols, int64_t m, int64_t n, const int64_t *near_ptr, const int64_t *near_idx,
                    const double *rhs, Csr **out, double *diag_out) {
    if (ntypes < 1) return set_error(ctx, INTI_ERR_ARG, "lsq: no element type");
    for (int t = 0; t < ntypes; ++t)
        if (nrows < 1 || ncols[t] < 1 || ne[t] < 0 || std::max(nrows, ncols[t]) > kMaxNs ||
            std::min(nrows, ncols[t]) > kMaxNc)
            return set_error(ctx, INTI_ERR_UNSUPPORTED,
                             "lsq: need 1 <= min(nrows, ncols) <= 24 and max(nrows, ncols) <= 32");
    const bool self = near_ptr == nullptr;
    if (self && m != n) return set_error(ctx, INTI_ERR_ARG, "lsq: the self pattern needs m == n");
    // per-type offsets
    std::vector<int64_t> a_off(ntypes + 1, 0), c_off(ntypes + 1, 0), pp_off(ntypes + 1, 0), pt_off(ntypes + 1, 0),
        tri_off(ntypes + 1, 0), el_off(ntypes + 1, 0);
    std::vector<int32_t> pair_ptr, pair_tgt, elems, tri_row, tri_col;
    int64_t nptr = 0, nidx = 0;
    for (int t = 0; t < ntypes; ++t) {
        const int nc = ncols[t];
        const int64_t *ct = cols + c_off[t];
        a_off[t + 1] = a_off[t] + ne[t] * nrows * nc;
        c_off[t + 1] = c_off[t] + (int64_t)nc * ne[t];
        for (int64_t q = 0; q < (int64_t)nc * ne[t]; ++q)
            if (ct[q] < 0 || ct[q] >= n) return set_error(ctx, INTI_ERR_ARG, "lsq: column index out of range");
        const int64_t *np = self ? nullptr : near_ptr + nptr;
        const int64_t *ni = self ? nullptr : near_idx + nidx;
        const int64_t p0 = (int64_t)pair_tgt.size();
        pair_ptr.push_back(0);
        for (int64_t e = 0; e < ne[t]; ++e) {
            const int64_t cnt = self ? nc : np[e + 1] - np[e];
            if (cnt < 0) return set_error(ctx, INTI_ERR_ARG, "lsq: near lists must be non-decreasing");
            for (int64_t q = 0; q < cnt; ++q) {
                const int64_t i = self ? ct[q + (int64_t)nc * e] : ni[np[e] + q];
                if (i < 0 || i >= m) return set_error(ctx, INTI_ERR_ARG, "lsq: target out of range");
                pair_tgt.push_back((int32_t)i);
            }
            pair_ptr.push_back((int32_t)((int64_t)pair_tgt.size() - p0));
            if (cnt > 0) elems.push_back((int32_t)e);
            for (int64_t q = 0; q < cnt; ++q)
                for (int c = 0; c < nc; ++c) {
                    tri_row.push_back(pair_tgt[p0 + pair_ptr[pp_off[t] + e] + q]);
                    tri_col.push_back((int32_t)ct[c + (int64_t)nc * e]);
                }
        }
        pp_off[t + 1] = pp_off[t] + ne[t] + 1;
        pt_off[t + 1] = (int64_t)pair_tgt.size();
        tri_off[t + 1] = (int64_t)tri_row.size();
        el_off[t + 1] = (int64_t)elems.size();
        if (tri_off[t + 1] >= (int64_t)INT32_MAX) return set_error(ctx, INTI_ERR_SIZE, "lsq: too many triplets");
        if (!self) {
            nidx += np[ne[t]];
            nptr += ne[t] + 1;
        }
    }
    const int64_t npairs = (int64_t)pair_tgt.size(), ntri = tri_off[ntypes], nel = (int64_t)elems.size();
    TripletPattern P;
    build_pattern(m, tri_row, tri_col, P);
    const int64_t nnz = (int64_t)P.colind.size();
    // device work
    hipStream_t s = ctx->stream;
    HipBuf dA, dpp, dpt, del, drhs, dtri, ddiag, dseg, dperm;
    int rc;
    if ((rc = dA.alloc(ctx, sizeof(double) * a_off[ntypes])) || (rc = dpp.alloc(ctx, sizeof(int32_t) * pair_ptr.size())) ||
        (rc = dpt.alloc(ctx, sizeof(int32_t) * npairs)) || (rc = del.alloc(ctx, sizeof(int32_t) * nel)) ||
        (rc = drhs.alloc(ctx, sizeof(double2) * m * nrows)) || (rc = dtri.alloc(ctx, sizeof(double2) * ntri)) ||
        (rc = ddiag.alloc(ctx, sizeof(double) * 2 * nel)) || (rc = dseg.alloc(ctx, sizeof(int32_t) * P.seg.size())) ||
        (rc = dperm.alloc(ctx, sizeof(int32_t) * P.perm.size())))
        return rc;
    if (a_off[ntypes]) INTI_HIP(ctx, hipMemcpyAsync(dA.p, A, sizeof(double) * a_off[ntypes], hipMemcpyHostToDevice, s));
    INTI_HIP(ctx, hipMemcpyAsync(dpp.p, pair_ptr.data(), sizeof(int32_t) * pair_ptr.size(), hipMemcpyHostToDevice, s));
    if (npairs) INTI_HIP(ctx, hipMemcpyAsync(dpt.p, pair_tgt.data(), sizeof(int32_t) * npairs, hipMemcpyHostToDevice, s));
    if (nel) INTI_HIP(ctx, hipMemcpyAsync(del.p, elems.data(), sizeof(int32_t) * nel, hipMemcpyHostToDevice, s));
    INTI_HIP(ctx, hipMemcpyAsync(dseg.p, P.seg.data(), sizeof(int32_t) * P.seg.size(), hipMemcpyHostToDevice, s));
    if (!P.perm.empty()) INTI_HIP(ctx, hipMemcpyAsync(dperm.p, P.perm.data(), sizeof(int32_t) * P.perm.size(), hipMemcpyHostToDevice, s));
    // right-hand sides as complex (m x nrows)
    std::vector<double> rc2((size_t)2 * m * nrows);
    for (int64_t q = 0; q < m * nrows; ++q) {
        rc2[2 * q] = cplx ? rhs[2 * q] : rhs[q];
        rc2[2 * q + 1] = cplx ? rhs[2 * q + 1] : 0.0;
    }
    INTI_HIP(ctx, hipMemcpyAsync(drhs.p, rc2.data(), sizeof(double) * rc2.size(), hipMemcpyHostToDevice, s));
    for (int t = 0; t < ntypes; ++t) {
        const int64_t cnt = el_off[t + 1] - el_off[t];
        if (cnt == 0) continue;
        LsqArgs a;
        a.nrows = nrows; a.ncols = ncols[t]; a.m = m;
        a.A = (const double *)dA.p + a_off[t];
        a.elems = (const int32_t *)del.p + el_off[t];
        a.pair_ptr = (const int32_t *)dpp.p + pp_off[t];
        a.pair_tgt = (const int32_t *)dpt.p + pt_off[t];
        a.rhs = (const double2 *)drhs.p;
        a.tri = (double2 *)dtri.p + tri_off[t];
        a.diag = (double *)ddiag.p + 2 * el_off[t];
        hipLaunchKernelGGL(k_lsq, dim3((unsigned)cnt), dim3(64), 0, s, a);
        INTI_HIP(ctx, hipGetLastError());
    }
    Csr *c = nullptr;
    rc = csr_alloc(ctx, m, n, nnz, cplx, &c);
    if (rc) return rc;
    if (hipMemcpyAsync(c->rowptr, P.rowptr.data(), sizeof(int32_t) * (m + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        (nnz && hipMemcpyAsync(c->colind, P.colind.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s) != hipSuccess)) {
        csr_free(c);
        return set_error(ctx, INTI_ERR_HIP, "lsq: CSR upload failed");
    }
    Gather1Args g;
    g.nnz = nnz; g.seg = (const int32_t *)dseg.p; g.perm = (const int32_t *)dperm.p;
    g.tri = (const double2 *)dtri.p; g.val = c->val; g.cplx = cplx; g.sign = 1.0;
    launch_gather1(g, s);
    std::vector<double> dg(2 * nel);
    if ((!dg.empty() && hipMemcpyAsync(dg.data(), ddiag.p, sizeof(double) * dg.size(), hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess) {
        csr_free(c);
        return set_error(ctx, INTI_ERR_HIP, "lsq: kernel execution failed");
    }
    if (diag_out) {
        double r = 0, w = 0;
        for (int64_t q = 0; q < nel; ++q) { r = std::max(r, dg[2 * q]); w = std::max(w, dg[2 * q + 1]); }
        diag_out[0] = r; diag_out[1] = w;
    }
    *out = c;
    return INTI_OK;
}

int lsq_build(Ctx *ctx, int cplx, int nrows, int ncols, int64_t ne, const double *A, const int64_t *cols,
              int64_t m, int64_t n, const int64_t *near_ptr, const int64_t *near_idx, const double *rhs,
              Csr **out, double *diag_out) {
    return lsq_build_types(ctx, cplx, nrows, 1, &ncols, &ne, A, cols, m, n, near_ptr, near_idx, rhs, out, diag_out);
}

}  // namespace inti
