// Batched small dense least squares for the density-interpolation
// corrections (src/bdim.jl:159-182, src/vdim.jl:96-117): one wavefront per
// element factorises its ns x nc matrix by Householder QR in LDS (lane per
// column for the reflector updates), then one lane per corrected target
// solves min ||A w - b||. The reflector convention is the generic unblocked
// QR of LinearAlgebra that qr!(transpose(M)) runs in bdim_correction:
// nu = copysign(||x||, Re x1), x1 += nu, x[1] = -nu, x[2:] /= x1,
// tau = x1 / nu; the solve applies Q^H then back-substitutes column by column.
#pragma once
#include <vector>

#include "common.hpp"
#include "pair.hpp"

namespace inti {

constexpr int kMaxNs = 32;  // rows of the least-squares matrices
constexpr int kMaxNc = 24;  // columns (unknowns)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {  // conj(a)*b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {  // Smith's division
    if (fabs(b.x) >= fabs(b.y)) {
        double t = b.y / b.x, den = b.x + b.y * t;
        return make_double2((a.x + a.y * t) / den, (a.y - a.x * t) / den);
    }
    double t = b.x / b.y, den = b.y + b.x * t;
    return make_double2((a.x * t + a.y) / den, (a.y * t - a.x) / den);
}

// In-place QR of A (ns x nc, column-major, LDS) by the 64 lanes of a
// one-wavefront block; tau[nc] receives the reflector scalars.
__device__ inline void wave_householder_qr(double2 *A, int ns, int nc, double2 *tau, double2 *s_xi1) {
    const int lane = threadIdx.x;
    for (int kk = 0; kk < nc; ++kk) {
        const int len = ns - kk;
        double2 *x = A + kk + ns * kk;
        if (lane == 0) {
            double s = 0.0;
            for (int q = 0; q < len; ++q) s += x[q].x * x[q].x + x[q].y * x[q].y;
            const double normu = sqrt(s);
            double2 t = make_double2(0.0, 0.0);
            if (normu != 0.0) {
                double2 xi1 = x[0];
                const double nu = copysign(normu, xi1.x);
                xi1.x += nu;
                x[0] = make_double2(-nu, 0.0);
                t = make_double2(xi1.x / nu, xi1.y / nu);
                *s_xi1 = xi1;
            }
            tau[kk] = t;
        }
        __syncthreads();
        const double2 t = tau[kk];
        if (t.x != 0.0 || t.y != 0.0) {
            const double2 xi1 = *s_xi1;
            for (int q = 1 + lane; q < len; q += 64) x[q] = cdiv(x[q], xi1);
        }
        __syncthreads();
        const int j = kk + 1 + lane;  // one lane per remaining column
        if (j < nc) {
            double2 *Aj = A + kk + ns * j;
            double2 v = Aj[0];
            for (int q = 1; q < len; ++q) {
                const double2 c = cconjmul(x[q], Aj[q]);
                v.x += c.x; v.y += c.y;
            }
            v = cconjmul(t, v);
            Aj[0].x -= v.x; Aj[0].y -= v.y;
            for (int q = 1; q < len; ++q) {
                const double2 c = cmul(x[q], v);
                Aj[q].x -= c.x; Aj[q].y -= c.y;
            }
        }
        __syncthreads();
    }
}

// b (ns) <- Q^H b, then R w = b[0:nc] (w overwrites b[0:nc]); one lane.
__device__ inline void lane_qr_solve(const double2 *A, int ns, int nc, const double2 *tau, double2 *b) {
    for (int kk = 0; kk < nc; ++kk) {
        const double2 *x = A + kk + ns * kk;
        double2 v = b[kk];
        for (int q = kk + 1; q < ns; ++q) {
            const double2 c = cconjmul(x[q - kk], b[q]);
            v.x += c.x; v.y += c.y;
        }
        v = cconjmul(tau[kk], v);
        b[kk].x -= v.x; b[kk].y -= v.y;
        for (int q = kk + 1; q < ns; ++q) {
            const double2 c = cmul(x[q - kk], v);
            b[q].x -= c.x; b[q].y -= c.y;
        }
    }
    for (int jj = nc - 1; jj >= 0; --jj) {
        const double2 xj = cdiv(b[jj], A[jj + ns * jj]);
        b[jj] = xj;
        for (int q = jj - 1; q >= 0; --q) {
            const double2 c = cmul(A[q + ns * jj], xj);
            b[q].x -= c.x; b[q].y -= c.y;
        }
    }
}

// Minimum-norm solution of M w = b for a wide M (nc x ns, nc <= ns) whose
// conjugate transpose was factorised, M^H = Q R, by wave_householder_qr (A is
// ns x nc): R^H z = b (forward substitution), then w = Q [z; 0]. b holds nc
// entries on entry and the ns entries of w on exit; one lane.
__device__ inline void lane_qr_minnorm(const double2 *A, int ns, int nc, const double2 *tau, double2 *b) {
    for (int i = 0; i < nc; ++i) {
        double2 v = b[i];
        for (int j = 0; j < i; ++j) {
            const double2 c = cconjmul(A[j + ns * i], b[j]);
            v.x -= c.x; v.y -= c.y;
        }
        const double2 rii = A[i + ns * i];
        b[i] = cdiv(v, make_double2(rii.x, -rii.y));
    }
    for (int q = nc; q < ns; ++q) b[q] = make_double2(0.0, 0.0);
    for (int kk = nc - 1; kk >= 0; --kk) {
        const double2 *x = A + kk + ns * kk;
        double2 v = b[kk];
        for (int q = kk + 1; q < ns; ++q) {
            const double2 c = cconjmul(x[q - kk], b[q]);
            v.x += c.x; v.y += c.y;
        }
        v = cmul(tau[kk], v);
        b[kk].x -= v.x; b[kk].y -= v.y;
        for (int q = kk + 1; q < ns; ++q) {
            const double2 c = cmul(x[q - kk], v);
            b[q].x -= c.x; b[q].y -= c.y;
        }
    }
}

// ratio of the extreme |R_ii| (a lower bound of the condition number)
__device__ inline double rdiag_ratio(const double2 *A, int ns, int nc) {
    double rmax = 0.0, rmin = INFINITY;
    for (int q = 0; q < nc; ++q) {
        const double2 v = A[q + ns * q];
        const double av = hypot(v.x, v.y);
        rmax = fmax(rmax, av);
        rmin = fmin(rmin, av);
    }
    return rmax / rmin;
}

// CSR pattern of a triplet list: rows sorted, columns sorted within a row,
// duplicates merged (the summing semantics of sparse(I, J, V)).
struct TripletPattern {
    std::vector<int32_t> rowptr, colind, seg, perm;
};
void build_pattern(int64_t m, const std::vector<int32_t> &tri_row, const std::vector<int32_t> &tri_col,
                   TripletPattern &P);

// val[q] = sum of the triplets gathered to CSR position q
struct Gather1Args {
    int64_t nnz;
    const int32_t *seg, *perm;
    const double2 *tri;
    double *val;
    int cplx;
    double sign;
};
void launch_gather1(const Gather1Args &g, hipStream_t s);

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

// ----------------------------------------------- DIM pieces (bdim.hip)
// Monopole centres of bdim.jl:79-92 around the box [lo, hi] (dim x ns,
// row-major by coordinate).
void monopoles_box(int dim, const double *lo, const double *hi, int ns, double radius_mult,
                   std::vector<double> &zs);

// Per-element Householder QR of M^T and the least-squares solve of each of
// the element's targets (k_ls, one wavefront per element).
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
void launch_dim_ls(const LsArgs &a, int64_t nel, hipStream_t s);

// Traces of every problem of a batch in one launch: node j of problem
// row2p[j] against that problem's monopoles zs + p * dim * ns with its
// kernel parameters kp[p]: B0 = G(y_j, z_k), B1 = dG/dn_j (y_j, z_k), and
// Theta's first term mu_j B0 (the targets are the sources). ntot x ns,
// column-major, complex.
struct TraceBatchArgs {
    int64_t ntot;
    int ns;
    const double *sx, *sn, *zs;
    const int32_t *row2p;
    const KParams *kp;
    const double *mu;
    double2 *B0, *B1, *theta;
};
// false when there is no kernel for (pde, dim)
bool launch_dim_traces_batch(int pde, int dim, const TraceBatchArgs &a, hipStream_t s);

// Theta += S B1 - D B0 for every problem of a batch (the S and D sums of
// each problem over its own nodes): the per-node records are packed in one
// launch, then one block per work item (problem, tile of 64 targets) sums
// over all the problem's sources, its wavefronts taking shares of them.
struct ThetaBatchArgs {
    int64_t ntot;
    int ns;
    const int32_t *items;  // nitems x (problem, tile), by decreasing problem size
    const int64_t *off;    // problem p's nodes are rows [off[p], off[p + 1])
    const int32_t *row2p;
    const KParams *kp;     // per problem: the kernel parameters of S and D
    const double *post;    // per problem: post factors (re, im) of S, then of D
    const double *coords, *normals, *weights;
    const double2 *B0, *B1;  // ntot x ns traces, column-major
    double *rec;             // ntot x record length, written by the pack
    double2 *theta;          // ntot x ns, column-major
};
// record length in doubles per node (to size `rec`), or 0 when there is no
// kernel for (pde, dim, ns)
int dim_theta_batch_record(int pde, int dim, int ns);
bool launch_dim_theta_batch(int pde, int dim, const ThetaBatchArgs &a, int64_t nitems, hipStream_t s);

}  // namespace inti
