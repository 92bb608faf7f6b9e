/*
 * CPU FP64 restatement of Inti.jl's corrected Nystrom hot path.
 * TEST INFRASTRUCTURE ONLY: the CPU reference for the GPU library, never shipped
 * or called by it. Each function cites the reference lines it follows.
 */
#include "inti_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* src/utils.jl:235 */
static const double SAME_POINT_TOLERANCE = 1.0e-14;
static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* Hankel functions of the first kind (the reference calls Bessels.hankelh1,
 * src/kernels.jl:344). Here: C library j0/j1/jn/y0/y1/yn.                 */
void ork_hankel1(int n, double x, double out[2]) {
    double j, y;
    if (n == 0) { j = j0(x); y = y0(x); }
    else if (n == 1) { j = j1(x); y = y1(x); }
    else { j = jn(n, x); y = yn(n, x); }
    out[0] = j; out[1] = y;
}

static inline cplx hankel(int n, double x) {
    double h[2];
    ork_hankel1(n, x, h);
    return h[0] + I * h[1];
}

static inline double dotn(int dim, const double *a, const double *b) {
    double s = 0.0;
    for (int d = 0; d < dim; ++d) s += a[d] * b[d];
    return s;
}

/* One kernel, point coordinates packed (x[dim]). kernels.jl:137-204,347-413 */
static cplx kernel_pt(int pde, int dim, int kind, double k, const double *x,
                      const double *nx, const double *y, const double *ny) {
    double r[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) r[d] = x[d] - y[d];
    double d2 = dotn(dim, r, r);
    double d = sqrt(d2);
    if (d <= SAME_POINT_TOLERANCE) return 0.0;
    if (pde == ORK_LAPLACE) {
        if (kind == ORK_SL) {
            return dim == 2 ? -1.0 / (2 * PI) * log(d) : 1.0 / (4 * PI) / d;
        } else if (kind == ORK_DL) {
            double rdn = dotn(dim, r, ny);
            return dim == 2 ? 1.0 / (2 * PI) / (d * d) * rdn
                            : 1.0 / (4 * PI) / (d * d * d) * rdn;
        } else if (kind == ORK_ADL) {
            double rdn = dotn(dim, r, nx);
            return dim == 2 ? -1.0 / (2 * PI) / (d * d) * rdn
                            : -1.0 / (4 * PI) / (d * d * d) * rdn;
        } else { /* HS */
            double id2 = 1.0 / (d * d);
            double nxdny = dotn(dim, nx, ny);
            double rdnx = dotn(dim, r, nx), rdny = dotn(dim, r, ny);
            if (dim == 2) return id2 / (2 * PI) * (nxdny - 2 * rdnx * rdny * id2);
            return id2 / (4 * PI * d) * (nxdny - 3 * rdnx * rdny * id2);
        }
    }
    /* Helmholtz */
    if (dim == 2) {
        double kd = k * d;
        if (kind == ORK_SL) return I / 4.0 * hankel(0, kd);
        if (kind == ORK_DL) return I * k / (4 * d) * hankel(1, kd) * dotn(dim, r, ny);
        if (kind == ORK_ADL) return -I * k / (4 * d) * hankel(1, kd) * dotn(dim, r, nx);
        {
            double rdnx = dotn(dim, r, nx), rdny = dotn(dim, r, ny);
            double nxdny = dotn(dim, nx, ny);
            cplx h1 = hankel(1, kd), h2 = hankel(2, kd);
            cplx a = -I * k * k / (4 * d * d) * h2;
            cplx b = I * k / (4 * d) * h1;
            return a * rdnx * rdny + b * nxdny;
        }
    }
    {
        cplx eikd = cexp(I * k * d);
        if (kind == ORK_SL) return 1.0 / (4 * PI) / d * eikd;
        if (kind == ORK_DL)
            return 1.0 / (4 * PI * d * d) * eikd * (-I * k + 1.0 / d) * dotn(dim, r, ny);
        if (kind == ORK_ADL)
            return -1.0 / (4 * PI * d * d) * eikd * (-I * k + 1.0 / d) * dotn(dim, r, nx);
        {
            double rdnx = dotn(dim, r, nx), rdny = dotn(dim, r, ny);
            double nxdny = dotn(dim, nx, ny);
            double id2 = 1.0 / (d * d);
            cplx b = eikd * id2 / (4 * PI) * (-I * k + 1.0 / d);
            cplx a = eikd * id2 * id2 / (4 * PI * d) * (3.0 * (d * I * k - 1.0) + d * d * k * k);
            return a * rdnx * rdny + b * nxdny;
        }
    }
}

/* Helmholtz far-field pattern kernels (no reference counterpart): xh a unit
 * direction; P e^{-ik xh.y} times 1 (SL) or -ik xh.n_y (DL);
 * P = 1/(4 pi) (3D), e^{i pi/4}/sqrt(8 pi k) (2D). */
static cplx ff_prefactor(int dim, double k) {
    return dim == 3 ? 1.0 / (4 * PI) : cexp(I * PI / 4) / sqrt(8 * PI * k);
}

static cplx kernel_ff(int dim, int kind, double k, const double *xh, const double *y,
                      const double *ny) {
    cplx e = ff_prefactor(dim, k) * cexp(-I * k * dotn(dim, xh, y));
    if (kind == ORK_FF_SL) return e;
    return e * (-I * k * dotn(dim, xh, ny));
}

static cplx kernel_any(int pde, int dim, int kind, double k, const double coefs[4],
                       const double *x, const double *nx, const double *y,
                       const double *ny) {
    if (kind == ORK_FF_SL || kind == ORK_FF_DL) return kernel_ff(dim, kind, k, x, y, ny);
    if (kind == ORK_FF_COMB) {
        cplx c0 = coefs[0] + I * coefs[1], c1 = coefs[2] + I * coefs[3];
        return c0 * kernel_ff(dim, ORK_FF_SL, k, x, y, ny) + c1 * kernel_ff(dim, ORK_FF_DL, k, x, y, ny);
    }
    if (kind == ORK_COMB_SD || kind == ORK_COMB_AH) {
        int ka = kind == ORK_COMB_SD ? ORK_SL : ORK_ADL;
        int kb = kind == ORK_COMB_SD ? ORK_DL : ORK_HS;
        cplx c0 = coefs[0] + I * coefs[1], c1 = coefs[2] + I * coefs[3];
        return c0 * kernel_pt(pde, dim, ka, k, x, nx, y, ny) +
               c1 * kernel_pt(pde, dim, kb, k, x, nx, y, ny);
    }
    return kernel_pt(pde, dim, kind, k, x, nx, y, ny);
}

static inline void gather(int dim, const double *soa, int64_t n, int64_t i, double *p) {
    for (int d = 0; d < dim; ++d) p[d] = soa ? soa[d * n + i] : 0.0;
}

void ork_kernel(int pde, int dim, int kind, double k, const double coefs[4],
                const double *x, const double *nx, const double *y,
                const double *ny, double out[2]) {
    static const double zero3[3] = {0, 0, 0};
    cplx v = kernel_any(pde, dim, kind, k, coefs, x, nx ? nx : zero3, y, ny ? ny : zero3);
    out[0] = creal(v); out[1] = cimag(v);
}

/* nystrom.jl:117-124: column-threaded dense fill, out[i,j] = K(X[i],Y[j]) w_j */
void ork_assemble(int pde, int dim, int kind, double k, const double coefs[4],
                  int64_t m, const double *tx, const double *tn,
                  int64_t n, const double *sx, const double *sn, const double *sw,
                  double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        double y[3], ny[3], x[3], nx[3];
        gather(dim, sx, n, j, y);
        gather(dim, sn, n, j, ny);
        for (int64_t i = 0; i < m; ++i) {
            gather(dim, tx, m, i, x);
            gather(dim, tn, m, i, nx);
            cplx v = kernel_any(pde, dim, kind, k, coefs, x, nx, y, ny) * sw[j];
            out[2 * (i + m * j)] = creal(v);
            out[2 * (i + m * j) + 1] = cimag(v);
        }
    }
}

/* Matrix-free counterpart of S*x (nystrom.jl:93-96 entries), rows threaded. */
void ork_matvec(int pde, int dim, int kind, double k, const double coefs[4],
                int64_t m, const double *tx, const double *tn,
                int64_t n, const double *sx, const double *sn, const double *sw,
                int nrhs, const double *xin, double *yout,
                int64_t row_lo, int64_t row_hi, int nthreads) {
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    int nt = 1; (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = row_lo; i < row_hi; ++i) {
        double x[3], nx[3], y[3], ny[3];
        gather(dim, tx, m, i, x);
        gather(dim, tn, m, i, nx);
        cplx acc[64];
        int r0 = 0;
        while (r0 < nrhs) {
            int nb = nrhs - r0 < 64 ? nrhs - r0 : 64;
            for (int r = 0; r < nb; ++r) acc[r] = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                gather(dim, sx, n, j, y);
                gather(dim, sn, n, j, ny);
                cplx kv = kernel_any(pde, dim, kind, k, coefs, x, nx, y, ny) * sw[j];
                for (int r = 0; r < nb; ++r) {
                    const double *xv = xin + 2 * (j + n * (int64_t)(r0 + r));
                    acc[r] += kv * (xv[0] + I * xv[1]);
                }
            }
            for (int r = 0; r < nb; ++r) {
                double *yv = yout + 2 * (i + m * (int64_t)(r0 + r));
                yv[0] = creal(acc[r]); yv[1] = cimag(acc[r]);
            }
            r0 += nb;
        }
    }
}

/* utils.jl:142-149 and utils.jl:158-171; centre and radius per bdim.jl:79-92 */
int ork_dim_monopoles(int dim, int64_t n, const double *sx, int nq,
                      double oversample, double radius_mult, double *zs) {
    int ns = (int)ceil(oversample * nq);
    double lo[3], hi[3];
    for (int d = 0; d < dim; ++d) {
        lo[d] = INFINITY; hi[d] = -INFINITY;
        for (int64_t j = 0; j < n; ++j) {
            double v = sx[d * n + j];
            lo[d] = v < lo[d] ? v : lo[d];
            hi[d] = v > hi[d] ? v : hi[d];
        }
    }
    double xc[3], diag2 = 0.0;
    for (int d = 0; d < dim; ++d) {
        xc[d] = (lo[d] + hi[d]) / 2;
        diag2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    }
    double R = radius_mult * sqrt(diag2) / 2;
    if (dim == 2) {
        for (int i = 0; i < ns; ++i) {
            double th = 2 * PI * i / ns;
            zs[0 * ns + i] = R * cos(th) + xc[0];
            zs[1 * ns + i] = R * sin(th) + xc[1];
        }
    } else {
        double phi = PI * (3 - sqrt(5.0));
        for (int i = 1; i <= ns; ++i) {
            double ytmp = 1 - ((double)(i - 1) / (ns - 1)) * 2;
            double radius = sqrt(1 - ytmp * ytmp);
            double theta = phi * i;
            zs[0 * ns + i - 1] = cos(theta) * radius * R + xc[0];
            zs[1 * ns + i - 1] = ytmp * R + xc[1];
            zs[2 * ns + i - 1] = sin(theta) * radius * R + xc[2];
        }
    }
    return ns;
}

/* Householder reflector, the generic unblocked QR path taken by
 * qr!(transpose(M)) in bdim.jl:163: nu = copysign(|x|, re x1),
 * x1 += nu, x[1] = -nu, x[2:] /= x1, tau = x1/nu.                       */
static cplx reflector(cplx *x, int len) {
    double s = 0.0;
    for (int i = 0; i < len; ++i) s += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    double normu = sqrt(s);
    if (normu == 0.0) return 0.0;
    cplx xi1 = x[0];
    double nu = copysign(normu, creal(xi1));
    xi1 += nu;
    x[0] = -nu;
    for (int i = 1; i < len; ++i) x[i] /= xi1;
    return xi1 / nu;
}

/* A is rows x cols, column-major (lda = rows). In-place QR, tau[cols]. */
static void qr_unblocked(cplx *A, int rows, int cols, cplx *tau) {
    for (int kk = 0; kk < cols && kk < rows; ++kk) {
        cplx *x = A + kk + (int64_t)rows * kk;
        int len = rows - kk;
        cplx t = reflector(x, len);
        tau[kk] = t;
        for (int j = kk + 1; j < cols; ++j) {
            cplx *Aj = A + kk + (int64_t)rows * j;
            cplx v = Aj[0];
            for (int i = 1; i < len; ++i) v += conj(x[i]) * Aj[i];
            v = conj(t) * v;
            Aj[0] -= v;
            for (int i = 1; i < len; ++i) Aj[i] -= x[i] * v;
        }
    }
}

/* Least-squares solve with the packed QR: b <- Q^H b, then R x = b[0:cols]
 * by column-oriented back substitution (the generic QR ldiv! path). */
static void qr_solve(const cplx *A, int rows, int cols, const cplx *tau, cplx *b) {
    for (int kk = 0; kk < cols && kk < rows; ++kk) {
        const cplx *x = A + kk + (int64_t)rows * kk;
        cplx v = b[kk];
        for (int i = kk + 1; i < rows; ++i) v += conj(x[i - kk]) * b[i];
        v = conj(tau[kk]) * v;
        b[kk] -= v;
        for (int i = kk + 1; i < rows; ++i) b[i] -= x[i - kk] * v;
    }
    for (int j = cols - 1; j >= 0; --j) {
        cplx xj = b[j] / A[j + (int64_t)rows * j];
        b[j] = xj;
        for (int i = j - 1; i >= 0; --i) b[i] -= A[i + (int64_t)rows * j] * xj;
    }
}

typedef struct { int64_t row, col; cplx s, d; } triplet;

static int cmp_triplet(const void *a, const void *b) {
    const triplet *p = (const triplet *)a, *q = (const triplet *)b;
    if (p->row != q->row) return p->row < q->row ? -1 : 1;
    if (p->col != q->col) return p->col < q->col ? -1 : 1;
    return 0;
}

/* bdim.jl:50-196 over element types: the monopole count uses the sum of nq
 * over the types (bdim.jl:79-80), the least squares run type by type
 * (bdim.jl:138-185). Type t's tags are the nq[t] x ne[t] block of qtags at
 * sum_{u<t} nq[u] ne[u]; its near lists the ne[t] + 1 entries of near_ptr at
 * sum_{u<t} (ne[u] + 1), indexing the block of near_idx after the earlier
 * types' lists. */
int64_t ork_bdim_types(int pde, int dim, double k, int derivative,
                       int64_t m, const double *tx, const double *tn,
                       int64_t n, const double *sx, const double *sn, const double *sw,
                       int ntypes, const int *nqs, const int64_t *nes, const int64_t *qtags_all,
                       const int64_t *near_ptr_all, const int64_t *near_idx_all,
                       const double *mu, double oversample, double radius_mult,
                       int64_t *rowptr, int64_t *colind, double *valS, double *valD,
                       double *diag) {
    const int self = near_ptr_all == NULL;
    if (self && m != n) return -1;
    if (ntypes < 1 || ntypes > 16) return -1;
    int qmax = 0, nqmax = 0;
    for (int t = 0; t < ntypes; ++t) {
        qmax += nqs[t];
        nqmax = nqs[t] > nqmax ? nqs[t] : nqmax;
    }
    int ns = (int)ceil(oversample * qmax);
    if (ns > 64 || ns < 2 * nqmax) return -1;
    /* per-type offsets into qtags, near_ptr, near_idx */
    int64_t tag_off[17], ptr_off[17], idx_off[17];
    tag_off[0] = ptr_off[0] = idx_off[0] = 0;
    for (int t = 0; t < ntypes; ++t) {
        tag_off[t + 1] = tag_off[t] + (int64_t)nqs[t] * nes[t];
        ptr_off[t + 1] = ptr_off[t] + nes[t] + 1;
        idx_off[t + 1] = idx_off[t] + (self ? 0 : near_ptr_all[ptr_off[t] + nes[t]]);
    }
    double *zs = (double *)malloc(sizeof(double) * dim * ns);
    ork_dim_monopoles(dim, n, sx, qmax, oversample, radius_mult, zs);
    const int kG = derivative ? ORK_ADL : ORK_SL; /* bdim.jl:112 */
    const int kS = derivative ? ORK_ADL : ORK_SL;
    const int kD = derivative ? ORK_HS : ORK_DL;

    /* traces (bdim.jl:94-103): B0[j,k] = G(y_j, z_k), B1[j,k] = ADL(y_j, z_k) */
    cplx *B0 = (cplx *)malloc(sizeof(cplx) * n * ns);
    cplx *B1 = (cplx *)malloc(sizeof(cplx) * n * ns);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        double y[3], ny[3], z[3];
        gather(dim, sx, n, j, y);
        gather(dim, sn, n, j, ny);
        for (int kk = 0; kk < ns; ++kk) {
            gather(dim, zs, ns, kk, z);
            B0[j + n * kk] = kernel_pt(pde, dim, ORK_SL, k, y, ny, z, NULL);
            B1[j + n * kk] = kernel_pt(pde, dim, ORK_ADL, k, y, ny, z, NULL);
        }
    }

    /* targets needing a correction, in type and element order */
    int64_t nlisted = self ? tag_off[ntypes] : idx_off[ntypes];
    /* local row of each global target (-1 = not corrected) */
    int64_t *loc = (int64_t *)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
    for (int64_t i = 0; i < m; ++i) loc[i] = -1;
    int64_t *glob = (int64_t *)malloc(sizeof(int64_t) * (nlisted > 0 ? nlisted : 1));
    int64_t nloc = 0;
    for (int ty = 0; ty < ntypes; ++ty) {
        const int nq = nqs[ty];
        const int64_t ne = nes[ty], *qtags = qtags_all + tag_off[ty];
        const int64_t *near_ptr = self ? NULL : near_ptr_all + ptr_off[ty];
        const int64_t *near_idx = self ? NULL : near_idx_all + idx_off[ty];
        for (int64_t e = 0; e < ne; ++e) {
            int64_t cnt = self ? nq : near_ptr[e + 1] - near_ptr[e];
            for (int64_t t = 0; t < cnt; ++t) {
                int64_t i = self ? qtags[t + (int64_t)nq * e] : near_idx[near_ptr[e] + t];
                if (loc[i] < 0) { loc[i] = nloc; glob[nloc++] = i; }
            }
        }
    }

    /* Theta (bdim.jl:106-118) for the listed targets:
     * Theta = mu*G(x,z) + S*B1 - D*B0 */
    cplx *Th = (cplx *)malloc(sizeof(cplx) * (nloc > 0 ? nloc : 1) * ns);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t l = 0; l < nloc; ++l) {
        int64_t i = glob[l];
        double x[3], nx[3], y[3], ny[3], z[3];
        gather(dim, tx, m, i, x);
        gather(dim, tn, m, i, nx);
        cplx acc[64];
        for (int kk = 0; kk < ns; ++kk) {
            gather(dim, zs, ns, kk, z);
            acc[kk] = mu[i] * kernel_pt(pde, dim, kG, k, x, nx, z, NULL);
        }
        cplx sacc[64], dacc[64];
        for (int kk = 0; kk < ns; ++kk) { sacc[kk] = 0; dacc[kk] = 0; }
        for (int64_t jj = 0; jj < n; ++jj) {
#ifdef ORK_REVERSE_SUMS
            const int64_t j = n - 1 - jj; /* the same sums, sources in reverse order */
#else
            const int64_t j = jj;
#endif
            gather(dim, sx, n, j, y);
            gather(dim, sn, n, j, ny);
            cplx Sij = kernel_pt(pde, dim, kS, k, x, nx, y, ny) * sw[j];
            cplx Dij = kernel_pt(pde, dim, kD, k, x, nx, y, ny) * sw[j];
            for (int kk = 0; kk < ns; ++kk) {
                sacc[kk] += Sij * B1[j + n * kk];
                dacc[kk] += Dij * B0[j + n * kk];
            }
        }
        for (int kk = 0; kk < ns; ++kk) Th[l + nloc * kk] = acc[kk] + sacc[kk] - dacc[kk];
    }

    /* per-element least squares (bdim.jl:137-185), type by type; the
     * triplets of type t start at slot tri_off[t] */
    int64_t tri_off[17];
    tri_off[0] = 0;
    for (int ty = 0; ty < ntypes; ++ty) {
        int64_t c = 0;
        for (int64_t e = 0; e < nes[ty]; ++e)
            c += (int64_t)nqs[ty] *
                 (self ? nqs[ty] : near_ptr_all[ptr_off[ty] + e + 1] - near_ptr_all[ptr_off[ty] + e]);
        tri_off[ty + 1] = tri_off[ty] + c;
    }
    int64_t nnz = tri_off[ntypes];
    triplet *trip = (triplet *)malloc(sizeof(triplet) * (nnz > 0 ? nnz : 1));
    double dg[4] = {0, 0, 0, 0};
    for (int ty = 0; ty < ntypes; ++ty) {
    const int nq = nqs[ty];
    const int64_t ne = nes[ty], *qtags = qtags_all + tag_off[ty];
    const int64_t *near_ptr = self ? NULL : near_ptr_all + ptr_off[ty];
    const int64_t *near_idx = self ? NULL : near_idx_all + idx_off[ty];
    triplet *trip_t = trip + tri_off[ty];
    const int nc = 2 * nq;
#pragma omp parallel
    {
        cplx *A = (cplx *)malloc(sizeof(cplx) * ns * nc);
        cplx *tau = (cplx *)malloc(sizeof(cplx) * nc);
        cplx *b = (cplx *)malloc(sizeof(cplx) * ns);
        double ldg[4] = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 8)
        for (int64_t e = 0; e < ne; ++e) {
            int64_t cnt = self ? nq : near_ptr[e + 1] - near_ptr[e];
            if (cnt == 0) continue;
            const int64_t *jg = qtags + (int64_t)nq * e;
            /* A = transpose(M), M = [B0[jg,:]; B1[jg,:]] (2nq x ns) */
            for (int kk = 0; kk < ns; ++kk)
                for (int r = 0; r < nq; ++r) {
                    A[kk + ns * r] = B0[jg[r] + n * kk];
                    A[kk + ns * (nq + r)] = B1[jg[r] + n * kk];
                }
            qr_unblocked(A, ns, nc, tau);
            double rmax = 0, rmin = INFINITY;
            for (int r = 0; r < nc; ++r) {
                double a = cabs(A[r + ns * r]);
                rmax = a > rmax ? a : rmax;
                rmin = a < rmin ? a : rmin;
            }
            if (rmax / rmin > ldg[0]) ldg[0] = rmax / rmin;
            for (int64_t t = 0; t < cnt; ++t) {
                int64_t i = self ? jg[t] : near_idx[near_ptr[e] + t];
                int64_t l = loc[i];
                double bn = 0;
                for (int kk = 0; kk < ns; ++kk) {
                    b[kk] = Th[l + nloc * kk];
                    bn += creal(b[kk]) * creal(b[kk]) + cimag(b[kk]) * cimag(b[kk]);
                }
                if (sqrt(bn) > ldg[3]) ldg[3] = sqrt(bn);
                qr_solve(A, ns, nc, tau, b);
                double wn = 0;
                for (int r = 0; r < nc; ++r) wn += creal(b[r]) * creal(b[r]) + cimag(b[r]) * cimag(b[r]);
                if (sqrt(wn) > ldg[2]) ldg[2] = sqrt(wn);
                int64_t slot0 = 0;
                if (self) slot0 = ((int64_t)nq * e + t) * nq;
                else slot0 = (near_ptr[e] + t) * nq;
                for (int r = 0; r < nq; ++r) {
                    triplet *tp = trip_t + slot0 + r;
                    tp->row = i;
                    tp->col = jg[r];
                    tp->s = -b[nq + r]; /* bdim.jl:179 */
                    tp->d = b[r];       /* bdim.jl:180 */
                }
            }
        }
#pragma omp critical
        {
            for (int q = 0; q < 4; ++q) dg[q] = ldg[q] > dg[q] ? ldg[q] : dg[q];
        }
        free(A); free(tau); free(b);
    }
    }
    /* sparse(Is, Js, V, m, n) semantics, emitted as CSR (sorted, summed) */
    qsort(trip, nnz, sizeof(triplet), cmp_triplet);
    int64_t out = 0;
    for (int64_t i = 0; i <= m; ++i) rowptr[i] = 0;
    for (int64_t p = 0; p < nnz; ++p) {
        if (out > 0 && trip[p].row == trip[out - 1].row && trip[p].col == trip[out - 1].col) {
            trip[out - 1].s += trip[p].s;
            trip[out - 1].d += trip[p].d;
        } else {
            trip[out++] = trip[p];
        }
    }
    for (int64_t p = 0; p < out; ++p) {
        rowptr[trip[p].row + 1]++;
        colind[p] = trip[p].col;
        valS[2 * p] = creal(trip[p].s); valS[2 * p + 1] = cimag(trip[p].s);
        valD[2 * p] = creal(trip[p].d); valD[2 * p + 1] = cimag(trip[p].d);
    }
    for (int64_t i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    if (diag) for (int q = 0; q < 4; ++q) diag[q] = dg[q];
    free(trip); free(Th); free(glob); free(loc); free(B0); free(B1); free(zs);
    return out;
}

/* bdim.jl:50-196 with one element type */
int64_t ork_bdim(int pde, int dim, double k, int derivative,
                 int64_t m, const double *tx, const double *tn,
                 int64_t n, const double *sx, const double *sn, const double *sw,
                 int nq, int64_t ne, const int64_t *qtags,
                 const int64_t *near_ptr, const int64_t *near_idx,
                 const double *mu, double oversample, double radius_mult,
                 int64_t *rowptr, int64_t *colind, double *valS, double *valD,
                 double *diag) {
    return ork_bdim_types(pde, dim, k, derivative, m, tx, tn, n, sx, sn, sw, 1, &nq, &ne, qtags,
                          near_ptr, near_idx, mu, oversample, radius_mult, rowptr, colind, valS, valD,
                          diag);
}

/* quadrature.jl:264-286 with a brute-force nearest neighbour */
int64_t ork_nearest_lists(int dim, int64_t m, const double *tx, int64_t n,
                          const double *sx, int nq, int64_t ne,
                          const int64_t *qtags, double maxdist,
                          int64_t *near_ptr, int64_t *near_idx) {
    int64_t *owner = (int64_t *)malloc(sizeof(int64_t) * n);
    for (int64_t e = 0; e < ne; ++e)
        for (int q = 0; q < nq; ++q) owner[qtags[q + (int64_t)nq * e]] = e;
    int64_t *nearest = (int64_t *)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        double best = INFINITY;
        int64_t bj = -1;
        for (int64_t j = 0; j < n; ++j) {
            double s = 0;
            for (int d = 0; d < dim; ++d) {
                double r = tx[d * m + i] - sx[d * n + j];
                s += r * r;
            }
            if (s < best) { best = s; bj = j; }
        }
        nearest[i] = sqrt(best) > maxdist ? -1 : bj;
    }
    /* per element: node order within the element, then target order */
    int64_t cnt = 0;
    near_ptr[0] = 0;
    for (int64_t e = 0; e < ne; ++e) {
        for (int q = 0; q < nq; ++q) {
            int64_t node = qtags[q + (int64_t)nq * e];
            for (int64_t i = 0; i < m; ++i)
                if (nearest[i] == node) near_idx[cnt++] = i;
        }
        near_ptr[e + 1] = cnt;
    }
    free(owner); free(nearest);
    return cnt;
}

void ork_csr_matvec(int64_t m, int64_t n, const int64_t *rowptr,
                    const int64_t *colind, const double *val, int nrhs,
                    const double *x, const double alpha[2], const double beta[2],
                    double *y) {
    cplx a = alpha[0] + I * alpha[1], bt = beta[0] + I * beta[1];
    for (int r = 0; r < nrhs; ++r) {
        for (int64_t i = 0; i < m; ++i) {
            cplx s = 0;
            for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                const double *xv = x + 2 * (colind[p] + n * (int64_t)r);
                s += (val[2 * p] + I * val[2 * p + 1]) * (xv[0] + I * xv[1]);
            }
            double *yv = y + 2 * (i + m * (int64_t)r);
            cplx yo = yv[0] + I * yv[1];
            cplx res = a * s + (bt == 0 ? 0 : bt * yo);
            yv[0] = creal(res); yv[1] = cimag(res);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Reverse mode of the direct sum (no reference counterpart). Derivatives of
 * one kernel written from its closed form: value K, dK/dy (gy), dK/dn_y
 * (gn), dK/dk (gk). dcoefs: dc0/dk, dc1/dk of the combined kinds. */
static void kernel_grad(int pde, int dim, int kind, double k, const double coefs[4],
                        const double dcoefs[4], const double *x, const double *y,
                        const double *ny, cplx *K, cplx gy[3], cplx gn[3], cplx *gk) {
    for (int d = 0; d < 3; ++d) gy[d] = gn[d] = 0.0;
    *K = *gk = 0.0;
    if (kind == ORK_COMB_SD || kind == ORK_FF_COMB) {
        int ka = kind == ORK_COMB_SD ? ORK_SL : ORK_FF_SL;
        int kb = kind == ORK_COMB_SD ? ORK_DL : ORK_FF_DL;
        cplx c0 = coefs[0] + I * coefs[1], c1 = coefs[2] + I * coefs[3];
        cplx d0 = dcoefs ? dcoefs[0] + I * dcoefs[1] : 0.0;
        cplx d1 = dcoefs ? dcoefs[2] + I * dcoefs[3] : 0.0;
        cplx Ka, Kb, gya[3], gyb[3], gna[3], gnb[3], gka, gkb;
        kernel_grad(pde, dim, ka, k, coefs, dcoefs, x, y, ny, &Ka, gya, gna, &gka);
        kernel_grad(pde, dim, kb, k, coefs, dcoefs, x, y, ny, &Kb, gyb, gnb, &gkb);
        *K = c0 * Ka + c1 * Kb;
        *gk = d0 * Ka + c0 * gka + d1 * Kb + c1 * gkb;
        for (int d = 0; d < dim; ++d) {
            gy[d] = c0 * gya[d] + c1 * gyb[d];
            gn[d] = c0 * gna[d] + c1 * gnb[d];
        }
        return;
    }
    if (kind == ORK_FF_SL || kind == ORK_FF_DL) {
        cplx P = ff_prefactor(dim, k);
        cplx Pk = dim == 3 ? 0.0 : -P / (2 * k);
        double xy = dotn(dim, x, y);
        cplx e = cexp(-I * k * xy);
        if (kind == ORK_FF_SL) {
            *K = P * e;
            *gk = Pk * e + P * (-I * xy) * e;
            for (int d = 0; d < dim; ++d) gy[d] = -I * k * x[d] * (*K);
        } else {
            double xn = dotn(dim, x, ny);
            *K = P * (-I * k * xn) * e;
            *gk = Pk * (-I * k * xn) * e + P * (-I * xn) * e + (*K) * (-I * xy);
            for (int d = 0; d < dim; ++d) {
                gy[d] = -I * k * x[d] * (*K);
                gn[d] = P * (-I * k * x[d]) * e;
            }
        }
        return;
    }
    double r[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) r[d] = x[d] - y[d];
    double d2 = dotn(dim, r, r);
    if (sqrt(d2) <= SAME_POINT_TOLERANCE) return;
    double dd = sqrt(d2);
    double rn = kind == ORK_DL ? dotn(dim, r, ny) : 0.0;
    /* K = F(d) (SL) or F(d) (r.n) (DL); dK/dy = -F'(d) r/d [* r.n] - F n (DL) */
    cplx F, Fp, Fk;
    if (pde == ORK_LAPLACE) {
        Fk = 0.0;
        if (dim == 2) {
            if (kind == ORK_SL) { F = -log(dd) / (2 * PI); Fp = -1.0 / (2 * PI * dd); }
            else { F = 1.0 / (2 * PI * d2); Fp = -1.0 / (PI * d2 * dd); }
        } else {
            if (kind == ORK_SL) { F = 1.0 / (4 * PI * dd); Fp = -1.0 / (4 * PI * d2); }
            else { F = 1.0 / (4 * PI * d2 * dd); Fp = -3.0 / (4 * PI * d2 * d2); }
        }
    } else if (dim == 3) {
        cplx e = cexp(I * k * dd) / (4 * PI);
        if (kind == ORK_SL) {
            F = e / dd;
            Fp = e * (I * k * dd - 1.0) / d2;
            Fk = I * e;
        } else {
            F = e * (1.0 - I * k * dd) / (d2 * dd);
            Fp = e * (k * k * d2 - 3.0 + 3.0 * I * k * dd) / (d2 * d2);
            Fk = e * k / dd;
        }
    } else {
        cplx H0 = hankel(0, k * dd), H1 = hankel(1, k * dd);
        if (kind == ORK_SL) {
            F = I / 4 * H0;
            Fp = -I * k / 4 * H1;
            Fk = -I / 4 * dd * H1;
        } else {
            F = I * k / 4 * H1 / dd;
            Fp = I * k / 4 * (k * H0 / dd - 2.0 * H1 / d2);
            Fk = I * k / 4 * H0;
        }
    }
    if (kind == ORK_SL) {
        *K = F;
        *gk = Fk;
        for (int d = 0; d < dim; ++d) gy[d] = -Fp * r[d] / dd;
    } else {
        *K = F * rn;
        *gk = Fk * rn;
        for (int d = 0; d < dim; ++d) {
            gy[d] = -Fp * rn * r[d] / dd - F * ny[d];
            gn[d] = F * r[d];
        }
    }
}

int ork_adjoint(int pde, int dim, int kind, double k, const double coefs[4],
                const double dcoefs[4], int64_t m, const double *tx, int64_t n,
                const double *sx, const double *sn, const double *sw, const double *ubar,
                const double *sigma, double *sigma_bar, double *ybar, double *nbar,
                double *wbar, double *xbar, double *kbar) {
    if (!(kind == ORK_SL || kind == ORK_DL || kind == ORK_COMB_SD || kind == ORK_FF_SL ||
          kind == ORK_FF_DL || kind == ORK_FF_COMB))
        return -1;
    double ksum = 0.0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : ksum)
    for (int64_t j = 0; j < n; ++j) {
        double y[3] = {0, 0, 0}, ny[3] = {0, 0, 0}, x[3] = {0, 0, 0};
        gather(dim, sx, n, j, y);
        gather(dim, sn, n, j, ny);
        cplx sig = sigma[2 * j] + I * sigma[2 * j + 1];
        cplx s = sw[j] * sig;
        cplx SK = 0.0, SY[3] = {0, 0, 0}, SN[3] = {0, 0, 0}, SKk = 0.0;
        for (int64_t i = 0; i < m; ++i) {
            gather(dim, tx, m, i, x);
            cplx c = ubar[2 * i] - I * ubar[2 * i + 1];
            cplx K, gy[3], gn[3], gk;
            kernel_grad(pde, dim, kind, k, coefs, dcoefs, x, y, ny, &K, gy, gn, &gk);
            SK += c * K;
            SKk += c * gk;
            for (int d = 0; d < dim; ++d) { SY[d] += c * gy[d]; SN[d] += c * gn[d]; }
        }
        cplx sb = sw[j] * conj(SK);
        if (sigma_bar) { sigma_bar[2 * j] = creal(sb); sigma_bar[2 * j + 1] = cimag(sb); }
        if (wbar) wbar[j] = creal(sig * SK);
        for (int d = 0; d < dim; ++d) {
            if (ybar) ybar[d * n + j] = creal(s * SY[d]);
            if (nbar) nbar[d * n + j] = creal(s * SN[d]);
        }
        ksum += creal(s * SKk);
    }
    if (kbar) *kbar = ksum;
    if (xbar) {
#pragma omp parallel for schedule(dynamic, 8)
        for (int64_t i = 0; i < m; ++i) {
            double x[3] = {0, 0, 0}, y[3] = {0, 0, 0}, ny[3] = {0, 0, 0};
            gather(dim, tx, m, i, x);
            cplx c = ubar[2 * i] - I * ubar[2 * i + 1];
            cplx T[3] = {0, 0, 0};
            for (int64_t j = 0; j < n; ++j) {
                gather(dim, sx, n, j, y);
                gather(dim, sn, n, j, ny);
                cplx s = sw[j] * (sigma[2 * j] + I * sigma[2 * j + 1]);
                cplx K, gy[3], gn[3], gk;
                kernel_grad(pde, dim, kind, k, coefs, dcoefs, x, y, ny, &K, gy, gn, &gk);
                for (int d = 0; d < dim; ++d) T[d] -= s * gy[d];
            }
            for (int d = 0; d < dim; ++d) xbar[d * m + i] = creal(c * T[d]);
        }
    }
    return 0;
}
