/*
 * CPU restatement of Inti.jl's corrected Nystrom hot path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the CPU reference the GPU library
 * is compared against (tests, smoke check, the CPU-baseline leg of the bench).
 * The product never links, loads or calls it.
 *
 * Reference files restated (Inti.jl):
 *   src/kernels.jl:137-204    Laplace SL/DL/ADL/HS, 2D and 3D
 *   src/kernels.jl:347-413    Helmholtz SL/DL/ADL/HS, 2D and 3D
 *   src/utils.jl:235          SAME_POINT_TOLERANCE = 1e-14
 *   src/nystrom.jl:93-96      IntegralOperator entry K(x_i, y_j) * w_j
 *   src/nystrom.jl:117-124    _assemble_matrix! (dense, column-major)
 *   src/nystrom.jl:196-207    _green_multiplier(:inside/:outside/:on)
 *   src/bdim.jl:7-10          DimParameters defaults
 *   src/bdim.jl:50-196        bdim_correction
 *   src/utils.jl:142-171      uniform_points_circle, fibonnaci_points_sphere
 *   src/quadrature.jl:249-286 etype_to_nearest_points
 *
 * Layouts: point sets are SoA, coords[d*n + i] for d < dim; normals alike;
 * complex values are interleaved (re, im) doubles; dense matrices are
 * column-major; all indices are 0-based.
 */
#ifndef INTI_ORACLE_H
#define INTI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORK_LAPLACE = 0, ORK_HELMHOLTZ = 1 };
/* kinds: 0 SL, 1 DL, 2 ADL, 3 HS,
 *        4 c0*SL + c1*DL   (combined field, e.g. DL - i*eta*SL),
 *        5 c0*ADL + c1*HS                                     */
enum { ORK_SL = 0, ORK_DL = 1, ORK_ADL = 2, ORK_HS = 3, ORK_COMB_SD = 4, ORK_COMB_AH = 5,
       ORK_FF_SL = 6, ORK_FF_DL = 7, ORK_FF_COMB = 8 };
/* 6-8: Helmholtz far-field patterns of SL, DL, c0*SL + c1*DL (targets are
 * unit directions; P e^{-ik xh.y} (1 or -ik xh.n_y), P = 1/(4 pi) in 3D,
 * e^{i pi/4}/sqrt(8 pi k) in 2D). */

/* One kernel value K(x, y); out[0..1] = (re, im). nx/ny may be NULL when
 * the kernel does not use them. */
void ork_kernel(int pde, int dim, int kind, double k, const double coefs[4],
                const double *x, const double *nx, const double *y,
                const double *ny, double out[2]);

/* Hankel functions of the first kind H_n(x), n = 0, 1, 2, real x > 0,
 * from the C library's Bessel functions j0/j1/jn/y0/y1/yn. */
void ork_hankel1(int n, double x, double out[2]);

/* Dense m x n matrix A[i + m*j] = K(x_i, y_j) * w_j (complex interleaved). */
void ork_assemble(int pde, int dim, int kind, double k, const double coefs[4],
                  int64_t m, const double *tx, const double *tn,
                  int64_t n, const double *sx, const double *sn, const double *sw,
                  double *out);

/* Matrix-free direct sum for rows [row_lo, row_hi):
 *   y[i + m*r] = sum_j K(x_i, y_j) w_j x[j + n*r]   (complex interleaved)
 * nthreads <= 0 uses every OpenMP thread. */
void ork_matvec(int pde, int dim, int kind, double k, const double coefs[4],
                int64_t m, const double *tx, const double *tn,
                int64_t n, const double *sx, const double *sn, const double *sw,
                int nrhs, const double *x, double *y,
                int64_t row_lo, int64_t row_hi, int nthreads);

/* Monopole centres of the density interpolation (bdim.jl:79-92).
 * Writes ns points SoA into zs[dim*ns]. Returns ns. */
int ork_dim_monopoles(int dim, int64_t n, const double *sx, int nq,
                      double oversample, double radius_mult, double *zs);

/* bdim_correction (bdim.jl:50-196) producing CSR (0-based, rows sorted,
 * columns sorted within a row, duplicates summed).
 *   derivative = 0: corrects (SL, DL); = 1: corrects (ADL, HS).
 *   qtags: nq x ne column-major node indices of each source element.
 *   near_ptr/near_idx: element -> target lists (CSR over elements); pass
 *     NULL for the self pattern (targets are the sources, each element
 *     corrects its own nodes).
 *   mu: Green multiplier per target (length m).
 *   Outputs: rowptr[m+1]; colind/valS/valD sized nnz = nq * sum |near list|;
 *   valS/valD complex interleaved.
 *   diag (optional, 4 doubles): max cond(M) estimate via R diagonal ratio,
 *     max LS residual, max |W|, max |Theta row|.
 * Returns nnz, or -1 on error. */
int64_t ork_bdim(int pde, int dim, double k, int derivative,
                 int64_t m, const double *tx, const double *tn,
                 int64_t n, const double *sx, const double *sn, const double *sw,
                 int nq, int64_t ne, const int64_t *qtags,
                 const int64_t *near_ptr, const int64_t *near_idx,
                 const double *mu, double oversample, double radius_mult,
                 int64_t *rowptr, int64_t *colind, double *valS, double *valD,
                 double *diag);

/* ork_bdim over ntypes element types (bdim.jl:79-80, 138-185): type t has
 * nqs[t] nodes per element and nes[t] elements; qtags_all holds the column-
 * major nq x ne blocks one after another; near_ptr_all holds nes[t] + 1
 * entries per type (each block starting at 0) and near_idx_all the target
 * blocks in the same order (NULL/NULL: self pattern). At most 16 types.
 * Outputs as ork_bdim (nnz = sum over types of nq * |near lists|). */
int64_t ork_bdim_types(int pde, int dim, double k, int derivative,
                       int64_t m, const double *tx, const double *tn,
                       int64_t n, const double *sx, const double *sn, const double *sw,
                       int ntypes, const int *nqs, const int64_t *nes, const int64_t *qtags_all,
                       const int64_t *near_ptr_all, const int64_t *near_idx_all,
                       const double *mu, double oversample, double radius_mult,
                       int64_t *rowptr, int64_t *colind, double *valS, double *valD,
                       double *diag);

/* Nearest-source-node assignment (quadrature.jl:264-286) by brute force:
 * each target goes to the element owning its nearest source node (lowest
 * node index on ties), dropped if the distance exceeds maxdist. Output is
 * element -> targets CSR (near_ptr[ne+1], near_idx[m]). Returns the number
 * of assigned targets. */
int64_t ork_nearest_lists(int dim, int64_t m, const double *tx, int64_t n,
                          const double *sx, int nq, int64_t ne,
                          const int64_t *qtags, double maxdist,
                          int64_t *near_ptr, int64_t *near_idx);

/* y = alpha * A_csr * x + beta * y, complex interleaved, nrhs columns. */
void ork_csr_matvec(int64_t m, int64_t n, const int64_t *rowptr,
                    const int64_t *colind, const double *val, int nrhs,
                    const double *x, const double alpha[2], const double beta[2],
                    double *y);

/* Reverse mode of y = K sigma for a cotangent ubar (all complex
 * interleaved; Laplace with zero imaginary parts): sigma_bar = A^H ubar,
 * ybar/nbar (dim x n SoA), wbar (n): gradients of Re<ubar, y> in the source
 * nodes, normals and weights; xbar (dim x m): in the target nodes; kbar: in
 * k. Kinds SL, DL, COMB_SD and FF_*; dcoefs (may be NULL) = dc0/dk, dc1/dk.
 * Any output may be NULL. Returns 0, or -1 for an unsupported kind. */
int ork_adjoint(int pde, int dim, int kind, double k, const double coefs[4],
                const double dcoefs[4], int64_t m, const double *tx, int64_t n,
                const double *sx, const double *sn, const double *sw, const double *ubar,
                const double *sigma, double *sigma_bar, double *ybar, double *nbar,
                double *wbar, double *xbar, double *kbar);

#ifdef __cplusplus
}
#endif
#endif
