/*
 * inti_amd: C-ABI of the MI355X (gfx950) backend for Inti.jl's corrected
 * Nystrom operators. Plain pointers and sizes only; every call returns an
 * int status (INTI_OK = 0) and records a message for inti_last_error().
 *
 * Conventions
 *   - Point sets are SoA: coords[d*n + i], d < dim; normals alike.
 *   - Helmholtz values and densities are complex, interleaved (re, im)
 *     doubles: the memory layout of Julia ComplexF64 and C double _Complex.
 *     Laplace values and densities are real doubles.
 *   - Matrices are column-major; indices are 0-based (the Julia shim adds 1).
 *   - Host pointers are never retained past a call; device memory handed out
 *     by inti_device_alloc is owned by the caller.
 *   - One context owns one HIP stream; calls on one context are serialised
 *     by a mutex, so any host thread may call in.
 *
 * Each entry point names the Inti.jl interface it replaces (file:line in the
 * Inti.jl sources).
 */
#ifndef INTI_AMD_H
#define INTI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct inti_ctx inti_ctx;
typedef struct inti_quad inti_quad;
typedef struct inti_op inti_op;
typedef struct inti_csr inti_csr;
typedef struct inti_batch inti_batch;

/* status codes */
#define INTI_OK 0
#define INTI_ERR_ARG 1         /* invalid argument (null pointer, bad id)   */
#define INTI_ERR_HIP 2         /* HIP runtime error                          */
#define INTI_ERR_UNSUPPORTED 3 /* kernel/option not available               */
#define INTI_ERR_NOMEM 4       /* device allocation failed                   */
#define INTI_ERR_SIZE 5        /* size mismatch between operands            */

/* PDEs (src/kernels.jl:124 Laplace, :322 Helmholtz) */
#define INTI_LAPLACE 0
#define INTI_HELMHOLTZ 1

/* kernels (src/kernels.jl:37-112): single layer, double layer, adjoint
 * double layer, hypersingular, and the combined fields
 * c0*SL + c1*DL and c0*ADL + c1*HS evaluated in one pass. */
#define INTI_SL 0
#define INTI_DL 1
#define INTI_ADL 2
#define INTI_HS 3
#define INTI_COMB_SD 4
#define INTI_COMB_AH 5
/* Helmholtz far-field patterns of the single- and double-layer potentials
 * and of c0*SL + c1*DL: targets are unit directions xh, K(xh, y) =
 * P e^{-i k xh.y} B with B = 1 (SL) or -i k xh.n_y (DL), P = 1/(4 pi) in
 * 3D and e^{i pi/4}/sqrt(8 pi k) in 2D (u_s ~ e^{ikr} r^{-(dim-1)/2} u_inf).
 * No counterpart in the reference; the objective of the shape-gradient
 * workload (squared far-field magnitude over a set of directions). */
#define INTI_FF_SL 6
#define INTI_FF_DL 7
#define INTI_FF_COMB 8

/* arithmetic of the direct sum */
#define INTI_FP64 0  /* FP64 throughout                                    */
#define INTI_MIXED 1 /* FP64 geometry/accumulation, FP32 transcendentals   */

/* ---------------------------------------------------------------- context */
/* Library version (major*10000 + minor*100 + patch). */
int inti_version(void);

/* Create a context on a device (module __init__ of the Julia shim). */
int inti_ctx_create(int device, inti_ctx **out);
int inti_ctx_destroy(inti_ctx *ctx);
/* Message of the last failed call on ctx ("" if none). ctx may be NULL for
 * failures of inti_ctx_create. */
const char *inti_last_error(const inti_ctx *ctx);
int inti_ctx_synchronize(inti_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that interleave their
 * own work. */
int inti_ctx_stream(inti_ctx *ctx, void **stream);
/* When on, the main kernel of every apply is bracketed by HIP events on the
 * context stream and inti_ctx_last_kernel_ms returns its duration. */
int inti_ctx_set_profiling(inti_ctx *ctx, int on);
int inti_ctx_last_kernel_ms(inti_ctx *ctx, double *ms);

/* Device buffers for callers keeping operands resident in HBM. */
int inti_device_alloc(inti_ctx *ctx, size_t bytes, void **dptr);
int inti_device_free(inti_ctx *ctx, void *dptr);
int inti_memcpy_h2d(inti_ctx *ctx, void *dst, const void *src, size_t bytes);
int inti_memcpy_d2h(inti_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel on the context stream
 * (bytes a multiple of 16). No Inti.jl counterpart: a kernel of exactly
 * known HBM traffic, used to calibrate the profiler's byte counters. */
int inti_device_copy(inti_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Issue rate of one instruction over the whole GPU, in lane-operations per
 * second (8 independent chains per lane, 32 waves per CU): op 0 v_fma_f64,
 * 1 v_fma_f32 (non-packed), 2 v_rsq_f32, 3 v_sin_f32, 4 v_cos_f32,
 * 5 v_log_f32 (with one v_add_f32 per log), 6 v_rcp_f32, 7 v_rsq_f64. No
 * Inti.jl counterpart: the measured VALU and transcendental roofs of the
 * direct-sum kernels' roofline. */
int inti_microbench(inti_ctx *ctx, int op, double *rate);

/* ------------------------------------------------------------- quadrature */
/* Upload a quadrature (replaces handing Inti.Quadrature's QuadratureNode
 * vector, src/quadrature.jl:7-11,65-70, to native code as the FMM ext does,
 * ext/IntiFMM3DExt.jl:14-23). normals may be NULL (point targets). */
int inti_quad_upload(inti_ctx *ctx, int dim, int64_t n, const double *coords,
                     const double *normals, const double *weights,
                     inti_quad **out);
int inti_quad_destroy(inti_quad *q);

/* --------------------------------------------------------------- operator */
/* Matrix-free IntegralOperator (src/nystrom.jl:67-89): entries
 * K(x_i, y_j) * w_j with K chosen by (pde, kind, dim, k). coefs (4 doubles,
 * c0 re/im, c1 re/im) are read only for the combined kinds. */
int inti_op_create(inti_ctx *ctx, int pde, int kind, int dim, double k,
                   const double *coefs, const inti_quad *trg,
                   const inti_quad *src, int precision, inti_op **out);
int inti_op_destroy(inti_op *op);
/* Adds cid * x to every apply (square operators only), e.g. the I/2 of
 * I/2 + D - i k S (docs/src/examples/helmholtz_scattering.jl:230). */
int inti_op_set_identity(inti_op *op, double re, double im);
/* Imaginary part of the wavenumber: Helmholtz(k = kr + i kimag) of
 * src/kernels.jl:322-345, whose kernels take any k::Number. 3D, FP64, the
 * SL/DL/ADL/HS and combined kinds; the sums run in one direction. Set it
 * before the first apply, assembly or DIM build; a nonzero kimag on another
 * operator returns INTI_ERR_UNSUPPORTED, and so does the adjoint of such an
 * operator. */
int inti_op_set_wavenumber_imag(inti_op *op, double kimag);
/* Launch shape: threads per block, targets per thread, source splits
 * (0 = automatic). Used by the tuner; defaults are sensible. */
int inti_op_set_launch(inti_op *op, int block, int tpt, int split);
/* Symmetric evaluation of a square operator whose target and source are the
 * same quadrature handle (single right-hand side): each unordered pair's
 * shared part (distance, e^{ikd}, Hankel values) is computed once and applied
 * in both directions. mode 0: off; 1 (default): where the work fills the
 * GPU (at least two blocks of tile pairs per CU, about 8,000 nodes on
 * MI355X); 2: wherever it applies. group: tiles of 256 nodes per work group
 * (1, 2 or 4; 0 = the smallest whose partial sums fit 4 GiB). Results are
 * deterministic in every mode. No counterpart in Inti.jl (its matrices are
 * assembled entry by entry). */
int inti_op_set_symmetric(inti_op *op, int mode, int group);

/* y = alpha*(A + dA)*x + beta*y for nrhs columns; host pointers
 * (mul!(y, A, x, alpha, beta) of LinearMaps, src/api.jl:182-185). */
int inti_op_apply(inti_op *op, int nrhs, const double *x, double *y,
                  const double *alpha, const double *beta);
/* The same with device pointers and leading dimensions (in elements);
 * asynchronous on the context stream. */
int inti_op_apply_device(inti_op *op, int nrhs, const double *x, int64_t ldx,
                         double *y, int64_t ldy, const double *alpha,
                         const double *beta);
/* inti_op_apply_device without any device allocation. No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). The
 * one-direction direct sum runs (never the symmetric one) with its partial
 * sums in a static buffer of INTI_NOALLOC_SCRATCH_BYTES that the library
 * holds on every GPU it runs on until the process ends; the source split is
 * capped so that the partials of a block of columns (16, 4 or 1) fit it.
 * Shapes whose partials do not fit even unsplit (m * 16 columns * 16 bytes
 * for complex operators with 16 or more right-hand sides, i.e. more than
 * 262,144 targets; 4,194,304 targets for one complex column) return
 * INTI_ERR_SIZE. The correction must already be attached (attaching
 * allocates). Every call on one GPU shares that buffer: calls that overlap
 * on one GPU (several contexts or host threads on the same device, or
 * streams running concurrently) race; calls on one context are serialised. */
#define INTI_NOALLOC_SCRATCH_BYTES (64u << 20)
int inti_op_apply_device_noalloc(inti_op *op, int nrhs, const double *x,
                                 int64_t ldx, double *y, int64_t ldy,
                                 const double *alpha, const double *beta);
/* Dense column-major m x n matrix of the operator, written to host memory
 * (assemble_matrix, src/nystrom.jl:103-124). */
int inti_op_assemble_dense(inti_op *op, double *out);
/* Pullback of y = A sigma (reverse mode; the rrule of `*` on an
 * IntegralOperator, for objectives J(y) with cotangent ubar = dJ/dRe y +
 * i dJ/dIm y). Kernel values are recomputed, never stored. Outputs (host,
 * NULL to skip):
 *   sigma_bar (n, complex for Helmholtz) = A^H ubar (corrections and the
 *     identity term included);
 *   src_xyz_bar (dim x n SoA), src_normal_bar (dim x n), src_weight_bar (n):
 *     dJ/d(source nodes, normals, weights);
 *   trg_xyz_bar (dim x m): dJ/d(target nodes) (not for far-field kinds);
 *   k_bar: dJ/dk. dcoefs = {Re, Im dc0/dk, Re, Im dc1/dk} for combined kinds
 *     whose coefficients depend on k (e.g. -i k), NULL for constants.
 * For an operator on one surface the gradient of a node is the sum of its
 * source and target parts; passing the same pointer as src_xyz_bar and
 * trg_xyz_bar (target == source) asks for that sum directly (fewer sums per
 * pair in the one-pass pullback). Kinds: SL, DL, COMB_SD and the far fields,
 * FP64.
 * Geometry and k gradients are of the direct sum: they are refused while a
 * correction is attached (its weights depend on the geometry).
 * No counterpart in the reference, which differentiates with ForwardDiff. */
int inti_op_adjoint(inti_op *op, const double *ubar, const double *sigma,
                    double *sigma_bar, double *src_xyz_bar,
                    double *src_normal_bar, double *src_weight_bar,
                    double *trg_xyz_bar, const double *dcoefs, double *k_bar);
/* Dense matrix of the operator written to device memory (column-major,
 * ldout = number of targets): assemble_matrix (src/nystrom.jl:103-124) kept
 * in HBM for the dense products below. */
int inti_op_assemble_dense_device(inti_op *op, double *out, int64_t ldout);
/* A += coef * csr on a dense device matrix (the axpy!(true, dS, Smat) of
 * src/api.jl:171-173). */
int inti_dense_add_csr(inti_ctx *ctx, const inti_csr *csr, double coef_re,
                       double coef_im, double *A, int64_t lda);
/* Y = alpha A X + beta Y for an assembled dense m x n matrix A and nrhs
 * columns, all device pointers, column-major, complex interleaved when
 * is_complex: the dense `S * sigma` of src/api.jl:171-173 on many densities
 * at once, a real GEMM on the FP64 matrix cores (complex A is multiplied in
 * its real form). */
int inti_dense_gemm(inti_ctx *ctx, int is_complex, int64_t m, int64_t n, int nrhs,
                    const double *A, int64_t lda, const double *X, int64_t ldx,
                    double *Y, int64_t ldy, const double *alpha, const double *beta);
/* Fuse coef * csr into every following apply of op (the sparse correction
 * of src/api.jl:171-185). Several attachments with the same pattern are
 * merged into one CSR. */
int inti_op_attach_correction(inti_op *op, const inti_csr *csr, double coef_re,
                              double coef_im);
int inti_op_detach_correction(inti_op *op);

/* ------------------------------------------------------------ batches */
/* Many independent problems (boundaries of uneven size, one wavenumber
 * each) as one block-diagonal operator: one persistent launch applies all
 * of them, one batched GMRES solves all (problem, right-hand side) systems.
 * The reference has no batch: this replaces a loop over problems of
 * IntegralOperator(K, Q_p, Q_p) (src/nystrom.jl:79-89) and `gmres(L_p, b_p)`
 * (docs/src/examples/helmholtz_scattering.jl:494-495).
 *   quads[p]: the problem's quadrature (target = source); k[p]: wavenumber
 *   (Helmholtz); coefs: 4 per problem for INTI_COMB_SD (NULL otherwise).
 *   Kinds: SL, DL, COMB_SD. Vectors are block vectors: ntot x nrhs,
 *   column-major, problem p in rows [offsets[p], offsets[p+1]). */
int inti_batch_create(inti_ctx *ctx, int pde, int kind, int dim, int nprob,
                      const inti_quad *const *quads, const double *k,
                      const double *coefs, int precision, inti_batch **out);
int inti_batch_destroy(inti_batch *batch);
/* ntot and offsets[nprob + 1] (either may be NULL) */
int inti_batch_info(const inti_batch *batch, int64_t *ntot, int64_t *offsets);
/* identity coefficients, 2 doubles (re, im) per problem */
int inti_batch_set_identity(inti_batch *batch, const double *cid);
/* The DIM corrections (dS, dD) of every problem of the batch, as
 * bdim_correction (src/bdim.jl:50-196) gives them for each problem's
 * quadrature with itself (the self pattern, one element type of nq nodes):
 * ne[p] elements per problem, qtags the column-major (nq, ne[p]) node
 * tables of the problems one after another (0-based within each problem),
 * mu the green multiplier of every node (ntot), params the DIM parameters
 * (oversampling, radius multiplier; NULL for 3, 1.5). dS[p], dD[p] receive
 * one CSR per problem. The batch's quadratures need normals (DL or combined
 * batches). Replaces nprob separate bdim_correction calls (the reference
 * loops over problems). */
int inti_batch_bdim_build(inti_batch *batch, int nq, const int64_t *ne, const int64_t *qtags,
                          const double *mu, const double *params, inti_csr **dS, inti_csr **dD);
/* add coef * csr (problem-local indices, n_p x n_p) to problem p */
int inti_batch_attach_correction(inti_batch *batch, int p, const inti_csr *csr,
                                 double coef_re, double coef_im);
/* source chunk length of a work item (0 = 2048) and persistent grid size
 * (0 = 4 workgroups per CU) */
int inti_batch_set_launch(inti_batch *batch, int64_t chunk_len, int grid);
/* y = alpha A x + beta y; device (leading dimensions >= ntot) or host */
int inti_batch_apply_device(inti_batch *batch, int nrhs, const double *x,
                            int64_t ldx, double *y, int64_t ldy,
                            const double *alpha, const double *beta);
int inti_batch_apply(inti_batch *batch, int nrhs, const double *x, double *y,
                     const double *alpha, const double *beta);
/* Time the apply of nrhs columns over chunk lengths {1024, 2048, 4096, 8192}
 * and grids {2, 4, 8} workgroups per CU (reps applies each) and keep the
 * fastest; reports the choice and its time per apply (ms). */
int inti_batch_autotune(inti_batch *batch, int nrhs, int reps, int64_t *chunk,
                        int *grid, double *ms);
/* Batched restarted GMRES on every system (p, r): stops each at
 * ||b - A x|| <= max(rtol ||b||, atol) of that system. b, x: block vectors
 * (leading dimension ntot), x in/out, host or device (on_device). nrhs
 * divides 256. iters: Arnoldi steps taken (shared by all systems); relres:
 * nprob * nrhs final relative residuals (system p * nrhs + r). */
int inti_batch_gmres(inti_batch *batch, int on_device, int nrhs, const double *b,
                     double *x, int restart, int maxiter, double rtol, double atol,
                     int use_graph, int *iters, double *relres);

/* ---------------------------------------------------------- DIM correction */
/* bdim_correction (src/bdim.jl:50-196). opS/opD are the SL/DL operators
 * (derivative = 0) or ADL/HS operators (derivative = 1) over the same
 * target and source quadratures; Theta = mu*G + S*B1 - D*B0 is computed
 * matrix-free on the GPU.
 *   qtags: nq x ne column-major source node indices per element.
 *   near_ptr/near_idx: element -> corrected targets (CSR over elements),
 *     or NULL/NULL for the self pattern (target == source,
 *     src/quadrature.jl:250-256).
 *   mu: green multiplier per target (src/nystrom.jl:196-207).
 *   params: {sources_oversample_factor, sources_radius_multiplier}
 *     (src/bdim.jl:7-10), NULL for the defaults {3, 1.5}.
 *   diag (optional, 4 doubles): max R-diagonal ratio, max LS residual,
 *     max |W|, max |Theta row| (the @debug report of src/bdim.jl:186-192).
 * Returns the two corrections as CSR (rows = targets, columns = sources). */
int inti_bdim_build(inti_op *opS, inti_op *opD, int derivative, int nq,
                    int64_t ne, const int64_t *qtags, const int64_t *near_ptr,
                    const int64_t *near_idx, const double *mu,
                    const double *params, inti_csr **dS, inti_csr **dD,
                    double *diag);
/* bdim_correction over a quadrature with several element types
 * (src/bdim.jl:79-80 sums nq over the types for the monopole count,
 * src/bdim.jl:138-185 runs the least squares type by type). Type t has nq[t]
 * nodes per element and ne[t] elements; qtags holds the nq[t] x ne[t]
 * column-major blocks one after another; near_ptr holds ne[t] + 1 entries per
 * type (each block starting at 0) and near_idx the blocks of targets they
 * index, in the same type order (or NULL/NULL for the self pattern, where the
 * types must cover every source node once). The other arguments are those of
 * inti_bdim_build; the CSR entries are summed over the types. */
int inti_bdim_build_types(inti_op *opS, inti_op *opD, int derivative,
                          int ntypes, const int *nq, const int64_t *ne,
                          const int64_t *qtags, const int64_t *near_ptr,
                          const int64_t *near_idx, const double *mu,
                          const double *params, inti_csr **dS,
                          inti_csr **dD, double *diag);

/* Batched per-element least-squares correction: the solve-and-scatter step
 * of vdim_correction (src/vdim.jl:84-127). Element e has a real
 * nrows x ncols matrix A_e (column-major, elements consecutive in A; in
 * vdim the interpolation basis at the element's nodes), and every target i
 * listed for e has the right-hand side row rhs[i, :] of the m x nrows
 * column-major rhs (complex interleaved if is_complex). The least-squares
 * solution w of A_e w = rhs[i, :] gives the CSR entries (i, cols[c, e]) =
 * w[c] (duplicates summed). The reference solves with an SVD; for a
 * full-rank A_e Householder QR gives the same solution: the least-squares one
 * when ncols <= nrows, the minimum-norm one (QR of A_e^T) when ncols > nrows.
 * Sizes: min(nrows, ncols) <= 24, max(nrows, ncols) <= 32.
 * near_ptr/near_idx: element -> targets, or NULL for
 * the self pattern (m == n, the targets of e are cols[:, e]).
 * diag (optional, 2 doubles): max R-diagonal ratio, max |w|. */
int inti_lsq_build(inti_ctx *ctx, int is_complex, int nrows, int ncols,
                   int64_t ne, const double *A, const int64_t *cols, int64_t m,
                   int64_t n, const int64_t *near_ptr, const int64_t *near_idx,
                   const double *rhs, inti_csr **out, double *diag);
/* inti_lsq_build over ntypes element types (vdim.jl:76-121 runs the least
 * squares type by type): type t has ncols[t] nodes per element and ne[t]
 * elements; A holds the nrows x ncols[t] matrices of its elements after those
 * of the earlier types, cols the ncols[t] x ne[t] column-major blocks one
 * after another, near_ptr ne[t] + 1 entries per type (each block from 0) and
 * near_idx the target blocks in the same order (NULL/NULL: self pattern). */
int inti_lsq_build_types(inti_ctx *ctx, int is_complex, int nrows, int ntypes,
                         const int *ncols, const int64_t *ne, const double *A,
                         const int64_t *cols, int64_t m, int64_t n,
                         const int64_t *near_ptr, const int64_t *near_idx,
                         const double *rhs, inti_csr **out, double *diag);

/* ----------------------------------------------------------------- solvers */
/* Restarted GMRES(restart) for the square operator op (with its attached
 * correction and identity term), the solve of user code such as
 * gmres(L, rhs; restart, maxiter) on L = I/2 + D - i k S
 * (docs/src/examples/helmholtz_scattering.jl:230,494-495). Stops when
 * ||b - A x|| <= max(rtol ||b||, atol) or after maxiter Arnoldi steps.
 * x holds the initial guess and receives the solution. on_device: b and x
 * are device pointers (else host). use_graph: replay one Arnoldi step as a
 * captured hipGraph. iters/relres (optional): steps taken, final
 * ||b - A x|| / ||b||. */
int inti_gmres(inti_op *op, int on_device, const double *b, double *x,
               int restart, int maxiter, double rtol, double atol,
               int use_graph, int *iters, double *relres);
/* Iterative refinement: residuals with op_hi (e.g. FP64), corrections by
 * GMRES on op_lo (e.g. MIXED) to inner_rtol relative to the current
 * residual, until ||b - A_hi x|| <= rtol ||b|| or max_outer sweeps. */
int inti_refine_solve(inti_op *op_hi, inti_op *op_lo, int on_device,
                      const double *b, double *x, int restart,
                      int maxiter_inner, double inner_rtol, int max_outer,
                      double rtol, int use_graph, int *outer,
                      int *inner_total, double *relres);
/* inti_gmres and inti_refine_solve with a right preconditioner prec (a
 * square operator of the same size and value type, with its own correction
 * and identity term; NULL: none): GMRES runs on A P y = r and updates
 * x += P y, so the residual it minimises is that of A itself (the
 * preconditioned solve of IterativeSolvers' gmres(A, b; Pr = P) that user
 * code on the reference would write, e.g. the Calderon preconditioner of a
 * first-kind single-layer equation). Under refinement P is applied in the
 * precision it was created with. */
int inti_gmres_prec(inti_op *op, inti_op *prec, int on_device, const double *b,
                    double *x, int restart, int maxiter, double rtol,
                    double atol, int use_graph, int *iters, double *relres);
int inti_refine_solve_prec(inti_op *op_hi, inti_op *op_lo, inti_op *prec,
                           int on_device, const double *b, double *x,
                           int restart, int maxiter_inner, double inner_rtol,
                           int max_outer, double rtol, int use_graph,
                           int *outer, int *inner_total, double *relres);

/* --------------------------------------------------------------------- CSR */
int inti_csr_info(const inti_csr *csr, int64_t *m, int64_t *n, int64_t *nnz,
                  int *is_complex);
/* Copy to host: rowptr[m+1], colind[nnz], val[nnz] (x2 if complex). The
 * Julia shim turns this into SparseMatrixCSC (src/bdim.jl:193-194). */
int inti_csr_export(const inti_csr *csr, int64_t *rowptr, int64_t *colind,
                    double *val);
int inti_csr_destroy(inti_csr *csr);

#ifdef __cplusplus
}
#endif
#endif /* INTI_AMD_H */
