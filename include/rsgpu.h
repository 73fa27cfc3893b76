/*
 * rsgpu: MI355X (gfx950) sparse Krylov path for the rs reservoir simulator.
 *
 * C ABI only: plain pointers, sizes and opaque handles.  Every entry point
 * returns an int status (RSGPU_OK = 0); reasons 1-3 map to MTJ's
 * IterativeSolverNotConvergedException.Reason, 10-14 to the runtime
 * exceptions the MTJ classes throw, negative values to HIP/runtime errors
 * whose text rsgpu_last_error() returns.
 *
 * Pointer conventions: arguments named *_host are host memory (a Java array
 * pinned by GetPrimitiveArrayCritical, a numpy buffer ...); arguments named
 * *_dev are device memory on the context's device.  Device-pointer calls
 * are enqueued on the context's stream; host-pointer calls block until
 * their result is back in host memory.
 *
 * Threading: one handle = one MTJ object; calls on one context must not
 * overlap (the reference classes are single-threaded too).  Different
 * contexts may be used from different threads.
 *
 * Each declaration cites the reference interface it replaces (paths
 * relative to the reference root, mtj = mtj/src/no/uib/cipr/matrix).
 */
#ifndef RSGPU_H
#define RSGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum rsgpu_status {
    RSGPU_OK = 0,
    RSGPU_BREAKDOWN = 1,        /* NotConvergedException.Reason.Breakdown  */
    RSGPU_DIVERGENCE = 2,       /* Reason.Divergence (incl. NaN residual)  */
    RSGPU_ITERATIONS = 3,       /* Reason.Iterations                       */
    RSGPU_ZERO_PIVOT = 10,      /* ILU.java:105-107 RuntimeException        */
    RSGPU_MISSING_DIAGONAL = 11,/* ILU.java:134-136 RuntimeException        */
    RSGPU_NOT_IN_PATTERN = 12,  /* CompRowMatrix.java:435 IndexOutOfBounds  */
    RSGPU_BAD_ARGUMENT = 13,    /* IllegalArgumentException                 */
    RSGPU_SINGULAR = 14,        /* MatrixSingularException (GMRES dtrtrs)   */
    RSGPU_HIP_ERROR = -1,
    RSGPU_OUT_OF_MEMORY = -2,
    RSGPU_NOT_BUILT = -3
};

typedef struct rsgpu_ctx rsgpu_ctx;
typedef struct rsgpu_csr rsgpu_csr;
typedef struct rsgpu_prec rsgpu_prec;
typedef struct rsgpu_solver rsgpu_solver;
typedef struct rsgpu_flux rsgpu_flux;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart (the JVM is the context); owns the HIP stream. */
int rsgpu_ctx_create(int device, rsgpu_ctx **out);
int rsgpu_ctx_destroy(rsgpu_ctx *ctx);
/* Use a caller-owned hipStream_t (e.g. torch's current stream); NULL
 * restores the context's own stream. */
int rsgpu_ctx_set_stream(rsgpu_ctx *ctx, void *hip_stream);
int rsgpu_ctx_synchronize(rsgpu_ctx *ctx);
const char *rsgpu_last_error(rsgpu_ctx *ctx);
int rsgpu_version(void);

/* ---- CompRowMatrix (mtj/sparse/CompRowMatrix.java) ----------------------- */
/* CompRowMatrix(int numRows, int numColumns, int[][] nz), :160-190.  rowptr
 * (n+1) and colind (nnz) describe the pattern; rows need not be sorted (they
 * are sorted as Arrays.sort does, :188); a column outside [0, ncols) gives
 * RSGPU_BAD_ARGUMENT.  Values start at zero. */
int rsgpu_csr_create(rsgpu_ctx *ctx, int n, int ncols, int nnz,
                     const int32_t *rowptr_host, const int32_t *colind_host,
                     rsgpu_csr **out);
int rsgpu_csr_destroy(rsgpu_csr *A);
/* getRowPointers/getColumnIndices (:261-273): the sorted pattern. */
int rsgpu_csr_get_pattern(const rsgpu_csr *A, int32_t *rowptr_host,
                          int32_t *colind_host);
/* Replaces the data[] array (getData(), :275-277) wholesale. */
int rsgpu_csr_set_values(rsgpu_csr *A, const double *data_host);
int rsgpu_csr_set_values_dev(rsgpu_csr *A, const double *data_dev);
int rsgpu_csr_get_values(const rsgpu_csr *A, double *data_host);
double *rsgpu_csr_values_dev(rsgpu_csr *A);
/* CompRowMatrix.zero() (:450-453). */
int rsgpu_csr_zero(rsgpu_csr *A);
/* CompRowMatrix.add(row, col, value) for a batch of triples, applied in
 * array order (:405-437); a triple outside the pattern makes the call
 * return RSGPU_NOT_IN_PATTERN with nothing applied. */
int rsgpu_csr_add_batch(rsgpu_csr *A, int count, const int32_t *rows_host,
                        const int32_t *cols_host, const double *vals_host);
/* Matrices.zeroRows(A, diagonal, rows...) (mtj/Matrices.java:631-652). */
int rsgpu_csr_zero_rows(rsgpu_csr *A, double diagonal,
                        const int32_t *rows_host, int nrows);
/* mult(x, y) (:301-333): y = A x, with y[i] = 0 when the row sum is 0. */
int rsgpu_csr_mult(rsgpu_csr *A, const double *x_dev, double *y_dev);
/* multAdd(alpha, x, y) (:336-353): y += alpha * A x. */
int rsgpu_csr_multadd(rsgpu_csr *A, double alpha, const double *x_dev,
                      double *y_dev);
/* transMult(x, y) (:356-372) and transMultAdd(alpha, x, y) (:375-394):
 * computed as a gather over a transposed copy (deterministic, no atomics). */
int rsgpu_csr_transmult(rsgpu_csr *A, const double *x_dev, double *y_dev);
int rsgpu_csr_transmultadd(rsgpu_csr *A, double alpha, const double *x_dev,
                           double *y_dev);
/* r = b - A x, the residual multAdd(-1, x, b) forms (:336-353), but with
 * the products and row sums carried in twice the working precision
 * (Ogita-Rump-Oishi Dot2) and rounded once, so r is exact to about one
 * rounding of its own size however much the row cancels.  No reference
 * counterpart: iterative refinement on it converges to the FP64-attainable
 * solution (the adjoint's forward and adjoint solves use it). */
int rsgpu_csr_residual_dd(rsgpu_csr *A, const double *b_dev,
                          const double *x_dev, double *r_dev);
/* Single-precision copy of the values (no reference counterpart; used by
 * the mixed-precision form) and y = A x in FP32. */
int rsgpu_csr_sync_f32(rsgpu_csr *A);
/* A new matrix holding A' (rows ascending per column, as transMult's
 * scatter visits them, CompRowMatrix.java:356-372); values copied now. */
int rsgpu_csr_transpose(const rsgpu_csr *A, rsgpu_csr **At);
int rsgpu_csr_mult_f32(rsgpu_csr *A, const float *x_dev, float *y_dev);
/* Sliced copy (no MTJ counterpart; the storage the solvers' SpMV streams
 * for stencil matrices).  Rows in chunks of 256, each chunk with the sorted
 * union of its rows' column offsets (column - row) as a table and per row a
 * mask of the slots it holds; values slot-major per chunk.  Built with the
 * pattern when every chunk's union has at most 32 offsets (nchunks = 0
 * otherwise; RSGPU_SLICED=0 in the environment leaves it out).  A 3x3 block
 * matrix (rsgpu_csr_create_bsr3) gets the block form instead: a chunk's
 * table lists block offsets (block column - block row, at most 10), row
 * 3I+a holds slots 3k+e for column 3(I+d_k)+e, masks per row over the
 * table's blocks.  nslots =
 * values per copy.  y = A x over it is bit-identical to rsgpu_csr_mult
 * (mult_sliced_f32: to rsgpu_csr_mult_f32).  The FP64 copy is refreshed by
 * rsgpu_csr_sync_sliced (and by every FP64 solve at its start), the FP32
 * copy by rsgpu_csr_sync_f32.  Every entry point that changes A's values
 * (set_values*, zero, add_batch, zero_rows, the assembly calls) makes both
 * copies stale: mult_sliced, mult_sliced_f32, mult_f32 and the mixed
 * solves then return RSGPU_BAD_ARGUMENT until the copy is synced again.
 * Writes through the rsgpu_csr_values_dev pointer are not tracked; sync
 * after them. */
int rsgpu_csr_sliced_info(const rsgpu_csr *A, int *nchunks, int64_t *nslots,
                          int *ntables);
int rsgpu_csr_sync_sliced(rsgpu_csr *A);
int rsgpu_csr_mult_sliced(rsgpu_csr *A, const double *x_dev, double *y_dev);
int rsgpu_csr_mult_sliced_f32(rsgpu_csr *A, const float *x_dev, float *y_dev);
/* 3x3 block rows (no MTJ counterpart: MTJ has no block-CSR storage).  The
 * block pattern browptr (nb+1) / bcolind (nnzb; sorted per block row as
 * Arrays.sort would, CompRowMatrix.java:188) defines the expanded scalar
 * matrix of 3 nb rows in which block row I is scalar rows 3I..3I+2, each
 * listing its blocks' three columns in ascending order.  The value array
 * (getData, rsgpu_csr_set_values) is that expanded CSR's, so value
 * (row a, block b, column c) of block row I sits at 9 browptr[I] +
 * 3 a (browptr[I+1] - browptr[I]) + 3 (b - browptr[I]) + c.  Every CSR entry
 * point accepts the matrix (mult, solvers and ILU act on the expanded
 * matrix); SpMV reads one block column index per block. */
int rsgpu_csr_create_bsr3(rsgpu_ctx *ctx, int nb, int nnzb,
                          const int32_t *browptr_host,
                          const int32_t *bcolind_host, rsgpu_csr **out);
/* Block size (1 or 3), block rows and block nonzeros (scalar ones for bs 1). */
int rsgpu_csr_block_info(const rsgpu_csr *A, int *bs, int *nb, int *nnzb);
/* Host-array conveniences (DenseVector.getData() round trip). */
int rsgpu_csr_mult_host(rsgpu_csr *A, const double *x_host, double *y_host);
int rsgpu_csr_transmult_host(rsgpu_csr *A, const double *x_host,
                             double *y_host);

/* ---- Preconditioner (mtj/sparse/Preconditioner.java:42-63) --------------- */
enum rsgpu_prec_kind {
    RSGPU_PREC_IDENTITY = 0,   /* AbstractIterativeSolver.IdentityPreconditioner */
    RSGPU_PREC_ILU0 = 1,       /* mtj/sparse/ILU.java                            */
    RSGPU_PREC_BLOCK_ILU0 = 2, /* distributed/BlockDiagonalPreconditioner + ILU  */
    RSGPU_PREC_DIAGONAL = 3,   /* mtj/sparse/DiagonalPreconditioner.java         */
    RSGPU_PREC_AMG = 4,        /* mtj/sparse/AMG.java (smoothed aggregation)     */
    RSGPU_PREC_SSOR = 5,       /* mtj/sparse/SSOR.java                           */
    RSGPU_PREC_ICC = 6,        /* mtj/sparse/ICC.java (FP64)                     */
    RSGPU_PREC_ILUT = 7        /* mtj/sparse/ILUT.java (FP64)                    */
};
enum rsgpu_precision { RSGPU_FP64 = 0, RSGPU_FP32 = 1 };

/* nblocks: contiguous equal row ranges (DistRowMatrix ownership,
 * distributed/DistRowMatrix.java:56-76) for BLOCK_ILU0, ignored otherwise.
 * precision: arithmetic of the factors and sweeps. */
int rsgpu_prec_create(rsgpu_ctx *ctx, int kind, int nblocks, int precision,
                      rsgpu_prec **out);
/* BLOCK_ILU0 only: the caller's block boundaries instead of equal ranges,
 * bstart_host[0] = 0 < bstart_host[1] < ... < bstart_host[nblocks] = n, as
 * the row ownership of a DistRowMatrix whose ranks own unequal row counts
 * (distributed/DistRowMatrix.java:56-76).  Blocks aligned to the systems of
 * a block-diagonal batch make the block-Jacobi ILU(0) exactly the ILU(0) of
 * each system.  Takes effect at the next rsgpu_prec_set_matrix. */
int rsgpu_prec_set_blocks(rsgpu_prec *M, int nblocks, const int32_t *bstart_host);
int rsgpu_prec_destroy(rsgpu_prec *M);
/* setMatrix(A) (ILU.java:80-84): copy values, factor.  The pattern analysis
 * (level schedule) is done on the first call and reused while A's pattern
 * is the same object.  bad_row_out (may be NULL) receives the reference's
 * row/entry number for ZERO_PIVOT / MISSING_DIAGONAL. */
int rsgpu_prec_set_matrix(rsgpu_prec *M, rsgpu_csr *A, int *bad_row_out);
/* apply(b, x) (ILU.java:64-70) and transApply(b, x) (:72-78). FP64 vectors;
 * an FP32 preconditioner converts on the way in and out. */
int rsgpu_prec_apply(rsgpu_prec *M, const double *b_dev, double *x_dev);
int rsgpu_prec_trans_apply(rsgpu_prec *M, const double *b_dev, double *x_dev);
int rsgpu_prec_apply_f32(rsgpu_prec *M, const float *b_dev, float *x_dev);
int rsgpu_prec_apply_host(rsgpu_prec *M, const double *b_host,
                          double *x_host);
int rsgpu_prec_trans_apply_host(rsgpu_prec *M, const double *b_host,
                                double *x_host);
/* Copy the factored LU values (CSR order of the preconditioner's pattern:
 * A's pattern for ILU0, its block-diagonal part for BLOCK_ILU0). */
int rsgpu_prec_get_factors(const rsgpu_prec *M, int32_t *rowptr_host,
                           int32_t *colind_host, double *lu_host, int *nnz);
int rsgpu_prec_info(const rsgpu_prec *M, int *nnz, int *nlevels_lower,
                    int *nlevels_upper);
/* Which sweep kernel ILU.apply runs (no MTJ counterpart; all three are
 * bit-identical to the sequential solves): 0 level schedules, 1 line
 * wavefront (every block Ly <= 64 lines of Lx rows coupled among 9-point
 * neighbours, at most 4 blocks per CU; RSGPU_ILU_BAND=0 in the environment
 * turns it off, RSGPU_ILU_BAND_APPLY=0/1 forces the choice), 2 3x3 cell
 * schedules, 3 3x3 cell schedules with the level records streamed through
 * LDS (2 or 3 neighbour blocks per triangle and a level's records within
 * 4 KiB; RSGPU_ILU_CELL_STREAM=0 turns it off). */
int rsgpu_prec_sweep_form(const rsgpu_prec *M, int *form);

/* AMG parameters: the constructor arguments of AMG(omegaPreF, omegaPreR,
 * omegaPostF, omegaPostR, nu1, nu2, gamma, min, omega)
 * (mtj/sparse/AMG.java:140-156, defaults :204-206) plus three of this
 * library's: block_rows (SSOR smoother blocks of that many contiguous rows;
 * couplings leaving a block use the vector from the start of the SSOR
 * application; 0 = one block per level), ordering (0: rows in natural order
 * as SSOR.java; 1: rows by greedy colour, multicolour SSOR, whose sweep is
 * one parallel step per colour), zero_guess (0: the cycle starts from x's
 * old content as AMG.apply does, AMG.java:208-216; 1: from zero, a fixed
 * linear preconditioner), max_coarse (largest dense coarsest level
 * accepted), coarsen_ratio (below).  block_rows = 0 with ordering = 0 is
 * SSOR.java exactly.  An FP32 AMG (rsgpu_prec_create precision FP32) builds
 * the hierarchy in FP64 and runs the cycle -- smoothers, residuals and
 * transfers -- in FP32 with the coarsest solve in FP64; coarsest levels
 * above 256 rows are solved with their explicit inverse (from the LU at
 * setup). */
typedef struct {
    double omega_pre_f, omega_pre_r, omega_post_f, omega_post_r;
    int nu1, nu2, gamma, min;
    double omega;
    int reverse;          /* 1 SSOR, 0 SOR (AMG.java:148,186) */
    int block_rows;
    int zero_guess;
    int max_coarse;
    int ordering;
    /* no MTJ counterpart: coarsening stops at a level whose aggregates keep
     * more than this fraction of its rows (0: never, AMG.java's rule of
     * coarsening down to `min` rows) */
    double coarsen_ratio;
} rsgpu_amg_params;
void rsgpu_amg_params_default(rsgpu_amg_params *p);
/* Takes effect at the next rsgpu_prec_set_matrix (AMG.setMatrix,
 * AMG.java:228-294, which rebuilds the hierarchy every call). */
int rsgpu_prec_set_amg_params(rsgpu_prec *M, const rsgpu_amg_params *p);
/* SSOR(F, reverse, omegaF, omegaR) (SSOR.java:80-130): apply(b, x) starts
 * from x's old content (SSOR.java:143-194); transApply = apply.  block_rows
 * and ordering as for the AMG smoother (0, 0 = SSOR.java exactly).
 * Takes effect at the next set_matrix. */
int rsgpu_prec_set_ssor(rsgpu_prec *M, double omega_f, double omega_r,
                        int reverse, int block_rows, int ordering);
/* ILUT(LU, tau, p) (ILUT.java:91-118; defaults tau = 1e-6, p = 25): drop
 * tolerance tau (relative to each row's 2-norm) and the entries kept per
 * side beyond A's own.  set_matrix factors on the device
 * (ILUT.setMatrix/factor, ILUT.java:136-208; errors: 11 "Missing diagonal
 * entry on row k", 10 "Zero diagonal entry on row k during ILU process");
 * apply / trans_apply are ILUT.apply / transApply (:120-134).  Every
 * factorisation starts from A's nonzero entries.  Takes effect at the next
 * set_matrix. */
int rsgpu_prec_set_ilut(rsgpu_prec *M, double tau, int p);
/* Hierarchy introspection (tests): level sizes / nonzeros; returns the
 * level count m (the last level is solved by a dense LU). */
int rsgpu_amg_levels(const rsgpu_prec *M, int *sizes, int *nnzs, int cap);
/* Level k's matrix (CSR), and for k < m-1 its aggregate of every row and
 * the smoothed prolongator P (CSR, n_k x n_{k+1}); pointers may be NULL. */
int rsgpu_amg_get_level(const rsgpu_prec *M, int k, int32_t *rowptr,
                        int32_t *colind, double *data, int32_t *aggregate,
                        int32_t *p_rowptr, int32_t *p_colind, double *p_data,
                        int *p_nnz);

/* ---- IterativeSolver (mtj/sparse/IterativeSolver.java:43-67) ------------- */
enum rsgpu_solver_kind {
    RSGPU_BICGSTAB = 0,  /* mtj/sparse/BiCGstab.java */
    RSGPU_GMRES = 1,     /* mtj/sparse/GMRES.java    */
    RSGPU_IR = 2,        /* mtj/sparse/IR.java       */
    RSGPU_CG = 3,        /* mtj/sparse/CG.java       */
    RSGPU_BICG = 4,      /* mtj/sparse/BiCG.java     */
    RSGPU_CGS = 5,       /* mtj/sparse/CGS.java      */
    RSGPU_QMR = 6        /* mtj/sparse/QMR.java (M1 = M2 = M) */
};

/* DefaultIterationMonitor(maxIter, rtol, atol, dtol)
 * (DefaultIterationMonitor.java:66-90, :133-156). */
typedef struct {
    int kind;
    int restart;          /* GMRES(m); GMRES.java:84-86 default 30 */
    int max_iter;
    double rtol, atol, dtol;
    /* Mixed precision (no reference counterpart): when mixed != 0 the solve
     * is an FP64 iterative refinement (IR.java:57-70 form) on the true
     * residual b - A x, whose correction is an FP32 solve of the given kind
     * with the FP32 preconditioner, stopped at inner_rtol or inner_max_iter.
     */
    int mixed;
    int inner_max_iter;
    double inner_rtol;
    int check_every;      /* host polls the device monitor every k steps */
    /* Transposed solve (no reference counterpart; the adjoint uses it): A
     * is the transpose made by rsgpu_csr_transpose and M the
     * preconditioner set on the original matrix, applied through
     * transApply (ILU.java:72-78), i.e. (LU)^-T. */
    int prec_transpose;
    /* GMRES orthogonalisation (no reference counterpart; MTJ uses modified
     * Gram-Schmidt, GMRES.java:150-153, the default 0): with cgs != 0 the
     * FP32 inner GMRES of the mixed form uses classical Gram-Schmidt (all
     * projections of a column from one pass) and the FP64 GMRES uses CGS2
     * (two classical passes, H(:,i) = h1 + h2): four kernels per step
     * instead of i + 2, same iterates up to round-off.  The FP64 refinement
     * of the mixed form still tests the true residual. */
    int cgs;
} rsgpu_solver_opts;

typedef struct {
    int iterations;        /* IterationMonitor.iterations()              */
    double residual;       /* IterationMonitor.residual()                */
    int status;            /* rsgpu_status                               */
    int hist_len;          /* residuals reported (OutputIterationReporter)*/
    int inner_iterations;  /* mixed: total inner FP32 iterations         */
    int outer_iterations;  /* mixed: FP64 refinement steps               */
    double true_residual;  /* ||b - A x||_2 in FP64 at exit              */
    double b_norm;         /* ||b||_2                                    */
} rsgpu_stats;

void rsgpu_solver_opts_default(rsgpu_solver_opts *opts);
/* BiCGstab(template)/GMRES(template, restart)/IR(template): n = size. */
int rsgpu_solver_create(rsgpu_ctx *ctx, const rsgpu_solver_opts *opts, int n,
                        rsgpu_solver **out);
int rsgpu_solver_destroy(rsgpu_solver *S);
/* solve(A, b, x) (BiCGstab.java:65-116, GMRES.java:127-176, IR.java:57-70):
 * x is the initial guess and the result.  M may be NULL (identity). */
int rsgpu_solver_solve(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M,
                       const double *b_dev, double *x_dev, rsgpu_stats *st);
int rsgpu_solver_solve_host(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M,
                            const double *b_host, double *x_host,
                            rsgpu_stats *st);
/* Residuals the monitor saw, in order (IterationReporter.monitor replay). */
int rsgpu_solver_history(const rsgpu_solver *S, double *out_host, int cap);
/* GMRES inner steps of the last solve enqueued as one fused launch (sliced
 * SpMV + line-wavefront block ILU(0) apply + first Gram-Schmidt pass; used
 * with the cgs option on full line blocks, RSGPU_GMRES_FUSED=0 in the
 * environment turns it off).  No reference counterpart. */
int rsgpu_solver_fused_steps(const rsgpu_solver *S);
/* One fused GMRES inner step on caller vectors, for timing and testing the
 * fused kernel (no reference counterpart): w = M^-1 (A v) with v = column i
 * of V (no scaling), and proj[k] = V(:,k)' w for k = 0..i (device, FP64
 * sums).  V is n x (i+1), column-major, leading dimension ldv (>= n, a
 * multiple of 4; V and w 16-byte aligned, w not aliasing V); i below the
 * solver's restart; A and M must meet the fused step's conditions.  The
 * FP32 form needs rsgpu_csr_sync_f32 and an FP32 preconditioner. */
int rsgpu_gmres_fused_step(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M,
                           const double *V_dev, size_t ldv, int i, double *w_dev,
                           double *proj_dev);
int rsgpu_gmres_fused_step_f32(rsgpu_solver *S, rsgpu_csr *A, rsgpu_prec *M,
                               const float *V_dev, size_t ldv, int i, float *w_dev,
                               double *proj_dev);

/* ---- Fully-implicit black-oil Jacobian (no reference counterpart) -------- */
/* rs is IMPES (numerics/Discretisation.java:208-225 solves pressure only);
 * this is the Newton Jacobian of a three-phase black-oil time step with
 * dissolved gas on a Cartesian grid: unknowns (p, Sw, Sg) per cell, TPFA
 * 7-point fluxes with phase upwinding and gravity, Corey relative
 * permeabilities in CoreyRockFluid's form (rock/CoreyRockFluid.java:
 * 129-152, integer exponents), linear shrinkage factors b = 1/B for water
 * and oil, b_g and Rs proportional to p.  Residual rows are the water, oil
 * and gas component balances in surface volumes. */
typedef struct {
    double dx, dy, dz;      /* cell sizes [m] */
    double dt;              /* time step [s] */
    double g;               /* gravity [m/s^2]; depth grows with k */
    double phi0, cr;        /* porosity at pref, rock compressibility [1/Pa] */
    double pref;            /* reference pressure [Pa] */
    double kz_ratio;        /* k_z / k_h */
    double bw0, cw;         /* bw = bw0 (1 + cw (p - pref)) */
    double bo0, co;         /* bo = bo0 (1 + co (p - pref)) */
    double bg0;             /* bg = bg0 p / pref */
    double rs0;             /* Rs = rs0 p / pref */
    double muw, muo, mug;   /* viscosities [Pa s] */
    double rhow_s, rhoo_s, rhog_s; /* surface densities [kg/m^3] */
    double swr, sor, sgr;   /* Corey residual saturations (Swi, Sor, Sgr) */
    double kw, ko, kg;      /* Corey end-point relative permeabilities */
    int nw, no, ng, pad;    /* Corey exponents (integers 0..16) */
} rsgpu_blackoil_params;
typedef struct rsgpu_blackoil rsgpu_blackoil;
/* nx*ny*nz cells numbered i + j nx + k nx ny; kh_host = horizontal
 * permeability per cell [m^2]. */
int rsgpu_blackoil_create(rsgpu_ctx *ctx, const rsgpu_blackoil_params *p,
                          int nx, int ny, int nz, const double *kh_host,
                          rsgpu_blackoil **out);
int rsgpu_blackoil_destroy(rsgpu_blackoil *B);
/* The Jacobian's matrix: 3x3 blocks on the 7-point pattern
 * (rsgpu_csr_create_bsr3 layout), values zero. */
int rsgpu_blackoil_matrix(rsgpu_blackoil *B, rsgpu_csr **J_out);
/* J = dR/du and R(u) at state u_dev (3 per cell: p, Sw, Sg) with the
 * previous time level u_old_dev; J or residual_dev may be NULL. */
int rsgpu_blackoil_assemble(rsgpu_blackoil *B, const double *u_dev,
                            const double *u_old_dev, rsgpu_csr *J,
                            double *residual_dev);

/* ---- Batched ensemble solves (no reference counterpart) ------------------ */
/* The reference solves one system per BiCGstab.solve call.  A batch solves
 * many independent systems in the same launches: the systems are the
 * diagonal blocks of one block-diagonal matrix, and each keeps its own
 * BiCGstab scalars and DefaultIterationMonitor (BiCGstab.java:65-116,
 * DefaultIterationMonitor.java:133-156), so each stops on its own. */
typedef struct rsgpu_batch rsgpu_batch;

/* The block-diagonal matrix diag(parts[0], ..., parts[count-1]): pattern
 * and current values copied (row and column offsets added). */
int rsgpu_csr_block_diag(rsgpu_ctx *ctx, int count, rsgpu_csr *const *parts,
                         rsgpu_csr **out);

enum rsgpu_batch_flags {
    RSGPU_BATCH_NO_GRAPH = 1,     /* enqueue iterations directly instead of
                                     replaying a captured HIP graph */
    RSGPU_BATCH_NO_AUTOTUNE = 2   /* default SpMV stage sizes, no timing */
};
/* nsys systems; system k owns rows sys_start_host[k] ..
 * sys_start_host[k+1]-1 (strictly rising from 0).  opts: kind must be
 * RSGPU_BICGSTAB; max_iter/rtol/atol/dtol are each system's monitor.  With
 * opts->mixed each system runs FP64 refinement on its true residual
 * (IR.java:57-70) around an FP32 BiCGstab stopped at inner_rtol /
 * inner_max_iter, independently of the other systems.  check_every =
 * iterations per captured graph (the host reads one flag per replay). */
int rsgpu_batch_create(rsgpu_ctx *ctx, const rsgpu_solver_opts *opts, int nsys,
                       const int32_t *sys_start_host, int flags,
                       rsgpu_batch **out);
int rsgpu_batch_destroy(rsgpu_batch *B);
/* Solve every system of A x = b from the initial guess x (device vectors
 * of the batch's total size).  M: NULL, DIAGONAL, ILU0 or BLOCK_ILU0 whose
 * blocks do not cross system boundaries (rsgpu_prec_set_blocks), FP64 for
 * the plain form and FP32 for the mixed form; set on A.  stats_host (may be
 * NULL) receives nsys records.  Returns RSGPU_OK when every system
 * converged, else the status of the first system that did not. */
int rsgpu_batch_solve(rsgpu_batch *B, rsgpu_csr *A, rsgpu_prec *M,
                      const double *b_dev, double *x_dev,
                      rsgpu_stats *stats_host);
/* The SpMV stage sizes (nonzeros per workgroup) the autotuner picked for
 * the last matrix pattern, their timed durations (microseconds; 0 without
 * tuning), and the graph replays of the last solve. */
int rsgpu_batch_info(const rsgpu_batch *B, int *spmv_stage_fp64,
                     int *spmv_stage_fp32, double *spmv_us_fp64,
                     double *spmv_us_fp32, int *graph_launches);

/* ---- MPFA-O assembly (rs geometry/flux/O_MPFA.java, LinearSolver.java) --- */
/* Cartesian grid, dim 2 or 3 (nz ignored for dim 2), K6_host per cell
 * (xx,yy,zz,xy,xz,yz).  Computes O_MPFA.calculateTransmissibilities
 * (O_MPFA.java:54-112) on the device, one interaction region per corner
 * point, and returns the per-connection lists in the order rs stores them:
 * md_ptr_host[nconn+1], md_elem_host / md_t_host (caller allocates nconn*18
 * in 3-D, nconn*6 in 2-D).  check_fail_host counts Connection.check
 * failures (Connection.java:80-107). */
int rsgpu_mpfa_o(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz, double dx,
                 double dy, double dz, const double *K6_host,
                 double continuity, int32_t *md_ptr_host,
                 int32_t *md_elem_host, double *md_t_host,
                 int *check_fail_host);
/* MPFA-O + LinearSolver.determineSparsityPattern (LinearSolver.java:111-168)
 * + PotentialCalculator.discretizeFlux (PotentialCalculator.java:319-333)
 * fully on the device: creates A (pattern with the diagonal, values
 * assembled in connection order then MD order). */
int rsgpu_mpfa_o_assemble(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz,
                          double dx, double dy, double dz,
                          const double *K6_host, double continuity,
                          rsgpu_csr **A_out, int *check_fail_host);

/* ---- Fluxes (rs numerics/PotentialCalculator.java) ----------------------- */
/* A device flux operator from per-connection MD lists (as rsgpu_mpfa_o
 * returns them, MD order kept) and each connection's here/there cells. */
int rsgpu_flux_create(rsgpu_ctx *ctx, int nconn, int ncells,
                      const int32_t *md_ptr_host, const int32_t *md_elem_host,
                      const double *md_t_host, const int32_t *here_host,
                      const int32_t *there_host, rsgpu_flux **out);
/* MPFA-O computed on the device straight into a flux operator. */
int rsgpu_mpfa_o_flux(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz,
                      double dx, double dy, double dz, const double *K6_host,
                      double continuity, rsgpu_flux **out,
                      int *check_fail_host);
int rsgpu_flux_destroy(rsgpu_flux *F);
int rsgpu_flux_info(const rsgpu_flux *F, int *nconn, int *ncells,
                    int *nentries);
/* calculateFlux(mesh, potential) (PotentialCalculator.java:265-280):
 * flux[c] = sum over c.MD in MD order of t_k * p[k]. */
int rsgpu_flux_apply(rsgpu_flux *F, const double *p_dev, double *flux_dev);
/* Net outflow of the listed cells (well rates of pinned cells, Source.java
 * FIXED semantics): sum over each cell's connections in ascending order of
 * +flux (cell is "here") or -flux ("there"). */
int rsgpu_flux_rates(rsgpu_flux *F, const double *flux_dev, int ncells_sel,
                     const int32_t *cells_host, double *rates_host);
/* No reference counterpart.  rsgpu_flux_rates without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving or optimisation loops that evaluate well rates
 * every step, harnesses that forbid hipMalloc): the selected cells and their
 * rates are staged in a static device array of the library (768 KiB per
 * GPU, held until the process ends).  At most 65536 cells per call; more
 * gives RSGPU_BAD_ARGUMENT.  Results are bit-identical to
 * rsgpu_flux_rates.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (from different threads or contexts) race. */
int rsgpu_flux_rates_noalloc(rsgpu_flux *F, const double *flux_dev,
                             int ncells_sel, const int32_t *cells_host,
                             double *rates_host);

/* ---- IMPES pressure Jacobian (rs numerics/Discretisation.java, field/CS.java)
 * The step before the solve in rs's pressure Newton loop
 * (Discretisation.java:284-289): control-surface secondaries (CS.java:94-170),
 * then assemblePressureJacobianFlux (Discretisation.java:459-561) and
 * assemblePressureJacobianDiagonal with its source terms (:365-453), values
 * added into A's pattern (LinearSolver.addToJacobian / addToResidual,
 * LinearSolver.java:226-235) in the reference's call order, so A and b are
 * bit-identical to the Java loops' (A and b start from zero, as
 * LinearSolver.solve leaves them, :287-288).  The control-volume
 * secondaries (flash, PVT, rock-fluid) are inputs.  Phases are water, oil,
 * gas (Phase.all()); per-phase arrays are phase-major [3][n]. */
typedef struct rsgpu_impes rsgpu_impes;

typedef struct rsgpu_impes_params {
    int thermal;     /* 1: convective heat and Fourier terms */
    int pad;
    double dt;       /* time step */
    double g;        /* gravity acceleration (CS.bodyForce) */
    double omega;    /* centrifuge rotational speed */
    double ax, ay;   /* centrifuge axis */
} rsgpu_impes_params;

/* Device pointers (double unless noted), one element per cell. */
typedef struct rsgpu_impes_state {
    const double *pressure;        /* [n] CV.getPressure */
    const double *phase_pressure;  /* [3][n] CV.getPhasePressure */
    const double *saturation;      /* [3][n] CV.getSaturation */
    const double *mass_density;    /* [3][n] CV.getMassDensity */
    const double *phase_mobility;  /* [3][n] CV.getPhaseMobility */
    const uint8_t *present;        /* [3][n] CV.isPhasePresent */
    const double *comp_mobility;   /* [3][ncomp][n] CV.getComponentMobility */
    const double *dRdN;            /* [ncomp][n] getResidualVolumeDerivativeMolarMass */
    const double *dRdp;            /* [n] getResidualVolumeDerivativePressure */
    const double *heat_capacity;   /* [n] thermal only */
    const double *dRdT;            /* [n] thermal only */
    const double *energy_mobility; /* [3][n] thermal only */
    const double *temperature;     /* [n] thermal only */
    const double *p0, *R0;         /* pressure and residual volume stored at the
                                      step start (storeCurrentState) */
    const double *cx, *cy, *cz;    /* element centres (CS.bodyForce) */
    /* explicit molar-mass update and temperature system */
    const double *mole_fraction;   /* [3][ncomp][n] the phases' EOS compositions */
    const double *molar_density;   /* [3][n] EquationOfStateData.getMolarDensity */
    const double *T0;              /* [n] temperature stored at the step start */
} rsgpu_impes_state;

/* F: the mesh's MD lists and connections (rsgpu_flux_create or
 * rsgpu_mpfa_o_flux); Fmf: the Fourier (MF) lists of the same connections,
 * required for thermal assembly, may be NULL otherwise. */
int rsgpu_impes_create(rsgpu_flux *F, const rsgpu_flux *Fmf, int ncomp,
                       rsgpu_impes **out);
int rsgpu_impes_destroy(rsgpu_impes *E);
/* Source terms in the order rs applies them (field.sources(), then each
 * source's elements): kind 0 = regular (q[t][0][nu] is the component rate,
 * Discretisation.java:405-424), kind 1 = outlet (q[t][phase][nu], summed over
 * phases, :429-452); qe[t] the energy rate (thermal).  Host arrays; nterms 0
 * clears them. */
int rsgpu_impes_set_sources(rsgpu_impes *E, int nterms, const int32_t *elem_host,
                            const int32_t *kind_host, const double *q_host,
                            const double *qe_host);
/* CS.calculateSecondaries for every connection: rho[3][nconn] (saturation-
 * weighted density, 0 where it is NaN), upstream[3][nconn] (here when the
 * phase's potential sum is positive, else there), darcy[3][nconn]
 * (upstream mobility times that sum).  Device arrays. */
int rsgpu_impes_control_surfaces(rsgpu_impes *E, const rsgpu_impes_params *P,
                                 const rsgpu_impes_state *S, double *rho_dev,
                                 int32_t *upstream_dev, double *darcy_dev);
/* A's values (pattern of LinearSolver.determineSparsityPattern) and b: the
 * flux part, then the diagonal part and the source terms.  Returns
 * RSGPU_NOT_IN_PATTERN when an entry is missing from A's pattern (A's
 * values are then unspecified). */
int rsgpu_impes_assemble(rsgpu_impes *E, rsgpu_csr *A, const rsgpu_impes_params *P,
                         const rsgpu_impes_state *S, const double *rho_dev,
                         const int32_t *upstream_dev, double *b_dev);

/* The explicit molar-mass update (MolarMassExplicit.solve,
 * Discretisation.java:768-919): dN from the connections' component fluxes
 * (upstream mole fraction x molar density x Darcy flux of each present
 * phase, CS secondaries from rsgpu_impes_control_surfaces) and the source
 * terms, zero in the locked cells, N_new = N_old + dN ([ncomp][n] device
 * arrays).  A negative N_new is set to 0 when truncate != 0 (the count goes
 * to *truncated), else the call returns RSGPU_BAD_ARGUMENT with the first
 * element * ncomp + component in the reference's order in *bad. */
int rsgpu_impes_molar_update(rsgpu_impes *E, const rsgpu_impes_params *P,
                             const rsgpu_impes_state *S, const double *darcy_dev,
                             const int32_t *upstream_dev, int nlocked,
                             const int32_t *locked_host, int truncate,
                             const double *N_old_dev, double *N_new_dev, int *truncated,
                             int *bad);
/* The temperature system (TemperatureNewtonRaphson.solve,
 * Discretisation.java:577-590): advective flux (residual only), Fourier
 * flux (requires the MF lists), diagonal part and the energy source terms,
 * into A (LinearSolver's pattern) and b in the reference's call order. */
int rsgpu_impes_assemble_temperature(rsgpu_impes *E, rsgpu_csr *A,
                                     const rsgpu_impes_params *P,
                                     const rsgpu_impes_state *S, const double *rho_dev,
                                     const int32_t *upstream_dev, double *b_dev);

/* ---- Adjoint (no reference counterpart: rs has no gradients) ------------ */
/* tbar[q] = (w[here(c)] - w[there(c)]) * p[md_elem[q]]: gradient of
 * sum_i w_i * (net outflow of cell i) with respect to the MD
 * transmissibilities at fixed potential p (layout of F's MD lists). */
int rsgpu_flux_vjp(rsgpu_flux *F, const double *p_dev, const double *w_dev,
                   double *tbar_dev);
/* Gradient through O_MPFA.calculateTransmissibilities: for K_e =
 * exp(m_e) K0_e, mbar[e] = sum_q tbar[q] d t_q / d m_e (F supplies the MD
 * lists of the same grid, as rsgpu_mpfa_o_flux built them). */
int rsgpu_mpfa_o_vjp(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz,
                     double dx, double dy, double dz, const double *K6_host,
                     double continuity, const rsgpu_flux *F,
                     const double *tbar_dev, double *mbar_dev);

#ifdef __cplusplus
}
#endif
#endif /* RSGPU_H */
