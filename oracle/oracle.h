/*
 * CPU oracle: a scalar C restatement of the MTJ sparse Krylov path and of the
 * rs MPFA-O transmissibility / CSR assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (ressim_amd/)
 * links, loads or calls this code.  It is the checker that every GPU result
 * is compared against, and the timed "MTJ-equivalent C restatement" CPU
 * baseline.  It is compiled with -O2 -ffp-contract=off so that every sum is
 * evaluated in the reference's left-to-right statement order.
 *
 * Reference files restated (paths relative to the reference root):
 *   mtj/src/no/uib/cipr/matrix/sparse/{CompRowMatrix,BiCGstab,GMRES,ILU,
 *     UnitLowerCompRowMatrix,UpperCompRowMatrix,DefaultIterationMonitor,
 *     AbstractIterationMonitor,IR,Arrays,AMG,SSOR,CompColMatrix,ILUT,
 *     SparseVector}.java,
 *   mtj/src/no/uib/cipr/matrix/DenseLU.java (amg.c)
 *   mtj/src/no/uib/cipr/matrix/{DenseVector,GivensRotation,Matrices,
 *     AbstractTriangDenseMatrix,DenseMatrix}.java
 *   mtj/src/no/uib/cipr/matrix/distributed/BlockDiagonalPreconditioner.java
 *   rs/src/no/uib/cipr/rs/geometry/flux/{O_MPFA,MPFA}.java
 *   rs/src/no/uib/cipr/rs/geometry/Connection.java
 *   rs/src/no/uib/cipr/rs/meshgen/structured/{CartesianTopology,
 *     CartesianTopology2D,CartesianTopology3D,NaturalIJKOrdering,HexCell,
 *     QuadCell,BilinearMap,TrilinearMap,StructuredGeometry2D,
 *     StructuredGeometry3D}.java
 *   rs/src/no/uib/cipr/rs/numerics/{LinearSolver,PotentialCalculator}.java
 *   rs/src/no/uib/cipr/rs/numerics/Discretisation.java, rs/.../field/CS.java
 *     (impes.c: the IMPES pressure Jacobian)
 *
 * Dense LAPACK/BLAS calls of the reference (dgesv, dtrtrs, dgemm) are
 * restated from the published reference-BLAS / LAPACK 3.0 algorithms
 * (dgetf2 + dgetrs, column-oriented dtrsm, j-l-i dgemm).  Which LAPACK the
 * reference ran is not fixed by its sources, so bit-level agreement of MPFA
 * values with a JVM run is unpinned; the MPFA functions are pinned by the
 * invariants listed in tests/test_oracle_mpfa.py instead.
 */
#ifndef RS_ORACLE_H
#define RS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: identical to include/rsgpu.h. */
enum {
    OR_OK = 0,
    OR_BREAKDOWN = 1,
    OR_DIVERGENCE = 2,
    OR_ITERATIONS = 3,
    OR_ZERO_PIVOT = 10,
    OR_MISSING_DIAGONAL = 11,
    OR_NOT_IN_PATTERN = 12,
    OR_BAD_ARGUMENT = 13,
    OR_SINGULAR = 14
};

/* ---- CompRowMatrix ------------------------------------------------------ */
int or_csr_construct(int n, int ncols, const int *rowlen, const int *cols,
                     int *rowptr, int *colind);
int or_binary_search(const int *index, int key, int begin, int end);
int or_csr_add(const int *rowptr, const int *colind, double *data, int row,
               int col, double value);
int or_csr_set(const int *rowptr, const int *colind, double *data, int row,
               int col, double value);
int or_zero_rows(int n, const int *rowptr, const int *colind, double *data,
                 double diagonal, const int *rows, int nrows);
void or_csr_mult(int n, const int *rowptr, const int *colind,
                 const double *data, const double *x, double *y);
void or_csr_multadd(int n, const int *rowptr, const int *colind,
                    const double *data, double alpha, const double *x,
                    double *y);
void or_csr_transmult(int n, int ncols, const int *rowptr, const int *colind,
                      const double *data, const double *x, double *y);
void or_csr_transmultadd(int n, int ncols, const int *rowptr,
                         const int *colind, const double *data, double alpha,
                         const double *x, double *y);
void or_csr_mult_f32(int n, const int *rowptr, const int *colind,
                     const float *data, const float *x, float *y);

/* ---- DenseVector BLAS-1 -------------------------------------------------- */
double or_dot(int n, const double *x, const double *y);
double or_norm2(int n, const double *x);

/* ---- ILU(0) -------------------------------------------------------------- */
int or_ilu_factor(int n, const int *rowptr, const int *colind, double *lu,
                  int *diagind, int *bad_row);
int or_ilu_factor_f32(int n, const int *rowptr, const int *colind, float *lu,
                      int *diagind, int *bad_row);
void or_ilu_apply(int n, const int *rowptr, const int *colind,
                  const double *lu, const int *diagind, const double *b,
                  double *x, double *y);
void or_ilu_apply_f32(int n, const int *rowptr, const int *colind,
                      const float *lu, const int *diagind, const float *b,
                      float *x, float *y);
void or_ilu_transapply(int n, const int *rowptr, const int *colind,
                       const double *lu, const int *diagind, const double *b,
                       double *x, double *y);
int or_block_filter_count(int n, const int *rowptr, const int *colind,
                          int nblocks, const int *bstart);
int or_block_filter(int n, const int *rowptr, const int *colind,
                    const double *data, int nblocks, const int *bstart,
                    int *browptr, int *bcolind, double *bdata, int *src_index);

/* ---- Smoothed-aggregation AMG (amg.c) ------------------------------------ */
typedef struct {
    double omega_pre_f, omega_pre_r, omega_post_f, omega_post_r;
    int nu1, nu2, gamma, min;
    double omega;       /* Jacobi damping of the tentative prolongator */
    int reverse;        /* SSOR (1) or SOR (0) smoothing */
    int mode;           /* 0 sequential aggregation as AMG.java, 1 parallel */
    int block_rows;     /* smoother blocks per level (0: one block) */
    int max_coarse;     /* largest dense coarsest level accepted */
    int zero_guess;     /* 1: each cycle starts from zero instead of x */
    int ordering;       /* SSOR row order: 0 natural, 1 greedy colour */
} or_amg_params;
typedef struct or_amg or_amg;
void or_amg_params_default(or_amg_params *p);
int or_amg_setup(int n, const int *rowptr, const int *colind,
                 const double *data, const or_amg_params *p, or_amg **out,
                 int *bad_row);
void or_amg_apply(or_amg *A, const double *b, double *x, int transpose);
void or_amg_free(or_amg *A);
int or_amg_levels(const or_amg *A, int *sizes, int *nnzs, int cap);
int or_amg_get_level(const or_amg *A, int k, int *rowptr, int *colind,
                     double *data, int *pt, int *Prp, int *Pci, double *Pv);
int or_amg_p_nnz(const or_amg *A, int k);
int or_amg_aggregate(int n, const int *rowptr, const int *colind,
                     const double *data, double eps, int mode, int *pt,
                     int *nagg, int *nroots);
void or_hashset_order(const int *cols, int cnt, int *out);
uint64_t or_amg_priority(int i);
int or_ssor_apply(int n, const int *rowptr, const int *colind,
                  const double *data, int block_rows, int ordering,
                  double omegaF, double omegaR, int reverse, const double *b,
                  double *x, int *order_out);

/* ---- Krylov solvers ------------------------------------------------------ */
typedef struct {
    int max_iter;
    double rtol, atol, dtol;
    int restart;        /* GMRES(m) */
    /* preconditioner: 0 identity, 1 ILU(0), 2 block-Jacobi ILU(0), 3 AMG,
     * 4 diagonal, 5 SSOR (parameters in amg: omega_pre_f/r, reverse,
     * block_rows, ordering), 6 ICC */
    int prec_kind;
    int nblocks;        /* for block-Jacobi, contiguous equal row ranges */
    or_amg_params amg;  /* for AMG */
} or_opts;

typedef struct {
    int iterations;
    double residual;    /* last monitored residual */
    int status;
    int bad_row;
    int hist_len;
} or_stats;

int or_bicgstab(int n, const int *rowptr, const int *colind,
                const double *data, const double *b, double *x,
                const or_opts *opts, or_stats *stats, double *hist,
                int hist_cap);
int or_gmres(int n, const int *rowptr, const int *colind, const double *data,
             const double *b, double *x, const or_opts *opts, or_stats *stats,
             double *hist, int hist_cap);
int or_cg(int n, const int *rowptr, const int *colind, const double *data,
          const double *b, double *x, const or_opts *opts, or_stats *stats,
          double *hist, int hist_cap);
int or_ir(int n, const int *rowptr, const int *colind, const double *data,
          const double *b, double *x, const or_opts *opts, or_stats *stats,
          double *hist, int hist_cap);
int or_icc_factor(int n, const int *rowptr, const int *colind, double *data, int *diagind,
                  int *bad_row);
void or_icc_apply(int n, const int *rowptr, const int *colind, const double *R,
                  const int *diagind, const double *b, double *x, double *y);
int or_bicg(int n, const int *rowptr, const int *colind, const double *data,
            const double *b, double *x, const or_opts *opts, or_stats *stats,
            double *hist, int hist_cap);
int or_cgs(int n, const int *rowptr, const int *colind, const double *data,
           const double *b, double *x, const or_opts *opts, or_stats *stats,
           double *hist, int hist_cap);
int or_qmr(int n, const int *rowptr, const int *colind, const double *data,
           const double *b, double *x, const or_opts *opts, or_stats *stats,
           double *hist, int hist_cap);
int or_block_starts(int n, int nblocks, int *bstart);
/* ILUT (ilut.c): the factor in CSR with diagonal positions; apply and
 * transApply are or_ilu_apply / or_ilu_transapply on it (the same sweeps as
 * ILUT.java:320-443).  or_ilut_bound: capacity for ocolind / oval. */
int or_ilut_factor(int n, const int *rowptr, const int *colind, const double *val, double tau,
                   int p, int *orowptr, int *ocolind, double *oval, int *odiag, int *bad_row);
long long or_ilut_bound(int n, const int *rowptr, const int *colind, int p);
/* All-host-cores GMRES(m) + block-Jacobi ILU(0) over opts->nblocks equal
 * row ranges (mtj_omp.c; the CPU baseline's multi-threaded leg): nthreads
 * OpenMP threads (0 = OpenMP's default). */
int or_gmres_bjilu_omp(int n, const int *rowptr, const int *colind, const double *data,
                       const double *b, double *x, const or_opts *opts, or_stats *stats,
                       int nthreads);

/* ---- Dense LU (LAPACK 3.0 dgetf2/dgetrs restatement) --------------------- */
int or_dgesv(int n, int nrhs, double *a, int lda, int *ipiv, double *b,
             int ldb);

/* ---- Cartesian mesh + MPFA-O -------------------------------------------- */
/* dim 2: nz ignored (0).  K6 holds per cell (xx,yy,zz,xy,xz,yz).
 * Outputs per connection c: md_ptr[c]..md_ptr[c+1] in md_elem / md_t;
 * caller allocates md_ptr[nconn+1] and nconn*18 (3D) / nconn*6 (2D) slots.
 * here/there: element index on either side of each connection. */
int or_mesh_counts(int dim, int nx, int ny, int nz, int *ncells, int *nconn,
                   int *npoints);
int or_mesh_connections(int dim, int nx, int ny, int nz, int *here,
                        int *there);
int or_mpfa_o(int dim, int nx, int ny, int nz, double dx, double dy,
              double dz, const double *K6, double continuity, int *md_ptr,
              int *md_elem, double *md_t, int *check_fail);
int or_mpfa_corner(int dim, int nx, int ny, int nz, double dx, double dy,
                   double dz, const double *K6, double continuity, int point,
                   int *nc, int *ne, double *A, double *B, double *C,
                   double *D, double *T, int *elems);
int or_element_centers(int dim, int nx, int ny, int nz, double dx, double dy,
                       double dz, double *centers);
int or_interface_geometry(int dim, int nx, int ny, int nz, double dx,
                          double dy, double dz, double *normal_area,
                          double *center);

/* rs LinearSolver.determineSparsityPattern (with_diag=1) and
 * PotentialCalculator.determineSparsityPattern (with_diag=0). */
int or_pattern_count(int n, int nconn, const int *here, const int *there,
                     const int *md_ptr, const int *md_elem, int with_diag,
                     int *rowlen);
int or_pattern(int n, int nconn, const int *here, const int *there,
               const int *md_ptr, const int *md_elem, int with_diag,
               int *rowptr, int *colind);
/* PotentialCalculator.discretizeFlux: A[i,k] += t, A[j,k] -= t. */
int or_assemble_flux(int n, int nconn, const int *here, const int *there,
                     const int *md_ptr, const int *md_elem, const double *md_t,
                     const int *rowptr, const int *colind, double *data);
/* PotentialCalculator.calculateFlux: F_c = sum_k t_k p_k. */
void or_connection_flux(int nconn, const int *md_ptr, const int *md_elem,
                        const double *md_t, const double *p, double *flux);
int or_mpfa_o_vjp(int dim, int nx, int ny, int nz, double dx, double dy,
                  double dz, const double *K6, double continuity,
                  const int *md_ptr, const int *md_elem, const double *tbar,
                  double *mbar);
void or_cell_rates(int nconn, const int *here, const int *there,
                   const double *flux, int nsel, const int *cells,
                   double *rates);

/* ---- Fully-implicit black-oil Jacobian (blackoil.c) ---------------------- */
typedef struct {
    double dx, dy, dz;      /* cell sizes [m] */
    double dt;              /* time step [s] */
    double g;               /* gravity [m/s^2] */
    double phi0, cr;        /* porosity at pref, rock compressibility [1/Pa] */
    double pref;            /* reference pressure [Pa] */
    double kz_ratio;        /* k_z / k_h */
    double bw0, cw;         /* bw = bw0 (1 + cw (p - pref)) */
    double bo0, co;         /* bo = bo0 (1 + co (p - pref)) */
    double bg0;             /* bg = bg0 p / pref */
    double rs0;             /* Rs = rs0 p / pref */
    double muw, muo, mug;   /* viscosities [Pa s] */
    double rhow_s, rhoo_s, rhog_s; /* surface densities [kg/m^3] */
    double swr, sor, sgr;   /* Corey residual saturations */
    double kw, ko, kg;      /* Corey end-point relative permeabilities */
    int nw, no, ng, pad;    /* Corey exponents */
} or_bo_params;
int or_bo_pattern(int nx, int ny, int nz, int *browptr, int *bcolind);
void or_bo_assemble(const or_bo_params *P, int nx, int ny, int nz, const double *kh,
                    const double *u, const double *u_old, double *val, double *R);
int or_bo_expand(int nb, const int *browptr, const int *bcolind, int *rowptr,
                 int *colind);

/* ---- IMPES pressure Jacobian (impes.c; rs numerics/Discretisation.java,
 * field/CS.java) ---------------------------------------------------------- */
/* Field names and layouts identical to rsgpu_impes_params / _state. */
typedef struct or_impes_params {
    int n, nconn, ncomp, thermal;
    double dt, g, omega, ax, ay;
} or_impes_params;

typedef struct or_impes_state {
    const double *pressure;        /* [n] CV.getPressure */
    const double *phase_pressure;  /* [3][n] water, oil, gas */
    const double *saturation;      /* [3][n] */
    const double *mass_density;    /* [3][n] */
    const double *phase_mobility;  /* [3][n] */
    const uint8_t *present;        /* [3][n] */
    const double *comp_mobility;   /* [3][ncomp][n] */
    const double *dRdN;            /* [ncomp][n] */
    const double *dRdp;            /* [n] */
    const double *heat_capacity;   /* [n] thermal only */
    const double *dRdT;            /* [n] thermal only */
    const double *energy_mobility; /* [3][n] thermal only */
    const double *temperature;     /* [n] thermal only */
    const double *p0, *R0;         /* pressure and residual volume stored at the step start */
    const double *cx, *cy, *cz;    /* element centres */
    const double *mole_fraction;   /* [3][ncomp][n] */
    const double *molar_density;   /* [3][n] */
    const double *T0;              /* [n] */
} or_impes_state;

void or_impes_cs(const or_impes_params *P, const or_impes_state *S, const int *here,
                 const int *there, const int *md_ptr, const int *md_elem, const double *md_t,
                 double *rho, int *upstream, double *darcy);
int or_impes_assemble(const or_impes_params *P, const or_impes_state *S, const int *here,
                      const int *there, const int *md_ptr, const int *md_elem,
                      const double *md_t, const int *mf_ptr, const int *mf_elem,
                      const double *mf_t, const double *rho, const int *upstream, int nterms,
                      const int *src_elem, const int *src_kind, const double *src_q,
                      const double *src_qe, const int *rowptr, const int *colind, double *val,
                      double *b);
int or_impes_molar_update(const or_impes_params *P, const or_impes_state *S, const int *here,
                          const int *there, const double *darcy, const int *upstream,
                          int nterms, const int *src_elem, const int *src_kind,
                          const double *src_q, const unsigned char *locked, int truncate,
                          const double *N_old, double *N_new, int *truncated, int *bad);
int or_impes_temperature(const or_impes_params *P, const or_impes_state *S, const int *here,
                         const int *there, const int *md_ptr, const int *md_elem,
                         const double *md_t, const int *mf_ptr, const int *mf_elem,
                         const double *mf_t, const double *rho, const int *upstream,
                         int nterms, const double *src_qe, const int *src_elem,
                         const int *rowptr, const int *colind, double *val, double *b);

#ifdef __cplusplus
}
#endif
#endif
