/*
 * oracle.h -- CPU restatement of Elmer's assemble-then-iterate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (elmerfem_amd, libehip.so) never links or calls it.
 *
 * Every function restates one routine of the reference (Elmer 7.0); the
 * citation (file:line, relative to the reference tree) is given beside it
 * in oracle.c.  Arrays are 1-based Fortran-CRS like Matrix_t
 * (fem/src/Types.src:155-228): rows[n+1], cols[nnz], diag[n] hold 1-based
 * indices; values[nnz] and rhs[n] are double.  Dense local matrices are
 * column-major (Fortran order): a[(j-1)*ld + (i-1)] is a(i,j).
 */
#ifndef ELMER_ORACLE_H
#define ELMER_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- element types (fem/src/elements.def) ---- */
#define OR_LINE2 202
#define OR_TRI3 303
#define OR_QUAD4 404
#define OR_QUAD9 409
#define OR_TET4 504
#define OR_TET10 510
#define OR_HEX8 808
#define OR_HEX27 827

/* Gauss-Legendre 1-D rule as ComputeGaussPoints1D builds it (n <= 3). */
int or_gauss1d(int n, double *points, double *weights);
/* Default rule of an element type (GaussPoints, Integration.src:1490). */
int or_gauss_points(int code, int np, double *u, double *v, double *w, double *s);
int or_default_gauss_n(int code);
int or_element_nodes(int code);

/* ElementInfo (ElementDescription.src:2586): basis, global gradients
 * (dbasisdx[i*3+j], row-major n x 3) and sqrt(det G).  x/y/z are the element
 * node coordinates.  cdim is the mesh coordinate dimension.  Returns 1 on
 * success, 0 for a degenerate element (ElementMetric returns FALSE). */
int or_element_info(int code, int cdim, const double *x, const double *y,
                    const double *z, double u, double v, double w,
                    double *basis, double *dbasisdx, double *detj);

/* ModelPDE bulk local matrix (modules/ModelPDE.src:75-166).  Per-node
 * coefficients (length n); velo is 3 x n column-major (velo[(j)*3+i]).
 * stiff is n x n column-major, force length n. */
int or_local_modelpde(int code, int cdim, const double *x, const double *y,
                      const double *z, const double *diff,
                      const double *conv, const double *react,
                      const double *velo, const double *load,
                      double *stiff, double *force);

/* StressCompose, 3-D isotropic, no mass/heat/prestress (Stress.src:59-521).
 * stiff is 3n x 3n column-major; force (3n) is zero. */
int or_local_elasticity(int code, const double *x, const double *y,
                        const double *z, const double *young,
                        const double *poisson, double *stiff, double *force);

/* StressBoundary traction part (Stress.src:814-1020): force(3n) from the
 * nodal loads load3 (3 x n column-major). */
int or_local_traction(int code, const double *x, const double *y,
                      const double *z, const double *load3, double *force);

/* ---- CRS topology (CreateMatrix chain, ElementUtils.src:698-935) ---- */
typedef struct or_pattern or_pattern;
/* elements: nelem x maxnpe 1-based node indices (row-major), npe[e] nodes
 * each.  All mesh nodes must be used (InitialPermutation = identity).
 * optimize != 0 runs OptimizeBandwidth (BandwidthOptimize.src:108-355). */
or_pattern *or_pattern_build(int nnodes, int nelem, int maxnpe,
                             const int *elements, const int *npe, int dofs,
                             int optimize);
int or_pattern_n(const or_pattern *p);
long long or_pattern_nnz(const or_pattern *p);
int or_pattern_bandwidth(const or_pattern *p, int *before, int *after,
                         int *accepted);
/* perm[nnodes] (1-based), rows[n+1], cols[nnz], diag[n] */
void or_pattern_copy(const or_pattern *p, int *perm, int *rows, int *cols,
                     int *diag);
void or_pattern_free(or_pattern *p);

/* CRS_GlueLocalMatrix + UpdateGlobalEquations RHS part
 * (CRSMatrix.src:507-590, SolverUtils.src:984-1053), one element.
 * idx: the element's node indices already mapped through Perm (1-based). */
int or_glue_local(int n, const int *rows, const int *cols, const int *diag,
                  int nen, int dofs, const int *idx, const double *local,
                  const double *force, double *values, double *rhs);

/* Whole-mesh assembly driven through the element loop in element order.
 * kind 0: ModelPDE (elem_coef = diffusion, constant per element, with
 *         conv (per element, 1 or 0), velo per element (3), react 0,
 *         load = source per element);
 * kind 1: elasticity (elem_coef = Young, poisson scalar).
 * coords: nnodes x 3 row-major. */
int or_assemble(int kind, int code, int cdim, int nnodes,
                const double *coords, int nelem, int npe, const int *elements,
                const double *elem_coef, const double *elem_conv,
                const double *elem_velo, const double *elem_load,
                double poisson, const int *perm, int n, const int *rows,
                const int *cols, const int *diag, int dofs, double *values,
                double *rhs);

/* Traction on boundary elements (kind 1 only): load (3) constant. */
int or_assemble_traction(int code, int nnodes, const double *coords,
                         int nbelem, int npe, const int *belements,
                         const double *load3, const int *perm, int dofs,
                         double *rhs);

/* DefaultDirichletBCs (DefUtils.src:3530-4030) for nodal Dirichlet rows:
 * scale (RHSscaling=F), per row ZeroRow+diag 1+b=val/D (or
 * CRS_SetSymmDirichlet when symmetric), back-scale.  bc_rows are 1-based
 * matrix rows in application order. */
int or_dirichlet(int n, const int *rows, const int *cols, const int *diag,
                 double *values, double *rhs, int nbc, const int *bc_rows,
                 const double *bc_vals, int symmetric, int scaling);

/* ---- linear algebra (the iterate half) ---- */
double or_ddot(int n, const double *x, const double *y);   /* blas/ddot.f */
double or_dnrm2(int n, const double *x);                  /* blas/dnrm2.f */
void or_spmv(int n, const int *rows, const int *cols, const double *values,
             const double *u, double *v);  /* CRS_MatrixVectorProd */

/* ---- ILU(n) preconditioner (CRS_IncompleteLU, CRSMatrix.src:2157-2500;
 * CRS_LUSolve :3293-3397), real non-Cholesky branch ---- */
/* Pattern of ILU(order): malloc'ed 1-based irows[n+1], icols, idiag[n]
 * (release with or_free).  Returns 0 on success. */
int or_ilu_pattern(int n, const int *rows, const int *cols, const int *diag,
                   int order, int **irows, int **icols, int **idiag);
void or_free(void *p);
/* Factors of values (pattern rows/cols) on the ILU pattern: unit lower L
 * strictly below idiag, U on and above it with U(i,i) stored inverted. */
int or_ilu_factor(int n, const int *rows, const int *cols,
                  const double *values, const int *irows, const int *icols,
                  const int *idiag, double *ilu);
/* b <- (LU)^-1 b */
void or_lu_solve(int n, const int *irows, const int *icols, const int *idiag,
                 const double *ilu, double *b);

/* method: 0 = CG, 1 = BiCGStab.  precond: 0 none, 1 diagonal, 2 + k ILU(k)
 * for k = 0..9 (the factors of the given values are computed first, as
 * IterSolver does).
 * stopc: HUTI stopping criterion code (1 = TRESID_SCALED_BYB).
 * hist (length maxit, may be NULL) receives the residual of each
 * iteration.  Returns HUTI_INFO, *iters = HUTI_ITERS. */
int or_huti_solve(int method, int precond, int n, const int *rows,
                  const int *cols, const int *diag, const double *values,
                  const double *b, double *x, double tol, int maxit,
                  int stopc, int *iters, double *hist);

/* SolveLinearSystem (SolverUtils.src:7151-7525) for the iterative branch:
 * zero-RHS exit, ScaleLinearSystem, x=1e-8, IterSolver, back-scale, NRM.
 * values are modified in place and restored by the back-scale round trip
 * like the reference.  Returns HUTI_INFO (or 1 with iters 0 when b = 0). */
int or_solve_linear_system(int method, int precond, int scaling, int n,
                           const int *rows, const int *cols, const int *diag,
                           double *values, double *b, double *x, double tol,
                           int maxit, int *iters, double *hist, double *norm);

/* Threads of or_spmv's row loop (0: keep); returns the count in use. */
int or_set_threads(int nthreads);

/* ComputeNorm, NormDim 2 (SolverUtils.src:4887-5006). */
double or_compute_norm(int n, const double *x);

#ifdef __cplusplus
}
#endif
#endif
