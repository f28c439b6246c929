/*
 * ref_driver.c -- drives the reference's own HUTI CG/BiCGStab, compiled from
 * the reference tree into oracle/_ref (see oracle/Makefile, target `ref`).
 *
 * TEST INFRASTRUCTURE ONLY.  It reproduces what IterSolver hands to HUTI
 * (fem/src/IterSolve.src:318-366, 630-669): ipar/dpar layout of
 * hutiter/src/huti_fdefs.h, user-supplied initial x, TRESID_SCALED_BYB,
 * matvec = CRS_MatrixVectorProd, pcondl = CRS_DiagPrecondition or the
 * dummy copy, pcondr = dummy copy (huti_sfe_c.src:117-124), and the
 * reference BLAS ddot/dnrm2 for dot and norm.  The dummy copy is the
 * reference's own huti_ddummy_pcondfun (hutiter/src/huti_aux_D.F90, built
 * beside the solvers, which also supplies huti_drandvec).  The matvec,
 * diagonal-preconditioner and ILU-preconditioner callbacks are this file's
 * restatement (the Fortran CRS module needs Elmer's generated build
 * configuration and is not built here); the Krylov loops, dot, norm and
 * dummy copy that run are the reference's compiled sources.  For ILU
 * (precond >= 2) the caller hands over the factors (ref_set_ilu) and the
 * preconditioner is CRS_LUPrecondition: u = v, then CRS_LUSolve
 * (fem/src/CRSMatrix.src:3259-3268, 3356-3374).
 *
 * The residual history is read off the norm callback: with stopc = 1 HUTI
 * calls normfun once for ||b|| and once per iteration for ||Ax - b||.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

extern void huti_dcgsolv_(int *, int *, double *, double *, int *, double *, double *, void (*)(),
                          void (*)(), void (*)(), double (*)(), double (*)(), double (*)());
extern void huti_dbicgstabsolv_(int *, int *, double *, double *, int *, double *, double *,
                                void (*)(), void (*)(), void (*)(), double (*)(), double (*)(),
                                double (*)());
extern double ddot_(int *, double *, int *, double *, int *);
extern double dnrm2_(int *, double *, int *);
extern void huti_ddummy_pcondfun_(double *u, double *v, int *ipar);

static struct {
  int n;
  const int *rows, *cols, *diag;
  const double *values;
  int precond;
  double *norms;
  int nnorm, cap;
  const int *irows, *icols, *idiag;
  const double *ilu;
} G;

/* Factors for precond >= 2 (1-based ILU pattern and values, U(i,i) inverted). */
void ref_set_ilu(const int *irows, const int *icols, const int *idiag, const double *ilu) {
  G.irows = irows;
  G.icols = icols;
  G.idiag = idiag;
  G.ilu = ilu;
}

static void cb_matvec(double *u, double *v, int *ipar) {
  (void)ipar;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < G.n; i++) {
    double s = 0.0;
    for (int j = G.rows[i]; j < G.rows[i + 1]; j++) s = s + G.values[j - 1] * u[G.cols[j - 1] - 1];
    v[i] = s;
  }
}

static void cb_pcond(double *u, double *v, int *ipar) {
  if (!G.precond) {
    huti_ddummy_pcondfun_(u, v, ipar);
    return;
  }
  if (G.precond >= 2) {
    for (int i = 0; i < G.n; i++) u[i] = v[i];
    for (int i = 1; i <= G.n; i++) {
      double s = u[i - 1];
      for (int j = G.irows[i - 1]; j < G.idiag[i - 1]; j++) s = s - G.ilu[j - 1] * u[G.icols[j - 1] - 1];
      u[i - 1] = s;
    }
    for (int i = G.n; i >= 1; i--) {
      double s = u[i - 1];
      for (int j = G.idiag[i - 1] + 1; j < G.irows[i]; j++) s = s - G.ilu[j - 1] * u[G.icols[j - 1] - 1];
      u[i - 1] = G.ilu[G.idiag[i - 1] - 1] * s;
    }
    return;
  }
  const double AEPS = 10.0 * DBL_EPSILON;
  for (int i = 0; i < G.n; i++) {
    double d = G.values[G.diag[i] - 1];
    u[i] = (G.precond && fabs(d) > AEPS) ? v[i] / d : v[i];
  }
}

static double cb_norm(int *n, double *x, int *inc) {
  double r = dnrm2_(n, x, inc);
  if (G.nnorm < G.cap) G.norms[G.nnorm] = r;
  G.nnorm++;
  return r;
}

static double cb_stopc(void) { return 0.0; }

/* Returns HUTI_INFO; *iters = HUTI_ITERS; hist[k] = residual after
 * iteration k+1 (for stopc = 1), length up to maxit. */
int ref_huti_solve(int method, int precond, int n, const int *rows, const int *cols,
                   const int *diag, const double *values, const double *b, double *x, double tol,
                   int maxit, int *iters, double *hist) {
  int ipar[50];
  double dpar[10];
  memset(ipar, 0, sizeof(ipar));
  memset(dpar, 0, sizeof(dpar));
  int wrk = method == 0 ? 4 : 8;
  ipar[2] = n;       /* HUTI_NDIM */
  ipar[3] = wrk;     /* HUTI_WRKDIM */
  ipar[4] = 0;       /* HUTI_DBUGLVL */
  ipar[9] = maxit;   /* HUTI_MAXIT */
  ipar[11] = 1;      /* HUTI_STOPC = TRESID_SCALED_BYB */
  ipar[13] = 1;      /* HUTI_INITIALX = USERSUPPLIEDX */
  dpar[0] = tol;     /* HUTI_TOLERANCE */
  double *work = calloc((size_t)n * wrk, sizeof(double));
  double *bb = malloc(sizeof(double) * n);
  memcpy(bb, b, sizeof(double) * n);
  G.n = n; G.rows = rows; G.cols = cols; G.diag = diag; G.values = values; G.precond = precond;
  G.cap = 2 * maxit + 4;
  G.norms = malloc(sizeof(double) * G.cap);
  G.nnorm = 0;
  int nd = n, wd = wrk;
  if (method == 0)
    huti_dcgsolv_(&nd, &wd, x, bb, ipar, dpar, work, (void (*)())cb_matvec, (void (*)())cb_pcond,
                  (void (*)())huti_ddummy_pcondfun_, (double (*)())ddot_, (double (*)())cb_norm,
                  (double (*)())cb_stopc);
  else
    huti_dbicgstabsolv_(&nd, &wd, x, bb, ipar, dpar, work, (void (*)())cb_matvec,
                        (void (*)())cb_pcond, (void (*)())huti_ddummy_pcondfun_, (double (*)())ddot_,
                        (double (*)())cb_norm, (double (*)())cb_stopc);
  *iters = ipar[30];
  if (G.nnorm > G.cap) G.nnorm = G.cap;
  if (hist) {
    /* CG: norms = [||b||, r1, r2, ...].  BiCGStab: per iteration ||s|| then
       ||Ax-b||, so residuals sit at the even positions after ||b||. */
    if (method == 0) {
      for (int k = 1; k < G.nnorm && k - 1 < maxit; k++) hist[k - 1] = G.norms[k] / G.norms[0];
    } else {
      for (int k = 2, it = 0; k < G.nnorm && it < maxit; k += 2, it++) hist[it] = G.norms[k] / G.norms[0];
    }
  }
  int info = ipar[29];
  free(work);
  free(bb);
  free(G.norms);
  return info;
}

/* The reference BLAS routines themselves, for direct known-answer checks. */
double ref_ddot(int n, const double *x, const double *y) {
  int one = 1;
  return ddot_(&n, (double *)x, &one, (double *)y, &one);
}
double ref_dnrm2(int n, const double *x) {
  int one = 1;
  return dnrm2_(&n, (double *)x, &one);
}
