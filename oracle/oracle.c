/*
 * oracle.c -- CPU restatement of Elmer 7.0's assemble-then-iterate path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never links or calls it.
 *
 * Each routine cites the reference routine (path:line relative to the
 * reference tree) whose arithmetic order it follows.  Build with
 * -ffp-contract=off: the reference is compiled for baseline x86-64 where no
 * fused multiply-add is formed, and the oracle keeps the same roundings.
 *
 * Pinning (tests/test_oracle_*.py):
 *  - or_gauss1d / or_gauss_points: closed-form Gauss-Legendre nodes and
 *    weights, exactness on polynomials, tet weights sum to 1/6.
 *  - or_element_info: partition of unity, Kronecker property at nodes,
 *    sum of gradients = 0, exact gradients of affine fields, detJ = volume.
 *  - or_local_*: closed-form unit-cube Hex8 Laplacian, affine Tet4
 *    stiffness, symmetry/zero row sums, elasticity rigid-body null space,
 *    integral of the source.
 *  - or_pattern_*: closed-form nnz for structured meshes, brute-force
 *    adjacency, permutation and bandwidth invariants of OptimizeBandwidth.
 *  - or_glue_local / or_assemble: brute-force dense assembly in numpy.
 *  - or_dirichlet / or_solve_linear_system: dense replay in numpy.
 *  - or_huti_solve, or_ddot, or_dnrm2: bit-exact against the reference's own
 *    HUTI CG/BiCGStab and BLAS sources compiled with flang into oracle/_ref
 *    (iteration counts, residual histories and solutions), and against
 *    scipy/numpy dense solves.
 */
#include "oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* ------------------------------------------------------------------ */
/* Gauss points                                                         */
/* ------------------------------------------------------------------ */

/* EvalPoly / DerivPoly / RefineRoots helpers of ComputeGaussPoints1D
 * (Integration.src:406-570). P holds coefficients highest degree first. */
static double eval_poly(int n, const double *P, double x) {
  double s = 0.0;
  for (int i = 0; i < n + 1; i++) s = s * x + P[i];
  return s;
}
static void deriv_poly(int n, double *Q, const double *P) {
  for (int i = 0; i < n; i++) Q[i] = P[i] * (double)(n - i);
}

/* ComputeGaussPoints1D (Integration.src:406-570) for n <= 3, where the
 * companion matrix handed to DGEEV is 1x1 and its eigenvalue is its only
 * entry, so the whole routine restates without LAPACK. */
int or_gauss1d(int n, double *Points, double *Weights) {
  if (n <= 1) {
    Points[0] = 0.0;
    Weights[0] = 2.0;
    return 0;
  }
  if (n > 3) return -1;
  double P[8] = {0}, P0[8] = {0}, P1[8] = {0}, Q[8] = {0};
  /* Legendre recurrence (i+1)P_{i+1} = (2i+1) x P_i - i P_{i-1} */
  P0[0] = 1;
  P1[0] = 1;
  P1[1] = 0;
  for (int i = 1; i <= n - 1; i++) {
    for (int k = 0; k < i + 1; k++) P[k] = (double)(2 * i + 1) * P1[k] / (double)(i + 1);
    for (int k = 2; k < i + 2; k++) P[k] = P[k] - (double)i * P0[k - 2] / (double)(i + 1);
    for (int k = 0; k < i + 1; k++) P0[k] = P1[k];
    for (int k = 0; k < i + 2; k++) P1[k] = P[k];
  }
  int np = n - (n % 2);
  np = np / 2;
  for (int i = 1; i <= np + 1; i++) P[i - 1] = P[2 * i - 2];
  /* np == 1: A(1,1) = -P(2)/P(1) is the eigenvalue */
  double eig = -P[1] / P[0];
  double Qs[8];
  for (int i = 0; i < np + 1; i++) Qs[i] = P[i];
  memset(P, 0, sizeof(P));
  for (int i = 1; i <= np + 1; i++) P[2 * i - 2] = Qs[i - 1];
  Points[0] = +sqrt(eig);
  Points[1] = -sqrt(eig);
  if (n % 2 == 1) Points[n - 1] = 0.0;
  deriv_poly(n, Q, P);
  /* RefineRoots */
  for (int i = 0; i < n; i++) {
    double x = Points[i], s;
    for (int j = 0; j < 100; j++) {
      s = eval_poly(n, P, x) / eval_poly(n - 1, Q, x);
      x = x - s;
      if (fabs(s) <= fabs(x) * DBL_EPSILON) break;
    }
    if (fabs(eval_poly(n, P, x)) < fabs(eval_poly(n, P, Points[i])))
      if (fabs(x - Points[i]) < 1.0e-8) Points[i] = x;
  }
  deriv_poly(n, Q, P);
  double sum = 0.0;
  for (int i = 0; i < n; i++) {
    double s = eval_poly(n - 1, Q, Points[i]);
    Weights[i] = 2.0 / ((1.0 - Points[i] * Points[i]) * (s * s));
  }
  for (int i = 0; i < n; i++) sum += Weights[i];
  for (int i = 0; i < n; i++) Weights[i] = 2.0 * Weights[i] / sum;
  return 0;
}

/* Tetrahedron 4-point rule table (Integration.src:281-295). */
static const double UT4[4] = {0.1757281246520584, 0.2445310270213291,
                              0.5556470949048655, 0.0240937534217468};
static const double VT4[4] = {0.5656137776620919, 0.0501800797762026,
                              0.1487681308666864, 0.2354380116950194};
static const double WT4[4] = {0.2180665126782654, 0.5635595064952189,
                              0.0350112499848832, 0.1833627308416330};

int or_default_gauss_n(int code) {
  switch (code) {
    case OR_LINE2: return 2;
    case OR_TRI3: return 3;
    case OR_QUAD4: return 4;
    case OR_QUAD9: return 9;
    case OR_TET4: return 4;
    case OR_TET10: return 27;
    case OR_HEX8: return 8;
    case OR_HEX27: return 27;
    default: return -1;
  }
}

/* GaussPointsQuad/Brick/Tetra (Integration.src:1355-1485, 896-985). */
int or_gauss_points(int code, int np, double *u, double *v, double *w, double *s) {
  double pts[3], wts[3];
  int fam = code / 100;
  if (fam == 2) {
    if (or_gauss1d(np, pts, wts)) return -1;
    for (int i = 0; i < np; i++) { u[i] = pts[i]; v[i] = 0; w[i] = 0; s[i] = wts[i]; }
    return np;
  }
  if (fam == 3 && np == 3) {
    /* GaussPointsTriangle, 3 points (Integration.src:84-92, 776-781) */
    static const double u3[3] = {0.16666666666666667, 0.66666666666666667, 0.16666666666666667};
    static const double v3[3] = {0.16666666666666667, 0.16666666666666667, 0.66666666666666667};
    for (int i = 0; i < 3; i++) { u[i] = u3[i]; v[i] = v3[i]; w[i] = 0.0; s[i] = 0.33333333333333333 / 2.0; }
    return 3;
  }
  if (fam == 4) {
    int n = (int)(sqrt((float)np) + 0.5f);
    if (or_gauss1d(n, pts, wts)) return -1;
    int t = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        u[t] = pts[j]; v[t] = pts[i]; w[t] = 0.0;
        s[t] = wts[i] * wts[j];
        t++;
      }
    return t;
  }
  if (fam == 8 || (fam == 5 && np != 4)) {
    int n = (int)(pow((double)np, 1.0 / 3.0) + 0.5);
    if (or_gauss1d(n, pts, wts)) return -1;
    int t = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
          u[t] = pts[k]; v[t] = pts[j]; w[t] = pts[i];
          s[t] = wts[i] * wts[j] * wts[k];
          t++;
        }
    if (fam == 5) { /* collapsed brick (Integration.src:944-963) */
      for (int i = 0; i < t; i++) {
        double sf = 0.5;
        u[i] = (u[i] + 1) * sf;
        v[i] = (v[i] + 1) * sf;
        w[i] = (w[i] + 1) * sf;
        s[i] = s[i] * (sf * sf * sf);
        sf = 1.0 - w[i];
        u[i] = u[i] * sf;
        v[i] = v[i] * sf;
        s[i] = s[i] * (sf * sf);
        sf = 1.0 - v[i] / sf;
        u[i] = u[i] * sf;
        s[i] = s[i] * sf;
      }
    }
    return t;
  }
  if (fam == 5 && np == 4) {
    for (int i = 0; i < 4; i++) { u[i] = UT4[i]; v[i] = VT4[i]; w[i] = WT4[i]; s[i] = 0.25 / 6.0; }
    return 4;
  }
  return -1;
}

/* ------------------------------------------------------------------ */
/* Element descriptions: monomial basis from the inverse Vandermonde     */
/* (AddElementDescription, ElementDescription.src:73-305 on elements.def) */
/* ------------------------------------------------------------------ */

typedef struct {
  int code, dim, n;
  double U[27], V[27], W[27];
  int p[27], q[27], r[27];
  double coeff[27 * 27]; /* coeff[term*n + node] */
  int ready;
} etype_t;

static etype_t g_types[9];

static const double HEX27_U[27] = {-1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 0, 0, 0};
static const double HEX27_V[27] = {-1, -1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, 0, 0, 0};
static const double HEX27_W[27] = {-1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, -1, 1, 0};
static const int HEX27_B[27] = {1, 2, 3, 5, 6, 7, 9, 10, 11, 17, 18, 19, 21, 22, 23, 25, 26, 27, 33, 34, 35, 37, 38, 39, 41, 42, 43};
static const double QUAD9_U[9] = {-1, 1, 1, -1, 0, 1, 0, -1, 0};
static const double QUAD9_V[9] = {-1, -1, 1, 1, -1, 0, 1, 0, 0};
static const int QUAD9_B[9] = {1, 2, 3, 5, 6, 7, 9, 10, 11};
static const double TET10_U[10] = {0, 1, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0};
static const double TET10_V[10] = {0, 0, 1, 0, 0, 0.5, 0.5, 0, 0, 0.5};
static const double TET10_W[10] = {0, 0, 0, 1, 0, 0, 0, 0.5, 0.5, 0.5};
static const int TET10_B[10] = {1, 2, 3, 5, 6, 9, 17, 18, 21, 33};

/* Gauss-Jordan inverse with partial pivoting (in place, column-major). */
static int invert(int n, double *a) {
  int piv[27];
  double b[27 * 27];
  memset(b, 0, sizeof(double) * n * n);
  for (int i = 0; i < n; i++) b[i * n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int pr = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(a[c * n + r]) > fabs(a[c * n + pr])) pr = r;
    if (a[c * n + pr] == 0.0) return -1;
    piv[c] = pr;
    if (pr != c)
      for (int k = 0; k < n; k++) {
        double t = a[k * n + c]; a[k * n + c] = a[k * n + pr]; a[k * n + pr] = t;
        t = b[k * n + c]; b[k * n + c] = b[k * n + pr]; b[k * n + pr] = t;
      }
    double d = a[c * n + c];
    for (int k = 0; k < n; k++) { a[k * n + c] /= d; b[k * n + c] /= d; }
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      double f = a[c * n + r];
      if (f == 0.0) continue;
      for (int k = 0; k < n; k++) {
        a[k * n + r] -= f * a[k * n + c];
        b[k * n + r] -= f * b[k * n + c];
      }
    }
  }
  (void)piv;
  memcpy(a, b, sizeof(double) * n * n);
  return 0;
}

static double ipow(double x, int p) {
  double r = 1.0;
  for (int i = 0; i < p; i++) r *= x;
  return r;
}

static etype_t *get_type(int code) {
  int slot;
  switch (code) {
    case OR_LINE2: slot = 0; break;
    case OR_QUAD4: slot = 1; break;
    case OR_QUAD9: slot = 2; break;
    case OR_TET4: slot = 3; break;
    case OR_TET10: slot = 4; break;
    case OR_HEX8: slot = 5; break;
    case OR_HEX27: slot = 6; break;
    case OR_TRI3: slot = 7; break;
    default: return NULL;
  }
  etype_t *t = &g_types[slot];
  if (t->ready) return t;
  memset(t, 0, sizeof(*t));
  t->code = code;
  const int *B = NULL;
  static const int B202[2] = {1, 2};
  static const int B404[4] = {1, 2, 5, 6};
  static const int B504[4] = {1, 2, 5, 17};
  static const int B808[8] = {1, 2, 5, 6, 17, 18, 21, 22};
  switch (code) {
    case OR_LINE2:
      t->dim = 1; t->n = 2; B = B202;
      t->U[0] = -1; t->U[1] = 1;
      break;
    case OR_TRI3: {
      static const int B303[3] = {1, 2, 5};
      t->dim = 2; t->n = 3; B = B303;
      t->U[1] = 1; t->V[2] = 1;
    } break;
    case OR_QUAD4: {
      static const double u[4] = {-1, 1, 1, -1}, v[4] = {-1, -1, 1, 1};
      t->dim = 2; t->n = 4; B = B404;
      for (int i = 0; i < 4; i++) { t->U[i] = u[i]; t->V[i] = v[i]; }
    } break;
    case OR_QUAD9:
      t->dim = 2; t->n = 9; B = QUAD9_B;
      for (int i = 0; i < 9; i++) { t->U[i] = QUAD9_U[i]; t->V[i] = QUAD9_V[i]; }
      break;
    case OR_TET4: {
      static const double u[4] = {0, 1, 0, 0}, v[4] = {0, 0, 1, 0}, w[4] = {0, 0, 0, 1};
      t->dim = 3; t->n = 4; B = B504;
      for (int i = 0; i < 4; i++) { t->U[i] = u[i]; t->V[i] = v[i]; t->W[i] = w[i]; }
    } break;
    case OR_TET10:
      t->dim = 3; t->n = 10; B = TET10_B;
      for (int i = 0; i < 10; i++) { t->U[i] = TET10_U[i]; t->V[i] = TET10_V[i]; t->W[i] = TET10_W[i]; }
      break;
    case OR_HEX8: {
      static const double u[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
      static const double v[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
      static const double w[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
      t->dim = 3; t->n = 8; B = B808;
      for (int i = 0; i < 8; i++) { t->U[i] = u[i]; t->V[i] = v[i]; t->W[i] = w[i]; }
    } break;
    case OR_HEX27:
      t->dim = 3; t->n = 27; B = HEX27_B;
      for (int i = 0; i < 27; i++) { t->U[i] = HEX27_U[i]; t->V[i] = HEX27_V[i]; t->W[i] = HEX27_W[i]; }
      break;
  }
  /* term number k (1-based) decodes with MaxDeg = 4 (ElementDescription.src:56-57) */
  for (int k = 0; k < t->n; k++) {
    int m = B[k] - 1;
    t->p[k] = m % 4;
    t->q[k] = (m / 4) % 4;
    t->r[k] = m / 16;
  }
  /* Vandermonde V(node, term); coefficients = V^{-1}: basis_i = sum_k C(k,i) m_k */
  double Vd[27 * 27];
  int n = t->n;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++)
      Vd[k * n + i] = ipow(t->U[i], t->p[k]) * ipow(t->V[i], t->q[k]) * ipow(t->W[i], t->r[k]);
  if (invert(n, Vd)) return NULL;
  /* Vd now holds V^{-1} column-major: Vinv(k, i) = Vd[i*n + k] */
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) t->coeff[k * n + i] = Vd[i * n + k];
  t->ready = 1;
  return t;
}

int or_element_nodes(int code) {
  etype_t *t = get_type(code);
  return t ? t->n : -1;
}

/* NodalBasisFunctions / NodalFirstDerivatives (ElementDescription.src:
 * 1399-1447, 1811-1869, 2012-2147): sums over terms in term order. */
static void basis_eval(const etype_t *t, double u, double v, double w, double *basis, double *dlb) {
  int n = t->n;
  for (int i = 0; i < n; i++) {
    double s = 0.0, du = 0.0, dv = 0.0, dw = 0.0;
    for (int k = 0; k < n; k++) {
      double c = t->coeff[k * n + i];
      if (c == 0.0) continue;
      int p = t->p[k], q = t->q[k], r = t->r[k];
      s += c * ipow(u, p) * ipow(v, q) * ipow(w, r);
      if (p >= 1) du += c * p * ipow(u, p - 1) * ipow(v, q) * ipow(w, r);
      if (q >= 1) dv += c * q * ipow(u, p) * ipow(v, q - 1) * ipow(w, r);
      if (r >= 1) dw += c * r * ipow(u, p) * ipow(v, q) * ipow(w, r - 1);
    }
    basis[i] = s;
    dlb[i * 3 + 0] = du;
    dlb[i * 3 + 1] = dv;
    dlb[i * 3 + 2] = dw;
  }
}

/* ElementInfo (ElementDescription.src:2586, gradients :3305-3314) with
 * ElementMetric (:6757-6884) and InvertMatrix3x3 (:8536-8556). */
int or_element_info(int code, int cdim, const double *x, const double *y, const double *z,
                    double u, double v, double w, double *basis, double *dbasisdx, double *detj) {
  etype_t *t = get_type(code);
  if (!t) return -1;
  int n = t->n, dim = t->dim;
  double dlb[27 * 3];
  basis_eval(t, u, v, w, basis, dlb);
  double dx[3][3] = {{0}}, G[3][3] = {{0}}, M[3][3] = {{0}}, L[3][3] = {{0}};
  const double *X[3] = {x, y, z};
  for (int i = 0; i < dim; i++)
    for (int c = 0; c < 3; c++) {
      double s = 0.0;
      for (int a = 0; a < n; a++) s += X[c][a] * dlb[a * 3 + i];
      dx[c][i] = s;
    }
  for (int i = 0; i < dim; i++)
    for (int j = 0; j < dim; j++) {
      double s = 0.0;
      for (int k = 0; k < cdim; k++) s = s + dx[k][i] * dx[k][j];
      G[i][j] = s;
    }
  double detg;
  if (dim == 1) {
    detg = G[0][0];
    if (detg <= DBL_MIN) return 0;
    M[0][0] = 1.0 / detg;
    detg = sqrt(detg);
  } else if (dim == 2) {
    detg = (G[0][0] * G[1][1] - G[0][1] * G[1][0]);
    if (detg <= DBL_MIN) return 0;
    M[0][0] = G[1][1] / detg;
    M[0][1] = -G[0][1] / detg;
    M[1][0] = -G[1][0] / detg;
    M[1][1] = G[0][0] / detg;
    detg = sqrt(detg);
  } else {
    detg = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) +
           G[0][1] * (G[1][2] * G[2][0] - G[1][0] * G[2][2]) +
           G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    if (detg <= DBL_MIN) return 0;
    double s = 1.0 / detg;
    M[0][0] = s * (G[1][1] * G[2][2] - G[2][1] * G[1][2]);
    M[1][0] = -s * (G[1][0] * G[2][2] - G[2][0] * G[1][2]);
    M[2][0] = s * (G[1][0] * G[2][1] - G[2][0] * G[1][1]);
    M[0][1] = -s * (G[0][1] * G[2][2] - G[2][1] * G[0][2]);
    M[1][1] = s * (G[0][0] * G[2][2] - G[2][0] * G[0][2]);
    M[2][1] = -s * (G[0][0] * G[2][1] - G[2][0] * G[0][1]);
    M[0][2] = s * (G[0][1] * G[1][2] - G[1][1] * G[0][2]);
    M[1][2] = -s * (G[0][0] * G[1][2] - G[1][0] * G[0][2]);
    M[2][2] = s * (G[0][0] * G[1][1] - G[1][0] * G[0][1]);
    detg = sqrt(detg);
  }
  for (int i = 0; i < cdim; i++)
    for (int j = 0; j < dim; j++) {
      double s = 0.0;
      for (int k = 0; k < dim; k++) s = s + dx[i][k] * M[k][j];
      L[i][j] = s;
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      if (j < cdim)
        for (int k = 0; k < dim; k++) s = s + dlb[i * 3 + k] * L[j][k];
      dbasisdx[i * 3 + j] = s;
    }
  *detj = detg;
  return 1;
}

/* ------------------------------------------------------------------ */
/* Local matrices                                                       */
/* ------------------------------------------------------------------ */

/* ModelPDE LocalMatrix (modules/ModelPDE.src:75-166). */
int or_local_modelpde(int code, int cdim, const double *x, const double *y, const double *z,
                      const double *diff, const double *conv, const double *react,
                      const double *velo, const double *load, double *stiff, double *force) {
  etype_t *t = get_type(code);
  if (!t) return -1;
  int n = t->n, dim = cdim;
  double gu[27], gv[27], gw[27], gs[27];
  int ngp = or_gauss_points(code, or_default_gauss_n(code), gu, gv, gw, gs);
  if (ngp < 0) return -1;
  memset(stiff, 0, sizeof(double) * n * n);
  memset(force, 0, sizeof(double) * n);
  double basis[27], db[27 * 3], detj;
  for (int g = 0; g < ngp; g++) {
    if (or_element_info(code, cdim, x, y, z, gu[g], gv[g], gw[g], basis, db, &detj) != 1) return -2;
    double LoadAtIP = 0, D = 0, C = 0, R = 0, a[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) LoadAtIP += basis[i] * load[i];
    for (int c = 0; c < 3; c++) {
      double s = 0.0;
      for (int i = 0; i < n; i++) s += velo[i * 3 + c] * basis[i];
      a[c] = s;
    }
    for (int i = 0; i < n; i++) D += basis[i] * diff[i];
    for (int i = 0; i < n; i++) C += basis[i] * conv[i];
    for (int i = 0; i < n; i++) R += basis[i] * react[i];
    double Weight = gs[g] * detj;
    double WD = Weight * D;
    for (int q = 0; q < n; q++)
      for (int p = 0; p < n; p++) {
        double m = 0.0;
        for (int k = 0; k < 3; k++) m += db[p * 3 + k] * db[q * 3 + k];
        stiff[q * n + p] = stiff[q * n + p] + WD * m;
      }
    for (int p = 0; p < n; p++)
      for (int q = 0; q < n; q++) {
        double s = 0.0;
        for (int k = 0; k < dim; k++) s += a[k] * db[q * 3 + k];
        stiff[q * n + p] = stiff[q * n + p] + Weight * C * s * basis[p];
        stiff[q * n + p] = stiff[q * n + p] + Weight * R * basis[q] * basis[p];
      }
    for (int p = 0; p < n; p++) force[p] = force[p] + Weight * LoadAtIP * basis[p];
  }
  return 0;
}

/* StressCompose, dim 3, isotropic (Stress.src:140-437). */
int or_local_elasticity(int code, const double *x, const double *y, const double *z,
                        const double *young, const double *poisson, double *stiff, double *force) {
  etype_t *t = get_type(code);
  if (!t) return -1;
  int n = t->n, N = 3 * n;
  double gu[27], gv[27], gw[27], gs[27];
  int ngp = or_gauss_points(code, or_default_gauss_n(code), gu, gv, gw, gs);
  if (ngp < 0) return -1;
  memset(stiff, 0, sizeof(double) * N * N);
  memset(force, 0, sizeof(double) * N);
  double basis[27], db[27 * 3], detj;
  for (int g = 0; g < ngp; g++) {
    if (or_element_info(code, 3, x, y, z, gu[g], gv[g], gw[g], basis, db, &detj) != 1) return -2;
    double s = detj * gs[g];
    double Poisson = 0, Young = 0;
    for (int i = 0; i < n; i++) Poisson += basis[i] * poisson[i];
    for (int i = 0; i < n; i++) Young += basis[i] * young[i];
    double C[6][6] = {{0}};
    C[0][0] = 1.0 - Poisson; C[0][1] = Poisson; C[0][2] = Poisson;
    C[1][0] = Poisson; C[1][1] = 1.0 - Poisson; C[1][2] = Poisson;
    C[2][0] = Poisson; C[2][1] = Poisson; C[2][2] = 1.0 - Poisson;
    C[3][3] = 0.5 - Poisson; C[4][4] = 0.5 - Poisson; C[5][5] = 0.5 - Poisson;
    double den = (1 + Poisson) * (1 - 2 * Poisson);
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) C[i][j] = C[i][j] * Young / den;
    for (int p = 0; p < n; p++) {
      double G[3][6] = {{0}}, GC[3][6];
      G[0][0] = db[p * 3 + 0];
      G[1][1] = db[p * 3 + 1];
      G[2][2] = db[p * 3 + 2];
      G[0][3] = db[p * 3 + 1];
      G[1][3] = db[p * 3 + 0];
      G[1][4] = db[p * 3 + 2];
      G[2][4] = db[p * 3 + 1];
      G[0][5] = db[p * 3 + 2];
      G[2][5] = db[p * 3 + 0];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 6; j++) {
          double a = 0.0;
          for (int k = 0; k < 6; k++) a += G[i][k] * C[k][j];
          GC[i][j] = a;
        }
      for (int q = 0; q < n; q++) {
        double B[6][3] = {{0}};
        B[0][0] = db[q * 3 + 0];
        B[1][1] = db[q * 3 + 1];
        B[2][2] = db[q * 3 + 2];
        B[3][0] = db[q * 3 + 1];
        B[3][1] = db[q * 3 + 0];
        B[4][1] = db[q * 3 + 2];
        B[4][2] = db[q * 3 + 1];
        B[5][0] = db[q * 3 + 2];
        B[5][2] = db[q * 3 + 0];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            double a = 0.0;
            for (int k = 0; k < 6; k++) a += GC[i][k] * B[k][j];
            int r = 3 * p + i, c = 3 * q + j;
            stiff[c * N + r] = stiff[c * N + r] + s * a;
          }
      }
    }
  }
  return 0;
}

/* StressBoundary traction (Stress.src:836-1000): FORCE(k) += s*Basis(q)*Load(i). */
int or_local_traction(int code, const double *x, const double *y, const double *z,
                      const double *load3, double *force) {
  etype_t *t = get_type(code);
  if (!t) return -1;
  int n = t->n;
  double gu[27], gv[27], gw[27], gs[27];
  int ngp = or_gauss_points(code, or_default_gauss_n(code), gu, gv, gw, gs);
  memset(force, 0, sizeof(double) * 3 * n);
  double basis[27], db[27 * 3], detj;
  for (int g = 0; g < ngp; g++) {
    if (or_element_info(code, 3, x, y, z, gu[g], gv[g], gw[g], basis, db, &detj) != 1) return -2;
    double s = detj * gs[g];
    double L[3];
    for (int i = 0; i < 3; i++) {
      double a = 0.0;
      for (int k = 0; k < n; k++) a += load3[k * 3 + i] * basis[k];
      L[i] = a;
    }
    for (int q = 0; q < n; q++)
      for (int i = 0; i < 3; i++) force[q * 3 + i] = force[q * 3 + i] + s * basis[q] * L[i];
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* CRS topology                                                         */
/* ------------------------------------------------------------------ */

struct or_pattern {
  int nnodes, dofs, n;
  long long nnz;
  int *perm, *rows, *cols, *diag;
  int bw_before, bw_after, accepted;
};

typedef struct {
  int *start; /* nnodes+1 */
  int *idx;   /* sorted unique adjacency incl. self, 1-based node ids */
} adj_t;

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* MakeListMatrix (ElementUtils.src:279-620) result: per node the sorted,
 * de-duplicated list of all nodes sharing an element (self included). */
static int build_adjacency(int nnodes, int nelem, int maxnpe, const int *el, const int *npe, adj_t *A) {
  long long *cnt = calloc((size_t)nnodes + 1, sizeof(long long));
  for (int e = 0; e < nelem; e++) {
    int m = npe[e];
    for (int a = 0; a < m; a++) cnt[el[(long long)e * maxnpe + a]] += m;
  }
  long long *off = malloc(sizeof(long long) * ((size_t)nnodes + 2));
  off[1] = 0;
  for (int i = 1; i <= nnodes; i++) off[i + 1] = off[i] + cnt[i];
  int *tmp = malloc(sizeof(int) * (size_t)(off[nnodes + 1] + 1));
  long long *fill = calloc((size_t)nnodes + 2, sizeof(long long));
  for (int e = 0; e < nelem; e++) {
    int m = npe[e];
    const int *nd = el + (long long)e * maxnpe;
    for (int a = 0; a < m; a++)
      for (int b = 0; b < m; b++) tmp[off[nd[a]] + fill[nd[a]]++] = nd[b];
  }
  A->start = malloc(sizeof(int) * ((size_t)nnodes + 2));
  long long total = 0;
  for (int i = 1; i <= nnodes; i++) {
    int *seg = tmp + off[i];
    long long m = fill[i];
    qsort(seg, (size_t)m, sizeof(int), cmp_int);
    long long u = 0;
    for (long long k = 0; k < m; k++)
      if (k == 0 || seg[k] != seg[k - 1]) seg[u++] = seg[k];
    fill[i] = u;
    total += u;
  }
  A->idx = malloc(sizeof(int) * (size_t)(total + 1));
  long long pos = 0;
  for (int i = 1; i <= nnodes; i++) {
    A->start[i] = (int)pos;
    memcpy(A->idx + pos, tmp + off[i], sizeof(int) * (size_t)fill[i]);
    pos += fill[i];
  }
  A->start[nnodes + 1] = (int)pos;
  free(cnt); free(off); free(tmp); free(fill);
  return 0;
}

/* ComputeBandwidth (BandwidthOptimize.src:60-90). */
static int compute_bandwidth(int n, const adj_t *A, const int *reorder) {
  int hb = 0;
  for (int i = 1; i <= n; i++)
    for (int c = A->start[i]; c < A->start[i + 1]; c++) {
      int k = A->idx[c];
      int d = reorder ? abs(reorder[i] - reorder[k]) : abs(i - k);
      if (d > hb) hb = d;
    }
  return hb;
}

/* Levelize (BandwidthOptimize.src:290-350): the reference's depth-first
 * walk over the list matrix with an explicit stack of list positions.  A
 * "list position" is an index into A->idx; -1 plays the null pointer. */
static void levelize(int nin, int levelin, const adj_t *A, int *level, char *done, int *maxlevel, int n) {
  int cap = 512, sp = 0;
  int *stack = malloc(sizeof(int) * cap);
  int node = nin, lev = levelin;
  int p = A->start[node] < A->start[node + 1] ? A->start[node] : -1;
  while (p >= 0) {
    if (sp >= cap) { cap *= 2; stack = realloc(stack, sizeof(int) * cap); }
    stack[sp++] = p;
    level[node] = lev;
    done[node] = 1;
    if (lev > *maxlevel) *maxlevel = lev;
    p = A->start[node] < A->start[node + 1] ? A->start[node] : -1;
    for (;;) {
      if (p >= 0) {
        node = A->idx[p];
        if (node <= n && !done[node]) { lev++; break; }
      } else if (sp >= 1) {
        p = stack[sp - 1];
        lev--;
        sp--;
      } else {
        break;
      }
      /* p => p % Next: the successor inside the owning list */
      if (p >= 0) {
        /* find the list owning position p: positions are contiguous per list,
           the successor is p+1 unless p is the last entry of its list */
        int owner_end = -1;
        /* binary search on start[] for the list containing p */
        int lo = 1, hi = n;
        while (lo < hi) {
          int mid = (lo + hi + 1) / 2;
          if (A->start[mid] <= p) lo = mid; else hi = mid - 1;
        }
        owner_end = A->start[lo + 1];
        p = (p + 1 < owner_end) ? p + 1 : -1;
      }
    }
  }
  free(stack);
}

/* OptimizeBandwidth (BandwidthOptimize.src:108-285). perm is 1-based over
 * nodes and enters as identity.  Returns the half bandwidth in use. */
static void optimize_bandwidth(int n, const adj_t *A, int *perm, int *before, int *after, int *accepted) {
  int *degree = malloc(sizeof(int) * (n + 1));
  int *level = calloc(n + 1, sizeof(int));
  char *done = calloc(n + 1, 1);
  for (int i = 1; i <= n; i++) degree[i] = A->start[i + 1] - A->start[i];
  int hb_before = compute_bandwidth(n, A, NULL) + 1;
  int start = 1, mindeg = degree[1];
  for (int i = 1; i <= n; i++) {
    if (degree[i] < mindeg) { start = i; mindeg = degree[i]; }
    level[i] = 0;
  }
  int maxlevel = 0;
  levelize(start, 0, A, level, done, &maxlevel, n);
  int newroot = 1;
  while (newroot) {
    newroot = 0;
    mindeg = degree[start];
    int k = start;
    for (int i = 1; i <= n; i++)
      if (level[i] == maxlevel && degree[i] < mindeg) { k = i; mindeg = degree[i]; }
    if (k != start) {
      int j = maxlevel;
      maxlevel = 0;
      memset(done, 0, n + 1);
      levelize(k, 0, A, level, done, &maxlevel, n);
      if (j > maxlevel) {
        newroot = 1;
        start = j; /* the reference assigns the previous depth here */
      }
    }
  }
  int *plocal = calloc(n + 2, sizeof(int));
  int *doneidx = calloc(n + 1, sizeof(int));
  int indx = 1;
  plocal[indx] = start;
  doneidx[start] = indx;
  indx++;
  for (int i = 1; i <= n; i++) {
    if (plocal[i] == 0) {
      for (int j = 1; j <= n; j++)
        if (doneidx[j] == 0) { plocal[indx] = j; doneidx[j] = indx; indx++; break; }
    }
    int node = plocal[i];
    for (int c = A->start[node]; c < A->start[node + 1]; c++) {
      int k = A->idx[c];
      if (k <= n && doneidx[k] == 0) { plocal[indx] = k; doneidx[k] = indx; indx++; }
    }
  }
  for (int i = 1; i <= n; i++) doneidx[plocal[i]] = n - i + 1;
  int *saved = malloc(sizeof(int) * (n + 1));
  memcpy(saved, perm, sizeof(int) * (n + 1));
  for (int i = 1; i <= n; i++) {
    int k = saved[i];
    perm[i] = k > 0 ? doneidx[k] : 0;
  }
  int hb_after = compute_bandwidth(n, A, perm) + 1;
  *before = hb_before;
  *after = hb_after;
  *accepted = 1;
  if (hb_before < hb_after) {
    memcpy(perm, saved, sizeof(int) * (n + 1));
    *accepted = 0;
  }
  free(degree); free(level); free(done); free(plocal); free(doneidx); free(saved);
}

/* CreateMatrix chain: InitialPermutation (identity, ElementUtils.src:765-773),
 * MakeListMatrix, OptimizeBandwidth, CRS_CreateMatrix + InitializeMatrix +
 * CRS_SortMatrix (CRSMatrix.src:816-869, 168-227). */
or_pattern *or_pattern_build(int nnodes, int nelem, int maxnpe, const int *elements, const int *npe,
                             int dofs, int optimize) {
  adj_t A;
  build_adjacency(nnodes, nelem, maxnpe, elements, npe, &A);
  or_pattern *P = calloc(1, sizeof(or_pattern));
  P->nnodes = nnodes;
  P->dofs = dofs;
  P->perm = malloc(sizeof(int) * (nnodes + 1));
  for (int i = 0; i <= nnodes; i++) P->perm[i] = i;
  int hb = compute_bandwidth(nnodes, &A, NULL) + 1;
  P->bw_before = P->bw_after = hb;
  P->accepted = 0;
  if (optimize) optimize_bandwidth(nnodes, &A, P->perm, &P->bw_before, &P->bw_after, &P->accepted);
  int n = nnodes * dofs;
  P->n = n;
  long long nnz = 0;
  for (int i = 1; i <= nnodes; i++) nnz += (long long)(A.start[i + 1] - A.start[i]) * dofs * dofs;
  P->nnz = nnz;
  P->rows = malloc(sizeof(int) * (n + 1));
  P->cols = malloc(sizeof(int) * (size_t)nnz);
  P->diag = malloc(sizeof(int) * n);
  int *inv = malloc(sizeof(int) * (nnodes + 1));
  for (int i = 1; i <= nnodes; i++) inv[P->perm[i]] = i;
  /* rows in new numbering: row block j = perm(node) */
  long long pos = 0;
  for (int j = 1; j <= nnodes; j++) {
    int node = inv[j];
    int deg = A.start[node + 1] - A.start[node];
    for (int l = 0; l < dofs; l++) {
      int r = dofs * (j - 1) + l; /* 0-based row */
      P->rows[r] = (int)(pos + 1);
      int *seg = P->cols + pos;
      int m = 0;
      for (int c = A.start[node]; c < A.start[node + 1]; c++) {
        int k = P->perm[A.idx[c]];
        for (int d = 1; d <= dofs; d++) seg[m++] = dofs * (k - 1) + d;
      }
      qsort(seg, m, sizeof(int), cmp_int);
      for (int c = 0; c < m; c++)
        if (seg[c] == r + 1) P->diag[r] = (int)(pos + c + 1);
      pos += (long long)deg * dofs;
    }
  }
  P->rows[n] = (int)(pos + 1);
  free(inv);
  free(A.start);
  free(A.idx);
  return P;
}

int or_pattern_n(const or_pattern *p) { return p->n; }
long long or_pattern_nnz(const or_pattern *p) { return p->nnz; }
int or_pattern_bandwidth(const or_pattern *p, int *before, int *after, int *accepted) {
  *before = p->bw_before;
  *after = p->bw_after;
  *accepted = p->accepted;
  return 0;
}
void or_pattern_copy(const or_pattern *p, int *perm, int *rows, int *cols, int *diag) {
  if (perm) memcpy(perm, p->perm + 1, sizeof(int) * p->nnodes);
  if (rows) memcpy(rows, p->rows, sizeof(int) * (p->n + 1));
  if (cols) memcpy(cols, p->cols, sizeof(int) * (size_t)p->nnz);
  if (diag) memcpy(diag, p->diag, sizeof(int) * p->n);
}
void or_pattern_free(or_pattern *p) {
  if (!p) return;
  free(p->perm); free(p->rows); free(p->cols); free(p->diag);
  free(p);
}

/* ------------------------------------------------------------------ */
/* Glue                                                                 */
/* ------------------------------------------------------------------ */

/* CRS_GlueLocalMatrix (CRSMatrix.src:507-590) + RHS of
 * UpdateGlobalEquations (SolverUtils.src:1042-1048). */
int or_glue_local(int n, const int *rows, const int *cols, const int *diag, int nen, int dofs,
                  const int *idx, const double *local, const double *force, double *values, double *rhs) {
  (void)n;
  int ld = nen * dofs;
  for (int i = 1; i <= nen; i++)
    for (int k = 0; k <= dofs - 1; k++) {
      if (idx[i - 1] <= 0) continue;
      int Row = dofs * idx[i - 1] - k;
      for (int j = 1; j <= nen; j++)
        for (int l = 0; l <= dofs - 1; l++) {
          if (idx[j - 1] <= 0) continue;
          int Col = dofs * idx[j - 1] - l;
          double val = local[(long long)(dofs * j - l - 1) * ld + (dofs * i - k - 1)];
          int c0, c1;
          if (Col >= Row) { c0 = diag[Row - 1]; c1 = rows[Row] - 1; }
          else { c0 = rows[Row - 1]; c1 = diag[Row - 1] - 1; }
          for (int c = c0; c <= c1; c++)
            if (cols[c - 1] == Col) { values[c - 1] = values[c - 1] + val; break; }
        }
    }
  if (rhs && force)
    for (int i = 1; i <= nen; i++) {
      if (idx[i - 1] <= 0) continue;
      for (int j = 1; j <= dofs; j++) {
        int k = dofs * (idx[i - 1] - 1) + j;
        rhs[k - 1] = rhs[k - 1] + force[dofs * (i - 1) + j - 1];
      }
    }
  return 0;
}

/* The module element loop (ModelPDE.src:36-65 / StressSolve BulkAssembly)
 * with DefaultInitialize's zeroing (SolverUtils.src:81). */
int or_assemble(int kind, int code, int cdim, int nnodes, const double *coords, int nelem, int npe,
                const int *elements, const double *elem_coef, const double *elem_conv,
                const double *elem_velo, const double *elem_load, double poisson, const int *perm,
                int n, const int *rows, const int *cols, const int *diag, int dofs, double *values,
                double *rhs) {
  (void)nnodes;
  long long nnz = rows[n] - 1;
  memset(values, 0, sizeof(double) * nnz);
  memset(rhs, 0, sizeof(double) * n);
  int ld = npe * dofs;
  double *stiff = malloc(sizeof(double) * ld * ld);
  double *force = malloc(sizeof(double) * ld);
  double x[27], y[27], z[27], c1[27], c2[27], c3[27], ld3[27], velo[27 * 3];
  int idx[27];
  for (int e = 0; e < nelem; e++) {
    const int *nd = elements + (long long)e * npe;
    for (int a = 0; a < npe; a++) {
      int g = nd[a] - 1;
      x[a] = coords[3 * (long long)g];
      y[a] = coords[3 * (long long)g + 1];
      z[a] = coords[3 * (long long)g + 2];
      idx[a] = perm[g];
    }
    int rc;
    if (kind == 0) {
      for (int a = 0; a < npe; a++) {
        c1[a] = elem_coef[e];
        c2[a] = elem_conv ? elem_conv[e] : 0.0;
        c3[a] = 0.0;
        ld3[a] = elem_load ? elem_load[e] : 0.0;
        for (int c = 0; c < 3; c++) velo[a * 3 + c] = elem_velo ? elem_velo[3 * (long long)e + c] : 0.0;
      }
      rc = or_local_modelpde(code, cdim, x, y, z, c1, c2, c3, velo, ld3, stiff, force);
    } else {
      for (int a = 0; a < npe; a++) { c1[a] = elem_coef[e]; c2[a] = poisson; }
      rc = or_local_elasticity(code, x, y, z, c1, c2, stiff, force);
    }
    if (rc) { free(stiff); free(force); return rc; }
    or_glue_local(n, rows, cols, diag, npe, dofs, idx, stiff, force, values, rhs);
  }
  free(stiff);
  free(force);
  return 0;
}

int or_assemble_traction(int code, int nnodes, const double *coords, int nbelem, int npe,
                         const int *belements, const double *load3, const int *perm, int dofs,
                         double *rhs) {
  (void)nnodes;
  double x[27], y[27], z[27], L[27 * 3], force[27 * 3];
  for (int a = 0; a < npe; a++)
    for (int c = 0; c < 3; c++) L[a * 3 + c] = load3[c];
  for (int e = 0; e < nbelem; e++) {
    const int *nd = belements + (long long)e * npe;
    for (int a = 0; a < npe; a++) {
      int g = nd[a] - 1;
      x[a] = coords[3 * (long long)g];
      y[a] = coords[3 * (long long)g + 1];
      z[a] = coords[3 * (long long)g + 2];
    }
    int rc = or_local_traction(code, x, y, z, L, force);
    if (rc) return rc;
    for (int i = 1; i <= npe; i++) {
      int id = perm[nd[i - 1] - 1];
      if (id <= 0) continue;
      for (int j = 1; j <= dofs; j++) {
        int k = dofs * (id - 1) + j;
        rhs[k - 1] = rhs[k - 1] + force[dofs * (i - 1) + j - 1];
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Scaling and Dirichlet                                                */
/* ------------------------------------------------------------------ */

/* ScaleLinearSystem (SolverUtils.src:6470-6640), real matrices. */
static void scale_system(int n, const int *rows, const int *cols, const int *diag, double *values,
                         double *b, double *x, int rhs_scaling, double *D, double *rhs_scale) {
  for (int i = 0; i < n; i++) D[i] = 0.0;
  for (int i = 0; i < n; i++) {
    int j = diag[i];
    if (j > 0) D[i] = values[j - 1];
  }
  int any = 0;
  for (int i = 0; i < n; i++)
    if (fabs(D[i]) <= DBL_MIN) any = 1;
  if (any) {
    for (int i = 0; i < n; i++) {
      if (fabs(D[i]) <= DBL_MIN) {
        double s = 0.0;
        for (int k = rows[i]; k < rows[i + 1]; k++) s += fabs(values[k - 1]);
        D[i] = s;
      } else {
        int j = diag[i];
        if (j > 0) D[i] = values[j - 1];
      }
    }
  }
  for (int i = 0; i < n; i++) D[i] = fabs(D[i]) > DBL_MIN ? 1.0 / sqrt(fabs(D[i])) : 1.0;
  for (int i = 0; i < n; i++)
    for (int j = rows[i]; j < rows[i + 1]; j++)
      values[j - 1] = values[j - 1] * (D[i] * D[cols[j - 1] - 1]);
  *rhs_scale = 1.0;
  if (b) {
    for (int i = 0; i < n; i++) b[i] = b[i] * D[i];
    double bnorm = 1.0;
    if (rhs_scaling) {
      double s = 0.0;
      for (int i = 0; i < n; i++) s += b[i] * b[i];
      bnorm = sqrt(s);
    }
    *rhs_scale = bnorm;
    for (int i = 0; i < n; i++) D[i] = D[i] * bnorm;
    for (int i = 0; i < n; i++) b[i] = b[i] / bnorm;
    if (x)
      for (int i = 0; i < n; i++) x[i] = x[i] / D[i];
  }
}

/* BackScaleLinearSystem (SolverUtils.src:6754-6866). */
static void backscale_system(int n, const int *rows, const int *cols, double *values, double *b,
                             double *x, double *D, double rhs_scale) {
  if (b) {
    if (x)
      for (int i = 0; i < n; i++) x[i] = x[i] * D[i];
    double bnorm = rhs_scale;
    for (int i = 0; i < n; i++) D[i] = D[i] / bnorm;
    for (int i = 0; i < n; i++) b[i] = b[i] / D[i] * bnorm;
  }
  for (int i = 0; i < n; i++)
    for (int j = rows[i]; j < rows[i + 1]; j++)
      values[j - 1] = values[j - 1] / (D[i] * D[cols[j - 1] - 1]);
}

/* CRS_SetSymmDirichlet (CRSMatrix.src:674-735). */
static void set_symm_dirichlet(const int *rows, const int *cols, const int *diag, double *values,
                               double *b, int nrow, double val) {
  for (int l = rows[nrow - 1]; l <= rows[nrow] - 1; l++) {
    int i = cols[l - 1];
    if (i == nrow) continue;
    int k1, k2;
    if (nrow > i) { k1 = diag[i - 1] + 1; k2 = rows[i] - 1; }
    else { k1 = rows[i - 1]; k2 = diag[i - 1] - 1; }
    for (int j = k1; j <= k2; j++) {
      if (cols[j - 1] == nrow) {
        b[i - 1] = b[i - 1] - values[j - 1] * val;
        values[j - 1] = 0.0;
        break;
      } else if (cols[j - 1] > nrow) {
        break;
      }
    }
  }
  for (int l = rows[nrow - 1]; l <= rows[nrow] - 1; l++) values[l - 1] = 0.0;
  b[nrow - 1] = val;
  values[diag[nrow - 1] - 1] = 1.0;
}

/* DefaultDirichletBCs (DefUtils.src:3530-4030) -> SetDirichletBoundaries ->
 * SetPointValues/SetSinglePoint (SolverUtils.src:2558-2657). */
int or_dirichlet(int n, const int *rows, const int *cols, const int *diag, double *values,
                 double *rhs, int nbc, const int *bc_rows, const double *bc_vals, int symmetric,
                 int scaling) {
  double *D = malloc(sizeof(double) * n);
  double rs = 1.0;
  if (scaling) {
    scale_system(n, rows, cols, diag, values, rhs, NULL, 0, D, &rs);
  } else {
    for (int i = 0; i < n; i++) D[i] = 1.0;
  }
  for (int t = 0; t < nbc; t++) {
    int k = bc_rows[t];
    double val = bc_vals[t];
    if (symmetric) {
      set_symm_dirichlet(rows, cols, diag, values, rhs, k, val / D[k - 1]);
    } else {
      for (int l = rows[k - 1]; l <= rows[k] - 1; l++) values[l - 1] = 0.0;
      values[diag[k - 1] - 1] = 1.0;
      rhs[k - 1] = val / D[k - 1];
    }
  }
  if (scaling) backscale_system(n, rows, cols, values, rhs, NULL, D, rs);
  free(D);
  return 0;
}

/* ------------------------------------------------------------------ */
/* BLAS-1 and SpMV                                                      */
/* ------------------------------------------------------------------ */

/* ddot (mathlibs/src/blas/ddot.f), unit stride: clean-up loop then
 * unrolled-by-5 groups accumulated left to right into one temporary. */
double or_ddot(int n, const double *dx, const double *dy) {
  double dtemp = 0.0;
  if (n <= 0) return 0.0;
  int m = n % 5;
  for (int i = 0; i < m; i++) dtemp = dtemp + dx[i] * dy[i];
  if (n < 5) return dtemp;
  for (int i = m; i < n; i += 5)
    dtemp = dtemp + dx[i] * dy[i] + dx[i + 1] * dy[i + 1] + dx[i + 2] * dy[i + 2] +
            dx[i + 3] * dy[i + 3] + dx[i + 4] * dy[i + 4];
  return dtemp;
}

/* dnrm2 (mathlibs/src/blas/dnrm2.f): scaled sum of squares. */
double or_dnrm2(int n, const double *x) {
  if (n < 1) return 0.0;
  if (n == 1) return fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; i++) {
    if (x[i] != 0.0) {
      double absxi = fabs(x[i]);
      if (scale < absxi) {
        double r = scale / absxi;
        ssq = 1.0 + ssq * (r * r);
        scale = absxi;
      } else {
        double r = absxi / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * sqrt(ssq);
}

/* CRS_MatrixVectorProd non-transposed branch (CRSMatrix.src:3565-3573). */
/* CRS_MatrixVectorProd (CRSMatrix.src:3515-3640): rows split over the
 * threads like its !$omp parallel do (:3565-3573); each row is summed left
 * to right from zero, so the bits do not depend on the thread count. */
void or_spmv(int n, const int *rows, const int *cols, const double *values, const double *u, double *v) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int j = rows[i]; j < rows[i + 1]; j++) s = s + values[j - 1] * u[cols[j - 1] - 1];
    v[i] = s;
  }
}

/* ------------------------------------------------------------------ */
/* ILU(n) preconditioner                                                */
/* ------------------------------------------------------------------ */

/* One round of InitializeILU1 (CRSMatrix.src:2387-2498): the pattern of
 * (rows, cols, diag) plus the fills that an original entry (k, j), k < i,
 * j > k, of an original lower entry (i, k) of row i brings; only the input
 * entries propagate within a round.  Returns the new 1-based arrays. */
static int ilu_fill_round(int n, const int *rows, const int *cols, const int *diag, int **orows, int **ocols,
                          int **odiag) {
  unsigned char *C = calloc((size_t)n + 1, 1);
  int *R = malloc(sizeof(int) * ((size_t)n + 1)), *D = malloc(sizeof(int) * (size_t)n);
  size_t cap = (size_t)rows[n] - 1 + 16, m = 0;
  int *K = malloc(sizeof(int) * cap);
  if (!C || !R || !D || !K) return -1;
  R[0] = 1;
  for (int i = 1; i <= n; i++) {
    for (int k = rows[i - 1]; k < rows[i]; k++) C[cols[k - 1]] = 1;
    int rmin = cols[rows[i - 1] - 1], rmax = cols[rows[i] - 2];
    for (int k = rmin; k <= i - 1; k++) {
      if (C[k] != 1) continue;
      for (int l = diag[k - 1] + 1; l < rows[k]; l++) {
        int j = cols[l - 1];
        if (C[j] == 0) {
          C[j] = 2;
          if (j > rmax) rmax = j;
        }
      }
    }
    for (int k = rmin; k <= rmax; k++) {
      if (C[k] == 0) continue;
      C[k] = 0;
      if (m == cap) {
        cap *= 2;
        K = realloc(K, sizeof(int) * cap);
        if (!K) return -1;
      }
      K[m++] = k;
      if (k == i) D[i - 1] = (int)m;
    }
    R[i] = (int)m + 1;
  }
  free(C);
  *orows = R;
  *ocols = K;
  *odiag = D;
  return 0;
}

/* The ILU(order) pattern: the matrix pattern for order 0, else order rounds
 * of InitializeILU1, each on the previous round's pattern (CRS_IncompleteLU,
 * CRSMatrix.src:2194-2219). */
int or_ilu_pattern(int n, const int *rows, const int *cols, const int *diag, int order, int **irows, int **icols,
                   int **idiag) {
  size_t nnz = (size_t)rows[n] - 1;
  int *R = malloc(sizeof(int) * ((size_t)n + 1)), *K = malloc(sizeof(int) * nnz), *D = malloc(sizeof(int) * n);
  if (!R || !K || !D) return -1;
  memcpy(R, rows, sizeof(int) * ((size_t)n + 1));
  memcpy(K, cols, sizeof(int) * nnz);
  memcpy(D, diag, sizeof(int) * (size_t)n);
  for (int r = 0; r < order; r++) {
    int *R2, *K2, *D2;
    if (ilu_fill_round(n, R, K, D, &R2, &K2, &D2)) return -1;
    free(R); free(K); free(D);
    R = R2; K = K2; D = D2;
  }
  *irows = R;
  *icols = K;
  *idiag = D;
  return 0;
}

void or_free(void *p) { free(p); }

/* CRS_IncompleteLU, real non-Cholesky branch (CRSMatrix.src:2312-2372):
 * row by row, the row of A scattered into a full work row S (flags C mark
 * the ILU pattern), eliminated with the finished rows k of its lower ILU
 * part in column order (S(k) = 0 skipped, division only when
 * |U(k,k)| > AEPS), stored back; then every U(i,i) is replaced by its
 * inverse (1 when |U(i,i)| < AEPS).  AEPS = 10*EPSILON(1.0d0) (Types.src:61). */
int or_ilu_factor(int n, const int *rows, const int *cols, const double *values, const int *irows,
                  const int *icols, const int *idiag, double *ilu) {
  const double AEPS = 10.0 * DBL_EPSILON;
  double *S = calloc((size_t)n + 1, sizeof(double));
  unsigned char *C = calloc((size_t)n + 1, 1);
  if (!S || !C) return -1;
  for (int i = 1; i <= n; i++) {
    for (int k = rows[i - 1]; k < rows[i]; k++) S[cols[k - 1]] = values[k - 1];
    for (int k = irows[i - 1]; k < irows[i]; k++) C[icols[k - 1]] = 1;
    for (int m = irows[i - 1]; m < idiag[i - 1]; m++) {
      int k = icols[m - 1];
      if (S[k] == 0.0) continue;
      if (fabs(ilu[idiag[k - 1] - 1]) > AEPS) S[k] = S[k] / ilu[idiag[k - 1] - 1];
      for (int l = idiag[k - 1] + 1; l < irows[k]; l++) {
        int j = icols[l - 1];
        if (C[j]) S[j] = S[j] - S[k] * ilu[l - 1];
      }
    }
    for (int k = irows[i - 1]; k < irows[i]; k++) {
      int j = icols[k - 1];
      if (C[j]) {
        ilu[k - 1] = S[j];
        S[j] = 0.0;
        C[j] = 0;
      }
    }
  }
  for (int i = 1; i <= n; i++) {
    double d = ilu[idiag[i - 1] - 1];
    ilu[idiag[i - 1] - 1] = fabs(d) < AEPS ? 1.0 : 1.0 / d;
  }
  free(S);
  free(C);
  return 0;
}

/* CRS_LUSolve, non-Cholesky (CRSMatrix.src:3356-3374): forward with the unit
 * lower factor, then backward with U and the inverted diagonal. */
void or_lu_solve(int n, const int *irows, const int *icols, const int *idiag, const double *ilu, double *b) {
  for (int i = 1; i <= n; i++) {
    double s = b[i - 1];
    for (int j = irows[i - 1]; j < idiag[i - 1]; j++) s = s - ilu[j - 1] * b[icols[j - 1] - 1];
    b[i - 1] = s;
  }
  for (int i = n; i >= 1; i--) {
    double s = b[i - 1];
    for (int j = idiag[i - 1] + 1; j < irows[i]; j++) s = s - ilu[j - 1] * b[icols[j - 1] - 1];
    b[i - 1] = ilu[idiag[i - 1] - 1] * s;
  }
}

/* ------------------------------------------------------------------ */
/* HUTI CG / BiCGStab                                                   */
/* ------------------------------------------------------------------ */

typedef struct {
  int n;
  const int *rows, *cols, *diag;
  const double *values;
  int precond;  /* 0 none, 1 diagonal, 2 + k ILU(k) */
  int *irows, *icols, *idiag;
  double *ilu;
} sysctx_t;

static void mv(const sysctx_t *c, const double *u, double *v) {
  or_spmv(c->n, c->rows, c->cols, c->values, u, v);
}

/* CRS_DiagPrecondition (CRSMatrix.src:1467-1511) or pcond_dummy
 * (IterSolve.src:111-118).  AEPS = 10*EPSILON(1.0d0) (Types.src:61). */
static void pcond(const sysctx_t *c, double *u, const double *v) {
  const double AEPS = 10.0 * DBL_EPSILON;
  if (!c->precond) {
    memcpy(u, v, sizeof(double) * c->n);
    return;
  }
  if (c->precond >= 2) {  /* CRS_LUPrecondition (CRSMatrix.src:3259-3268): u = v, LU solve */
    memcpy(u, v, sizeof(double) * c->n);
    or_lu_solve(c->n, c->irows, c->icols, c->idiag, c->ilu, u);
    return;
  }
  for (int i = 0; i < c->n; i++) {
    double d = c->values[c->diag[i] - 1];
    u[i] = fabs(d) > AEPS ? v[i] / d : v[i];
  }
}

#define HUTI_CONVERGENCE 1
#define HUTI_MAXITER 2
#define HUTI_CG_RHO 20
#define HUTI_BICGSTAB_RHO 35
#define HUTI_BICGSTAB_SNORM 36
#define HUTI_BICGSTAB_OMEGA 37

static double huti_residual(const sysctx_t *c, int stopc, const double *x, const double *b,
                            const double *r, double *tmp, double rhsnorm) {
  int n = c->n;
  switch (stopc) {
    case 1:
      mv(c, x, tmp);
      for (int i = 0; i < n; i++) tmp[i] = tmp[i] - b[i];
      return or_dnrm2(n, tmp) / rhsnorm;
    case 2:
      return or_dnrm2(n, r);
    case 3:
      return or_dnrm2(n, r) / rhsnorm;
    default:
      mv(c, x, tmp);
      for (int i = 0; i < n; i++) tmp[i] = tmp[i] - b[i];
      return or_dnrm2(n, tmp);
  }
}

/* huti_dcgsolv (hutiter/src/huti_cg.src:73-244; huti_cg_D.F90). */
static int huti_cg(const sysctx_t *c, const double *B, double *X, double tol, int maxit, int stopc,
                   int *iters, double *hist) {
  int n = c->n;
  double *Z = calloc(n, sizeof(double)), *P = calloc(n, sizeof(double));
  double *Q = calloc(n, sizeof(double)), *R = calloc(n, sizeof(double));
  double alpha, beta, rho, oldrho = 0, residual = 0, rhsnorm = 1.0;
  int it = 1, info = 0;
  if (stopc == 1 || stopc == 3) rhsnorm = or_dnrm2(n, B);
  mv(c, X, R);
  for (int i = 0; i < n; i++) R[i] = B[i] - R[i];
  for (;;) {
    pcond(c, Q, R);
    memcpy(Z, Q, sizeof(double) * n); /* pcondr = dummy copy */
    rho = or_ddot(n, R, Z);
    if (rho == 0) { info = HUTI_CG_RHO; break; }
    if (it == 1) {
      memcpy(P, Z, sizeof(double) * n);
    } else {
      beta = rho / oldrho;
      for (int i = 0; i < n; i++) P[i] = Z[i] + beta * P[i];
    }
    mv(c, P, Q);
    alpha = rho / or_ddot(n, P, Q);
    for (int i = 0; i < n; i++) X[i] = X[i] + alpha * P[i];
    for (int i = 0; i < n; i++) R[i] = R[i] - alpha * Q[i];
    residual = huti_residual(c, stopc, X, B, R, Z, rhsnorm);
    if (hist) hist[it - 1] = residual;
    if (residual < tol) { info = HUTI_CONVERGENCE; break; }
    oldrho = rho;
    it++;
    if (it > maxit) { info = HUTI_MAXITER; break; }
  }
  *iters = it;
  free(Z); free(P); free(Q); free(R);
  return info;
}

/* huti_dbicgstabsolv (hutiter/src/huti_bicgstab.src:85-273). */
static int huti_bicgstab(const sysctx_t *c, const double *B, double *X, double tol, int maxit,
                         int stopc, int *iters, double *hist) {
  int n = c->n;
  double *RTLD = calloc(n, sizeof(double)), *P = calloc(n, sizeof(double));
  double *T1V = calloc(n, sizeof(double)), *V = calloc(n, sizeof(double));
  double *S = calloc(n, sizeof(double)), *T2V = calloc(n, sizeof(double));
  double *T = calloc(n, sizeof(double)), *R = calloc(n, sizeof(double));
  double rho, oldrho, alpha, beta, omega, residual = 0, rhsnorm = 1.0;
  const double HUTI_EPSILON = (double)1.17549435E-38f;
  int it = 1, info = 0;
  if (stopc == 1 || stopc == 3) rhsnorm = or_dnrm2(n, B);
  mv(c, X, R);
  for (int i = 0; i < n; i++) R[i] = B[i] - R[i];
  memcpy(RTLD, R, sizeof(double) * n);
  oldrho = 1; omega = 1; alpha = 0;
  for (;;) {
    rho = or_ddot(n, RTLD, R);
    if (rho == 0) { info = HUTI_BICGSTAB_RHO; break; }
    beta = (rho * alpha) / (oldrho * omega);
    for (int i = 0; i < n; i++) P[i] = R[i] + beta * (P[i] - omega * V[i]);
    pcond(c, V, P);
    memcpy(T1V, V, sizeof(double) * n);
    mv(c, T1V, V);
    alpha = rho / or_ddot(n, RTLD, V);
    for (int i = 0; i < n; i++) S[i] = R[i] - alpha * V[i];
    residual = or_dnrm2(n, S);
    if (residual < HUTI_EPSILON) {
      for (int i = 0; i < n; i++) X[i] = X[i] + alpha * T1V[i];
      info = HUTI_BICGSTAB_SNORM;
      break;
    }
    pcond(c, T, S);
    memcpy(T2V, T, sizeof(double) * n);
    mv(c, T2V, T);
    omega = (or_ddot(n, T, S)) / (or_ddot(n, T, T));
    for (int i = 0; i < n; i++) X[i] = X[i] + alpha * T1V[i] + omega * T2V[i];
    for (int i = 0; i < n; i++) R[i] = S[i] - omega * T[i];
    if (stopc == 1 || stopc == 0) {
      mv(c, X, T2V);
      for (int i = 0; i < n; i++) T1V[i] = T2V[i] - B[i];
      residual = or_dnrm2(n, T1V);
      if (stopc == 1) residual = residual / rhsnorm;
    } else if (stopc == 2) {
      residual = or_dnrm2(n, R);
    } else {
      residual = or_dnrm2(n, R) / rhsnorm;
    }
    if (hist) hist[it - 1] = residual;
    if (residual < tol) { info = HUTI_CONVERGENCE; break; }
    if (omega == 0) { info = HUTI_BICGSTAB_OMEGA; break; }
    oldrho = rho;
    it++;
    if (it > maxit) { info = HUTI_MAXITER; break; }
  }
  *iters = it;
  free(RTLD); free(P); free(T1V); free(V); free(S); free(T2V); free(T); free(R);
  return info;
}

int or_huti_solve(int method, int precond, int n, const int *rows, const int *cols, const int *diag,
                  const double *values, const double *b, double *x, double tol, int maxit, int stopc,
                  int *iters, double *hist) {
  sysctx_t c = {n, rows, cols, diag, values, precond, NULL, NULL, NULL, NULL};
  if (precond >= 2) {
    /* IterSolver factorizes the (scaled) matrix it is given before HUTI
     * starts (IterSolve.src:418-520, CRS_IncompleteLU) */
    if (precond > 11) return -1;
    if (or_ilu_pattern(n, rows, cols, diag, precond - 2, &c.irows, &c.icols, &c.idiag)) return -1;
    c.ilu = malloc(sizeof(double) * ((size_t)c.irows[n] - 1));
    if (!c.ilu || or_ilu_factor(n, rows, cols, values, c.irows, c.icols, c.idiag, c.ilu)) return -1;
  }
  int info = method == 0 ? huti_cg(&c, b, x, tol, maxit, stopc, iters, hist)
                         : huti_bicgstab(&c, b, x, tol, maxit, stopc, iters, hist);
  free(c.irows);
  free(c.icols);
  free(c.idiag);
  free(c.ilu);
  return info;
}

/* Threads of the row-parallel SpMV (the reference's OMP_NUM_THREADS). */
int or_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* ComputeNorm NormDim 2 (SolverUtils.src:4887-5006). */
double or_compute_norm(int n, const double *x) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += x[i] * x[i];
  return sqrt(s / (double)n);
}

/* SolveLinearSystem iterative branch (SolverUtils.src:7151-7525) with
 * IterSolver's x = 1e-8 start (IterSolve.src:358) and TRESID_SCALED_BYB. */
int or_solve_linear_system(int method, int precond, int scaling, int n, const int *rows,
                           const int *cols, const int *diag, double *values, double *b, double *x,
                           double tol, int maxit, int *iters, double *hist, double *norm) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += b[i] * b[i];
  double bnorm = sqrt(s);
  if (bnorm <= DBL_MIN) {
    for (int i = 0; i < n; i++) x[i] = 0.0;
    *iters = 0;
    *norm = 0.0;
    return HUTI_CONVERGENCE;
  }
  double *D = malloc(sizeof(double) * n);
  double rs = 1.0;
  if (scaling) scale_system(n, rows, cols, diag, values, b, x, bnorm != 0.0, D, &rs);
  int allzero = 1;
  for (int i = 0; i < n; i++)
    if (x[i] != 0.0) { allzero = 0; break; }
  if (allzero)
    for (int i = 0; i < n; i++) x[i] = 1.0e-8;
  int info = or_huti_solve(method, precond, n, rows, cols, diag, values, b, x, tol, maxit, 1, iters, hist);
  if (scaling) backscale_system(n, rows, cols, values, b, x, D, rs);
  free(D);
  *norm = or_compute_norm(n, x);
  return info;
}
