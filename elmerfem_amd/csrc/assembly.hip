// Element matrices and the deterministic CRS gather of libehip.
//
// Element side: ElementInfo (fem/src/ElementDescription.src:2586-3505) at
// the default Gauss points (fem/src/Integration.src:1490-1596) of each
// element type, with the ModelPDE local matrix (modules/ModelPDE.src:75-166)
// or the isotropic StressCompose matrix (fem/src/Stress.src:59-521).
// Reference-element basis tables come from the same monomial description as
// elements.def (inverse Vandermonde, AddElementDescription :73-305).
//
// Glue side: instead of CRS_GlueLocalMatrix's searched atomic adds
// (CRSMatrix.src:507-590) every CRS slot gathers its contributions from a
// precomputed list sorted by element index, so Values are the reference's
// serial element-order sums and identical run to run (no float atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ehip_internal.h"
#include "linalg_kernels.h"
#include "pattern.h"

extern "C" int ehip_internal_upload_pattern(ehip_system *s);

namespace ehip {

// ---------------------------------------------------------------- element descriptions

struct EType {
  int code = 0, dim = 0, npe = 0, ngp = 0;
  std::vector<double> U, V, W;
  std::vector<int> p, q, r;
  std::vector<double> coeff;  // coeff[term*npe + node]
  std::vector<double> gu, gv, gw, gs;
  std::vector<double> basis;  // [ngp][npe]
  std::vector<double> dlb;    // [ngp][npe][3]
};

// ComputeGaussPoints1D (Integration.src:406-570) for n <= 3, where the
// companion matrix is 1x1; Newton refinement and weight renormalisation as
// in the reference.
static void gauss1d(int n, double *pts, double *wts) {
  if (n <= 1) { pts[0] = 0.0; wts[0] = 2.0; return; }
  double P[8] = {0}, P0[8] = {0}, P1[8] = {0}, Q[8] = {0};
  P0[0] = 1; P1[0] = 1; P1[1] = 0;
  for (int i = 1; i <= n - 1; i++) {
    for (int k = 0; k < i + 1; k++) P[k] = (double)(2 * i + 1) * P1[k] / (double)(i + 1);
    for (int k = 2; k < i + 2; k++) P[k] = P[k] - (double)i * P0[k - 2] / (double)(i + 1);
    for (int k = 0; k < i + 1; k++) P0[k] = P1[k];
    for (int k = 0; k < i + 2; k++) P1[k] = P[k];
  }
  int np = (n - (n % 2)) / 2;
  for (int i = 1; i <= np + 1; i++) P[i - 1] = P[2 * i - 2];
  const double eig = -P[1] / P[0];
  double Qs[8];
  for (int i = 0; i < np + 1; i++) Qs[i] = P[i];
  std::memset(P, 0, sizeof(P));
  for (int i = 1; i <= np + 1; i++) P[2 * i - 2] = Qs[i - 1];
  pts[0] = +std::sqrt(eig);
  pts[1] = -std::sqrt(eig);
  if (n % 2 == 1) pts[n - 1] = 0.0;
  auto ev = [](int m, const double *c, double x) { double s = 0; for (int i = 0; i < m + 1; i++) s = s * x + c[i]; return s; };
  for (int i = 0; i < n; i++) Q[i] = P[i] * (double)(n - i);
  for (int i = 0; i < n; i++) {
    double x = pts[i], s;
    for (int j = 0; j < 100; j++) {
      s = ev(n, P, x) / ev(n - 1, Q, x);
      x = x - s;
      if (std::fabs(s) <= std::fabs(x) * DBL_EPSILON) break;
    }
    if (std::fabs(ev(n, P, x)) < std::fabs(ev(n, P, pts[i])) && std::fabs(x - pts[i]) < 1.0e-8) pts[i] = x;
  }
  double sum = 0;
  for (int i = 0; i < n; i++) {
    const double s = ev(n - 1, Q, pts[i]);
    wts[i] = 2.0 / ((1.0 - pts[i] * pts[i]) * (s * s));
  }
  for (int i = 0; i < n; i++) sum += wts[i];
  for (int i = 0; i < n; i++) wts[i] = 2.0 * wts[i] / sum;
}

static bool make_etype(int code, EType &t) {
  static const double H27U[27] = {-1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 0, 0, 0};
  static const double H27V[27] = {-1, -1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, 0, 0, 0};
  static const double H27W[27] = {-1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, -1, 1, 0};
  static const int H27B[27] = {1, 2, 3, 5, 6, 7, 9, 10, 11, 17, 18, 19, 21, 22, 23, 25, 26, 27, 33, 34, 35, 37, 38, 39, 41, 42, 43};
  static const double Q9U[9] = {-1, 1, 1, -1, 0, 1, 0, -1, 0}, Q9V[9] = {-1, -1, 1, 1, -1, 0, 1, 0, 0};
  static const int Q9B[9] = {1, 2, 3, 5, 6, 7, 9, 10, 11};
  static const int B404[4] = {1, 2, 5, 6}, B504[4] = {1, 2, 5, 17}, B808[8] = {1, 2, 5, 6, 17, 18, 21, 22}, B202[2] = {1, 2};
  static const int B303[3] = {1, 2, 5};
  static const double T10U[10] = {0, 1, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0};
  static const double T10V[10] = {0, 0, 1, 0, 0, 0.5, 0.5, 0, 0, 0.5};
  static const double T10W[10] = {0, 0, 0, 1, 0, 0, 0, 0.5, 0.5, 0.5};
  static const int B510[10] = {1, 2, 3, 5, 6, 9, 17, 18, 21, 33};
  const int *B = nullptr;
  t.code = code;
  switch (code) {
    case 202: t.dim = 1; t.npe = 2; B = B202; t.U = {-1, 1}; t.V = {0, 0}; t.W = {0, 0}; break;
    case 303: t.dim = 2; t.npe = 3; B = B303; t.U = {0, 1, 0}; t.V = {0, 0, 1}; t.W = {0, 0, 0}; break;
    case 404: t.dim = 2; t.npe = 4; B = B404; t.U = {-1, 1, 1, -1}; t.V = {-1, -1, 1, 1}; t.W = {0, 0, 0, 0}; break;
    case 409: t.dim = 2; t.npe = 9; B = Q9B; t.U.assign(Q9U, Q9U + 9); t.V.assign(Q9V, Q9V + 9); t.W.assign(9, 0.0); break;
    case 504: t.dim = 3; t.npe = 4; B = B504; t.U = {0, 1, 0, 0}; t.V = {0, 0, 1, 0}; t.W = {0, 0, 0, 1}; break;
    case 808:
      t.dim = 3; t.npe = 8; B = B808;
      t.U = {-1, 1, 1, -1, -1, 1, 1, -1}; t.V = {-1, -1, 1, 1, -1, -1, 1, 1}; t.W = {-1, -1, -1, -1, 1, 1, 1, 1};
      break;
    case 510:
      t.dim = 3; t.npe = 10; B = B510;
      t.U.assign(T10U, T10U + 10); t.V.assign(T10V, T10V + 10); t.W.assign(T10W, T10W + 10);
      break;
    case 827:
      t.dim = 3; t.npe = 27; B = H27B;
      t.U.assign(H27U, H27U + 27); t.V.assign(H27V, H27V + 27); t.W.assign(H27W, H27W + 27);
      break;
    default: return false;
  }
  const int n = t.npe;
  t.p.resize(n); t.q.resize(n); t.r.resize(n);
  for (int k = 0; k < n; k++) { int m = B[k] - 1; t.p[k] = m % 4; t.q[k] = (m / 4) % 4; t.r[k] = m / 16; }
  auto ipow = [](double x, int e) { double v = 1; for (int i = 0; i < e; i++) v *= x; return v; };
  // Vandermonde V(node, term), inverted by Gauss-Jordan with partial pivoting
  std::vector<double> a((size_t)n * n), inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) a[(size_t)i * n + k] = ipow(t.U[i], t.p[k]) * ipow(t.V[i], t.q[k]) * ipow(t.W[i], t.r[k]);
  for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int pr = c;
    for (int r = c + 1; r < n; r++) if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)pr * n + c])) pr = r;
    if (a[(size_t)pr * n + c] == 0.0) return false;
    if (pr != c)
      for (int k = 0; k < n; k++) { std::swap(a[(size_t)c * n + k], a[(size_t)pr * n + k]); std::swap(inv[(size_t)c * n + k], inv[(size_t)pr * n + k]); }
    const double d = a[(size_t)c * n + c];
    for (int k = 0; k < n; k++) { a[(size_t)c * n + k] /= d; inv[(size_t)c * n + k] /= d; }
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; k++) { a[(size_t)r * n + k] -= f * a[(size_t)c * n + k]; inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k]; }
    }
  }
  // V * C = I with C(term, node) = inv(term, node)
  t.coeff.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) t.coeff[(size_t)k * n + i] = inv[(size_t)k * n + i];
  // default Gauss rule
  double pts[3], wts[3];
  if (t.dim == 1) {
    gauss1d(2, pts, wts);
    for (int i = 0; i < 2; i++) { t.gu.push_back(pts[i]); t.gv.push_back(0); t.gw.push_back(0); t.gs.push_back(wts[i]); }
  } else if (code == 303) {
    // GaussPointsTriangle, 3 points (Integration.src:84-92, 776-781)
    static const double u3[3] = {0.16666666666666667, 0.66666666666666667, 0.16666666666666667};
    static const double v3[3] = {0.16666666666666667, 0.16666666666666667, 0.66666666666666667};
    for (int i = 0; i < 3; i++) { t.gu.push_back(u3[i]); t.gv.push_back(v3[i]); t.gw.push_back(0); t.gs.push_back(0.33333333333333333 / 2.0); }
  } else if (t.dim == 2) {
    const int m = code == 409 ? 3 : 2;
    gauss1d(m, pts, wts);
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) { t.gu.push_back(pts[j]); t.gv.push_back(pts[i]); t.gw.push_back(0); t.gs.push_back(wts[i] * wts[j]); }
  } else if (code == 504) {
    static const double u4[4] = {0.1757281246520584, 0.2445310270213291, 0.5556470949048655, 0.0240937534217468};
    static const double v4[4] = {0.5656137776620919, 0.0501800797762026, 0.1487681308666864, 0.2354380116950194};
    static const double w4[4] = {0.2180665126782654, 0.5635595064952189, 0.0350112499848832, 0.1833627308416330};
    for (int i = 0; i < 4; i++) { t.gu.push_back(u4[i]); t.gv.push_back(v4[i]); t.gw.push_back(w4[i]); t.gs.push_back(0.25 / 6.0); }
  } else {
    const int m = (code == 827 || code == 510) ? 3 : 2;
    gauss1d(m, pts, wts);
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++)
        for (int k = 0; k < m; k++) {
          double u = pts[k], v = pts[j], w = pts[i], sw = wts[i] * wts[j] * wts[k];
          if (code == 510) {
            // Tet10's 27 points: the brick rule collapsed onto the tetrahedron
            // (GaussPointsTetra, Integration.src:944-963)
            double f = 0.5;
            u = (u + 1) * f; v = (v + 1) * f; w = (w + 1) * f;
            sw = sw * (f * f * f);
            f = 1.0 - w;
            u = u * f; v = v * f;
            sw = sw * (f * f);
            f = 1.0 - v / f;
            u = u * f;
            sw = sw * f;
          }
          t.gu.push_back(u); t.gv.push_back(v); t.gw.push_back(w); t.gs.push_back(sw);
        }
  }
  t.ngp = (int)t.gu.size();
  t.basis.assign((size_t)t.ngp * n, 0.0);
  t.dlb.assign((size_t)t.ngp * n * 3, 0.0);
  for (int g = 0; g < t.ngp; g++)
    for (int i = 0; i < n; i++) {
      double s = 0, du = 0, dv = 0, dw = 0;
      const double u = t.gu[g], v = t.gv[g], w = t.gw[g];
      for (int k = 0; k < n; k++) {
        const double c = t.coeff[(size_t)k * n + i];
        if (c == 0.0) continue;
        const int P = t.p[k], Qd = t.q[k], R = t.r[k];
        s += c * ipow(u, P) * ipow(v, Qd) * ipow(w, R);
        if (P >= 1) du += c * P * ipow(u, P - 1) * ipow(v, Qd) * ipow(w, R);
        if (Qd >= 1) dv += c * Qd * ipow(u, P) * ipow(v, Qd - 1) * ipow(w, R);
        if (R >= 1) dw += c * R * ipow(u, P) * ipow(v, Qd) * ipow(w, R - 1);
      }
      t.basis[(size_t)g * n + i] = s;
      t.dlb[((size_t)g * n + i) * 3 + 0] = du;
      t.dlb[((size_t)g * n + i) * 3 + 1] = dv;
      t.dlb[((size_t)g * n + i) * 3 + 2] = dw;
    }
  return true;
}

// ElementInfo's metric for one Gauss point (ElementMetric :6757-6884 and
// InvertMatrix3x3 :8536-8556).  dx[c][k] = sum_a X_c(a) dLB(a,k);
// G = dx^T dx over cdim; returns LtoG (cdim x dim) and sqrt(det G).
template <int DIM, int CDIM>
__device__ __forceinline__ bool metric(const double (&dx)[3][3], double (&L)[3][3], double &detj) {
  double G[3][3], M[3][3];
#pragma unroll
  for (int i = 0; i < DIM; i++)
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < CDIM; k++) s = s + dx[k][i] * dx[k][j];
      G[i][j] = s;
    }
  double detg;
  if constexpr (DIM == 1) {
    detg = G[0][0];
    if (!(detg > DBL_MIN)) return false;
    M[0][0] = 1.0 / detg;
  } else if constexpr (DIM == 2) {
    detg = G[0][0] * G[1][1] - G[0][1] * G[1][0];
    if (!(detg > DBL_MIN)) return false;
    M[0][0] = G[1][1] / detg;
    M[0][1] = -G[0][1] / detg;
    M[1][0] = -G[1][0] / detg;
    M[1][1] = G[0][0] / detg;
  } else {
    detg = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) + G[0][1] * (G[1][2] * G[2][0] - G[1][0] * G[2][2]) +
           G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    if (!(detg > DBL_MIN)) return false;
    const double s = 1.0 / detg;
    M[0][0] = s * (G[1][1] * G[2][2] - G[2][1] * G[1][2]);
    M[1][0] = -s * (G[1][0] * G[2][2] - G[2][0] * G[1][2]);
    M[2][0] = s * (G[1][0] * G[2][1] - G[2][0] * G[1][1]);
    M[0][1] = -s * (G[0][1] * G[2][2] - G[2][1] * G[0][2]);
    M[1][1] = s * (G[0][0] * G[2][2] - G[2][0] * G[0][2]);
    M[2][1] = -s * (G[0][0] * G[2][1] - G[2][0] * G[0][1]);
    M[0][2] = s * (G[0][1] * G[1][2] - G[1][1] * G[0][2]);
    M[1][2] = -s * (G[0][0] * G[1][2] - G[1][0] * G[0][2]);
    M[2][2] = s * (G[0][0] * G[1][1] - G[1][0] * G[0][1]);
  }
  detj = sqrt(detg);
#pragma unroll
  for (int i = 0; i < CDIM; i++)
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; k++) s = s + dx[i][k] * M[k][j];
      L[i][j] = s;
    }
  return true;
}

struct ElemArgs {
  int nelem;
  const int *el;        // 0-based node ids, nelem*npe
  const double *xyz;    // nnodes*3
  const double *basis;  // [ngp][npe]
  const double *dlb;    // [ngp][npe][3]
  const double *gs;     // [ngp]
  const double *coef, *conv, *velo, *load;  // per element (may be null)
  double poisson;
  double *emat;         // element blocks: k_element, ld x ld row-major one after another; k_element_reg,
                        // row-major or packed upper triangles, component-major (entry k at k * nelem + e)
  double *evec;         // nelem * ld, laid out as emat
  int *bad;             // degenerate element counter
};

// One workgroup handles EPB elements.  Phase 1: per-Gauss-point metric and
// material values; phase 2: global basis gradients; phase 3: one lane per
// (p, q) node pair accumulates over the Gauss points in Gauss-point order.
template <int NPE, int NGP, int DIM, int CDIM, int PHYS, int EPB, int BS>
__global__ __launch_bounds__(BS) void k_element(ElemArgs a) {
  constexpr int DOFS = PHYS == EHIP_PHYS_ELASTIC ? 3 : 1;
  constexpr int LD = NPE * DOFS;
  __shared__ double X[EPB][NPE][3];
  __shared__ double Lg[EPB][NGP][3][3];
  __shared__ double Wg[EPB][NGP];
  __shared__ double Mat[EPB][NGP][6];  // D, C, a0, a1, a2, load  (ModelPDE) | Young, Poisson (elasticity)
  __shared__ double dB[EPB][NGP][NPE][3];
  __shared__ double Bs[NGP][NPE];
  for (int k = threadIdx.x; k < NGP * NPE; k += BS) Bs[k / NPE][k % NPE] = a.basis[k];
  // a fixed grid walks the element batches: far fewer workgroups to
  // dispatch than one per batch
  for (int e0 = blockIdx.x * EPB; e0 < a.nelem; e0 += gridDim.x * EPB) {
  __syncthreads();  // the previous batch's readers are done with the LDS tiles
  for (int k = threadIdx.x; k < EPB * NPE; k += BS) {
    const int le = k / NPE, nd = k % NPE, e = e0 + le;
    if (e < a.nelem) {
      const int g = a.el[(size_t)e * NPE + nd];
      X[le][nd][0] = a.xyz[3 * (size_t)g];
      X[le][nd][1] = a.xyz[3 * (size_t)g + 1];
      X[le][nd][2] = a.xyz[3 * (size_t)g + 2];
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < EPB * NGP; k += BS) {
    const int le = k / NGP, gp = k % NGP, e = e0 + le;
    if (e >= a.nelem) continue;
    double dx[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int nd = 0; nd < NPE; nd++) s += X[le][nd][c] * a.dlb[((size_t)gp * NPE + nd) * 3 + i];
        dx[c][i] = s;
      }
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, detj = 0.0;
    if (!metric<DIM, CDIM>(dx, L, detj)) {
      atomicAdd(a.bad, 1);
      detj = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) Lg[le][gp][i][j] = L[i][j];
    Wg[le][gp] = a.gs[gp] * detj;
    // interpolate per-element (nodal-constant) coefficients with the basis, as GetReal + SUM(Basis*...)
    double sb[6] = {0, 0, 0, 0, 0, 0};
    const double c0 = a.coef ? a.coef[e] : 0.0;
    const double c1 = PHYS == EHIP_PHYS_ELASTIC ? a.poisson : (a.conv ? a.conv[e] : 0.0);
    const double v0 = a.velo ? a.velo[3 * (size_t)e] : 0.0, v1 = a.velo ? a.velo[3 * (size_t)e + 1] : 0.0,
                 v2 = a.velo ? a.velo[3 * (size_t)e + 2] : 0.0;
    const double ld = a.load ? a.load[e] : 0.0;
    for (int nd = 0; nd < NPE; nd++) {
      const double b = Bs[gp][nd];
      sb[0] += b * c0;
      sb[1] += b * c1;
      sb[2] += v0 * b;
      sb[3] += v1 * b;
      sb[4] += v2 * b;
      sb[5] += b * ld;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) Mat[le][gp][i] = sb[i];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < EPB * NGP * NPE; k += BS) {
    const int le = k / (NGP * NPE), rem = k % (NGP * NPE), gp = rem / NPE, nd = rem % NPE;
    const double *d = a.dlb + ((size_t)gp * NPE + nd) * 3;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      if (j < CDIM)
#pragma unroll
        for (int kk = 0; kk < DIM; kk++) s = s + d[kk] * Lg[le][gp][j][kk];
      dB[le][gp][nd][j] = s;
    }
  }
  __syncthreads();
  if constexpr (PHYS == EHIP_PHYS_MODELPDE) {
    for (int k = threadIdx.x; k < EPB * NPE * NPE; k += BS) {
      const int le = k / (NPE * NPE), rem = k % (NPE * NPE), p = rem / NPE, q = rem % NPE, e = e0 + le;
      if (e >= a.nelem) continue;
      double acc = 0.0;
      for (int gp = 0; gp < NGP; gp++) {
        const double W = Wg[le][gp];
        const double WD = W * Mat[le][gp][0];
        const double m = dB[le][gp][p][0] * dB[le][gp][q][0] + dB[le][gp][p][1] * dB[le][gp][q][1] +
                         dB[le][gp][p][2] * dB[le][gp][q][2];
        acc = acc + WD * m;
        const double C = Mat[le][gp][1];
        if (C != 0.0) {
          double sa = 0.0;
#pragma unroll
          for (int kk = 0; kk < CDIM; kk++) sa += Mat[le][gp][2 + kk] * dB[le][gp][q][kk];
          acc = acc + W * C * sa * Bs[gp][p];
        }
      }
      a.emat[(size_t)e * LD * LD + p * LD + q] = acc;
    }
    for (int k = threadIdx.x; k < EPB * NPE; k += BS) {
      const int le = k / NPE, p = k % NPE, e = e0 + le;
      if (e >= a.nelem) continue;
      double f = 0.0;
      for (int gp = 0; gp < NGP; gp++) f = f + Wg[le][gp] * Mat[le][gp][5] * Bs[gp][p];
      a.evec[(size_t)e * LD + p] = f;
    }
  } else {
    for (int k = threadIdx.x; k < EPB * NPE * NPE; k += BS) {
      const int le = k / (NPE * NPE), rem = k % (NPE * NPE), p = rem / NPE, q = rem % NPE, e = e0 + le;
      if (e >= a.nelem) continue;
      double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int gp = 0; gp < NGP; gp++) {
        const double s = Wg[le][gp];
        const double E = Mat[le][gp][0], nu = Mat[le][gp][1];
        const double den = (1 + nu) * (1 - 2 * nu);
        const double c11 = (1.0 - nu) * E / den, c12 = nu * E / den, c44 = (0.5 - nu) * E / den;
        const double px = dB[le][gp][p][0], py = dB[le][gp][p][1], pz = dB[le][gp][p][2];
        const double qx = dB[le][gp][q][0], qy = dB[le][gp][q][1], qz = dB[le][gp][q][2];
        // G_p C (rows i = 1..3 over the 6 Voigt strains), then times B_q
        const double g00 = px * c11, g01 = px * c12, g02 = px * c12, g03 = py * c44, g05 = pz * c44;
        const double g10 = py * c12, g11 = py * c11, g12 = py * c12, g13 = px * c44, g14 = pz * c44;
        const double g20 = pz * c12, g21 = pz * c12, g22 = pz * c11, g24 = py * c44, g25 = px * c44;
        A[0][0] = A[0][0] + s * (g00 * qx + g03 * qy + g05 * qz);
        A[0][1] = A[0][1] + s * (g01 * qy + g03 * qx);
        A[0][2] = A[0][2] + s * (g02 * qz + g05 * qx);
        A[1][0] = A[1][0] + s * (g10 * qx + g13 * qy);
        A[1][1] = A[1][1] + s * (g11 * qy + g13 * qx + g14 * qz);
        A[1][2] = A[1][2] + s * (g12 * qz + g14 * qy);
        A[2][0] = A[2][0] + s * (g20 * qx + g25 * qz);
        A[2][1] = A[2][1] + s * (g21 * qy + g24 * qz);
        A[2][2] = A[2][2] + s * (g22 * qz + g24 * qy + g25 * qx);
      }
      double *out = a.emat + (size_t)e * LD * LD;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) out[(3 * p + i) * LD + 3 * q + j] = A[i][j];
    }
    for (int k = threadIdx.x; k < EPB * LD; k += BS) {
      const int le = k / LD, i = k % LD, e = e0 + le;
      if (e < a.nelem) a.evec[(size_t)e * LD + i] = 0.0;
    }
  }
  }  // element batches
}

// Thread-per-element ModelPDE blocks for the first-order elements (Tri3,
// Quad4, Tet4, Hex8): coordinates, every Gauss point's metric and global
// basis gradients, and the NPE x NPE block stay in registers; the element
// tables are read with wave-uniform loads.  The arithmetic is k_element's,
// term for term and in the same order.  SYM (no convection term): only
// p <= q is accumulated and the block is stored as its packed upper
// triangle (row p holds q = p..NPE-1).
template <int NPE>
__host__ __device__ constexpr int tri_index(int p, int q) {
  return p * NPE - p * (p - 1) / 2 + (q - p);
}

// The element tables by value: they travel in the kernel arguments, so every
// read is a scalar load shared by the wave (2.1 KB for Hex8).
template <int NPE, int NGP>
struct RegTables {
  double basis[NGP * NPE];
  double dlb[NGP * NPE * 3];
  double gs[NGP];
};

template <int NPE, int NGP, int DIM, int CDIM, bool SYM>
__global__ __launch_bounds__(256) void k_element_reg(ElemArgs a, RegTables<NPE, NGP> t) {
  constexpr int NK = SYM ? NPE * (NPE + 1) / 2 : NPE * NPE;
  // the element's node coordinates, loaded once into this thread's own LDS
  // column (no barrier: no thread reads another's) and read again at every
  // Gauss point -- held in registers they would spill the block, the
  // gradients and the metric; 48 KB for Hex8, within what the register-
  // limited occupancy leaves per workgroup
  __shared__ double xs[NPE * 3][256];
  const int tid = (int)threadIdx.x;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < a.nelem; e += gridDim.x * blockDim.x) {
    {
      int gid[NPE];
#pragma unroll
      for (int nd = 0; nd < NPE; nd++) gid[nd] = a.el[(size_t)e * NPE + nd];
#pragma unroll
      for (int nd = 0; nd < NPE; nd++)
#pragma unroll
        for (int c = 0; c < 3; c++) xs[nd * 3 + c][tid] = a.xyz[3 * (size_t)gid[nd] + c];
    }
    const double c0 = a.coef ? a.coef[e] : 0.0;
    const double c1 = (!SYM && a.conv) ? a.conv[e] : 0.0;
    const double v0 = (!SYM && a.velo) ? a.velo[3 * (size_t)e] : 0.0;
    const double v1 = (!SYM && a.velo) ? a.velo[3 * (size_t)e + 1] : 0.0;
    const double v2 = (!SYM && a.velo) ? a.velo[3 * (size_t)e + 2] : 0.0;
    const double ld = a.load ? a.load[e] : 0.0;
    double K[NK], F[NPE];
#pragma unroll
    for (int k = 0; k < NK; k++) K[k] = 0.0;
#pragma unroll
    for (int p = 0; p < NPE; p++) F[p] = 0.0;
#pragma unroll 1
    for (int gp = 0; gp < NGP; gp++) {
      double dx[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double X[NPE];
#pragma unroll
        for (int nd = 0; nd < NPE; nd++) X[nd] = xs[nd * 3 + c][tid];
#pragma unroll
        for (int i = 0; i < DIM; i++) {
          double sdx = 0.0;
#pragma unroll
          for (int nd = 0; nd < NPE; nd++) sdx += X[nd] * t.dlb[((size_t)gp * NPE + nd) * 3 + i];
          dx[c][i] = sdx;
        }
      }
      double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, detj = 0.0;
      if (!metric<DIM, CDIM>(dx, L, detj)) {
        atomicAdd(a.bad, 1);
        detj = 0.0;
      }
      const double W = t.gs[gp] * detj;
      double sb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int nd = 0; nd < NPE; nd++) {
        const double b = t.basis[gp * NPE + nd];
        sb[0] += b * c0;
        sb[1] += b * c1;
        sb[2] += v0 * b;
        sb[3] += v1 * b;
        sb[4] += v2 * b;
        sb[5] += b * ld;
      }
      double dB[NPE][3];
#pragma unroll
      for (int nd = 0; nd < NPE; nd++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          double sv = 0.0;
          if (j < CDIM)
#pragma unroll
            for (int kk = 0; kk < DIM; kk++) sv = sv + t.dlb[((size_t)gp * NPE + nd) * 3 + kk] * L[j][kk];
          dB[nd][j] = sv;
        }
      const double WD = W * sb[0];
      const double C = sb[1];
#pragma unroll
      for (int p = 0; p < NPE; p++) {
#pragma unroll
        for (int q = SYM ? p : 0; q < NPE; q++) {
          const double m = dB[p][0] * dB[q][0] + dB[p][1] * dB[q][1] + dB[p][2] * dB[q][2];
          double acc = K[SYM ? tri_index<NPE>(p, q) : p * NPE + q];
          acc = acc + WD * m;
          if (!SYM && C != 0.0) {
            double sa = 0.0;
#pragma unroll
            for (int kk = 0; kk < CDIM; kk++) sa += sb[2 + kk] * dB[q][kk];
            acc = acc + W * C * sa * t.basis[gp * NPE + p];
          }
          K[SYM ? tri_index<NPE>(p, q) : p * NPE + q] = acc;
        }
        F[p] = F[p] + W * sb[5] * t.basis[gp * NPE + p];
      }
    }
    // packed (SYM) or row-major blocks, stored component-major: entry k of
    // element e at k * nelem + e, so each store of a wave covers consecutive
    // addresses (and so do the gather's loads when neighbouring rows take
    // neighbouring elements).  The full row-major block with the lower
    // triangle mirrored measured the same in total (element kernel +31%,
    // gather -19% at 128^3 Hex8) with 470 MB more traffic.
    const size_t ne = (size_t)a.nelem;
#pragma unroll
    for (int k = 0; k < NK; k++) a.emat[k * ne + e] = K[k];
#pragma unroll
    for (int p = 0; p < NPE; p++) a.evec[p * ne + e] = F[p];
  }
}

// Gather (the additions of DefaultUpdateEquations / CRS_GlueLocalMatrix,
// CRSMatrix.src:507-590) over the row-owner maps: one thread per matrix row
// adds, for every element touching the row in element order, that
// element's block row into the row's slots at the column positions of the
// maps, and its force entry into the right-hand side -- every slot receives
// its contributions in element order, starting from zero.  A workgroup's
// BS rows have consecutive slots: when they fit in CAP doubles they are
// accumulated in LDS (zeroed, added, written out with coalesced stores),
// otherwise in place in Values.
// SOA: blocks stored component-major (entry k of element e at k * ne + e,
// force entry i at i * ne + e: k_element_reg), else block after block
// (k_element).
template <int LD, bool SYM, bool SOA, typename OffT>
__device__ __forceinline__ double row_accumulate(int r, double *dst, const int *__restrict__ mbase,
                                                 const int *__restrict__ rcontrib, const OffT *__restrict__ roff,
                                                 const double *__restrict__ emat, const double *__restrict__ evec,
                                                 size_t ne) {
  constexpr int NK = SYM ? LD * (LD + 1) / 2 : LD * LD;
  // the block row in pieces of CH columns, and the entries of small blocks
  // in groups of G: a piece's loads (of every entry in the group) go out
  // before its additions, which stay in entry order, within an entry in
  // column order
  constexpr int CH = LD <= 16 ? LD : (LD % 9 == 0 ? 9 : (LD % 10 == 0 ? 10 : (LD % 8 == 0 ? 8 : 1)));
  constexpr int G = LD <= 8 ? 4 : 1;
  // one word for an entry's column positions when they are the bytes of a
  // whole aligned word (the map is 256-byte aligned, an entry LD bytes)
  constexpr bool kWord = sizeof(OffT) == 1 && CH == LD && (LD == 4 || LD == 8);
  double f = 0.0;
  // slice-major maps: entry k of slice row l at b + 64 k + l; the slice's
  // entry count is the same for its 64 rows (shorter rows padded with -1)
  const int sl = r >> 6, l = r & 63;
  const int b = mbase[sl], cnt = (mbase[sl + 1] - b) >> 6;
  for (int kg = 0; kg < cnt; kg += G) {
    int code[G];
#pragma unroll
    for (int g = 0; g < G; ++g) code[g] = kg + g < cnt ? rcontrib[b + (kg + g) * 64 + l] : -1;
    for (int l0 = 0; l0 < LD; l0 += CH) {
      double v[G][CH];
      int o[G][CH];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int cg = code[g] >= 0 ? code[g] : 0;  // a padding entry loads element 0 and adds nothing
        const int e = cg / LD, li = cg - e * LD;
        const OffT *off = roff + ((size_t)b + (size_t)min(kg + g, cnt - 1) * 64 + l) * LD;
        uint64_t ow = 0;
        if constexpr (kWord) {
          if constexpr (LD == 8) ow = *reinterpret_cast<const uint64_t *>(off);
          else ow = *reinterpret_cast<const uint32_t *>(off);
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int lj = l0 + u;
          const int idx = SYM ? (li <= lj ? tri_index<LD>(li, lj) : tri_index<LD>(lj, li)) : li * LD + lj;
          v[g][u] = SOA ? emat[(size_t)idx * ne + e] : emat[(size_t)e * NK + idx];
          if constexpr (kWord) o[g][u] = (int)((ow >> (8 * u)) & 0xff);
          else o[g][u] = (int)off[lj];
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (code[g] >= 0)
#pragma unroll
          for (int u = 0; u < CH; ++u) dst[o[g][u]] += v[g][u];
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (code[g] >= 0) {
        const int e = code[g] / LD, li = code[g] - e * LD;
        f = f + (SOA ? evec[(size_t)li * ne + e] : evec[code[g]]);
      }
  }
  return f;
}

template <int LD, bool SYM, bool SOA, typename OffT, int BS, int CAP>
__global__ __launch_bounds__(BS) void k_row_gather(int n, const int *__restrict__ rows,
                                                    const int *__restrict__ mbase, const int *__restrict__ rcontrib,
                                                    const OffT *__restrict__ roff, const double *__restrict__ emat,
                                                    const double *__restrict__ evec, int nelem, double *vals,
                                                    double *rhs) {
  const size_t ne = (size_t)nelem;
  __shared__ double acc[CAP];
  const XcdWalk w((n + BS - 1) / BS);
  for (int rb = w.first; rb < w.last; rb += w.step) {
    const int r0 = rb * BS;
    const int r1 = min(r0 + BS, n);
    const int base = rows[r0], len = rows[r1] - base;
    const int r = r0 + (int)threadIdx.x;
    if (len <= CAP) {
      for (int j = threadIdx.x; j < len; j += BS) acc[j] = 0.0;
      __syncthreads();
      if (r < r1)
        rhs[r] = row_accumulate<LD, SYM, SOA, OffT>(r, acc + (rows[r] - base), mbase, rcontrib, roff, emat, evec,
                                                    ne);
      __syncthreads();
      for (int j = threadIdx.x; j < len; j += BS) vals[base + j] = acc[j];
      __syncthreads();
    } else if (r < r1) {
      for (int j = rows[r]; j < rows[r + 1]; ++j) vals[j] = 0.0;
      rhs[r] = row_accumulate<LD, SYM, SOA, OffT>(r, vals + rows[r], mbase, rcontrib, roff, emat, evec, ne);
    }
  }
}

// Slice-major gather maps from the row-major ones: one thread per slice
// row (padding rows of the last slice included), entry k of slice row l
// written at base[s] + 64 k + l, rows shorter than the slice's longest
// padded with code -1 and zero positions.
__global__ void k_slice_count(int n, int nslice, const int *__restrict__ rs, int *cnt) {
  for (int sl = blockIdx.x * blockDim.x + threadIdx.x; sl < nslice; sl += gridDim.x * blockDim.x) {
    int m = 0;
    for (int r = sl * 64; r < min(n, sl * 64 + 64); ++r) m = max(m, rs[r + 1] - rs[r]);
    cnt[sl] = m;
  }
}

template <typename OffT>
__global__ void k_slice_maps(int n, int nslice, int ld, const int *__restrict__ rs, const int *__restrict__ rc,
                             const OffT *__restrict__ ro, const int *__restrict__ base, int *c2, OffT *o2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nslice * 64; i += gridDim.x * blockDim.x) {
    const int sl = i >> 6, l = i & 63;
    const int b = base[sl], cnt = (base[sl + 1] - b) >> 6;
    const int k0 = i < n ? rs[i] : 0, len = i < n ? rs[i + 1] - k0 : 0;
    for (int k = 0; k < cnt; ++k) {
      const int64_t d = (int64_t)b + (int64_t)k * 64 + l;
      c2[d] = k < len ? rc[k0 + k] : -1;
      for (int j = 0; j < ld; ++j) o2[d * ld + j] = k < len ? ro[(int64_t)(k0 + k) * ld + j] : (OffT)0;
    }
  }
}

int slice_maps(ehip_system *s) {
  const int n = s->n, S = (n + 63) / 64, ld = s->npe * s->dofs;
  const size_t ob = (size_t)s->row_off_bytes;
  int nent = 0;
  EHIP_HIP_TRY(hipMemcpy(&nent, s->rhs_start + n, sizeof(int), hipMemcpyDeviceToHost));
  int *cnt = nullptr;
  int rc;
  if ((rc = dmalloc(s, (void **)&cnt, sizeof(int) * S))) return rc;
  hipLaunchKernelGGL(k_slice_count, dim3(std::min((S + 255) / 256, 4096)), dim3(256), 0, s->stream, n, S,
                     s->rhs_start, cnt);
  std::vector<int> h((size_t)S), base((size_t)S + 1);
  EHIP_HIP_TRY(hipMemcpyAsync(h.data(), cnt, sizeof(int) * S, hipMemcpyDeviceToHost, s->stream));
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
  dfree(s, cnt, sizeof(int) * S);
  int64_t tot = 0;
  for (int sl = 0; sl < S; ++sl) {
    base[sl] = (int)tot;
    tot += 64 * (int64_t)h[sl];
    if (tot * ld >= (int64_t)INT32_MAX) {
      set_error("ehip_system_from_mesh: padded gather maps exceed 2^31 entries");
      return EHIP_ECAP;
    }
  }
  base[S] = (int)tot;
  int *mb = nullptr, *c2 = nullptr;
  void *o2 = nullptr;
  if ((rc = dmalloc(s, (void **)&mb, sizeof(int) * (S + 1))) ||
      (rc = dmalloc(s, (void **)&c2, sizeof(int) * std::max<int64_t>(tot, 1))) ||
      (rc = dmalloc(s, &o2, std::max<size_t>((size_t)tot * ld * ob, 8)))) {
    dfree(s, mb, sizeof(int) * (S + 1));
    dfree(s, c2, sizeof(int) * std::max<int64_t>(tot, 1));
    return rc;
  }
  EHIP_HIP_TRY(hipMemcpyAsync(mb, base.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice, s->stream));
  const int grid = std::min((S * 64 + 255) / 256, 16384);
  if (ob == 1)
    hipLaunchKernelGGL(k_slice_maps<uint8_t>, dim3(grid), dim3(256), 0, s->stream, n, S, ld, s->rhs_start,
                       s->rhs_contrib, (const uint8_t *)s->row_off, mb, c2, (uint8_t *)o2);
  else
    hipLaunchKernelGGL(k_slice_maps<uint16_t>, dim3(grid), dim3(256), 0, s->stream, n, S, ld, s->rhs_start,
                       s->rhs_contrib, (const uint16_t *)s->row_off, mb, c2, (uint16_t *)o2);
  EHIP_HIP_TRY(hipGetLastError());
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
  dfree(s, s->rhs_start, sizeof(int) * ((size_t)n + 1));
  dfree(s, s->rhs_contrib, sizeof(int) * std::max(nent, 1));
  dfree(s, s->row_off, (size_t)nent * ld * ob);
  s->rhs_start = nullptr;
  s->map_base = mb;
  s->rhs_contrib = c2;
  s->row_off = o2;
  s->map_entries = tot;
  return EHIP_OK;
}

struct DevTables {
  double *basis = nullptr, *dlb = nullptr, *gs = nullptr;
};

__global__ void k_zero_int(int *p) { *p = 0; }

// Rows per workgroup and LDS doubles: 128 rows in 32 KB for blocks of up to
// 8 rows (rows of up to 32 entries on average stay in LDS), 64 rows in 64 KB
// for blocks of up to 30 rows, 16 rows in 64 KB for the 81-row Hex27
// elasticity blocks (rows of up to 375 entries)
template <int LD, bool SYM, bool SOA>
static void row_gather_as(ehip_system *s) {
  constexpr int BS = LD <= 8 ? 128 : (LD <= 30 ? 64 : 16);
  constexpr int CAP = LD <= 8 ? 4096 : 8192;
  const int grid = (std::min((s->n + BS - 1) / BS, 32768) + 7) / 8 * 8;  // a multiple of 8 (XcdWalk)
  if (s->row_off_bytes == 1)
    hipLaunchKernelGGL((k_row_gather<LD, SYM, SOA, uint8_t, BS, CAP>), dim3(grid), dim3(BS), 0, s->stream, s->n,
                       s->rows, s->map_base, s->rhs_contrib, (const uint8_t *)s->row_off, s->emat, s->evec,
                       s->nelem, s->vals, s->rhs);
  else
    hipLaunchKernelGGL((k_row_gather<LD, SYM, SOA, uint16_t, BS, CAP>), dim3(grid), dim3(BS), 0, s->stream, s->n,
                       s->rows, s->map_base, s->rhs_contrib, (const uint16_t *)s->row_off, s->emat, s->evec,
                       s->nelem, s->vals, s->rhs);
}

// The gathers compiled: block row lengths of the supported element types
// (Tri3, Quad4/Tet4, Hex8 scalar from k_element_reg, component-major;
// Tet10, Hex27 scalar and Hex8, Tet10, Hex27 with 3 dofs from k_element,
// block after block)
static int launch_row_gather(ehip_system *s, int ld, bool packed) {
  if (packed) {
    switch (ld) {
      case 3: row_gather_as<3, true, true>(s); return EHIP_OK;
      case 4: row_gather_as<4, true, true>(s); return EHIP_OK;
      case 8: row_gather_as<8, true, true>(s); return EHIP_OK;
    }
  } else {
    switch (ld) {
      case 3: row_gather_as<3, false, true>(s); return EHIP_OK;
      case 4: row_gather_as<4, false, true>(s); return EHIP_OK;
      case 8: row_gather_as<8, false, true>(s); return EHIP_OK;
      case 10: row_gather_as<10, false, false>(s); return EHIP_OK;
      case 24: row_gather_as<24, false, false>(s); return EHIP_OK;
      case 27: row_gather_as<27, false, false>(s); return EHIP_OK;
      case 30: row_gather_as<30, false, false>(s); return EHIP_OK;
      case 81: row_gather_as<81, false, false>(s); return EHIP_OK;
    }
  }
  set_error("ehip_assemble: no gather for element blocks of %d rows", ld);
  return EHIP_EUNSUP;
}

template <int NPE, int NGP, int DIM, int CDIM, int PHYS, int EPB, int BS>
static void launch_element(ehip_system *s, const ElemArgs &a) {
  const int grid = std::min((a.nelem + EPB - 1) / EPB, 8192);
  hipLaunchKernelGGL((k_element<NPE, NGP, DIM, CDIM, PHYS, EPB, BS>), dim3(grid), dim3(BS), 0, s->stream, a);
}

template <int NPE, int NGP, int DIM, int CDIM>
static void launch_element_reg(ehip_system *s, const ElemArgs &a, const EType &et, bool *packed) {
  const int grid = std::min((a.nelem + 255) / 256, 16384);
  RegTables<NPE, NGP> t;
  std::copy(et.basis.begin(), et.basis.end(), t.basis);
  std::copy(et.dlb.begin(), et.dlb.end(), t.dlb);
  std::copy(et.gs.begin(), et.gs.end(), t.gs);
  *packed = a.conv == nullptr && a.velo == nullptr;
  if (*packed)
    hipLaunchKernelGGL((k_element_reg<NPE, NGP, DIM, CDIM, true>), dim3(grid), dim3(256), 0, s->stream, a, t);
  else
    hipLaunchKernelGGL((k_element_reg<NPE, NGP, DIM, CDIM, false>), dim3(grid), dim3(256), 0, s->stream, a, t);
}

// Element blocks: first-order ModelPDE elements one thread per element
// (packed symmetric blocks without convection), the rest one workgroup per
// batch of elements.  *packed tells the gather which block layout it reads.
static int dispatch_element(ehip_system *s, int physics, int code, int cdim, const ElemArgs &a, const EType &et,
                            bool *packed) {
  *packed = false;
  if (physics == EHIP_PHYS_MODELPDE) {
    if (code == 808 && cdim == 3) { launch_element_reg<8, 8, 3, 3>(s, a, et, packed); return EHIP_OK; }
    if (code == 504 && cdim == 3) { launch_element_reg<4, 4, 3, 3>(s, a, et, packed); return EHIP_OK; }
    if (code == 404 && cdim == 2) { launch_element_reg<4, 4, 2, 2>(s, a, et, packed); return EHIP_OK; }
    if (code == 303 && cdim == 2) { launch_element_reg<3, 3, 2, 2>(s, a, et, packed); return EHIP_OK; }
    if (code == 510 && cdim == 3) { launch_element<10, 27, 3, 3, 0, 2, 256>(s, a); return EHIP_OK; }
    if (code == 827 && cdim == 3) { launch_element<27, 27, 3, 3, 0, 1, 256>(s, a); return EHIP_OK; }
  } else {
    if (code == 808) { launch_element<8, 8, 3, 3, 1, 8, 256>(s, a); return EHIP_OK; }
    if (code == 827) { launch_element<27, 27, 3, 3, 1, 1, 256>(s, a); return EHIP_OK; }
    if (code == 510) { launch_element<10, 27, 3, 3, 1, 2, 256>(s, a); return EHIP_OK; }
  }
  set_error("ehip_assemble: element %d / physics %d / cdim %d not supported", code, physics, cdim);
  return EHIP_EUNSUP;
}

}  // namespace ehip

using namespace ehip;

// The host build (pattern.cpp, OpenMP) -- used when EHIP_PATTERN_BUILD=host.
static int from_mesh_host(int nnodes, int nelem, int npe, const int *elements, int nbnd, int npe_bnd,
                          const int *bnd_elements, int dofs, int optimize, ehip_system **out) {
  HostPattern P;
  int prc = build_host_pattern(nnodes, nelem, npe, elements, bnd_elements ? nbnd : 0, npe_bnd, bnd_elements, dofs,
                               optimize, P);
  if (prc) {
    set_error("ehip_system_from_mesh: pattern build failed (%d): node ids out of range, unused nodes or nnz >= 2^31", prc);
    return prc == -3 ? EHIP_ECAP : EHIP_EARG;
  }
  std::vector<int> rstart, rcontrib;
  std::vector<uint16_t> roff;
  int max_len = 0;
  prc = build_row_maps(P, nelem, npe, elements, rstart, rcontrib, roff, &max_len);
  if (prc) {
    set_error("ehip_system_from_mesh: element blocks exceed int32 or a row is longer than 65536 (%d)", prc);
    return EHIP_ECAP;
  }
  ehip_system *s = new ehip_system();
  EHIP_HIP_TRY(hipGetDevice(&s->dev));
  EHIP_HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  s->n = P.n;
  s->nnz = P.nnz;
  s->dofs = dofs;
  s->nnodes = nnodes;
  s->h_rows = std::move(P.rows);
  s->h_cols = std::move(P.cols);
  s->h_diag = std::move(P.diag);
  s->h_perm = std::move(P.perm);
  s->bw_before = P.bw_before;
  s->bw_after = P.bw_after;
  s->bw_accepted = P.accepted;
  int rc = ehip_internal_upload_pattern(s);
  if (rc) { ehip_system_free(s); return rc; }
  s->nelem = nelem;
  s->npe = npe;
  const int ld = npe * dofs;
  std::vector<int> el0((size_t)nelem * npe);
  for (size_t k = 0; k < el0.size(); ++k) el0[k] = elements[k] - 1;
  // column positions in one byte when every row has at most 256 entries
  s->row_off_bytes = max_len <= 256 ? 1 : 2;
  std::vector<uint8_t> roff8;
  if (s->row_off_bytes == 1) {
    roff8.resize(roff.size());
    for (size_t k = 0; k < roff.size(); ++k) roff8[k] = (uint8_t)roff[k];
  }
  const size_t off_bytes = roff.size() * (size_t)s->row_off_bytes;
  if ((rc = dmalloc(s, (void **)&s->elements, sizeof(int) * el0.size())) ||
      (rc = dmalloc(s, (void **)&s->rhs_start, sizeof(int) * rstart.size())) ||
      (rc = dmalloc(s, (void **)&s->rhs_contrib, sizeof(int) * rcontrib.size())) ||
      (rc = dmalloc(s, &s->row_off, off_bytes)) ||
      (rc = dmalloc(s, (void **)&s->emat, sizeof(double) * (size_t)nelem * ld * ld)) ||
      (rc = dmalloc(s, (void **)&s->evec, sizeof(double) * (size_t)nelem * ld))) {
    ehip_system_free(s);
    return rc;
  }
  EHIP_HIP_TRY(hipMemcpy(s->elements, el0.data(), sizeof(int) * el0.size(), hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(s->rhs_start, rstart.data(), sizeof(int) * rstart.size(), hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(s->rhs_contrib, rcontrib.data(), sizeof(int) * rcontrib.size(), hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(s->row_off, s->row_off_bytes == 1 ? (const void *)roff8.data() : (const void *)roff.data(),
                         off_bytes, hipMemcpyHostToDevice));
  if ((rc = slice_maps(s))) {
    ehip_system_free(s);
    return rc;
  }
  *out = s;
  return EHIP_OK;
}

// CreateMatrix on the device (pattern_gpu.hip): the same arrays as the host
// build, bit for bit.
static int from_mesh_device(int nnodes, int nelem, int npe, const int *elements, int nbnd, int npe_bnd,
                            const int *bnd_elements, int dofs, int optimize, ehip_system **out) {
  for (int64_t k = 0; k < (int64_t)nelem * npe; ++k)
    if (elements[k] < 1 || elements[k] > nnodes) {
      set_error("ehip_system_from_mesh: node id %d out of range", elements[k]);
      return EHIP_EARG;
    }
  ehip_system *s = new ehip_system();
  EHIP_HIP_TRY(hipGetDevice(&s->dev));
  EHIP_HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  s->dofs = dofs;
  s->nnodes = nnodes;
  s->nelem = nelem;
  s->npe = npe;
  const int ld = npe * dofs;
  int rc;
  {
    std::vector<int> el0((size_t)nelem * npe);
    for (size_t k = 0; k < el0.size(); ++k) el0[k] = elements[k] - 1;
    if ((rc = dmalloc(s, (void **)&s->elements, sizeof(int) * el0.size()))) {
      ehip_system_free(s);
      return rc;
    }
    EHIP_HIP_TRY(hipMemcpy(s->elements, el0.data(), sizeof(int) * el0.size(), hipMemcpyHostToDevice));
  }
  int max_len = 0;
  rc = build_device_pattern(s, nnodes, nelem, npe, s->elements, bnd_elements ? nbnd : 0, npe_bnd, bnd_elements, dofs,
                            optimize, &max_len);
  if (rc) { ehip_system_free(s); return rc; }
  s->pattern_on_device = true;
  if ((rc = ehip_internal_upload_pattern(s)) || (rc = slice_maps(s)) ||
      (rc = dmalloc(s, (void **)&s->emat, sizeof(double) * (size_t)nelem * ld * ld)) ||
      (rc = dmalloc(s, (void **)&s->evec, sizeof(double) * (size_t)nelem * ld))) {
    ehip_system_free(s);
    return rc;
  }
  *out = s;
  return EHIP_OK;
}

extern "C" int ehip_system_from_mesh(int nnodes, int nelem, int npe, const int *elements, int nbnd, int npe_bnd,
                                     const int *bnd_elements, int dofs, int optimize, ehip_system **out) {
  if (!elements || !out || nnodes <= 0 || nelem <= 0 || npe <= 0 || npe > 27 || dofs < 1 || dofs > 3) {
    set_error("ehip_system_from_mesh: bad arguments");
    return EHIP_EARG;
  }
  const char *how = std::getenv("EHIP_PATTERN_BUILD");
  if (how && std::strcmp(how, "host") == 0)
    return from_mesh_host(nnodes, nelem, npe, elements, nbnd, npe_bnd, bnd_elements, dofs, optimize, out);
  return from_mesh_device(nnodes, nelem, npe, elements, nbnd, npe_bnd, bnd_elements, dofs, optimize, out);
}

namespace ehip {

// Element tables of one type, uploaded once per system and element type.
static int ensure_tables(ehip_system *s, int code, const EType &t, DevTables &d) {
  const size_t nb = t.basis.size(), nd = t.dlb.size(), ng = t.gs.size();
  if (s->asm_code != code || !s->asm_tab) {
    if (s->asm_tab) dfree(s, s->asm_tab, 0);
    s->asm_tab = nullptr;
    s->asm_code = 0;
    int rc;
    if ((rc = dmalloc(s, (void **)&s->asm_tab, sizeof(double) * (nb + nd + ng)))) return rc;
    std::vector<double> h;
    h.reserve(nb + nd + ng);
    h.insert(h.end(), t.basis.begin(), t.basis.end());
    h.insert(h.end(), t.dlb.begin(), t.dlb.end());
    h.insert(h.end(), t.gs.begin(), t.gs.end());
    EHIP_HIP_TRY(hipMemcpy(s->asm_tab, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    s->asm_code = code;
  }
  d.basis = s->asm_tab;
  d.dlb = s->asm_tab + nb;
  d.gs = s->asm_tab + nb + nd;
  return EHIP_OK;
}

// The degenerate-metric count of an asynchronous assembly, once it is done.
int assembly_check(ehip_system *s) {
  if (!s->asm_pending) return EHIP_OK;
  s->asm_pending = 0;
  int hbad = 0;
  EHIP_HIP_TRY(hipMemcpyAsync(&hbad, s->asm_bad, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
  if (hbad) {
    set_error("ehip_assemble: %d degenerate Gauss-point metrics (ElementMetric)", hbad);
    return EHIP_EARG;
  }
  return EHIP_OK;
}

}  // namespace ehip

extern "C" int ehip_assemble(ehip_system *s, int physics, int elem_code, int cdim, const double *coords,
                             const double *elem_coef, const double *elem_conv, const double *elem_velo,
                             const double *elem_load, double poisson, double *t_ms) {
  if (!s || !s->emat) {
    set_error("ehip_assemble: system was not created from a mesh");
    return EHIP_ESTATE;
  }
  if (!coords && !s->asm_xyz) {
    set_error("ehip_assemble: no coordinates (none given now or before)");
    return EHIP_EARG;
  }
  if ((physics == EHIP_PHYS_ELASTIC) != (s->dofs == 3) || (physics != EHIP_PHYS_ELASTIC && s->dofs != 1)) {
    set_error("ehip_assemble: physics %d does not match %d dofs per node", physics, s->dofs);
    return EHIP_EARG;
  }
  EType t;
  if (!make_etype(elem_code, t) || t.npe != s->npe) {
    set_error("ehip_assemble: element type %d does not match the mesh (%d nodes)", elem_code, s->npe);
    return EHIP_EUNSUP;
  }
  int rc;
  if ((rc = assembly_check(s))) return rc;  // an earlier asynchronous call's verdict
  const int ne = s->nelem;
  DevTables tab;
  if ((rc = ensure_tables(s, elem_code, t, tab))) return rc;
  // inputs: device buffers sized once per system, refilled on the stream
  if (!s->asm_xyz && (rc = dmalloc(s, (void **)&s->asm_xyz, sizeof(double) * (size_t)s->nnodes * 3))) return rc;
  if (!s->asm_in && (rc = dmalloc(s, (void **)&s->asm_in, sizeof(double) * (size_t)ne * 6))) return rc;
  if (!s->asm_bad && (rc = dmalloc(s, (void **)&s->asm_bad, sizeof(int)))) return rc;
  if (coords)
    EHIP_HIP_TRY(hipMemcpyAsync(s->asm_xyz, coords, sizeof(double) * (size_t)s->nnodes * 3, hipMemcpyHostToDevice,
                                s->stream));
  auto up = [&](const double *h, double *d, size_t cnt) -> const double * {
    if (!h) return nullptr;
    (void)hipMemcpyAsync(d, h, sizeof(double) * cnt, hipMemcpyHostToDevice, s->stream);
    return d;
  };
  ElemArgs a{};
  a.nelem = ne;
  a.el = s->elements;
  a.xyz = s->asm_xyz;
  a.basis = tab.basis;
  a.dlb = tab.dlb;
  a.gs = tab.gs;
  a.coef = up(elem_coef, s->asm_in, ne);
  a.conv = up(elem_conv, s->asm_in + ne, ne);
  a.load = up(elem_load, s->asm_in + 2 * (size_t)ne, ne);
  a.velo = up(elem_velo, s->asm_in + 3 * (size_t)ne, (size_t)ne * 3);
  a.poisson = poisson;
  a.emat = s->emat;
  a.evec = s->evec;
  a.bad = s->asm_bad;
  EHIP_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s->stream, s->asm_bad);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (t_ms) {
    EHIP_HIP_TRY(hipEventCreate(&e0));
    EHIP_HIP_TRY(hipEventCreate(&e1));
    EHIP_HIP_TRY(hipEventRecord(e0, s->stream));
  }
  bool packed = false;
  rc = dispatch_element(s, physics, elem_code, cdim, a, t, &packed);
  if (!rc) rc = launch_row_gather(s, s->npe * s->dofs, packed);
  if (!rc) touch_vals(s, true, false);
  if (t_ms) {
    EHIP_HIP_TRY(hipEventRecord(e1, s->stream));
    EHIP_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *t_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  EHIP_HIP_TRY(hipGetLastError());
  if (rc) return rc;
  s->has_values = true;
  s->asm_pending = 1;
  // with t_ms the call waits, and reports degenerate elements now; without,
  // the next call that waits on the system reports them
  return t_ms ? assembly_check(s) : EHIP_OK;
}

// StressBoundary traction part (Stress.src:836-1000), evaluated on the host
// in boundary-element order and added to the device RHS.
extern "C" int ehip_assemble_traction(ehip_system *s, int bnd_code, const double *coords, int nb, int npe,
                                      const int *belems, const double *load3) {
  if (!s || !coords || !belems || !load3 || s->dofs != 3) return EHIP_EARG;
  EType t;
  if (!make_etype(bnd_code, t) || t.npe != npe || t.dim != 2) return EHIP_EUNSUP;
  std::vector<double> rhs(s->n);
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
  EHIP_HIP_TRY(hipMemcpy(rhs.data(), s->rhs, sizeof(double) * s->n, hipMemcpyDeviceToHost));
  std::vector<double> force(3 * npe);
  for (int e = 0; e < nb; ++e) {
    const int *nd = belems + (size_t)e * npe;
    std::fill(force.begin(), force.end(), 0.0);
    for (int g = 0; g < t.ngp; ++g) {
      double dx[3][2] = {{0, 0}, {0, 0}, {0, 0}};
      for (int i = 0; i < 2; ++i)
        for (int c = 0; c < 3; ++c) {
          double sm = 0;
          for (int a = 0; a < npe; ++a) sm += coords[3 * (size_t)(nd[a] - 1) + c] * t.dlb[((size_t)g * npe + a) * 3 + i];
          dx[c][i] = sm;
        }
      double G[2][2];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
          double sm = 0;
          for (int k = 0; k < 3; ++k) sm = sm + dx[k][i] * dx[k][j];
          G[i][j] = sm;
        }
      const double detg = G[0][0] * G[1][1] - G[0][1] * G[1][0];
      if (!(detg > DBL_MIN)) return EHIP_EARG;
      const double sw = std::sqrt(detg) * t.gs[g];
      double L[3];
      for (int i = 0; i < 3; ++i) {
        double sm = 0;
        for (int k = 0; k < npe; ++k) sm += load3[i] * t.basis[(size_t)g * npe + k];
        L[i] = sm;
      }
      for (int q = 0; q < npe; ++q)
        for (int i = 0; i < 3; ++i) force[3 * q + i] = force[3 * q + i] + sw * t.basis[(size_t)g * npe + q] * L[i];
    }
    for (int a = 0; a < npe; ++a) {
      const int id = s->h_perm[nd[a] - 1];
      for (int j = 0; j < 3; ++j) {
        const int k = 3 * (id - 1) + j;
        rhs[k] = rhs[k] + force[3 * a + j];
      }
    }
  }
  EHIP_HIP_TRY(hipMemcpy(s->rhs, rhs.data(), sizeof(double) * s->n, hipMemcpyHostToDevice));
  return EHIP_OK;
}
