// Host CRS topology for libehip: node adjacency, bandwidth ordering, CRS
// rows/cols/diag and the element-to-slot gather maps.  Integer work on the
// host, parallel over nodes with OpenMP; runs once per mesh.
#include "pattern.h"

#include <algorithm>
#include <cstdlib>
#include <numeric>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace ehip {

namespace {

struct Adjacency {
  std::vector<int64_t> start;  // nnodes+2, entries for node i (1-based) at [start[i], start[i+1])
  std::vector<int> idx;        // sorted unique neighbour nodes, self included (1-based)
};

// node -> elements incidence (elements in ascending order)
struct Incidence {
  std::vector<int64_t> start;  // nnodes+2
  std::vector<int> elem;       // element ids (0-based), ascending per node
  std::vector<int> local;      // local node index inside that element
};

Incidence incidence(int nnodes, int nelem, int npe, const int *el) {
  Incidence I;
  I.start.assign((size_t)nnodes + 2, 0);
  for (int64_t k = 0; k < (int64_t)nelem * npe; ++k) I.start[el[k] + 1]++;
  for (int i = 1; i <= nnodes + 1; ++i) I.start[i] += I.start[i - 1];
  I.elem.resize((size_t)nelem * npe);
  I.local.resize((size_t)nelem * npe);
  std::vector<int64_t> fill(I.start.begin(), I.start.end());
  for (int e = 0; e < nelem; ++e)
    for (int a = 0; a < npe; ++a) {
      const int nd = el[(int64_t)e * npe + a];
      const int64_t pos = fill[nd]++;
      I.elem[pos] = e;
      I.local[pos] = a;
    }
  // start is indexed by 1-based node id: node i occupies [start[i], start[i+1])
  return I;
}

// MakeListMatrix (ElementUtils.src:279-620): every pair of nodes sharing an
// element, each list sorted and without duplicates (List_GetMatrixIndex).
Adjacency adjacency(int nnodes, int nelem, int npe, const int *el, int nbnd, int npe_b, const int *bel) {
  Incidence I = incidence(nnodes, nelem, npe, el);
  Incidence B;
  if (nbnd > 0) B = incidence(nnodes, nbnd, npe_b, bel);
  Adjacency A;
  std::vector<std::vector<int>> lists((size_t)nnodes + 1);
  std::vector<int64_t> deg((size_t)nnodes + 2, 0);
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int i = 1; i <= nnodes; ++i) {
      buf.clear();
      for (int64_t k = I.start[i]; k < I.start[i + 1]; ++k) {
        const int *nd = el + (int64_t)I.elem[k] * npe;
        buf.insert(buf.end(), nd, nd + npe);
      }
      if (nbnd > 0)
        for (int64_t k = B.start[i]; k < B.start[i + 1]; ++k) {
          const int *nd = bel + (int64_t)B.elem[k] * npe_b;
          buf.insert(buf.end(), nd, nd + npe_b);
        }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      lists[i] = buf;
      deg[i + 1] = (int64_t)buf.size();
    }
  }
  A.start.assign((size_t)nnodes + 2, 0);
  for (int i = 1; i <= nnodes; ++i) A.start[i + 1] = A.start[i] + deg[i + 1];
  A.idx.resize((size_t)A.start[nnodes + 1]);
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= nnodes; ++i) std::copy(lists[i].begin(), lists[i].end(), A.idx.begin() + A.start[i]);
  return A;
}

// ComputeBandwidth (BandwidthOptimize.src:60-90)
int bandwidth(int n, const Adjacency &A, const std::vector<int> *reorder) {
  int hb = 0;
#pragma omp parallel for reduction(max : hb) schedule(static)
  for (int i = 1; i <= n; ++i)
    for (int64_t c = A.start[i]; c < A.start[i + 1]; ++c) {
      const int k = A.idx[c];
      const int d = reorder ? std::abs((*reorder)[i] - (*reorder)[k]) : std::abs(i - k);
      hb = std::max(hb, d);
    }
  return hb;
}

// Levelize (BandwidthOptimize.src:290-350): the reference's depth-first walk
// with an explicit stack of list positions; `level` records the depth at
// which each node is first reached and maxlevel the deepest one.
void levelize(int root, const Adjacency &A, std::vector<int> &level, std::vector<char> &done, int &maxlevel) {
  struct Pos { int64_t p, end; };
  std::vector<Pos> stack;
  stack.reserve(512);
  int node = root, lev = 0;
  int64_t end = A.start[node + 1];
  int64_t p = A.start[node] < end ? A.start[node] : -1;
  while (p >= 0) {
    stack.push_back({p, end});
    level[node] = lev;
    done[node] = 1;
    maxlevel = std::max(maxlevel, lev);
    end = A.start[node + 1];
    p = A.start[node] < end ? A.start[node] : -1;
    for (;;) {
      if (p >= 0) {
        node = A.idx[p];
        if (!done[node]) { ++lev; break; }
      } else if (!stack.empty()) {
        p = stack.back().p;
        end = stack.back().end;
        stack.pop_back();
        --lev;
      } else {
        break;
      }
      p = (p + 1 < end) ? p + 1 : -1;
    }
  }
}

// OptimizeBandwidth (BandwidthOptimize.src:108-285): minimum-degree start,
// pseudo-peripheral root search over Levelize depths, Cuthill-McKee
// numbering over the sorted lists, reversal, and rejection when the half
// bandwidth grows.  Also reproduces the reference's assignment of the old
// depth to StartNode when a better root is found (:170-176).
void optimize_bandwidth(int n, const Adjacency &A, std::vector<int> &perm, int &before, int &after, int &accepted) {
  std::vector<int> degree(n + 1), level(n + 1, 0);
  std::vector<char> done(n + 1, 0);
  for (int i = 1; i <= n; ++i) degree[i] = (int)(A.start[i + 1] - A.start[i]);
  before = bandwidth(n, A, nullptr) + 1;
  int start = 1, mindeg = degree[1];
  for (int i = 1; i <= n; ++i)
    if (degree[i] < mindeg) { start = i; mindeg = degree[i]; }
  int maxlevel = 0;
  levelize(start, A, level, done, maxlevel);
  bool newroot = true;
  while (newroot) {
    newroot = false;
    mindeg = degree[start];
    int k = start;
    for (int i = 1; i <= n; ++i)
      if (level[i] == maxlevel && degree[i] < mindeg) { k = i; mindeg = degree[i]; }
    if (k != start) {
      const int j = maxlevel;
      maxlevel = 0;
      std::fill(done.begin(), done.end(), 0);
      levelize(k, A, level, done, maxlevel);
      if (j > maxlevel) { newroot = true; start = j; }
    }
  }
  std::vector<int> order(n + 2, 0), doneidx(n + 1, 0);
  int indx = 1;
  order[indx] = start;
  doneidx[start] = indx++;
  int scan = 1;  // first node that may still be unnumbered
  for (int i = 1; i <= n; ++i) {
    if (order[i] == 0) {
      while (scan <= n && doneidx[scan] != 0) ++scan;
      if (scan <= n) { order[indx] = scan; doneidx[scan] = indx++; }
    }
    const int node = order[i];
    for (int64_t c = A.start[node]; c < A.start[node + 1]; ++c) {
      const int k = A.idx[c];
      if (doneidx[k] == 0) { order[indx] = k; doneidx[k] = indx++; }
    }
  }
  for (int i = 1; i <= n; ++i) doneidx[order[i]] = n - i + 1;
  std::vector<int> saved(perm);
  for (int i = 1; i <= n; ++i) perm[i] = saved[i] > 0 ? doneidx[saved[i]] : 0;
  after = bandwidth(n, A, &perm) + 1;
  accepted = 1;
  if (before < after) {
    perm = saved;
    accepted = 0;
  }
}

}  // namespace

void bandwidth_permutation(int n, const int64_t *start, const int *idx, std::vector<int> &perm, int &before,
                           int &after, int &accepted) {
  Adjacency A;
  A.start.assign(start, start + n + 2);
  A.idx.assign(idx, idx + start[n + 1]);
  perm.resize((size_t)n + 1);
  std::iota(perm.begin(), perm.end(), 0);
  before = after = bandwidth(n, A, nullptr) + 1;
  accepted = 0;
  optimize_bandwidth(n, A, perm, before, after, accepted);
}

int build_host_pattern(int nnodes, int nelem, int npe, const int *el, int nbnd, int npe_b, const int *bel, int dofs,
                       int optimize, HostPattern &P) {
  if (nnodes <= 0 || nelem <= 0 || npe <= 0 || dofs <= 0) return -1;
  for (int64_t k = 0; k < (int64_t)nelem * npe; ++k)
    if (el[k] < 1 || el[k] > nnodes) return -1;
  for (int64_t k = 0; k < (int64_t)nbnd * npe_b; ++k)
    if (bel[k] < 1 || bel[k] > nnodes) return -1;
  Adjacency A = adjacency(nnodes, nelem, npe, el, nbnd, npe_b, bel);
  for (int i = 1; i <= nnodes; ++i)
    if (A.start[i + 1] == A.start[i]) return -2;  // node outside every element: not supported
  P.nnodes = nnodes;
  P.dofs = dofs;
  P.n = nnodes * dofs;
  std::vector<int> perm(nnodes + 1);
  std::iota(perm.begin(), perm.end(), 0);
  P.bw_before = P.bw_after = bandwidth(nnodes, A, nullptr) + 1;
  P.accepted = 0;
  if (optimize) optimize_bandwidth(nnodes, A, perm, P.bw_before, P.bw_after, P.accepted);
  P.perm.assign(perm.begin() + 1, perm.end());
  std::vector<int> inv(nnodes + 1);
  for (int i = 1; i <= nnodes; ++i) inv[perm[i]] = i;
  const int n = P.n;
  P.rows.assign((size_t)n + 1, 0);
  int64_t pos = 0;
  for (int j = 1; j <= nnodes; ++j) {
    const int node = inv[j];
    const int64_t deg = (A.start[node + 1] - A.start[node]) * dofs;
    for (int l = 0; l < dofs; ++l) {
      P.rows[(size_t)dofs * (j - 1) + l] = (int)(pos + 1);
      pos += deg;
    }
  }
  P.rows[n] = (int)(pos + 1);
  P.nnz = pos;
  if (P.nnz >= INT32_MAX) return -3;
  P.cols.resize((size_t)P.nnz);
  P.diag.resize((size_t)n);
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int j = 1; j <= nnodes; ++j) {
      const int node = inv[j];
      buf.clear();
      for (int64_t c = A.start[node]; c < A.start[node + 1]; ++c) {
        const int k = perm[A.idx[c]];
        for (int d = 1; d <= dofs; ++d) buf.push_back(dofs * (k - 1) + d);
      }
      std::sort(buf.begin(), buf.end());
      for (int l = 0; l < dofs; ++l) {
        const int r = dofs * (j - 1) + l;  // 0-based row
        const int64_t base = P.rows[r] - 1;
        std::copy(buf.begin(), buf.end(), P.cols.begin() + base);
        const auto it = std::lower_bound(buf.begin(), buf.end(), r + 1);
        P.diag[r] = (int)(base + (it - buf.begin()) + 1);
      }
    }
  }
  return 0;
}

int build_row_maps(const HostPattern &P, int nelem, int npe, const int *el, std::vector<int> &row_start,
                   std::vector<int> &row_contrib, std::vector<uint16_t> &row_off, int *max_len) {
  const int dofs = P.dofs, n = P.n, nnodes = P.nnodes;
  const int ld = npe * dofs;
  if ((int64_t)nelem * ld * ld >= INT32_MAX) return -3;
  Incidence I = incidence(nnodes, nelem, npe, el);
  std::vector<int> inv(nnodes + 1);
  for (int i = 0; i < nnodes; ++i) inv[P.perm[i]] = i + 1;
  // one entry per (row, touching element): counts, then offsets
  row_start.assign((size_t)n + 1, 0);
  int lmax = 0;
  for (int j = 1; j <= nnodes; ++j) {
    const int a = inv[j];
    for (int l = 0; l < dofs; ++l) {
      const int r = dofs * (j - 1) + l;
      row_start[r + 1] = (int)(I.start[a + 1] - I.start[a]);
      lmax = std::max(lmax, P.rows[r + 1] - P.rows[r]);
    }
  }
  for (int r = 0; r < n; ++r) row_start[r + 1] += row_start[r];
  *max_len = lmax;
  if (lmax > 65536) return -3;
  row_contrib.resize((size_t)row_start[n]);
  row_off.resize((size_t)row_start[n] * ld);
  // fill: the incidence lists are ascending, so every row's elements -- and
  // with them every slot's contributions -- come in element order
#pragma omp parallel for schedule(dynamic, 1024)
  for (int j = 1; j <= nnodes; ++j) {
    const int a = inv[j];
    for (int l = 0; l < dofs; ++l) {
      const int r = dofs * (j - 1) + l;
      const int64_t b0 = P.rows[r] - 1, b1 = P.rows[r + 1] - 1;
      const int *cb = P.cols.data() + b0;
      const int len = (int)(b1 - b0);
      int64_t k = row_start[r];
      for (int64_t q = I.start[a]; q < I.start[a + 1]; ++q, ++k) {
        const int e = I.elem[q];
        const int li = I.local[q] * dofs + l;
        const int *nd = el + (int64_t)e * npe;
        row_contrib[k] = e * ld + li;
        for (int bl = 0; bl < npe; ++bl) {
          const int kb = P.perm[nd[bl] - 1];
          for (int m = 1; m <= dofs; ++m) {
            const int col = dofs * (kb - 1) + m;
            const int pos = (int)(std::lower_bound(cb, cb + len, col) - cb);
            row_off[(size_t)k * ld + bl * dofs + (m - 1)] = (uint16_t)pos;
          }
        }
      }
    }
  }
  return 0;
}

// Per slice, the row offsets are merged into the running union one row at a
// time (rows of one slice mostly repeat the union, so a merge is a short
// walk); a second walk per row sets its mask bits.  With D > 1 a row's
// entries are taken D at a time (whole D-column node blocks) and the
// offsets are node offsets: column / D - row / D.
bool slice_offsets(int n, const int *rows1, const int *c0, int max_len, std::vector<int> &len,
                   std::vector<int> &tab, std::vector<unsigned> &mask, int D) {
  if (max_len > 32) max_len = 32;
  const int S = (n + 63) / 64;
  len.assign((size_t)S, 0);
  tab.assign((size_t)S * max_len, 0);
  mask.assign((size_t)n, 0u);
  int fail = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : fail)
  for (int sl = 0; sl < S; ++sl) {
    if (fail) continue;
    int u[33], t[33];
    int nu = 0;
    const int r0 = sl * 64, r1 = std::min(n, r0 + 64);
    bool ok = true;
    for (int r = r0; r < r1 && ok; ++r) {
      const int j1 = rows1[r + 1] - 1, node = r / D;
      int a = 0, b = rows1[r] - 1, nt = 0;
      while (a < nu || b < j1) {
        const int va = a < nu ? u[a] : INT32_MAX;
        const int vb = b < j1 ? c0[b] / D - node : INT32_MAX;
        const int v = std::min(va, vb);
        a += va == v;
        b += vb == v ? D : 0;
        if (nt == max_len) {
          ok = false;
          break;
        }
        t[nt++] = v;
      }
      std::copy(t, t + nt, u);
      nu = nt;
    }
    if (!ok) {
      fail = 1;
      continue;
    }
    len[sl] = nu;
    std::copy(u, u + nu, tab.begin() + (size_t)sl * max_len);
    for (int r = r0; r < r1; ++r) {
      unsigned m = 0;
      int k = 0;
      const int node = r / D;
      for (int b = rows1[r] - 1; b < rows1[r + 1] - 1; b += D) {
        const int v = c0[b] / D - node;
        while (u[k] != v) ++k;  // v is in the union, which is sorted
        m |= 1u << k;
      }
      mask[r] = m;
    }
  }
  return !fail;
}

}  // namespace ehip
