// ILU(k) preconditioner of libehip: CRS_IncompleteLU's real non-Cholesky
// branch (fem/src/CRSMatrix.src:2157-2500) and CRS_LUSolve
// (:3293-3397), as CRS_LUPrecondition applies it (:3259-3268).
//
// The factorization and both triangular solves are row recurrences: row i
// needs the finished rows its lower (forward) or upper (backward) part
// names.  The host sorts the rows into dependency levels once per pattern;
// each level is one launch with one lane per row, and every lane walks its
// row in column order exactly as the reference does, so the factors and the
// solves are the reference's bits (FP64, no contraction).  A level's rows
// never read one another, so the solves run in place.
//
// The ILU(k) pattern for k >= 1 is k rounds of InitializeILU1's fill
// (:2387-2498): each round adds, to row i, the upper entries (c, j) of the
// rows c that row i's lower part names, taking only entries of the round's
// input pattern.  Fills start at zero in the factors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "ehip_internal.h"
#include "launch.h"

namespace ehip {

constexpr double kAeps = 10.0 * DBL_EPSILON;  // AEPS, fem/src/Types.src:61
constexpr int kIluBlock = 256;

// ---------------------------------------------------------------- kernels

__global__ void k_ilu_scatter(int64_t nnz, const double *__restrict__ vals, const int *__restrict__ amap,
                              double *__restrict__ ilu) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * blockDim.x)
    ilu[amap[j]] = vals[j];
}

// One level of the factorization: row i is eliminated with the finished
// rows k of its lower part in column order (a zero entry is skipped, the
// division only when |U(k,k)| > AEPS), each subtracting S(k) U(k, j) from
// the entries j of row i's pattern that row k's upper part shares.
__global__ __launch_bounds__(kIluBlock) void k_ilu_rows(const int *__restrict__ rows, const int *__restrict__ cols,
                                                      const int *__restrict__ diag, double *ilu,
                                                      const int *__restrict__ order, int begin, int end) {
#pragma clang fp contract(off)
  const int t = begin + blockIdx.x * kIluBlock + threadIdx.x;
  if (t >= end) return;
  const int i = order[t];
  const int r1 = rows[i + 1], di = diag[i];
  for (int m = rows[i]; m < di; ++m) {
    const int k = cols[m];
    double sk = ilu[m];
    if (sk == 0.0) continue;
    const double ukk = ilu[diag[k]];
    if (fabs(ukk) > kAeps) {
      sk = sk / ukk;
      ilu[m] = sk;
    }
    int p = m + 1;  // row i's entries beyond k, merged with row k's upper part
    const int l1 = rows[k + 1];
    for (int l = diag[k] + 1; l < l1; ++l) {
      const int j = cols[l];
      while (p < r1 && cols[p] < j) ++p;
      if (p == r1) break;
      if (cols[p] == j) ilu[p] = ilu[p] - sk * ilu[l];
    }
  }
}

// The LU solve's prescaled diagonal: 1/U(i,i), or 1 when |U(i,i)| < AEPS.
__global__ void k_ilu_invdiag(int n, const int *__restrict__ diag, double *ilu) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double d = ilu[diag[i]];
    ilu[diag[i]] = fabs(d) < kAeps ? 1.0 : 1.0 / d;
  }
}

// s - sum over k of a(k) x(cols(k)) for the row part [j0, j1), subtracted
// in order k = j0, j0+1, ... (the reference's loop): the part's values and
// column indices are loaded kLuU at a time before any x gather, and the x
// gathers before any subtraction, so a level costs a few memory latencies
// instead of two per entry (indices past the part are clamped into it and
// their terms skipped).
constexpr int kLuU = 16;
__device__ __forceinline__ double lu_part(double s, int j0, int j1, const int *__restrict__ cols,
                                          const double *__restrict__ ilu, const double *x) {
#pragma clang fp contract(off)
  const int len = j1 - j0;
  for (int k0 = 0; k0 < len; k0 += kLuU) {
    double a[kLuU];
    int c[kLuU];
#pragma unroll
    for (int u = 0; u < kLuU; ++u) {
      const int j = j0 + min(k0 + u, len - 1);
      a[u] = ilu[j];
      c[u] = cols[j];
    }
    double xv[kLuU];
#pragma unroll
    for (int u = 0; u < kLuU; ++u) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < kLuU; ++u) {
      const double t = s - a[u] * xv[u];
      s = k0 + u < len ? t : s;
    }
  }
  return s;
}

// Forward level: dst(i) = src(i) - sum over the lower part of L(i,j) dst(j)
__global__ __launch_bounds__(kIluBlock) void k_lu_fwd(const int *__restrict__ rows, const int *__restrict__ cols,
                                                    const int *__restrict__ diag, const double *__restrict__ ilu,
                                                    const double *src, double *dst, const int *__restrict__ order,
                                                    int begin, int end) {
  const int t = begin + blockIdx.x * kIluBlock + threadIdx.x;
  if (t >= end) return;
  const int i = order[t];
  dst[i] = lu_part(src[i], rows[i], diag[i], cols, ilu, dst);
}

// Backward level: dst(i) = U(i,i)^-1 (dst(i) - sum over the upper part of U(i,j) dst(j))
__global__ __launch_bounds__(kIluBlock) void k_lu_bwd(const int *__restrict__ rows, const int *__restrict__ cols,
                                                    const int *__restrict__ diag, const double *__restrict__ ilu,
                                                    double *dst, const int *__restrict__ order, int begin, int end) {
#pragma clang fp contract(off)
  const int t = begin + blockIdx.x * kIluBlock + threadIdx.x;
  if (t >= end) return;
  const int i = order[t];
  const int di = diag[i];
  dst[i] = ilu[di] * lu_part(dst[i], di + 1, rows[i + 1], cols, ilu, dst);
}

// ---------------------------------------------------------------- dependency-driven sweeps
// One launch per triangular solve instead of one per level.  Every output
// entry starts as kPending, a signalling-NaN bit pattern no arithmetic
// produces (a NaN result is the quiet canonical one); a row's lane reloads
// the entries its part names (agent-scope atomic loads: the writer may sit
// on another XCD) until none is pending, subtracts them in the reference's
// order exactly as lu_part does -- the same bits as the level launches --
// and publishes its value with one agent-scope atomic store, the value
// being its own ready flag.  Rows are taken in level order by workgroups
// numbered in the order they start (an atomic ticket), so a row waits only
// for rows of workgroups already running or done, and a lane finishes its
// row inside the wait loop, so lanes of one wave never wait for each
// other's exit.  Every wait is bounded: past kSpinLimit rounds, or once any
// lane has given up, a lane sets the error flag (sync[2]) and publishes a
// NaN, and the solve reports the failure instead of hanging.
constexpr unsigned long long kPending = 0x7ff4c0ffee5eed00ull;
constexpr int kSpinLimit = 1 << 20;
constexpr int kSweepGroups = 256;

__device__ __forceinline__ unsigned long long ld_agent(const double *p) {
  return __hip_atomic_load(reinterpret_cast<unsigned long long *>(const_cast<double *>(p)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_lu_pending(int n, double *y, double *z, int *sync) {
  const double p = __longlong_as_double((long long)kPending);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    y[i] = p;
    z[i] = p;
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) sync[threadIdx.x] = 0;  // the two sweeps' tickets
}

// One row of a sweep.  FWD: dst(i) = src(i) - sum over the lower part of
// L(i,j) dst(j); else dst(i) = U(i,i)^-1 (src(i) - sum over the upper part
// of U(i,j) dst(j)).  The lane first polls only the part's nearest entry
// (column i - 1 or i + 1 as a rule, the dependency finished last), then
// loads kLuU entries at a time; all of it in one loop, publishing the row
// inside it.  The loop ends for the whole wave at once (when every lane is
// done): with a per-lane exit the compiler may move the publishing store
// after the loop, where a lane publishes only once every lane of its wave is
// done -- and a lane waiting for another lane of its wave would never be.
template <bool FWD>
__device__ __forceinline__ void sweep_row(int i, const int *__restrict__ rows, const int *__restrict__ cols,
                                          const int *__restrict__ diag, const double *__restrict__ ilu,
                                          const double *__restrict__ src, double *dst, int *sync) {
#pragma clang fp contract(off)
  const int j0 = FWD ? rows[i] : diag[i] + 1;
  const int len = (FWD ? diag[i] : rows[i + 1]) - j0;
  double s = src[i];
  int k0 = 0, spins = 0;
  bool near = len == 0, done = false;
  // the part's factors and columns do not change during the sweep: the
  // current chunk's are loaded before the wait, so a dependency that turns
  // ready costs one more trip (its values), not two
  double a[kLuU];
  int c[kLuU];
  auto load_chunk = [&](int k) {
#pragma unroll
    for (int u = 0; u < kLuU; ++u) {
      const int j = j0 + min(k + u, len - 1);  // clamped entries are dependencies too
      a[u] = ilu[j];
      c[u] = cols[j];
    }
  };
  int cn = 0;
  if (len > 0) {
    load_chunk(0);
    cn = cols[FWD ? j0 + len - 1 : j0];
  }
  for (;;) {
    if (__all(done)) break;  // wave-uniform exit
    if (done) continue;
    bool ready = true;
    if (!near) ready = near = ld_agent(dst + cn) != kPending;
    if (near && k0 < len) {
      unsigned long long xb[kLuU];
#pragma unroll
      for (int u = 0; u < kLuU; ++u) xb[u] = ld_agent(dst + c[u]);
#pragma unroll
      for (int u = 0; u < kLuU; ++u) ready = ready && xb[u] != kPending;
      if (ready) {
#pragma unroll
        for (int u = 0; u < kLuU; ++u) {
          const double v = s - a[u] * __longlong_as_double((long long)xb[u]);
          s = k0 + u < len ? v : s;
        }
        k0 += kLuU;
        if (k0 < len) load_chunk(k0);
      }
    }
    bool fail = false;
    if (!ready) {
      if (++spins > kSpinLimit || __hip_atomic_load(&sync[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        fail = true;
      else
        __builtin_amdgcn_s_sleep(1);
    }
    if ((near && k0 >= len) || fail) {
      if (fail) {
        __hip_atomic_store(&sync[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s = __longlong_as_double(0x7ff8000000000000ll);
      } else if (!FWD) {
        s = ilu[diag[i]] * s;
      }
      __hip_atomic_store(reinterpret_cast<unsigned long long *>(dst + i), (unsigned long long)__double_as_longlong(s),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      done = true;
    }
  }
}

// A sweep over all rows in level order (`order`): each workgroup takes the
// next kIluBlock rows by ticket until none is left, so the grid bounds the
// rows in flight.
template <bool FWD>
__global__ __launch_bounds__(kIluBlock) void k_lu_sweep(const int *__restrict__ rows, const int *__restrict__ cols,
                                                      const int *__restrict__ diag, const double *__restrict__ ilu,
                                                      const double *__restrict__ src, double *dst,
                                                      const int *__restrict__ order, int n, int *sync) {
  __shared__ int ticket;
  for (;;) {
    __syncthreads();  // every lane has read the previous ticket
    if (threadIdx.x == 0) ticket = atomicAdd(&sync[FWD ? 0 : 1], 1);
    __syncthreads();
    const int t0 = ticket * kIluBlock;
    if (t0 >= n) return;
    const int t = t0 + (int)threadIdx.x;
    if (t < n) sweep_row<FWD>(order[t], rows, cols, diag, ilu, src, dst, sync);
  }
}

// ---------------------------------------------------------------- host: pattern and levels

namespace {

// One fill round on a 0-based sorted pattern (R, K, D) -> (R2, K2, D2).
void fill_round(int n, const std::vector<int> &R, const std::vector<int> &K, const std::vector<int> &D,
                std::vector<int> &R2, std::vector<int> &K2, std::vector<int> &D2) {
  std::vector<unsigned char> mark((size_t)n, 0);
  std::vector<int> row;
  R2.assign((size_t)n + 1, 0);
  D2.assign((size_t)n, 0);
  K2.clear();
  K2.reserve(K.size() + K.size() / 2);
  for (int i = 0; i < n; ++i) {
    row.clear();
    for (int k = R[i]; k < R[i + 1]; ++k) {
      mark[K[k]] = 1;
      row.push_back(K[k]);
    }
    for (int k = R[i]; k < D[i]; ++k) {  // lower entries c of row i: input entries only
      const int c = K[k];
      for (int l = D[c] + 1; l < R[c + 1]; ++l) {
        const int j = K[l];
        if (!mark[j]) {
          mark[j] = 2;
          row.push_back(j);
        }
      }
    }
    std::sort(row.begin(), row.end());
    for (int c : row) {
      mark[c] = 0;
      if (c == i) D2[i] = (int)K2.size();
      K2.push_back(c);
    }
    R2[i + 1] = (int)K2.size();
  }
}

// Rows grouped by dependency level (forward: over the lower part in row
// order; backward: over the upper part in reverse row order), ascending row
// index inside a level.
void levels(int n, const std::vector<int> &R, const std::vector<int> &K, const std::vector<int> &D, bool lower,
            std::vector<int> &order, std::vector<int> &off) {
  std::vector<int> lev((size_t)n, 0);
  int top = 0;
  if (lower) {
    for (int i = 0; i < n; ++i) {
      int l = 0;
      for (int j = R[i]; j < D[i]; ++j) l = std::max(l, lev[K[j]] + 1);
      lev[i] = l;
      top = std::max(top, l);
    }
  } else {
    for (int i = n - 1; i >= 0; --i) {
      int l = 0;
      for (int j = D[i] + 1; j < R[i + 1]; ++j) l = std::max(l, lev[K[j]] + 1);
      lev[i] = l;
      top = std::max(top, l);
    }
  }
  off.assign((size_t)top + 2, 0);
  for (int i = 0; i < n; ++i) off[lev[i] + 1]++;
  for (int l = 0; l <= top; ++l) off[l + 1] += off[l];
  order.assign((size_t)n, 0);
  std::vector<int> pos(off.begin(), off.end() - 1);
  for (int i = 0; i < n; ++i) order[pos[lev[i]]++] = i;
}

}  // namespace

void ilu_release(ehip_system *s) {
  if (s->ilu_order > 0) {
    dfree(s, s->ilu_rows, sizeof(int) * ((size_t)s->n + 1));
    dfree(s, s->ilu_cols, sizeof(int) * (size_t)s->ilu_nnz);
    dfree(s, s->ilu_diag, sizeof(int) * (size_t)s->n);
    dfree(s, s->ilu_amap, sizeof(int) * (size_t)s->nnz);
  }
  dfree(s, s->ilu, sizeof(double) * (size_t)s->ilu_nnz);
  dfree(s, s->ilu_lord, sizeof(int) * (size_t)s->n);
  dfree(s, s->ilu_uord, sizeof(int) * (size_t)s->n);
  dfree(s, s->ilu_y, sizeof(double) * (size_t)s->n);
  dfree(s, s->ilu_sync, sizeof(int) * 4);
  s->ilu_y = nullptr;
  s->ilu_sync = nullptr;
  for (int k = 0; k < 2; ++k)
    if (s->ilu_graph[k]) {
      (void)hipGraphExecDestroy(s->ilu_graph[k]);
      s->ilu_graph[k] = nullptr;
    }
  s->ilu_rows = s->ilu_cols = s->ilu_diag = s->ilu_amap = s->ilu_lord = s->ilu_uord = nullptr;
  s->ilu = nullptr;
  s->ilu_loff.clear();
  s->ilu_uoff.clear();
  s->ilu_order = -1;
  s->ilu_nnz = 0;
}

int ilu_setup(ehip_system *s, int order) {
  if (order == s->ilu_order) return EHIP_OK;
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));  // nothing queued may still use the old arrays
  ilu_release(s);
  const int n = s->n;
  std::vector<int> R((size_t)n + 1), K((size_t)s->nnz), D((size_t)n);
  for (int i = 0; i <= n; ++i) R[i] = s->h_rows[i] - 1;
  for (int64_t j = 0; j < s->nnz; ++j) K[j] = s->h_cols[j] - 1;
  for (int i = 0; i < n; ++i) D[i] = s->h_diag[i] - 1;
  for (int r = 0; r < order; ++r) {
    std::vector<int> R2, K2, D2;
    fill_round(n, R, K, D, R2, K2, D2);
    if ((int64_t)K2.size() >= (int64_t)INT32_MAX) {
      set_error("ILU(%d) pattern exceeds 2^31 entries", order);
      return EHIP_ECAP;
    }
    R.swap(R2);
    K.swap(K2);
    D.swap(D2);
  }
  std::vector<int> lord, loff, uord, uoff;
  levels(n, R, K, D, true, lord, loff);
  levels(n, R, K, D, false, uord, uoff);
  int rc;
  s->ilu_nnz = (int64_t)K.size();
  if ((rc = dmalloc(s, (void **)&s->ilu, sizeof(double) * (size_t)s->ilu_nnz)) ||
      (rc = dmalloc(s, (void **)&s->ilu_lord, sizeof(int) * (size_t)n)) ||
      (rc = dmalloc(s, (void **)&s->ilu_uord, sizeof(int) * (size_t)n)) ||
      (rc = dmalloc(s, (void **)&s->ilu_y, sizeof(double) * (size_t)n)) ||
      (rc = dmalloc(s, (void **)&s->ilu_sync, sizeof(int) * 4)))
    return rc;
  EHIP_HIP_TRY(hipMemcpy(s->ilu_lord, lord.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(s->ilu_uord, uord.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemset(s->ilu_sync, 0, sizeof(int) * 4));
  if (order > 0) {
    // the matrix's nonzeros inside the ILU rows (a subset of them, both sorted)
    std::vector<int> amap((size_t)s->nnz);
    for (int i = 0; i < n; ++i) {
      int p = R[i];
      for (int j = s->h_rows[i] - 1; j < s->h_rows[i + 1] - 1; ++j) {
        while (K[p] < s->h_cols[j] - 1) ++p;
        amap[j] = p;
      }
    }
    if ((rc = dmalloc(s, (void **)&s->ilu_rows, sizeof(int) * ((size_t)n + 1))) ||
        (rc = dmalloc(s, (void **)&s->ilu_cols, sizeof(int) * K.size())) ||
        (rc = dmalloc(s, (void **)&s->ilu_diag, sizeof(int) * (size_t)n)) ||
        (rc = dmalloc(s, (void **)&s->ilu_amap, sizeof(int) * (size_t)s->nnz)))
      return rc;
    EHIP_HIP_TRY(hipMemcpy(s->ilu_rows, R.data(), sizeof(int) * R.size(), hipMemcpyHostToDevice));
    EHIP_HIP_TRY(hipMemcpy(s->ilu_cols, K.data(), sizeof(int) * K.size(), hipMemcpyHostToDevice));
    EHIP_HIP_TRY(hipMemcpy(s->ilu_diag, D.data(), sizeof(int) * D.size(), hipMemcpyHostToDevice));
    EHIP_HIP_TRY(hipMemcpy(s->ilu_amap, amap.data(), sizeof(int) * amap.size(), hipMemcpyHostToDevice));
  } else {
    s->ilu_rows = s->rows;
    s->ilu_cols = s->cols;
    s->ilu_diag = s->diag;
  }
  s->ilu_loff = std::move(loff);
  s->ilu_uoff = std::move(uoff);
  s->ilu_order = order;
  return EHIP_OK;
}

static inline int level_grid(int cnt) { return (cnt + kIluBlock - 1) / kIluBlock; }

int ilu_factor(ehip_system *s) {
  const int n = s->n;
  if (s->ilu_order > 0) {
    EHIP_HIP_TRY(hipMemsetAsync(s->ilu, 0, sizeof(double) * (size_t)s->ilu_nnz, s->stream));
    hipLaunchKernelGGL(k_ilu_scatter, dim3(2048), dim3(256), 0, s->stream, s->nnz, s->vals, s->ilu_amap, s->ilu);
  } else {
    EHIP_HIP_TRY(hipMemcpyAsync(s->ilu, s->vals, sizeof(double) * (size_t)s->nnz, hipMemcpyDeviceToDevice,
                                s->stream));
  }
  const std::vector<int> &off = s->ilu_loff;
  for (size_t l = 0; l + 1 < off.size(); ++l) {
    const int cnt = off[l + 1] - off[l];
    if (cnt > 0)
      hipLaunchKernelGGL(k_ilu_rows, dim3(level_grid(cnt)), dim3(kIluBlock), 0, s->stream, s->ilu_rows, s->ilu_cols,
                         s->ilu_diag, s->ilu, s->ilu_lord, off[l], off[l + 1]);
  }
  hipLaunchKernelGGL(k_ilu_invdiag, dim3(std::min(2048, (n + 255) / 256)), dim3(256), 0, s->stream, n, s->ilu_diag,
                     s->ilu);
  EHIP_HIP_TRY(hipGetLastError());
  return EHIP_OK;
}

// The LU solve as two dependency-driven sweeps (the default), or as one
// launch per level (EHIP_ILU_SWEEP=levels); both give the same bits.
static bool level_launches() {
  const char *v = std::getenv("EHIP_ILU_SWEEP");
  return v && v[0] == 'l';
}

// Workgroups of a sweep (EHIP_ILU_SWEEP_WG overrides).
static int sweep_groups(int g) {
  const char *v = std::getenv("EHIP_ILU_SWEEP_WG");
  const int w = v && *v ? std::atoi(v) : kSweepGroups;
  return std::max(1, std::min(g, w));
}

static void ilu_sweeps(ehip_system *s, const double *src, double *dst) {
  if (!level_launches()) {
    const int n = s->n, g = (n + kIluBlock - 1) / kIluBlock, w = sweep_groups(g);
    launch_k(k_lu_pending, dim3(std::min(2048, g)), dim3(kIluBlock), s->stream, n, s->ilu_y, dst, s->ilu_sync);
    launch_k(k_lu_sweep<true>, dim3(w), dim3(kIluBlock), s->stream, s->ilu_rows, s->ilu_cols, s->ilu_diag, s->ilu,
             src, s->ilu_y, s->ilu_lord, n, s->ilu_sync);
    launch_k(k_lu_sweep<false>, dim3(w), dim3(kIluBlock), s->stream, s->ilu_rows, s->ilu_cols, s->ilu_diag, s->ilu,
             (const double *)s->ilu_y, dst, s->ilu_uord, n, s->ilu_sync);
    return;
  }
  const std::vector<int> &lo = s->ilu_loff, &uo = s->ilu_uoff;
  for (size_t l = 0; l + 1 < lo.size(); ++l) {
    const int cnt = lo[l + 1] - lo[l];
    if (cnt > 0)
      launch_k(k_lu_fwd, dim3(level_grid(cnt)), dim3(kIluBlock), s->stream, s->ilu_rows, s->ilu_cols, s->ilu_diag,
               s->ilu, src, dst, s->ilu_lord, lo[l], lo[l + 1]);
  }
  for (size_t l = 0; l + 1 < uo.size(); ++l) {
    const int cnt = uo[l + 1] - uo[l];
    if (cnt > 0)
      launch_k(k_lu_bwd, dim3(level_grid(cnt)), dim3(kIluBlock), s->stream, s->ilu_rows, s->ilu_cols, s->ilu_diag,
               s->ilu, dst, s->ilu_uord, uo[l], uo[l + 1]);
  }
}

int ilu_apply(ehip_system *s, const double *src, double *dst) {
  if (!graphs_enabled()) {
    ilu_sweeps(s, src, dst);
    EHIP_HIP_TRY(take_launch_error());
    return EHIP_OK;
  }
  // the sweeps of one (src, dst) pair recorded once and replayed: one
  // launch instead of one per level
  int slot = -1;
  for (int k = 0; k < 2; ++k)
    if (s->ilu_graph[k] && s->ilu_graph_src[k] == src && s->ilu_graph_dst[k] == dst) slot = k;
  if (slot < 0) {
    slot = s->ilu_graph[0] ? 1 : 0;
    if (s->ilu_graph[slot]) {
      EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
      (void)hipGraphExecDestroy(s->ilu_graph[slot]);
      s->ilu_graph[slot] = nullptr;
    }
    GraphRec rec;
    EHIP_HIP_TRY(hipGraphCreate(&rec.graph, 0));
    tl_graph_rec = &rec;
    ilu_sweeps(s, src, dst);
    tl_graph_rec = nullptr;
    hipError_t e = rec.err;
    if (e == hipSuccess) e = hipGraphInstantiate(&s->ilu_graph[slot], rec.graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(rec.graph);
    EHIP_HIP_TRY(e);
    s->ilu_graph_src[slot] = src;
    s->ilu_graph_dst[slot] = dst;
  }
  EHIP_HIP_TRY(hipGraphLaunch(s->ilu_graph[slot], s->stream));
  return EHIP_OK;
}

int ilu_check(ehip_system *s) {
  if (!s->ilu_sync) return EHIP_OK;
  int flag = 0;
  EHIP_HIP_TRY(hipMemcpyAsync(&flag, s->ilu_sync + 2, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
  if (flag) {
    EHIP_HIP_TRY(hipMemset(s->ilu_sync + 2, 0, sizeof(int)));
    set_error("ILU solve: a row waited past the spin limit for the rows it depends on");
    return EHIP_EHIP;
  }
  return EHIP_OK;
}

}  // namespace ehip

using namespace ehip;

extern "C" int ehip_ilu_factor(ehip_system *s, int order, int64_t *ilu_nnz, int *irows, int *icols, int *idiag,
                               double *ilu) {
  if (!s || order < 0 || order > 9) {
    set_error("ehip_ilu_factor: bad arguments (order 0..9)");
    return EHIP_EARG;
  }
  if (!s->has_values) {
    set_error("ehip_ilu_factor: the system has no values");
    return EHIP_ESTATE;
  }
  int rc;
  if ((rc = ilu_setup(s, order)) || (rc = ilu_factor(s))) return rc;
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
  if (ilu_nnz) *ilu_nnz = s->ilu_nnz;
  const int n = s->n;
  if (irows || icols || idiag) {
    std::vector<int> R((size_t)n + 1), K((size_t)s->ilu_nnz), D((size_t)n);
    EHIP_HIP_TRY(hipMemcpy(R.data(), s->ilu_rows, sizeof(int) * R.size(), hipMemcpyDeviceToHost));
    EHIP_HIP_TRY(hipMemcpy(K.data(), s->ilu_cols, sizeof(int) * K.size(), hipMemcpyDeviceToHost));
    EHIP_HIP_TRY(hipMemcpy(D.data(), s->ilu_diag, sizeof(int) * D.size(), hipMemcpyDeviceToHost));
    if (irows)
      for (int i = 0; i <= n; ++i) irows[i] = R[i] + 1;
    if (icols)
      for (int64_t j = 0; j < s->ilu_nnz; ++j) icols[j] = K[j] + 1;
    if (idiag)
      for (int i = 0; i < n; ++i) idiag[i] = D[i] + 1;
  }
  if (ilu) EHIP_HIP_TRY(hipMemcpy(ilu, s->ilu, sizeof(double) * (size_t)s->ilu_nnz, hipMemcpyDeviceToHost));
  return EHIP_OK;
}

extern "C" int ehip_lu_solve(ehip_system *s, const double *b, double *x) {
  if (!s || !b || !x) return EHIP_EARG;
  if (s->ilu_order < 0) {
    set_error("ehip_lu_solve: no factors (ehip_ilu_factor first)");
    return EHIP_ESTATE;
  }
  int rc;
  if ((rc = ensure_work(s, 2))) return rc;
  const int n = s->n;
  double *src = s->work, *dst = s->work + n;
  EHIP_HIP_TRY(hipMemcpyAsync(src, b, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
  if ((rc = ilu_apply(s, src, dst))) return rc;
  EHIP_HIP_TRY(hipMemcpyAsync(x, dst, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
  EHIP_HIP_TRY(hipStreamSynchronize(s->stream));
  return ilu_check(s);
}

extern "C" int ehip_ilu_info(const ehip_system *s, int *order, int64_t *ilu_nnz, int *levels_forward,
                             int *levels_backward) {
  if (!s) return EHIP_EARG;
  if (order) *order = s->ilu_order;
  if (ilu_nnz) *ilu_nnz = s->ilu_nnz;
  if (levels_forward) *levels_forward = s->ilu_loff.empty() ? 0 : (int)s->ilu_loff.size() - 1;
  if (levels_backward) *levels_backward = s->ilu_uoff.empty() ? 0 : (int)s->ilu_uoff.size() - 1;
  return EHIP_OK;
}
