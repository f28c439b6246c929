// Multi-system Krylov engine of libehip: batching and continuous-batching
// serving.
//
// Many independent systems run in the same kernel launches.  The work list
// is a device-resident array of (slot, row chunk) items walked by a fixed
// grid; every slot carries its own Krylov scalars and a phase, so systems
// of very different size and iteration count share one launch sequence.
// That sequence (one iteration of every active slot) is recorded once in a
// hipGraph and replayed; admitting or retiring a system only rewrites the
// work list and the slot table between replays, so the graph never changes.
//
// Per-slot reductions are deterministic: each chunk publishes its partial
// sums with plain stores, and a per-slot finalize launch (bk_fin) adds the
// slot's chunk partials in chunk order and runs the scalar update.
//
// FP64 faithful mode runs HUTI's CG / BiCGStab loop order per slot (true
// residual every iteration, hutiter/src/huti_cg.src:145-227,
// huti_bicgstab.src:150-262).  FP64 fast mode runs the recurrence residual
// and confirms on the true residual, restarting the Krylov sequence from it
// when it misses (also on HUTI's S-norm, omega and rho exits).  Mixed mode
// runs FP32 inner Krylov (FP32 values, vectors and accumulation; recurrence
// residual) inside per-slot FP64 iterative refinement (true residual,
// correction and Jacobi diagonal in FP64).
//
// Chunk size (nonzeros per work item) is per system, from a process-wide
// table keyed by solver form and size class and filled by a timed schedule
// search; the workgroup count is per form.  Neither depends on the other
// systems of the batch, so a system's bits are the same alone and in a mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "ehip_internal.h"
#include "launch.h"
#include "linalg_kernels.h"
#include "spmv_core.h"

extern "C" int ehip_internal_report(ehip_system *s, const double *b0, double *res_true, double *norm);

namespace ehip {

enum Phase : int {
  PH_IDLE = 0,
  PH_INIT = 1,    // FP64: r = b - A x
  PH_START = 2,   // FP64: first search direction
  PH_ITER = 3,    // FP64 Krylov iterations
  PH_RES = 4,     // mixed: FP64 true residual
  PH_PREP = 5,    // mixed: normalised FP32 residual, inner start
  PH_INNER = 6,   // mixed: FP32 inner iterations
  PH_CORR = 7,    // mixed: x += scale * d
  PH_DONE = 8,
};

struct Chunk {
  int slot, r0, r1, k;  // rows [r0, r1) of the slot's system; k-th chunk of that slot
};

constexpr int NV64 = 10, NV32 = 11;

struct SlotPtrs {
  const int *rows, *cols;
  const double *vals;
  const float *vals32;
  double *b, *x;      // scaled right-hand side and solution (the system's own buffers)
  double *w64;        // NV64 vectors of n doubles
  float *w32;         // NV32 vectors of n floats
  const double *du;   // the scaling's D when the unscaled residual is tested, else null
  int n, ck0, nck, method;
  int nb;             // the system's chunk size (nonzeros); a longer chunk is one long row
  // the system's slice-shared offset layout (spmv_sho_body) when it has one
  // and uses it: SpMV ops then take one thread per row instead of staging
  // the chunk's products (null: the CRS stream)
  const int *sho_off, *sho_tab;
  const unsigned *sho_mask;
  const double *sv64;
  const float *sv32;
};

struct BSys {
  double rho, rho_new, alpha, omega, rhsnorm, res, scale, tol, snorm, scale_prev;
  float rho32, rho_new32, alpha32, omega32, beta32, rhsnorm32, res32, itol;
  int iter, info, done, phase, maxit, outer, inner_iter, inner_total, inner_maxit, snorm_exit;
  int inner_bad;   // the inner solve ended on a non-finite residual: its correction is dropped
  int bad_count;   // consecutive such inner solves
  int restarts;    // FP64 fast mode: restarts from the true residual so far
  // unscaled residual test (tol_u > 0): ||b - A x|| / ||b|| = ||r_s / D|| / ||b_s / D||
  double tol_u, bb_u, res_u;
};

// Records the slot's unscaled residual from the sum of squares of r_s / D and
// tells whether it meets tol_u (true when the test is off).
__device__ __forceinline__ bool slot_unscaled_ok(BSys &st, double sum_u) {
  if (!(st.tol_u > 0.0)) return true;
  st.res_u = sqrt(sum_u / st.bb_u);
  return st.res_u < st.tol_u;
}

struct BatchArgs {
  const Chunk *chunks;
  const int *nchunks;
  const SlotPtrs *slots;
  BSys *st;
  double *p64;   // chunk partials, 4 per chunk
  float *p32;
  int *active;   // number of slots not yet done
};

// FP64 vector slots: CG r z p q t dg ; BiCGStab r rtld p t1v v s t2v t dg z
template <typename T>
__device__ __forceinline__ Vecs<T> slot_vecs(T *w, int n, int method, T *x, T *b) {
  Vecs<T> v{};
  const size_t N = (size_t)n;
  v.x = x;
  v.b = b;
  if (method == EHIP_CG) {
    v.r = w; v.z = w + N; v.p = w + 2 * N; v.q = w + 3 * N; v.t = w + 4 * N; v.dg = w + 5 * N;
  } else {
    v.r = w; v.rtld = w + N; v.p = w + 2 * N; v.t1v = w + 3 * N; v.v = w + 4 * N;
    v.sv = w + 5 * N; v.t2v = w + 6 * N; v.t = w + 7 * N; v.dg = w + 8 * N; v.z = w + 9 * N;
  }
  return v;
}

__device__ __forceinline__ void slot_done(BatchArgs &a, BSys &st, int info) {
  st.info = info;
  st.done = 1;
  st.phase = PH_DONE;
  atomicSub(a.active, 1);
}

// One chunk's SpMV through LDS: ITEMS nonzeros per thread, streamed
// branch-free (indices past the chunk are clamped to its last nonzero, so
// every load is unconditional and in bounds; the extra products land in LDS
// slots no row reads), then every row's products summed left to right.
template <class Op, int BS, int ITEMS>
__device__ __forceinline__ void chunk_spmv(const Op &op, typename Op::Ctx &cx, const SlotPtrs &sp, const Chunk &ck,
                                           int j0, int len, typename Op::T *prod,
                                           typename Op::T (&acc)[Op::NV > 0 ? Op::NV : 1]) {
  using T = typename Op::T;
  const T *xin = op.xin(cx);
  const auto *vals = op.vals(sp);
  using VT = typename std::remove_cv<typename std::remove_pointer<decltype(vals)>::type>::type;
  VT vv[ITEMS];
  int cc[ITEMS];
  T xv[ITEMS];
  const int rr = ck.r0 + (int)threadIdx.x;
  const int rl = min(rr, ck.r1 - 1);
  const int ra = sp.rows[rl], rb = sp.rows[rl + 1];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int j = lane_nz<BS>(j0, len, k);
    vv[k] = ntload(vals + j);
    cc[k] = ntload(sp.cols + j);
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) xv[k] = xin[cc[k]];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) prod[threadIdx.x + k * BS] = (T)vv[k] * xv[k];
  lds_barrier();
  if (rr < ck.r1) {
    T s = (T)0;
    for (int k = ra - j0; k < rb - j0; ++k) s = s + prod[k];
    op.row(cx, rr, s, acc);
  }
  for (int r = rr + BS; r < ck.r1; r += BS) {
    const int s0 = sp.rows[r] - j0, s1 = sp.rows[r + 1] - j0;
    T s = (T)0;
    for (int k = s0; k < s1; ++k) s = s + prod[k];
    op.row(cx, r, s, acc);
  }
  lds_barrier();
}

// SpMV of a chunk on the slice-shared offset layout: thread t takes rows
// r0 + t, r0 + t + BS, ... -- the rows, and so the partial sums, that the
// staged path gives it -- and sums each row in CRS order as spmv_sho_body
// does (FP64: the row's own nonzeros by its mask; FP32: the absent entries
// as +0 times x), so the results are the staged path's bits.
template <class Op, int BS>
__device__ __forceinline__ void chunk_spmv_sho(const Op &op, typename Op::Ctx &cx, const SlotPtrs &sp,
                                               const Chunk &ck, typename Op::T (&acc)[Op::NV > 0 ? Op::NV : 1]) {
  using T = typename Op::T;
  using VT = typename std::remove_cv<typename std::remove_pointer<decltype(op.vals(sp))>::type>::type;
  constexpr bool kF32 = sizeof(VT) == 4;
  constexpr int U = 9;
  const VT *vals;
  if constexpr (kF32) vals = sp.sv32; else vals = sp.sv64;
  const T *xin = op.xin(cx);
  const int n = sp.n;
  for (int r = ck.r0 + (int)threadIdx.x; r < ck.r1; r += BS) {
    const int sl = r >> 6, l = r & 63;
    const unsigned m = kF32 ? 0u : sp.sho_mask[r];
    const int p0 = sp.sho_off[sl];
    const int wd = (sp.sho_off[sl + 1] - p0) >> 6;
    const int *t = sp.sho_tab + (p0 >> 6);
    T sum = (T)0;
    for (int k0 = 0; k0 < wd; k0 += U) {
      VT vv[U];
      int cc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kk = min(k0 + u, wd - 1);
        vv[u] = ntload(vals + p0 + kk * 64 + l);
        cc[u] = min(max(r + t[kk], 0), n - 1);
      }
      T xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = xin[cc[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma clang fp contract(off)
        const T tv = sum + (T)vv[u] * xv[u];
        if constexpr (kF32) {
          sum = (k0 + u < wd) ? tv : sum;
        } else {
          const int k = k0 + u;
          sum = (k < 32 && ((m >> (k & 31)) & 1u)) ? tv : sum;
        }
      }
    }
    op.row(cx, r, sum, acc);
  }
}

// One generic kernel: every work item whose slot the op accepts is handled by
// one workgroup -- SpMV ops stage the chunk's products in LDS and give each
// row its left-to-right sum, vector ops visit the rows -- and the op's finish
// runs in the per-slot finalize launch (bk_fin).  NB is the largest chunk
// size among the systems in the work list.
template <class Op, int BS, int NB>
__global__ __launch_bounds__(BS) void bk_run(BatchArgs a, Op op) {
  using T = typename Op::T;
  constexpr int NV = Op::NV;
  __shared__ T prod[Op::SPMV ? NB : 1];
  __shared__ T sh[16 * 4];
  __shared__ int amlast;
  __shared__ int act[BS];  // this round's active chunks, in walk order
  __shared__ int wcount[BS / 64];
  const int nch = *a.nchunks;
  XcdWalk w(nch);
  // The workgroup's chunks are w.first + i w.step.  Most ops are active for
  // only some slots (their phase), so the walk first checks BS chunks at
  // once -- one chunk per thread, one memory latency for all of them -- and
  // compacts the active ones; inactive chunks then cost nothing more.
  const int m = w.first < w.last ? (w.last - w.first + w.step - 1) / w.step : 0;
  for (int i0 = 0; i0 < m; i0 += BS) {
    bool on = false;
    const int ci = w.first + (i0 + (int)threadIdx.x) * w.step;
    if (i0 + (int)threadIdx.x < m) on = op.active(a.st[a.chunks[ci].slot]);
    const unsigned long long bal = __ballot(on);
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    if (lane == 0) wcount[wv] = __popcll(bal);
    __syncthreads();
    int off = 0, nact = 0;
    for (int q = 0; q < BS / 64; ++q) {
      off += q < wv ? wcount[q] : 0;
      nact += wcount[q];
    }
    if (on) act[off + __popcll(bal & ((1ull << lane) - 1))] = ci;
    __syncthreads();
    for (int ai = 0; ai < nact; ++ai) {
      const int c = act[ai];
      const Chunk ck = a.chunks[c];
      const SlotPtrs sp = a.slots[ck.slot];
      BSys &st = a.st[ck.slot];
      typename Op::Ctx cx = op.begin(st, sp);
      T acc[NV > 0 ? NV : 1];
#pragma unroll
      for (int k = 0; k < (NV > 0 ? NV : 1); ++k) acc[k] = (T)0;
      if constexpr (Op::SPMV) {
        const int j0 = sp.rows[ck.r0], j1 = sp.rows[ck.r1];
        const int len = j1 - j0;
        // the branch-free stream's depth follows the chunk's length, not the
        // kernel's capacity NB: every depth gives each row the same
        // left-to-right sum, so a system's results depend on its chunking only
        if (sp.sho_off) {
          chunk_spmv_sho<Op, BS>(op, cx, sp, ck, acc);
        } else if (len > sp.nb) {
          // one row longer than the system's chunk size: strided partial sums
          for (int r = ck.r0; r < ck.r1; ++r) {
            T part = (T)0;
            const auto *vals = op.vals(sp);
            const T *xin = op.xin(cx);
            for (int j = sp.rows[r] + threadIdx.x; j < sp.rows[r + 1]; j += BS)
              part += (T)ntload(vals + j) * xin[ntload(sp.cols + j)];
            T s = block_sum(part, sh);
            if (threadIdx.x == 0) op.row(cx, r, s, acc);
            __syncthreads();
          }
        } else if (len <= BS * 4) {
          chunk_spmv<Op, BS, 4>(op, cx, sp, ck, j0, len, prod, acc);
        } else if (NB >= BS * 8 && len <= BS * 8) {
          chunk_spmv<Op, BS, (NB >= BS * 8 ? 8 : 4)>(op, cx, sp, ck, j0, len, prod, acc);
        } else {
          chunk_spmv<Op, BS, (NB >= BS * 16 ? 16 : 4)>(op, cx, sp, ck, j0, len, prod, acc);
        }
      } else {
        for (int r = ck.r0 + threadIdx.x; r < ck.r1; r += BS) op.row(cx, r, (T)0, acc);
      }
      // publish this chunk's partials (plain stores; the finalize launch reads them)
      if constexpr (NV > 0) {
        T bs[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) bs[k] = block_sum(acc[k], sh + 16 * k);
        T *part = Op::template partials<T>(a);
        if (threadIdx.x == 0) {
#pragma unroll
          for (int k = 0; k < NV; ++k) part[(size_t)(sp.ck0 + ck.k) * 4 + k] = bs[k];
        }
      }
      __syncthreads();
    }
  }
  (void)amlast;
}

// Per-slot finalize: one workgroup per slot sums the slot's chunk partials
// in chunk order and runs the op's scalar update (phase changes included).
template <class Op>
__global__ __launch_bounds__(256) void bk_fin(BatchArgs a, Op op, int nslots) {
  using T = typename Op::T;
  constexpr int NV = Op::NV;
  __shared__ T sh[16 * 4];
  for (int slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
    BSys &st = a.st[slot];
    const SlotPtrs sp = a.slots[slot];
    if (sp.nck == 0 || !op.active(st)) continue;
    T tot[NV > 0 ? NV : 1];
    if constexpr (NV > 0) {
      const T *part = Op::template partials<T>(a);
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        T s = (T)0;
        for (int i = threadIdx.x; i < sp.nck; i += blockDim.x) s += part[(size_t)(sp.ck0 + i) * 4 + k];
        tot[k] = block_sum(s, sh + 16 * k);
      }
    }
    if (threadIdx.x == 0) op.finish(a, st, sp, tot);
    __syncthreads();
  }
}

template <typename TT>
struct OpBase {
  using T = TT;
  template <typename U>
  __device__ static U *partials(BatchArgs &a) {
    if constexpr (sizeof(U) == 8) return reinterpret_cast<U *>(a.p64);
    else return reinterpret_cast<U *>(a.p32);
  }
};

// ---------------------------------------------------------------- FP64 HUTI loop per slot

struct Ctx64 {
  Vecs<double> v;
  double alpha, beta, omega;
  const double *du;
  int restart;  // OpInit: a restart (the norms of b are known)
};

__device__ __forceinline__ Ctx64 ctx64(const BSys &st, const SlotPtrs &sp) {
  Ctx64 c;
  c.v = slot_vecs<double>(sp.w64, sp.n, sp.method, sp.x, sp.b);
  c.du = sp.du;
  c.alpha = st.alpha;
  c.omega = st.omega;
  c.beta = 0.0;
  c.restart = st.restarts > 0;
  return c;
}

// FP64 fast mode: the recurrence residual met the tolerance, or the
// recurrence broke down -- the next iteration recomputes the true residual
// from x (OpInit), ends the solve if it meets the tolerance and otherwise
// restarts the Krylov sequence from it (the single-system fast mode's
// confirmation and restart).
__device__ __forceinline__ void slot_restart(BSys &st) {
  st.restarts += 1;
  st.snorm_exit = 0;
  st.phase = PH_INIT;
}

// r = b - A x ; z = r/dg ; r.z, r.r, b.b, (b/D).(b/D).  On a restart (fast
// mode) b's norms are known and the third partial is (r/D).(r/D) instead:
// the true residual is tested before the Krylov sequence restarts from it.
struct OpInit : OpBase<double> {
  static constexpr bool SPMV = true;
  static constexpr int NV = 4;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INIT; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ const double *xin(const Ctx &c) const { return c.v.x; }
  __device__ const double *vals(const SlotPtrs &sp) const { return sp.vals; }
  __device__ void row(Ctx &c, int i, double s, double (&acc)[4]) const {
    const double bi = c.v.b[i], ri = bi - s;
    c.v.r[i] = ri;
    const double zi = ri / c.v.dg[i];
    c.v.z[i] = zi;
    acc[0] += ri * zi;
    acc[1] += ri * ri;
    const double wi = c.restart ? ri : bi;
    const double ui = c.du ? wi / c.du[i] : wi;
    acc[2] += c.restart ? ui * ui : bi * bi;
    acc[3] += c.restart ? 0.0 : ui * ui;
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &sp, double (&tot)[4]) const {
    if (st.restarts == 0) {
      st.rhsnorm = sqrt(tot[2]);
      st.bb_u = tot[3];
    } else {
      // st.iter is one past the iterations done
      const double res = sqrt(tot[1]) / st.rhsnorm;
      st.res = res;
      if (res < st.tol) {
        if (slot_unscaled_ok(st, tot[2])) {
          st.iter -= 1;
          slot_done(a, st, 1);
          return;
        }
        // the scaled test holds but the unscaled residual misses its
        // tolerance: tighten the recurrence test by the observed ratio
        st.tol = res * (st.tol_u / st.res_u);
      }
      if (st.iter > st.maxit) {
        slot_done(a, st, 2);
        return;
      }
    }
    st.rho = sp.method == EHIP_CG ? tot[0] : tot[1];
    st.phase = PH_START;
  }
};

// CG: p = z.  BiCGStab: rtld = r, p = r, v = 0, t1v = r/dg; alpha = 0, omega = 1
struct OpStart : OpBase<double> {
  static constexpr bool SPMV = false;
  static constexpr int NV = 0;
  using Ctx = Ctx64;
  int method;
  __device__ bool active(const BSys &st) const { return st.phase == PH_START; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&)[1]) const {
    if (method == EHIP_CG) {
      c.v.p[i] = c.v.z[i];
    } else {
      const double ri = c.v.r[i];
      c.v.rtld[i] = ri;
      c.v.p[i] = ri;
      c.v.v[i] = 0.0;
      c.v.t1v[i] = ri / c.v.dg[i];
    }
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&)[1]) const {
    st.alpha = 0.0;
    st.omega = 1.0;
    if (st.restarts == 0) st.iter = 1;  // a restart keeps counting
    st.phase = PH_ITER;
    if (st.rho == 0.0) slot_done(a, st, method == EHIP_CG ? 20 : 35);
  }
};

// ---- CG, faithful
struct OpCgQ : OpBase<double> {  // q = A p ; alpha = rho / p.q
  static constexpr bool SPMV = true;
  static constexpr int NV = 1;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ const double *xin(const Ctx &c) const { return c.v.p; }
  __device__ const double *vals(const SlotPtrs &sp) const { return sp.vals; }
  __device__ void row(Ctx &c, int i, double s, double (&acc)[1]) const {
    c.v.q[i] = s;
    acc[0] += c.v.p[i] * s;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&tot)[1]) const {
    st.alpha = st.rho / tot[0];
  }
};

struct OpCgDa : OpBase<double> {  // x += alpha p ; r -= alpha q ; z = r/dg ; rho_new
  static constexpr bool SPMV = false;
  static constexpr int NV = 1;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&acc)[1]) const {
    c.v.x[i] = c.v.x[i] + c.alpha * c.v.p[i];
    const double ri = c.v.r[i] - c.alpha * c.v.q[i];
    c.v.r[i] = ri;
    const double zi = ri / c.v.dg[i];
    c.v.z[i] = zi;
    acc[0] += ri * zi;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&tot)[1]) const { st.rho_new = tot[0]; }
};

struct OpCgRes : OpBase<double> {  // ||A x - b|| / ||b|| ; p = z + (rho_new/rho) p
  static constexpr bool SPMV = true;
  static constexpr int NV = 2;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const {
    Ctx c = ctx64(st, sp);
    c.beta = st.rho_new / st.rho;
    return c;
  }
  __device__ const double *xin(const Ctx &c) const { return c.v.x; }
  __device__ const double *vals(const SlotPtrs &sp) const { return sp.vals; }
  __device__ void row(Ctx &c, int i, double s, double (&acc)[2]) const {
    const double ti = s - c.v.b[i];
    acc[0] += ti * ti;
    const double ui = c.du ? ti / c.du[i] : ti;
    acc[1] += ui * ui;
    c.v.p[i] = c.v.z[i] + c.beta * c.v.p[i];
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&tot)[2]) const {
    const double res = sqrt(tot[0]) / st.rhsnorm;
    st.res = res;
    if (res < st.tol && slot_unscaled_ok(st, tot[1])) { slot_done(a, st, 1); return; }
    st.rho = st.rho_new;
    st.iter += 1;
    if (st.iter > st.maxit) slot_done(a, st, 2);
    else if (st.rho_new == 0.0) slot_done(a, st, 20);
  }
};

// ---- CG, fast: recurrence residual, true residual only to confirm (OpInit)
struct OpCgDaF : OpBase<double> {  // x += alpha p ; r -= alpha q ; z = r/dg ; rho_new, r.r
  static constexpr bool SPMV = false;
  static constexpr int NV = 2;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&acc)[2]) const {
    c.v.x[i] = c.v.x[i] + c.alpha * c.v.p[i];
    const double ri = c.v.r[i] - c.alpha * c.v.q[i];
    c.v.r[i] = ri;
    const double zi = ri / c.v.dg[i];
    c.v.z[i] = zi;
    acc[0] += ri * zi;
    acc[1] += ri * ri;
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&tot)[2]) const {
    st.rho_new = tot[0];
    st.res = sqrt(tot[1]) / st.rhsnorm;
    st.iter += 1;
    if (st.res < st.tol || tot[0] == 0.0) slot_restart(st);
    else if (st.iter > st.maxit) slot_done(a, st, 2);
  }
};

struct OpCgP : OpBase<double> {  // p = z + (rho_new/rho) p
  static constexpr bool SPMV = false;
  static constexpr int NV = 0;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const {
    Ctx c = ctx64(st, sp);
    c.beta = st.rho_new / st.rho;
    return c;
  }
  __device__ void row(Ctx &c, int i, double, double (&)[1]) const { c.v.p[i] = c.v.z[i] + c.beta * c.v.p[i]; }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&)[1]) const { st.rho = st.rho_new; }
};

// ---- BiCGStab, faithful
template <typename TT>
struct CtxT {
  Vecs<TT> v;
  TT alpha, beta, omega;
};

struct OpBiV : OpBase<double> {  // v = A t1v ; alpha = rho / rtld.v
  static constexpr bool SPMV = true;
  static constexpr int NV = 1;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ const double *xin(const Ctx &c) const { return c.v.t1v; }
  __device__ const double *vals(const SlotPtrs &sp) const { return sp.vals; }
  __device__ void row(Ctx &c, int i, double s, double (&acc)[1]) const {
    c.v.v[i] = s;
    acc[0] += c.v.rtld[i] * s;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&tot)[1]) const { st.alpha = st.rho / tot[0]; }
};

struct OpBiS : OpBase<double> {  // s = r - alpha v ; t2v = s/dg ; ||s||
  static constexpr bool SPMV = false;
  static constexpr int NV = 1;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&acc)[1]) const {
    const double si = c.v.r[i] - c.alpha * c.v.v[i];
    c.v.sv[i] = si;
    c.v.t2v[i] = si / c.v.dg[i];
    acc[0] += si * si;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&tot)[1]) const {
    st.snorm = sqrt(tot[0]);
    if (st.snorm < (double)1.17549435E-38f) st.snorm_exit = 1;  // HUTI_EPSILON
  }
};

struct OpBiSnorm : OpBase<double> {  // s-norm exit: x += alpha t1v (fast mode: then restart)
  static constexpr bool SPMV = false;
  static constexpr int NV = 0;
  using Ctx = Ctx64;
  int fast;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER && st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&)[1]) const { c.v.x[i] = c.v.x[i] + c.alpha * c.v.t1v[i]; }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&)[1]) const {
    if (!fast) {
      slot_done(a, st, 36);
      return;
    }
    st.iter += 1;
    if (st.iter > st.maxit) slot_done(a, st, 2);
    else slot_restart(st);
  }
};

struct OpBiT : OpBase<double> {  // t = A t2v ; omega = t.s / t.t
  static constexpr bool SPMV = true;
  static constexpr int NV = 2;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER && !st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ const double *xin(const Ctx &c) const { return c.v.t2v; }
  __device__ const double *vals(const SlotPtrs &sp) const { return sp.vals; }
  __device__ void row(Ctx &c, int i, double s, double (&acc)[2]) const {
    c.v.t[i] = s;
    acc[0] += s * c.v.sv[i];
    acc[1] += s * s;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&tot)[2]) const { st.omega = tot[0] / tot[1]; }
};

struct OpBiXr : OpBase<double> {  // x += alpha t1v + omega t2v ; r = s - omega t ; rho_new = rtld.r
  static constexpr bool SPMV = false;
  static constexpr int NV = 1;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER && !st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&acc)[1]) const {
    c.v.x[i] = c.v.x[i] + c.alpha * c.v.t1v[i] + c.omega * c.v.t2v[i];
    const double ri = c.v.sv[i] - c.omega * c.v.t[i];
    c.v.r[i] = ri;
    acc[0] += c.v.rtld[i] * ri;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&tot)[1]) const { st.rho_new = tot[0]; }
};

struct OpBiRes : OpBase<double> {  // ||A x - b||/||b|| ; p = r + beta (p - omega v) ; t1v = p/dg
  static constexpr bool SPMV = true;
  static constexpr int NV = 2;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER && !st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const {
    Ctx c = ctx64(st, sp);
    c.beta = (st.rho_new * st.alpha) / (st.rho * st.omega);
    return c;
  }
  __device__ const double *xin(const Ctx &c) const { return c.v.x; }
  __device__ const double *vals(const SlotPtrs &sp) const { return sp.vals; }
  __device__ void row(Ctx &c, int i, double s, double (&acc)[2]) const {
    const double ti = s - c.v.b[i];
    acc[0] += ti * ti;
    const double ui = c.du ? ti / c.du[i] : ti;
    acc[1] += ui * ui;
    const double pi = c.v.r[i] + c.beta * (c.v.p[i] - c.omega * c.v.v[i]);
    c.v.p[i] = pi;
    c.v.t1v[i] = pi / c.v.dg[i];
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&tot)[2]) const {
    const double res = sqrt(tot[0]) / st.rhsnorm;
    st.res = res;
    if (res < st.tol && slot_unscaled_ok(st, tot[1])) { slot_done(a, st, 1); return; }
    if (st.omega == 0.0) { slot_done(a, st, 37); return; }
    st.iter += 1;
    if (st.iter > st.maxit) slot_done(a, st, 2);
    else if (st.rho_new == 0.0) slot_done(a, st, 35);
    else st.rho = st.rho_new;
  }
};

// ---- BiCGStab, fast: recurrence residual, true residual only to confirm (OpInit)
struct OpBiXrF : OpBase<double> {  // x += alpha t1v + omega t2v ; r = s - omega t ; rho_new, r.r
  static constexpr bool SPMV = false;
  static constexpr int NV = 2;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER && !st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctx64(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&acc)[2]) const {
    c.v.x[i] = c.v.x[i] + c.alpha * c.v.t1v[i] + c.omega * c.v.t2v[i];
    const double ri = c.v.sv[i] - c.omega * c.v.t[i];
    c.v.r[i] = ri;
    acc[0] += c.v.rtld[i] * ri;
    acc[1] += ri * ri;
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&tot)[2]) const {
    st.rho_new = tot[0];
    st.res = sqrt(tot[1]) / st.rhsnorm;
    st.iter += 1;
    if (st.res < st.tol || st.omega == 0.0 || tot[0] == 0.0) slot_restart(st);
    else if (st.iter > st.maxit) slot_done(a, st, 2);
  }
};

struct OpBiPF : OpBase<double> {  // p = r + beta (p - omega v) ; t1v = p/dg
  static constexpr bool SPMV = false;
  static constexpr int NV = 0;
  using Ctx = Ctx64;
  __device__ bool active(const BSys &st) const { return st.phase == PH_ITER && !st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const {
    Ctx c = ctx64(st, sp);
    c.beta = (st.rho_new * st.alpha) / (st.rho * st.omega);
    return c;
  }
  __device__ void row(Ctx &c, int i, double, double (&)[1]) const {
    const double pi = c.v.r[i] + c.beta * (c.v.p[i] - c.omega * c.v.v[i]);
    c.v.p[i] = pi;
    c.v.t1v[i] = pi / c.v.dg[i];
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, double (&)[1]) const { st.rho = st.rho_new; }
};

// ---------------------------------------------------------------- mixed: FP64 refinement per slot

struct CtxM {
  Vecs<float> v;   // inner FP32 vectors; v.x = d (the correction), v.b = normalised residual
  double *t64, *x64, *b64, *dg64;
  const double *du;
  double scale;
  double ratio;  // previous residual scale / current one (0 before the first replacement)
  float alpha, beta, omega;
  int bad;       // the inner solve broke down: no correction
};

__device__ __forceinline__ CtxM ctxm(const BSys &st, const SlotPtrs &sp) {
  CtxM c;
  const int nvi = sp.method == EHIP_CG ? 6 : 10;
  c.v = slot_vecs<float>(sp.w32, sp.n, sp.method, sp.w32 + (size_t)nvi * sp.n, nullptr);
  c.v.b = c.v.r;
  c.t64 = sp.w64;
  c.dg64 = sp.w64 + (size_t)sp.n;
  c.x64 = sp.x;
  c.b64 = sp.b;
  c.du = sp.du;
  c.scale = st.scale;
  c.ratio = st.scale_prev > 0.0 ? st.scale_prev / st.scale : 0.0;
  c.alpha = st.alpha32;
  c.beta = st.beta32;
  c.omega = st.omega32;
  c.bad = st.inner_bad;
  return c;
}

// t = b - A x (FP64), the refinement's true residual
struct OpRes : OpBase<double> {
  static constexpr bool SPMV = true;
  static constexpr int NV = 4;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_RES; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ const double *xin(const Ctx &c) const { return c.x64; }
  __device__ const double *vals(const SlotPtrs &sp) const { return sp.vals; }
  __device__ void row(Ctx &c, int i, double s, double (&acc)[4]) const {
    const double bi = c.b64[i], ti = bi - s;
    c.t64[i] = ti;
    acc[0] += ti * ti;
    acc[1] += bi * bi;
    const double d = c.du ? c.du[i] : 1.0;
    const double ui = ti / d, vi = bi / d;
    acc[2] += ui * ui;
    acc[3] += vi * vi;
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&tot)[4]) const {
    if (st.outer == 0) {
      st.rhsnorm = sqrt(tot[1]);
      st.bb_u = tot[3];
    }
    const double nt = sqrt(tot[0]);
    const double res = nt / st.rhsnorm;
    st.res = res;
    st.outer += 1;
    if (res < st.tol && slot_unscaled_ok(st, tot[2])) slot_done(a, st, 1);
    else if (!(res == res)) slot_done(a, st, 3);
    else if (st.inner_total >= st.maxit) slot_done(a, st, 2);
    else {
      st.scale_prev = st.outer > 1 ? st.scale : 0.0;
      st.scale = nt;
      st.phase = PH_PREP;
    }
  }
};

// r32 = t/||t|| ; d = 0 ; BiCGStab: rtld = p = r, v = 0, t1v = r/dg (restart).
// CG (reliable update): z = r/dg, and the search direction survives the
// residual replacement -- p and q = A p are carried over, rescaled from the
// previous residual's scale to this one, and beta continues the sequence
// (beta = r.z / (r.z before the last step, rescaled)).  First time: p = q = 0.
// (Reliable updates for BiCGStab -- rtld kept, p and v rescaled -- were
// measured on the convection-diffusion batch: about twice the FP64
// iteration count like the restarts, slower at every replacement threshold,
// and divergent at 0.1.)
struct OpPrep : OpBase<float> {
  static constexpr bool SPMV = false;
  static constexpr int NV = 2;
  using Ctx = CtxM;
  int method;
  __device__ bool active(const BSys &st) const { return st.phase == PH_PREP; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ void row(Ctx &c, int i, float, float (&acc)[2]) const {
    const float ri = (float)(c.t64[i] / c.scale);
    c.v.r[i] = ri;
    c.v.x[i] = 0.f;
    if (method == EHIP_CG) {
      const float zi = ri / c.v.dg[i];
      c.v.z[i] = zi;
      const float ratio = (float)c.ratio;
      c.v.p[i] = ratio > 0.f ? c.v.p[i] * ratio : 0.f;
      c.v.q[i] = ratio > 0.f ? c.v.q[i] * ratio : 0.f;
      acc[0] += ri * zi;
    } else {
      c.v.rtld[i] = ri;
      c.v.p[i] = ri;
      c.v.v[i] = 0.f;
      c.v.t1v[i] = ri / c.v.dg[i];
      acc[0] += ri * ri;
    }
    acc[1] += ri * ri;
  }
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, float (&tot)[2]) const {
    const double ratio = st.scale_prev > 0.0 ? st.scale_prev / st.scale : 0.0;
    st.beta32 = 0.f;
    if (method == EHIP_CG && ratio > 0.0 && st.rho_new32 != 0.f)
      st.beta32 = (float)((double)tot[0] / ((double)st.rho_new32 * ratio * ratio));
    st.rho32 = tot[0];
    st.rhsnorm32 = sqrtf(tot[1]);
    st.alpha32 = 0.f;
    st.omega32 = 1.f;
    st.inner_iter = 0;
    st.snorm_exit = 0;
    if (tot[0] == 0.f) slot_done(a, st, method == EHIP_CG ? 20 : 35);
    else st.phase = PH_INNER;
  }
};

__device__ __forceinline__ void inner_step_done(BSys &st, float res, bool breakdown) {
  st.res32 = res;
  st.inner_iter += 1;
  st.inner_total += 1;
  const bool finite = res == res && res <= 3.4e38f;
  if (!finite) st.inner_bad = 1;
  if (res < st.itol || breakdown || !finite || st.inner_iter >= st.inner_maxit || st.inner_total >= st.maxit)
    st.phase = PH_CORR;
}

// inner CG step 1: q = A z + beta q ; p = z + beta p ; alpha = rho / p.q
struct OpCgF1 : OpBase<float> {
  static constexpr bool SPMV = true;
  static constexpr int NV = 1;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ const float *xin(const Ctx &c) const { return c.v.z; }
  __device__ const float *vals(const SlotPtrs &sp) const { return sp.vals32; }
  __device__ void row(Ctx &c, int i, float s, float (&acc)[1]) const {
    const float qi = s + c.beta * c.v.q[i];
    const float pi = c.v.z[i] + c.beta * c.v.p[i];
    c.v.q[i] = qi;
    c.v.p[i] = pi;
    acc[0] += pi * qi;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, float (&tot)[1]) const { st.alpha32 = st.rho32 / tot[0]; }
};

// inner CG step 2: d += alpha p ; r -= alpha q ; z = r/dg ; rho, ||r||
struct OpCgF2 : OpBase<float> {
  static constexpr bool SPMV = false;
  static constexpr int NV = 2;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ void row(Ctx &c, int i, float, float (&acc)[2]) const {
    c.v.x[i] = c.v.x[i] + c.alpha * c.v.p[i];
    const float ri = c.v.r[i] - c.alpha * c.v.q[i];
    c.v.r[i] = ri;
    const float zi = ri / c.v.dg[i];
    c.v.z[i] = zi;
    acc[0] += ri * zi;
    acc[1] += ri * ri;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, float (&tot)[2]) const {
    const float rho_old = st.rho32;
    st.rho_new32 = rho_old;  // kept for the next residual replacement
    st.rho32 = tot[0];
    st.beta32 = tot[0] / rho_old;
    inner_step_done(st, sqrtf(tot[1]) / st.rhsnorm32, tot[0] == 0.f);
  }
};

// inner BiCGStab (recurrence residual)
struct OpBiV32 : OpBase<float> {
  static constexpr bool SPMV = true;
  static constexpr int NV = 1;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ const float *xin(const Ctx &c) const { return c.v.t1v; }
  __device__ const float *vals(const SlotPtrs &sp) const { return sp.vals32; }
  __device__ void row(Ctx &c, int i, float s, float (&acc)[1]) const {
    c.v.v[i] = s;
    acc[0] += c.v.rtld[i] * s;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, float (&tot)[1]) const { st.alpha32 = st.rho32 / tot[0]; }
};

struct OpBiS32 : OpBase<float> {
  static constexpr bool SPMV = false;
  static constexpr int NV = 1;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ void row(Ctx &c, int i, float, float (&acc)[1]) const {
    const float si = c.v.r[i] - c.alpha * c.v.v[i];
    c.v.sv[i] = si;
    c.v.t2v[i] = si / c.v.dg[i];
    acc[0] += si * si;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, float (&tot)[1]) const {
    if (sqrtf(tot[0]) < 1.17549435E-38f) st.snorm_exit = 1;
  }
};

struct OpBiSnorm32 : OpBase<float> {  // d += alpha t1v, end of the inner solve
  static constexpr bool SPMV = false;
  static constexpr int NV = 0;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER && st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ void row(Ctx &c, int i, float, float (&)[1]) const { c.v.x[i] = c.v.x[i] + c.alpha * c.v.t1v[i]; }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, float (&)[1]) const {
    st.inner_iter += 1;
    st.inner_total += 1;
    st.phase = PH_CORR;
  }
};

struct OpBiT32 : OpBase<float> {
  static constexpr bool SPMV = true;
  static constexpr int NV = 2;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER && !st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ const float *xin(const Ctx &c) const { return c.v.t2v; }
  __device__ const float *vals(const SlotPtrs &sp) const { return sp.vals32; }
  __device__ void row(Ctx &c, int i, float s, float (&acc)[2]) const {
    c.v.t[i] = s;
    acc[0] += s * c.v.sv[i];
    acc[1] += s * s;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, float (&tot)[2]) const { st.omega32 = tot[0] / tot[1]; }
};

struct OpBiXr32 : OpBase<float> {
  static constexpr bool SPMV = false;
  static constexpr int NV = 2;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER && !st.snorm_exit; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ void row(Ctx &c, int i, float, float (&acc)[2]) const {
    c.v.x[i] = c.v.x[i] + c.alpha * c.v.t1v[i] + c.omega * c.v.t2v[i];
    const float ri = c.v.sv[i] - c.omega * c.v.t[i];
    c.v.r[i] = ri;
    acc[0] += c.v.rtld[i] * ri;
    acc[1] += ri * ri;
  }
  __device__ void finish(BatchArgs &, BSys &st, const SlotPtrs &, float (&tot)[2]) const {
    const float rho_new = tot[0];
    const bool brk = st.omega32 == 0.f || rho_new == 0.f;
    st.beta32 = (rho_new * st.alpha32) / (st.rho32 * st.omega32);
    st.rho32 = rho_new;
    inner_step_done(st, sqrtf(tot[1]) / st.rhsnorm32, brk);
  }
};

struct OpBiP32 : OpBase<float> {  // p = r + beta (p - omega v) ; t1v = p/dg
  static constexpr bool SPMV = false;
  static constexpr int NV = 0;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_INNER; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ void row(Ctx &c, int i, float, float (&)[1]) const {
    const float pi = c.v.r[i] + c.beta * (c.v.p[i] - c.omega * c.v.v[i]);
    c.v.p[i] = pi;
    c.v.t1v[i] = pi / c.v.dg[i];
  }
  __device__ void finish(BatchArgs &, BSys &, const SlotPtrs &, float (&)[1]) const {}
};

// x += scale * d, then back to the FP64 residual
struct OpCorr : OpBase<double> {
  static constexpr bool SPMV = false;
  static constexpr int NV = 0;
  using Ctx = CtxM;
  __device__ bool active(const BSys &st) const { return st.phase == PH_CORR; }
  __device__ Ctx begin(const BSys &st, const SlotPtrs &sp) const { return ctxm(st, sp); }
  __device__ void row(Ctx &c, int i, double, double (&)[1]) const {
    if (!c.bad) c.x64[i] = c.x64[i] + c.scale * (double)c.v.x[i];
  }
  // An inner solve that broke down with a non-finite residual leaves x as it
  // was; the next refinement step restarts from the FP64 residual with the
  // inner iteration cap halved below the breakdown point (the restart is
  // deterministic, so the same cap would break down the same way).  Three in
  // a row end the solve as diverged.
  __device__ void finish(BatchArgs &a, BSys &st, const SlotPtrs &, double (&)[1]) const {
    if (st.inner_bad) {
      st.inner_bad = 0;
      st.inner_maxit = st.inner_iter / 2 > 1 ? st.inner_iter / 2 : 1;
      st.scale = 0.0;  // so the next replacement starts CG afresh: no search direction carried over
      if (++st.bad_count >= 3) {
        slot_done(a, st, 3);
        return;
      }
    } else {
      st.bad_count = 0;
    }
    st.phase = PH_RES;
  }
};

// ---------------------------------------------------------------- per-slot setup kernels

__global__ void k_fill_x(int n, double *x, double v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

__global__ void k_inc(int *c) { *c += 1; }

// a slot's initial state, written in stream order behind any replay queued
__global__ void k_set_slot(BSys *dst, BSys v) { *dst = v; }

__global__ void k_to_f32(int64_t nnz, const double *v, float *v32) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * blockDim.x)
    v32[j] = (float)v[j];
}

// dg64 / dg32 from the scaled diagonal (CRS_DiagPrecondition's divisor)
__global__ void bk_dg(int n, const int *diag, const double *vals, double *dg64, float *dg32, int precond) {
  const double AEPS = 10.0 * DBL_EPSILON;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double d = vals[diag[i]];
    const double g = (precond && fabs(d) > AEPS) ? d : 1.0;
    if (dg64) dg64[i] = g;
    if (dg32) dg32[i] = (float)g;
  }
}

}  // namespace ehip

using namespace ehip;

// ================================================================ host engine

struct ehip_batch {
  int dev = 0;
  hipStream_t stream = nullptr;
  ehip_params p{};
  int capacity = 0;              // slots
  int64_t chunk_cap = 0;         // work items
  // device
  Chunk *chunks = nullptr;
  int *nchunks = nullptr;
  SlotPtrs *slots = nullptr;
  BSys *st = nullptr;
  double *p64 = nullptr;
  float *p32 = nullptr;
  int *active = nullptr;
  double *bnorm = nullptr;       // per slot, ||b|| of the right-hand-side scaling (device)
  int *h_active = nullptr;       // pinned
  BSys *h_st = nullptr;          // pinned mirror
  // host view
  struct Slot {
    ehip_system *s = nullptr;
    bool used = false;
    std::vector<Chunk> ck;       // with slot/k filled, ck0 set at list build
    double *w64 = nullptr;       // NV64*n (+ unscaled b copy)
    float *w32 = nullptr;
    double bscale = 1.0;
    int64_t bytes = 0;
    int nb = 2048;               // chunk size (nonzeros) of this system
    int64_t first_launch = 0;    // the first replay whose results include this system
    bool asm_check = false;      // its assembly's degenerate-element count is read at retirement
    bool sho = false;            // SpMV on the system's slice-shared offset layout
  };
  std::vector<Slot> slot;
  bool graph_ok = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_iters = 0;
  int mixed = 0, method = 0;
  int fast = 0;                  // FP64: recurrence residual with true-residual restarts
  int refine_every = 1;          // mixed: the refinement ops run every refine_every-th iteration
  int nbmax = 2048;              // largest chunk size in the work list (selects bk_run's capacity)
  int grid = 2048;               // workgroups of the multi-system launches
  // serving: admissions only enqueue (the worklist is rebuilt at the next
  // step), and a step launches the next replay without waiting for it
  hipEvent_t ev_admit = nullptr;
  bool dirty = false;            // the worklist no longer matches the slots
  bool in_flight = false;        // a replay (and its state copy) is queued
  int64_t nlaunch = 0;           // replays launched
  int tuning = 0;                // ehip_batch_set_tuning: 0 off, 1 untuned classes, 2 everything
  int tune_replays = 3;
  int64_t bytes = 0;
  int64_t budget = 0;              // serving: HBM budget for solver workspace (0 = none)
  std::vector<int> batch_members;  // for ehip_batch_* : slot of each system (-1: concurrent lane)
  struct Lane *lane = nullptr;     // serving: concurrent lane for large systems (null = off)
  // ehip_batch_solve: systems of at least lane_min_dofs unknowns are solved
  // whole on lane_workers concurrent threads instead of in the multi-system
  // launches (ehip_batch_set_lane)
  int64_t lane_min_dofs = 40000;
  int lane_workers = 0;
};

// Serving's concurrent lane: a system of at least min_dofs unknowns is not
// batched -- one iteration of a large system fills the GPU on its own -- but
// solved by ehip_solve on one of `workers` host threads, each on the
// system's own stream, so large solves overlap one another and the batched
// small ones on the device.  Admitted systems queue in the lane (up to
// kLaneDepth per worker) so that a worker starts its next solve the moment
// it finishes one, not at the caller's next server step; a request waits
// (EHIP_ECAP) when the lane queue is full or the HBM budget would not hold
// its workspace.
constexpr int kLaneDepth = 4;
struct Lane {
  std::vector<std::thread> th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ehip_system *> q;
  std::vector<std::pair<ehip_system *, ehip_report>> done;
  std::map<ehip_system *, int64_t> need;
  int running = 0;  // queued or being solved
  int64_t bytes = 0;
  int64_t min_dofs = 0;
  int workers = 0;
  bool stop = false;
  ehip_params p{};
};

namespace ehip {

// Chunk sizes (nonzeros per work item) the multi-system kernels are built
// for, and the workgroup counts the schedule search tries.  The workgroup
// size stays 256: it fixes which rows each thread's partial sums collect,
// so it is kept the same for every system.
constexpr int kChunkSizes[3] = {1024, 2048, 4096};
constexpr int kGrids[3] = {1024, 2048, 4096};
constexpr int kDefaultNB = 2048;
constexpr int kSizeClasses = 6;
constexpr int kRefineEvery = 2;

// Size class of a system by its nonzeros: < 2^15, 2^17, 2^19, 2^21, 2^23, more.
static int size_class(int64_t nnz) {
  int c = 0;
  for (int64_t lim = (int64_t)1 << 15; c < kSizeClasses - 1 && nnz >= lim; lim <<= 2) ++c;
  return c;
}

// The schedule table: per (precision form, method, fast mode) the chunk
// size of each size class and the workgroup count; 0 = not tuned (defaults).
// Process-wide, filled by the schedule search (ehip_batch_set_tuning) or set
// directly (ehip_batch_set_schedule).  A system's results depend only on the
// chunk size of its class, never on the other systems of its batch.
struct SchedTable {
  std::mutex mu;
  int nb[2][2][2][kSizeClasses] = {};
  int grid[2][2][2] = {};
};
static SchedTable g_sched;

static int sched_nb(int mixed, int method, int fast, int cls) {
  std::lock_guard<std::mutex> lk(g_sched.mu);
  const int v = g_sched.nb[mixed][method][fast][cls];
  return v ? v : kDefaultNB;
}
static int sched_grid(int mixed, int method, int fast) {
  std::lock_guard<std::mutex> lk(g_sched.mu);
  const int v = g_sched.grid[mixed][method][fast];
  return v ? v : max_grid();
}

static std::vector<Chunk> make_chunks(const ehip_system *s, int nb) {
  std::vector<Chunk> out;
  const int *rows = s->h_rows.data();
  int r = 0, k = 0;
  while (r < s->n) {
    const int start = rows[r];
    int e = r + 1;
    if (rows[e] - start > nb) {
      out.push_back({0, r, e, k++});
      r = e;
      continue;
    }
    int lo = e, hi = s->n;
    while (lo < hi) {
      int mid = lo + (hi - lo + 1) / 2;
      if (rows[mid] - start <= nb) lo = mid; else hi = mid - 1;
    }
    out.push_back({0, r, lo, k++});
    r = lo;
  }
  return out;
}

static BatchArgs args_of(ehip_batch *b) {
  return BatchArgs{b->chunks, b->nchunks, b->slots, b->st, b->p64, b->p32, b->active};
}

template <class Op>
static void launch_op(ehip_batch *b, const Op &op) {
  const int grid = b->grid;
  if (b->nbmax <= 1024)
    launch_k((bk_run<Op, 256, 1024>), dim3(grid), dim3(256), b->stream, args_of(b), op);
  else if (b->nbmax <= 2048)
    launch_k((bk_run<Op, 256, 2048>), dim3(grid), dim3(256), b->stream, args_of(b), op);
  else
    launch_k((bk_run<Op, 256, 4096>), dim3(grid), dim3(256), b->stream, args_of(b), op);
  const int fgrid = std::min(b->capacity, 1024);
  launch_k((bk_fin<Op>), dim3(fgrid), dim3(256), b->stream, args_of(b), op, b->capacity);
}

// one iteration of every active slot.  Mixed precision: the refinement ops
// (correction, FP64 residual, FP32 restart) are enqueued only when `refine`
// is set; a slot whose inner solve ended in between waits in PH_CORR,
// doing nothing, until the next iteration that has them.  Every op acts on
// a slot only in its own phase and idle iterations count nowhere, so a
// slot's arithmetic -- and its bits -- do not depend on where they fall.
static void enqueue_iteration(ehip_batch *b, bool refine) {
  if (!b->mixed) {
    launch_op(b, OpInit{});
    launch_op(b, OpStart{{}, b->method});
    if (b->method == EHIP_CG) {
      launch_op(b, OpCgQ{});
      if (b->fast) {
        launch_op(b, OpCgDaF{});
        launch_op(b, OpCgP{});
      } else {
        launch_op(b, OpCgDa{});
        launch_op(b, OpCgRes{});
      }
    } else {
      launch_op(b, OpBiV{});
      launch_op(b, OpBiS{});
      launch_op(b, OpBiSnorm{{}, b->fast});
      launch_op(b, OpBiT{});
      if (b->fast) {
        launch_op(b, OpBiXrF{});
        launch_op(b, OpBiPF{});
      } else {
        launch_op(b, OpBiXr{});
        launch_op(b, OpBiRes{});
      }
    }
  } else {
    if (refine) {
      launch_op(b, OpCorr{});
      launch_op(b, OpRes{});
      launch_op(b, OpPrep{{}, b->method});
    }
    if (b->method == EHIP_CG) {
      launch_op(b, OpCgF1{});
      launch_op(b, OpCgF2{});
    } else {
      launch_op(b, OpBiV32{});
      launch_op(b, OpBiS32{});
      launch_op(b, OpBiSnorm32{});
      launch_op(b, OpBiT32{});
      launch_op(b, OpBiXr32{});
      launch_op(b, OpBiP32{});
    }
  }
}

static int build_graph(ehip_batch *b, int iters) {
  // nothing recorded earlier may still be queued when its graph goes
  EHIP_HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  // built node by node, not by stream capture, which would make the runtime
  // reject legacy-stream work from other threads (the concurrent lane's)
  GraphRec rec;
  EHIP_HIP_TRY(hipGraphCreate(&rec.graph, 0));
  b->graph = rec.graph;
  tl_graph_rec = &rec;
  for (int k = 0; k < iters; ++k) enqueue_iteration(b, k % b->refine_every == 0);
  tl_graph_rec = nullptr;
  EHIP_HIP_TRY(rec.err);
  EHIP_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_iters = iters;
  b->graph_ok = true;
  return EHIP_OK;
}

// Rewrites the device work list from the used slots (between replays only).
static int upload_worklist(ehip_batch *b) {
  std::vector<Chunk> all;
  std::vector<SlotPtrs> tab(b->capacity);
  int nbmax = 1024;
  for (int i = 0; i < b->capacity; ++i) {
    auto &sl = b->slot[i];
    SlotPtrs sp{};
    if (sl.used) {
      ehip_system *s = sl.s;
      sp.rows = s->rows; sp.cols = s->cols; sp.vals = s->vals; sp.vals32 = s->vals32;
      sp.b = s->rhs; sp.x = s->x; sp.w64 = sl.w64; sp.w32 = sl.w32;
      sp.du = (b->p.true_residual_stop && b->p.scaling) ? s->dscale : nullptr;
      sp.n = s->n; sp.method = b->method;
      sp.nb = sl.nb;
      if (sl.sho) {
        sp.sho_off = s->sho_off; sp.sho_tab = s->sho_tab; sp.sho_mask = s->sho_mask;
        sp.sv64 = s->sho64; sp.sv32 = s->sho32;
      }
      nbmax = std::max(nbmax, sl.nb);
      sp.ck0 = (int)all.size();
      sp.nck = (int)sl.ck.size();
      for (Chunk c : sl.ck) {
        c.slot = i;
        all.push_back(c);
      }
    }
    tab[i] = sp;
  }
  if ((int64_t)all.size() > b->chunk_cap) {
    // grow the work list and the partials
    if (b->chunks) (void)hipFree(b->chunks);
    if (b->p64) (void)hipFree(b->p64);
    if (b->p32) (void)hipFree(b->p32);
    b->chunk_cap = (int64_t)all.size() * 2 + 1024;
    EHIP_HIP_TRY(hipMalloc(&b->chunks, sizeof(Chunk) * b->chunk_cap));
    EHIP_HIP_TRY(hipMalloc(&b->p64, sizeof(double) * 4 * b->chunk_cap));
    EHIP_HIP_TRY(hipMalloc(&b->p32, sizeof(float) * 4 * b->chunk_cap));
    b->graph_ok = false;  // kernel arguments changed
  }
  // a larger chunk size than the recorded kernels take needs new ones (a
  // smaller one runs in the recorded kernels unchanged)
  if (nbmax > b->nbmax) b->graph_ok = false;
  if (nbmax > b->nbmax || !b->graph_ok) b->nbmax = nbmax;
  const int nch = (int)all.size();
  if (nch) EHIP_HIP_TRY(hipMemcpyAsync(b->chunks, all.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, b->stream));
  EHIP_HIP_TRY(hipMemcpyAsync(b->nchunks, &nch, sizeof(int), hipMemcpyHostToDevice, b->stream));
  EHIP_HIP_TRY(hipMemcpyAsync(b->slots, tab.data(), sizeof(SlotPtrs) * b->capacity, hipMemcpyHostToDevice, b->stream));
  EHIP_HIP_TRY(hipStreamSynchronize(b->stream));
  return EHIP_OK;
}

// Iterations between the mixed form's refinement points in the multi-system
// launches (EHIP_BATCH_REFINE_EVERY, 1 to 8).
static int refine_every() {
  const char *v = std::getenv("EHIP_BATCH_REFINE_EVERY");
  const int k = v ? std::atoi(v) : kRefineEvery;
  return k < 1 ? 1 : (k > 8 ? 8 : k);
}

static int engine_create(int capacity, const ehip_params *p, ehip_batch **out) {
  ehip_batch *b = new ehip_batch();
  EHIP_HIP_TRY(hipGetDevice(&b->dev));
  EHIP_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  b->p = *p;
  b->capacity = capacity;
  b->slot.resize(capacity);
  b->mixed = p->precision == EHIP_MIXED_IR;
  b->method = p->method;
  b->fast = !b->mixed && p->mode == EHIP_FAST;
  b->grid = sched_grid(b->mixed, b->method == EHIP_CG ? 0 : 1, b->fast);
  b->refine_every = refine_every();
  EHIP_HIP_TRY(hipMalloc(&b->nchunks, sizeof(int)));
  EHIP_HIP_TRY(hipMalloc(&b->slots, sizeof(SlotPtrs) * capacity));
  EHIP_HIP_TRY(hipMalloc(&b->st, sizeof(BSys) * capacity));
  EHIP_HIP_TRY(hipMalloc(&b->active, sizeof(int)));
  EHIP_HIP_TRY(hipMalloc(&b->bnorm, sizeof(double) * capacity));
  EHIP_HIP_TRY(hipHostMalloc((void **)&b->h_active, sizeof(int), hipHostMallocDefault));
  EHIP_HIP_TRY(hipHostMalloc((void **)&b->h_st, sizeof(BSys) * capacity, hipHostMallocDefault));
  std::vector<BSys> zero(capacity);
  std::memset(zero.data(), 0, sizeof(BSys) * capacity);
  for (auto &z : zero) { z.phase = PH_IDLE; z.done = 1; }
  EHIP_HIP_TRY(hipMemcpy(b->st, zero.data(), sizeof(BSys) * capacity, hipMemcpyHostToDevice));
  const int zi = 0;
  EHIP_HIP_TRY(hipMemcpy(b->active, &zi, sizeof(int), hipMemcpyHostToDevice));
  *out = b;
  return upload_worklist(b);
}

static int64_t slot_bytes(const ehip_system *s, bool mixed) {
  int64_t copies = 0;  // slice-shared offset value copies not allocated yet
  if (s->sho_ok && s->sho_d == 1 && s->fmt64 == 2 && (!mixed || s->fmt32 == 2))
    copies = (s->sho64 ? 0 : 8 * s->sho_nnz) + (mixed && !s->sho32 ? 4 * s->sho_nnz : 0);
  return (int64_t)s->n * (8 * (NV64 + 1) + (mixed ? 4 * (NV32 + 1) : 0)) + (mixed && !s->vals32 ? 4 * s->nnz : 0) +
         copies;
}

// Scales the system, prepares its vectors and state, and puts it into a free
// slot.  Returns the slot or a negative status.
static int engine_admit(ehip_batch *b, ehip_system *s) {
  int slot = -1;
  for (int i = 0; i < b->capacity; ++i)
    if (!b->slot[i].used) { slot = i; break; }
  if (slot < 0) {
    set_error("batch: no free slot (capacity %d)", b->capacity);
    return EHIP_ECAP;
  }
  if (!s->has_values) {
    set_error("batch: system has no values");
    return EHIP_ESTATE;
  }
  auto &sl = b->slot[slot];
  // an assembly still running is not waited for: its degenerate-element
  // count is read when the system retires
  bool asm_later = false;
  if (s->asm_pending) {
    if (hipStreamQuery(s->stream) == hipSuccess) {
      if (int rc0 = assembly_check(s)) return rc0;
    } else {
      asm_later = true;
    }
  }
  const int n = s->n;
  int rc;
  if ((rc = ensure_state(s))) return rc;
  // unscaled b kept for the report residual
  const size_t n64 = (size_t)n * (NV64 + 1), n32 = b->mixed ? (size_t)n * (NV32 + 1) : 0;
  if (s->bw64_n < n64) {
    if (s->bw64) EHIP_HIP_TRY(hipFree(s->bw64));
    s->bw64 = nullptr;
    s->bw64_n = 0;
    EHIP_HIP_TRY(hipMalloc(&s->bw64, sizeof(double) * n64));
    s->bw64_n = n64;
  }
  if (s->bw32_n < n32) {
    if (s->bw32) EHIP_HIP_TRY(hipFree(s->bw32));
    s->bw32 = nullptr;
    s->bw32_n = 0;
    EHIP_HIP_TRY(hipMalloc(&s->bw32, sizeof(float) * n32));
    s->bw32_n = n32;
  }
  sl.w64 = s->bw64;
  sl.w32 = b->mixed ? s->bw32 : nullptr;
  sl.bytes = slot_bytes(s, b->mixed);
  double *b0 = sl.w64 + (size_t)NV64 * n;
  EHIP_HIP_TRY(hipMemcpyAsync(b0, s->rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
  // x = 0 initial guess -> 1e-8 after scaling (IterSolve.src:358)
  EHIP_HIP_TRY(hipMemsetAsync(s->x, 0, sizeof(double) * n, s->stream));
  if (b->p.scaling) {
    if ((rc = scale_system(s, s->rhs, true, s->x, &sl.bscale, b->mixed != 0, b->bnorm + slot))) return rc;
  } else if (b->mixed) {
    if (!s->vals32)
      if ((rc = dmalloc(s, (void **)&s->vals32, sizeof(float) * s->nnz))) return rc;
    hipLaunchKernelGGL(k_to_f32, dim3(2048), dim3(256), 0, s->stream, s->nnz, s->vals, s->vals32);
    touch_vals(s, false, true);
  }
  hipLaunchKernelGGL(k_fill_x, dim3(vec_grid(n)), dim3(kVecBlock), 0, s->stream, n, s->x, 1.0e-8);
  // the slice-shared offset copies of the scaled values, when the system
  // uses that layout (gathered on the system's stream, as the scaling)
  sl.sho = s->sho_ok && s->sho_d == 1 && s->fmt64 == 2 && (!b->mixed || s->fmt32 == 2);
  if (sl.sho && (!sho_values64(s) || (b->mixed && !sho_values32(s)))) return EHIP_ENOMEM;
  // Jacobi divisors
  const int nvi = b->method == EHIP_CG ? 6 : 10;
  double *dg64 = b->mixed ? sl.w64 + (size_t)n : sl.w64 + (size_t)(b->method == EHIP_CG ? 5 : 8) * n;
  float *dg32 = b->mixed ? sl.w32 + (size_t)(b->method == EHIP_CG ? 5 : 8) * n : nullptr;
  (void)nvi;
  hipLaunchKernelGGL(bk_dg, dim3(vec_grid(n)), dim3(kVecBlock), 0, s->stream, n, s->diag, s->vals, dg64, dg32,
                     b->p.precond);
  EHIP_HIP_TRY(hipGetLastError());
  // the engine's stream waits for this preparation on the system's stream
  // (no host wait: a replay may be running)
  if (!b->ev_admit) EHIP_HIP_TRY(hipEventCreateWithFlags(&b->ev_admit, hipEventDisableTiming));
  EHIP_HIP_TRY(hipEventRecord(b->ev_admit, s->stream));
  EHIP_HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_admit, 0));
  // slot state
  BSys z;
  std::memset(&z, 0, sizeof(z));
  z.tol = b->p.tol;
  z.tol_u = (b->p.true_residual_stop && b->p.scaling) ? b->p.tol : 0.0;
  z.res_u = -1.0;
  z.maxit = b->p.maxit;
  // CG: residual drop between replacements; BiCGStab: restarted inner tolerance
  z.itol = (float)(b->p.inner_tol > 0 ? b->p.inner_tol : 1e-2);
  z.inner_maxit = b->p.inner_maxit > 0 ? b->p.inner_maxit : 1000;
  z.phase = b->mixed ? PH_RES : PH_INIT;
  z.omega = 1.0;
  hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(1), 0, b->stream, &b->st[slot], z);
  sl.s = s;
  sl.used = true;
  sl.asm_check = asm_later;
  sl.first_launch = b->nlaunch + 1;
  b->dirty = true;
  sl.nb = sched_nb(b->mixed, b->method == EHIP_CG ? 0 : 1, b->fast, size_class(s->nnz));
  sl.ck = make_chunks(s, sl.nb);
  b->bytes += sl.bytes;
  // the active counter counts this slot
  hipLaunchKernelGGL(k_inc, dim3(1), dim3(1), 0, b->stream, b->active);
  EHIP_HIP_TRY(hipGetLastError());
  return slot;
}

// Back-scales the slot's system, fills its report, frees the slot.
static int engine_retire(ehip_batch *b, int slot, ehip_report *r) {
  auto &sl = b->slot[slot];
  ehip_system *s = sl.s;
  const BSys &h = b->h_st[slot];
  ehip_report rep{};
  rep.info = h.info;
  rep.iters = b->mixed ? h.inner_total : h.iter;
  rep.outer_iters = b->mixed ? h.outer : h.restarts;
  rep.res_scaled = h.res;
  int rc;
  if (b->p.scaling)
    if ((rc = backscale_system(s, s->rhs, s->x, 0.0, b->bnorm + slot))) return rc;
  // unscaled FP64 residual and ComputeNorm
  double *b0 = sl.w64 + (size_t)NV64 * s->n;
  double rt = 0, nrm = 0;
  if ((rc = ehip_internal_report(s, b0, &rt, &nrm))) return rc;
  rep.res_true = rt;
  rep.norm = nrm;
  if (sl.asm_check && assembly_check(s) != EHIP_OK) rep.info = EHIP_HUTI_DIVERGENCE;  // degenerate elements
  s->has_x = true;
  if (r) *r = rep;
  b->bytes -= sl.bytes;  // the workspace stays with the system for its next admission
  sl = ehip_batch::Slot();
  return EHIP_OK;
}

// Replays the iteration graph once (graph_iters iterations of every slot),
// or launches the same kernels one by one when graphs are off.
static int engine_launch(ehip_batch *b) {
  const int iters = b->p.check_every > 0 ? b->p.check_every : 8;
  if (graphs_enabled()) {
    if (!b->graph_ok) {
      int rc = build_graph(b, iters);
      if (rc) return rc;
    }
    EHIP_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  } else {
    for (int k = 0; k < iters; ++k) enqueue_iteration(b, k % b->refine_every == 0);
    EHIP_HIP_TRY(take_launch_error());
  }
  EHIP_HIP_TRY(hipMemcpyAsync(b->h_active, b->active, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  EHIP_HIP_TRY(hipMemcpyAsync(b->h_st, b->st, sizeof(BSys) * b->capacity, hipMemcpyDeviceToHost, b->stream));
  ++b->nlaunch;
  return EHIP_OK;
}

static int engine_step(ehip_batch *b) {
  int rc = engine_launch(b);
  if (rc) return rc;
  EHIP_HIP_TRY(hipStreamSynchronize(b->stream));
  return EHIP_OK;
}

// Schedule search over the admitted slots: for every size class present
// (smallest first) each chunk size is timed over `replays` replays of the
// iteration graph from the admitted state, the others held at their current
// choice; then the workgroup count.  Between candidates every slot goes back
// to the admitted state (its Krylov state and x; everything else is
// recomputed from them), and after the search the slots are left there with
// the chosen schedule, which the table keeps for later batches and servers.
// retune = 0 searches only what the table does not hold yet.
static int engine_tune(ehip_batch *b, int replays, bool retune) {
  const int mi = b->method == EHIP_CG ? 0 : 1;
  std::vector<int> used;
  int64_t xn = 0;
  for (int i = 0; i < b->capacity; ++i)
    if (b->slot[i].used) {
      used.push_back(i);
      xn += b->slot[i].s->n;
    }
  if (used.empty()) return EHIP_OK;
  bool todo[kSizeClasses] = {};
  bool any = false;
  {
    std::lock_guard<std::mutex> lk(g_sched.mu);
    for (int i : used) {
      const int c = size_class(b->slot[i].s->nnz);
      if (retune || !g_sched.nb[b->mixed][mi][b->fast][c]) todo[c] = any = true;
    }
    if (retune || !g_sched.grid[b->mixed][mi][b->fast]) any = true;
  }
  if (!any) return EHIP_OK;
  // the admitted state
  EHIP_HIP_TRY(hipStreamSynchronize(b->stream));
  std::vector<BSys> st0(b->capacity);
  int active0 = 0;
  EHIP_HIP_TRY(hipMemcpy(st0.data(), b->st, sizeof(BSys) * b->capacity, hipMemcpyDeviceToHost));
  EHIP_HIP_TRY(hipMemcpy(&active0, b->active, sizeof(int), hipMemcpyDeviceToHost));
  double *x0 = nullptr;
  EHIP_HIP_TRY(hipMalloc(&x0, sizeof(double) * xn));
  std::vector<int64_t> xoff(b->capacity, 0);
  {
    int64_t o = 0;
    for (int i : used) {
      xoff[i] = o;
      const ehip_system *s = b->slot[i].s;
      EHIP_HIP_TRY(hipMemcpy(x0 + o, s->x, sizeof(double) * s->n, hipMemcpyDeviceToDevice));
      o += s->n;
    }
  }
  auto restore = [&]() -> int {
    EHIP_HIP_TRY(hipStreamSynchronize(b->stream));
    for (int i : used) {
      auto &sl = b->slot[i];
      EHIP_HIP_TRY(hipMemcpyAsync(sl.s->x, x0 + xoff[i], sizeof(double) * sl.s->n, hipMemcpyDeviceToDevice,
                                  b->stream));
      sl.ck = make_chunks(sl.s, sl.nb);
    }
    EHIP_HIP_TRY(hipMemcpyAsync(b->st, st0.data(), sizeof(BSys) * b->capacity, hipMemcpyHostToDevice, b->stream));
    EHIP_HIP_TRY(hipMemcpyAsync(b->active, &active0, sizeof(int), hipMemcpyHostToDevice, b->stream));
    b->graph_ok = false;
    return upload_worklist(b);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EHIP_HIP_TRY(hipEventCreate(&e0));
  EHIP_HIP_TRY(hipEventCreate(&e1));
  auto timed = [&](double *ms) -> int {
    int rc;
    if ((rc = restore())) return rc;
    if (graphs_enabled() && (rc = build_graph(b, b->p.check_every > 0 ? b->p.check_every : 8))) return rc;
    EHIP_HIP_TRY(hipEventRecord(e0, b->stream));
    for (int k = 0; k < replays; ++k)
      if ((rc = engine_step(b))) return rc;
    EHIP_HIP_TRY(hipEventRecord(e1, b->stream));
    EHIP_HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    EHIP_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    return EHIP_OK;
  };
  int rc = EHIP_OK;
  for (int c = 0; c < kSizeClasses && rc == EHIP_OK; ++c) {
    if (!todo[c]) continue;
    double best = 0;
    int pick = kDefaultNB;
    for (int nb : kChunkSizes) {
      for (int i : used)
        if (size_class(b->slot[i].s->nnz) == c) b->slot[i].nb = nb;
      double ms = 0;
      if ((rc = timed(&ms))) break;
      if (nb == kChunkSizes[0] || ms < best) { best = ms; pick = nb; }
    }
    for (int i : used)
      if (size_class(b->slot[i].s->nnz) == c) b->slot[i].nb = pick;
    std::lock_guard<std::mutex> lk(g_sched.mu);
    if (rc == EHIP_OK) g_sched.nb[b->mixed][mi][b->fast][c] = pick;
  }
  bool do_grid;
  {
    std::lock_guard<std::mutex> lk(g_sched.mu);
    do_grid = retune || !g_sched.grid[b->mixed][mi][b->fast];
  }
  if (rc == EHIP_OK && do_grid) {
    double best = 0;
    int pick = b->grid;
    for (int g : kGrids) {
      b->grid = g;
      double ms = 0;
      if ((rc = timed(&ms))) break;
      if (g == kGrids[0] || ms < best) { best = ms; pick = g; }
    }
    b->grid = pick;
    std::lock_guard<std::mutex> lk(g_sched.mu);
    if (rc == EHIP_OK) g_sched.grid[b->mixed][mi][b->fast] = pick;
  }
  if (rc == EHIP_OK) rc = restore();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(x0);
  return rc;
}

static void lane_stop(ehip_batch *b) {
  Lane *L = b->lane;
  if (!L) return;
  {
    std::lock_guard<std::mutex> lk(L->mu);
    L->stop = true;
  }
  L->cv_work.notify_all();
  for (auto &t : L->th) t.join();
  delete L;
  b->lane = nullptr;
}

static void lane_worker(Lane *L) {
  for (;;) {
    ehip_system *s = nullptr;
    {
      std::unique_lock<std::mutex> lk(L->mu);
      L->cv_work.wait(lk, [&] { return L->stop || !L->q.empty(); });
      if (L->q.empty()) return;  // stop requested and nothing left
      s = L->q.front();
      L->q.pop_front();
    }
    ehip_report r{};
    if (hipSetDevice(s->dev) != hipSuccess || ehip_solve(s, &L->p, nullptr, &r, nullptr) != EHIP_OK) r.info = 3;
    {
      std::lock_guard<std::mutex> lk(L->mu);
      L->done.push_back({s, r});
    }
    L->cv_done.notify_all();
  }
}

// workspace a single-system mixed or FP64 solve keeps on the device
// (plus the format copies of the values it has not allocated yet)
static int64_t lane_bytes(const ehip_system *s, bool mixed) {
  const int64_t n = s->n, nnz = s->nnz;
  int64_t copies = 0;
  if (s->fmt64 == 2 && s->sho_ok && !s->sho64) copies += 8 * s->sho_nnz;
  if (s->fmt64 == 1 && s->sell_ok && !s->sell64) copies += 8 * (s->sell2_ok ? s->sell2_nnz : s->sell_nnz);
  if (!mixed) return n * 8 * 12 + copies;
  if (s->fmt32 == 2 && s->sho_ok && !s->sho32) copies += 4 * s->sho_nnz;
  if (s->fmt32 == 1 && s->sell_ok && !s->sell32) copies += 4 * s->sell_nnz;
  return n * (4 * 8 + 8 * 2) + (s->vals32 ? 0 : 4 * nnz) + copies;
}

static void engine_free(ehip_batch *b) {
  if (!b) return;
  lane_stop(b);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->ev_admit) (void)hipEventDestroy(b->ev_admit);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void *ptrs[] = {b->chunks, b->nchunks, b->slots, b->st, b->p64, b->p32, b->active, b->bnorm};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->h_active) (void)hipHostFree(b->h_active);
  if (b->h_st) (void)hipHostFree(b->h_st);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace ehip

// ================================================================ C ABI: batching

extern "C" int ehip_batch_create(int nsys, ehip_system **systems, ehip_batch **out) {
  if (nsys <= 0 || !systems || !out) return EHIP_EARG;
  for (int i = 0; i < nsys; ++i)
    if (!systems[i]) return EHIP_EARG;
  // the engine is created per solve (its graph depends on the method and
  // precision); the batch handle records the member systems
  ehip_batch *b = new ehip_batch();
  b->capacity = nsys;
  b->batch_members.resize(nsys);
  b->slot.resize(nsys);
  for (int i = 0; i < nsys; ++i) b->slot[i].s = systems[i];
  *out = b;
  return EHIP_OK;
}

extern "C" int ehip_batch_free(ehip_batch *b) {
  if (!b) return EHIP_OK;
  delete b;
  return EHIP_OK;
}

// Systems of at least lane_min_dofs unknowns do not join the multi-system
// launches: an iteration of one of them already fills the GPU, and in a
// shared launch sequence every op would also walk their chunks for the
// phases they are not in.  They are solved whole by ehip_solve on
// lane_workers threads (largest first), each on the system's own stream,
// concurrently with the multi-system launches of the small ones; each such
// result is exactly the system's ehip_solve result.
extern "C" int ehip_batch_solve(ehip_batch *bh, const ehip_params *p, ehip_report *reports, double *t_ms) {
  if (!bh || !p) return EHIP_EARG;
  if (p->precond >= EHIP_PRECOND_ILU0) {
    set_error("ehip_batch_solve: the multi-system engine takes no ILU preconditioning");
    return EHIP_EUNSUP;
  }
  const int nsys = bh->capacity;
  std::vector<ehip_system *> sys(nsys);
  for (int i = 0; i < nsys; ++i) sys[i] = bh->slot[i].s;
  std::vector<int> lane_idx, eng_idx;
  for (int i = 0; i < nsys; ++i)
    (bh->lane_workers > 0 && sys[i]->n >= bh->lane_min_dofs ? lane_idx : eng_idx).push_back(i);
  for (int i : lane_idx)
    if (!sys[i]->has_values) {
      set_error("batch: system has no values");
      return EHIP_ESTATE;
    }
  ehip_batch *e = nullptr;
  int rc = engine_create(std::max(1, (int)eng_idx.size()), p, &e);
  if (rc) return rc;
  hipEvent_t e0, e1;
  EHIP_HIP_TRY(hipEventCreate(&e0));
  EHIP_HIP_TRY(hipEventCreate(&e1));
  auto t0 = std::chrono::steady_clock::now();
  Lane *L = nullptr;
  if (!lane_idx.empty()) {
    std::sort(lane_idx.begin(), lane_idx.end(), [&](int a, int b) { return sys[a]->n > sys[b]->n; });
    L = new Lane();
    L->min_dofs = bh->lane_min_dofs;
    L->workers = std::min(bh->lane_workers, (int)lane_idx.size());
    L->p = *p;
    L->p.record_history = 0;
    L->p.profile_every = 0;
    for (int i : lane_idx) {
      L->q.push_back(sys[i]);
      L->running++;
      bh->batch_members[i] = -1;
    }
    e->lane = L;  // engine_free joins the workers
    for (int w = 0; w < L->workers; ++w) L->th.emplace_back(lane_worker, L);
  }
  for (int i : eng_idx) {
    int slot = engine_admit(e, sys[i]);
    if (slot < 0) { engine_free(e); return slot; }
    bh->batch_members[i] = slot;
  }
  if ((rc = upload_worklist(e))) { engine_free(e); return rc; }
  if (bh->tuning && !eng_idx.empty())
    if ((rc = engine_tune(e, bh->tune_replays, bh->tuning == 2))) { engine_free(e); return rc; }
  EHIP_HIP_TRY(hipEventRecord(e0, e->stream));
  if (!eng_idx.empty()) {
    // every replay runs graph_iters iterations of every live slot
    const int gi = p->check_every > 0 ? p->check_every : 8;
    // mixed: up to maxit inner steps, one refinement per step at worst, each
    // waiting up to refine_every - 1 idle iterations
    const int64_t cap = (int64_t)p->maxit * (p->precision == EHIP_MIXED_IR ? 1 + e->refine_every : 1) / gi + 64;
    int64_t replays = 0;
    for (;;) {
      if ((rc = engine_step(e))) { engine_free(e); return rc; }
      ++replays;
      if (*e->h_active == 0 || replays > cap) break;
    }
  }
  EHIP_HIP_TRY(hipEventRecord(e1, e->stream));
  EHIP_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  std::map<ehip_system *, ehip_report> lane_rep;
  if (L) {
    std::unique_lock<std::mutex> lk(L->mu);
    L->cv_done.wait(lk, [&] { return L->done.size() == lane_idx.size(); });
    for (auto &d : L->done) lane_rep[d.first] = d.second;
  }
  const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < nsys; ++i) {
    const int slot = bh->batch_members[i];
    ehip_report r{};
    if (slot < 0) {
      r = lane_rep[sys[i]];
    } else {
      if (!e->h_st[slot].done) {
        BSys &h = e->h_st[slot];
        h.info = 2;  // replay cap reached before the slot finished
      }
      if ((rc = engine_retire(e, slot, &r))) { engine_free(e); return rc; }
      r.t_krylov_ms = ms;
      r.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (reports) reports[i] = r;
  }
  if (t_ms) *t_ms = L ? std::max((double)ms, wall) : ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  engine_free(e);
  return EHIP_OK;
}

extern "C" int ehip_batch_set_lane(ehip_batch *b, int64_t min_dofs, int workers) {
  if (!b || workers < 0 || workers > 64 || min_dofs < 0) return EHIP_EARG;
  b->lane_min_dofs = min_dofs;
  b->lane_workers = workers;
  return EHIP_OK;
}

extern "C" int ehip_batch_set_tuning(ehip_batch *b, int mode, int replays) {
  if (!b || mode < 0 || mode > 2 || replays < 0 || replays > 64) return EHIP_EARG;
  b->tuning = mode;
  b->tune_replays = replays > 0 ? replays : 3;
  return EHIP_OK;
}

static int sched_key(const ehip_params *p, int *mixed, int *mi, int *fast) {
  if (!p || (p->method != EHIP_CG && p->method != EHIP_BICGSTAB)) return EHIP_EARG;
  *mixed = p->precision == EHIP_MIXED_IR;
  *mi = p->method == EHIP_CG ? 0 : 1;
  *fast = !*mixed && p->mode == EHIP_FAST;
  return EHIP_OK;
}

extern "C" int ehip_batch_get_schedule(const ehip_params *p, int64_t nnz, int *chunk_nnz, int *grid) {
  int mixed, mi, fast;
  if (nnz < 0 || sched_key(p, &mixed, &mi, &fast)) return EHIP_EARG;
  if (chunk_nnz) *chunk_nnz = sched_nb(mixed, mi, fast, size_class(nnz));
  if (grid) *grid = sched_grid(mixed, mi, fast);
  return EHIP_OK;
}

extern "C" int ehip_batch_set_schedule(const ehip_params *p, int64_t nnz, int chunk_nnz, int grid) {
  int mixed, mi, fast;
  if (nnz < 0 || sched_key(p, &mixed, &mi, &fast)) return EHIP_EARG;
  bool nb_ok = chunk_nnz == 0 || chunk_nnz == -1;
  for (int c : kChunkSizes) nb_ok = nb_ok || chunk_nnz == c;
  if (!nb_ok || grid < -1 || grid > 65536 || (grid > 0 && grid % 8)) {
    set_error("ehip_batch_set_schedule: chunk size must be -1, 0, 1024, 2048 or 4096 and the grid -1, 0 or a "
              "multiple of 8 up to 65536");
    return EHIP_EARG;
  }
  std::lock_guard<std::mutex> lk(g_sched.mu);
  if (chunk_nnz >= 0) g_sched.nb[mixed][mi][fast][size_class(nnz)] = chunk_nnz;
  if (grid >= 0) g_sched.grid[mixed][mi][fast] = grid;
  return EHIP_OK;
}

extern "C" int ehip_batch_get_solution(ehip_batch *b, int i, double *x) {
  if (!b || i < 0 || i >= b->capacity || !x) return EHIP_EARG;
  return ehip_get_solution(b->slot[i].s, x);
}

// ================================================================ C ABI: serving

extern "C" int ehip_server_create(const ehip_params *p, int max_slots, int64_t mem_budget, ehip_batch **out) {
  if (!p || max_slots <= 0 || !out) return EHIP_EARG;
  if (p->precond >= EHIP_PRECOND_ILU0) {
    set_error("ehip_server_create: the multi-system engine takes no ILU preconditioning");
    return EHIP_EUNSUP;
  }
  ehip_batch *b = nullptr;
  int rc = engine_create(max_slots, p, &b);
  if (rc) return rc;
  b->chunk_cap = 0;
  b->bytes = 0;
  b->graph_ok = false;
  b->p.divtol = 0;
  b->budget = mem_budget;
  *out = b;
  return EHIP_OK;
}

// Admits a system into a free slot if a slot and the HBM budget allow;
// EHIP_ECAP otherwise (the caller keeps it queued).
extern "C" int ehip_server_submit(ehip_batch *b, ehip_system *s, int *slot_out) {
  if (!b || !s) return EHIP_EARG;
  const int64_t budget = b->budget;
  Lane *L = b->lane;
  if (L && s->n >= L->min_dofs) {
    if (!s->has_values) {
      set_error("server: the system has no values");
      return EHIP_ESTATE;
    }
    const int64_t need = lane_bytes(s, b->mixed != 0);
    std::lock_guard<std::mutex> lk(L->mu);
    if (L->running >= L->workers * kLaneDepth) {
      set_error("server: the concurrent lane's queue is full");
      return EHIP_ECAP;
    }
    if (budget > 0 && b->bytes + L->bytes + need > budget) {
      set_error("server: HBM budget %lld bytes would be exceeded", (long long)budget);
      return EHIP_ECAP;
    }
    L->q.push_back(s);
    L->need[s] = need;
    L->bytes += need;
    L->running++;
    L->cv_work.notify_one();
    if (slot_out) *slot_out = -1;
    return EHIP_OK;
  }
  const int64_t need = slot_bytes(s, b->mixed != 0);
  const int64_t lane_b = L ? L->bytes : 0;
  if (budget > 0 && b->bytes + lane_b + need > budget) {
    set_error("server: HBM budget %lld bytes would be exceeded", (long long)budget);
    return EHIP_ECAP;
  }
  int slot = engine_admit(b, s);
  if (slot < 0) return slot;
  if (slot_out) *slot_out = slot;
  return EHIP_OK;
}

// Runs one graph replay (graph_iters iterations of every live system), then
// retires the finished ones: their systems get back-scaled solutions and
// reports.  done_systems/done_reports receive up to max_done entries.
// Collects finished concurrent-lane solves into the output arrays; with
// `wait`, blocks up to wait_ms for one when none is ready.
static void lane_collect(ehip_batch *b, int *ndone, ehip_system **done_systems, ehip_report *done_reports,
                         int max_done, bool wait, int wait_ms) {
  Lane *L = b->lane;
  if (!L) return;
  std::unique_lock<std::mutex> lk(L->mu);
  if (wait && L->done.empty() && L->running > 0)
    L->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !L->done.empty(); });
  size_t k = 0;
  for (; k < L->done.size() && *ndone < max_done; ++k) {
    auto &d = L->done[k];
    if (done_systems) done_systems[*ndone] = d.first;
    if (done_reports) done_reports[*ndone] = d.second;
    (*ndone)++;
    L->running--;
    L->bytes -= L->need[d.first];
    L->need.erase(d.first);
  }
  L->done.erase(L->done.begin(), L->done.begin() + (long)k);
}

// Pipelined: a step first takes the results of the replay the previous step
// launched (waiting for it), retires the finished systems, rebuilds the
// worklist if systems came or went, and launches the next replay without
// waiting for it -- so the caller's admissions (assembly, Dirichlet rows,
// scaling on each system's own stream) overlap the running replay.  A
// system admitted while a replay runs joins the next one.
extern "C" int ehip_server_step(ehip_batch *b, int *ndone, ehip_system **done_systems, ehip_report *done_reports,
                                int max_done) {
  if (!b || !ndone) return EHIP_EARG;
  *ndone = 0;
  int rc;
  bool changed = false;
  if (b->in_flight) {
    EHIP_HIP_TRY(hipStreamSynchronize(b->stream));
    b->in_flight = false;
    for (int i = 0; i < b->capacity; ++i) {
      auto &sl = b->slot[i];
      // results of replay nlaunch: only systems that were in it
      if (!sl.used || sl.first_launch > b->nlaunch || !b->h_st[i].done) continue;
      if (*ndone >= max_done) break;
      ehip_system *s = sl.s;
      ehip_report r{};
      if ((rc = engine_retire(b, i, &r))) return rc;
      if (done_systems) done_systems[*ndone] = s;
      if (done_reports) done_reports[*ndone] = r;
      (*ndone)++;
      changed = true;
    }
  }
  int live = 0;
  for (auto &sl : b->slot) live += sl.used;
  if (changed || b->dirty) {
    if ((rc = upload_worklist(b))) return rc;
    b->dirty = false;
  }
  if (live) {
    if ((rc = engine_launch(b))) return rc;
    b->in_flight = true;
    lane_collect(b, ndone, done_systems, done_reports, max_done, false, 0);
  } else {
    // nothing batched: wait briefly for a concurrent-lane solve instead of spinning
    lane_collect(b, ndone, done_systems, done_reports, max_done, *ndone == 0, 2);
  }
  return EHIP_OK;
}

extern "C" int ehip_server_status(const ehip_batch *b, int *live, int64_t *bytes) {
  if (!b) return EHIP_EARG;
  int l = 0;
  for (auto &sl : b->slot) l += sl.used;
  int64_t by = b->bytes;
  if (b->lane) {
    std::lock_guard<std::mutex> lk(b->lane->mu);
    l += b->lane->running;
    by += b->lane->bytes;
  }
  if (live) *live = l;
  if (bytes) *bytes = by;
  return EHIP_OK;
}

extern "C" int ehip_server_set_lane(ehip_batch *b, int64_t min_dofs, int workers) {
  if (!b || workers < 0 || workers > 64 || min_dofs < 0) return EHIP_EARG;
  if (b->lane) {
    std::lock_guard<std::mutex> lk(b->lane->mu);
    if (b->lane->running) {
      set_error("server: the concurrent lane is busy");
      return EHIP_ESTATE;
    }
  }
  lane_stop(b);
  if (workers == 0) return EHIP_OK;
  Lane *L = new Lane();
  L->min_dofs = min_dofs;
  L->workers = workers;
  L->p = b->p;
  L->p.mode = EHIP_FAST;
  L->p.record_history = 0;
  L->p.profile_every = 0;
  for (int i = 0; i < workers; ++i) L->th.emplace_back(lane_worker, L);
  b->lane = L;
  return EHIP_OK;
}

extern "C" int ehip_server_free(ehip_batch *b) {
  engine_free(b);
  return EHIP_OK;
}
