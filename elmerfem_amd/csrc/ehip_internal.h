// Internal declarations of libehip (MI355X / gfx950).  Not part of the ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/ehip.h"

namespace ehip {

void set_error(const char *fmt, ...);

#define EHIP_HIP_TRY(call)                                                          \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::ehip::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return EHIP_EHIP;                                                             \
    }                                                                               \
  } while (0)

// SpMV schedule: rows are cut into chunks of at most `nnz_per_block`
// nonzeros (a longer row gets a chunk of its own).  One workgroup streams a
// chunk's values/cols coalesced into LDS as products, then one lane per row
// sums its row left to right (the reference's summation order).
struct SpmvPlan {
  int block = 256;          // workgroup size
  int nb = 2048;            // nnz per chunk (LDS capacity in elements)
  int nchunks = 0;
  int grid = 0;             // launched workgroups (multiple of 8, XCD-aware)
  int *chunk_row = nullptr; // device, nchunks+1 (0-based first row of each chunk)
};

// The SpMV schedules (workgroup size, nonzeros per chunk) compiled and
// offered to the autotuner.
#define EHIP_SPMV_CONFIGS(X) X(256, 2048) X(256, 1024) X(512, 2048)

// Per-system partial-sum buffer (entries): SpMV epilogues publish at
// [0, 4 * kEpiStride), vector kernels whose producer's partials are still
// being read in the same launch publish at [kVecPartOff, kPartials).
constexpr int kPartials = 16 * 2048;
constexpr int kVecPartOff = 8 * 2048;
// Stride of an SpMV epilogue's partials (value k of workgroup b at
// k * kEpiStride + b): at least the largest SpMV grid (sell_grid, plan
// grids), and 4 values fit below kVecPartOff.
constexpr int kEpiStride = 4096;
static_assert(4 * kEpiStride <= kVecPartOff, "epilogue partials overlap the vector-step partials");

// Device-resident Krylov scalars.  Kernels publish per-workgroup partial
// sums; a one-workgroup finalize launch adds them in a fixed order and
// updates these scalars, so results are run-to-run deterministic.
// ru: reliable-update threshold of the mixed-precision CG (0 = off): the
// FP32 iteration stops for an FP64 residual replacement once its recurrence
// residual falls below it.
// tol_u: tolerance of the unscaled residual ||b - A x|| / ||b|| (0 = not
// tested), evaluated on the scaled system as ||r_s / D|| / ||b_s / D||;
// bb_u = ||b_s / D||^2; res_u: the last such residual.
template <typename T>
struct KStateT {
  T rho, rho_new, alpha, beta, omega, pq, rr, rhsnorm, res, ts, tt, snorm, res_true, scale, ru;
  int iter, info, done, maxit, snorm_exit, rec_conv, pad0, pad1;
  T tol, divtol, tol_u, bb_u, res_u;
  unsigned int ticket[4];
  // folded FP32 CG steps: rho of the last vector step, carried from the SpMV
  // step that sums it to the vector step that uses it; fresh = 1 when rho and
  // beta are current (start, residual replacement, end of a block)
  T rho_next;
  int fresh;
  // a folded SpMV step's stop, made `done` by the vector step that follows
  int stop_pending;
};
using KState = KStateT<double>;
using KState32 = KStateT<float>;

struct ehip_system_impl;

}  // namespace ehip

struct ehip_system {
  int dev = 0;
  hipStream_t stream = nullptr;
  int n = 0;
  int64_t nnz = 0;
  int dofs = 1;
  int nnodes = 0;
  // host copy of the pattern, 1-based like Matrix_t
  std::vector<int> h_rows, h_cols, h_diag, h_perm;
  int bw_before = 0, bw_after = 0, bw_accepted = 0;
  // device CRS, 0-based
  int *rows = nullptr, *cols = nullptr, *diag = nullptr;
  // compressed columns for SpMV (see ColsU16): 16-bit offsets from a base
  // per segment of 2048 nonzeros, when every segment's span fits
  unsigned short *cols16 = nullptr;
  int *seg_base = nullptr;
  int nseg = 0;
  bool c16 = false;
  // SELL-64 copy for SpMV (see spmv_sell_body): slice s holds slice rows
  // [64 s, 64 s + 64) column-major at [sell_off[s], sell_off[s+1]), padded to
  // its longest row.  Slice row i is row i, or sell_row[i] when rows were
  // sorted by length inside windows (sell_row[i] >= n: padding lane).  Built
  // at pattern upload when the padding stays small; its values are refreshed
  // from Values / the FP32 copy when those change.
  int nslice = 0;
  int *sell_row = nullptr;            // nslice * 64, or null
  int64_t sell_nnz = 0;
  int *sell_off = nullptr;            // nslice + 1
  int *sell_c32 = nullptr;            // sell_nnz column indices, or
  unsigned short *sell_c16 = nullptr; // 16-bit offsets from sell_base[slice]
  int *sell_base = nullptr;           // nslice
  // FP64 values use the pair-packed layout (spmv_sell2_body) when the 16-bit
  // offsets fit: slice s at [sell2_off[s], sell2_off[s+1]), with its own
  // offsets array; sell64 then holds sell2_nnz values in that layout
  int64_t sell2_nnz = 0;
  int *sell2_off = nullptr;
  unsigned short *sell2_c16 = nullptr;
  bool sell2_ok = false;
  // node-blocked SELL (spmv_sellb_body) when the pattern is made of whole
  // dofs x dofs blocks and a slice's node span fits 16 bits: sellb_d = dofs,
  // one offset per block in sellb_c16 (sell_nnz / dofs entries) from
  // sellb_base[slice]; the values then use the plain SELL layout
  int sellb_d = 0;
  unsigned short *sellb_c16 = nullptr;
  int *sellb_base = nullptr;
  int sell_sigma = 0;                 // window of the length sort (0: row order)
  int sell_u = 8;                     // nonzeros per lane in flight: kSellU or 9
  int sell2_u = 4;                    // pairs per lane in flight (pair-packed FP64): kSell2U or 7
  double *sell64 = nullptr;
  float *sell32 = nullptr;
  bool sell_ok = false;
  // SELL-64 with slice-shared column offsets (spmv_sho_body), built when
  // every slice's rows take their columns from one short list of offsets
  // from the row index (stencil-like patterns: meshes of regular topology
  // in natural node order): slice s holds |list| entries per lane at
  // [sho_off[s], sho_off[s+1]), the list itself at sho_tab[sho_off[s] / 64],
  // an absent entry holding value 0; bit k of sho_mask[row] says entry k is
  // one of the row's nonzeros (the row's nonzeros in CRS order)
  bool sho_ok = false;
  // the SELL-64 and 16-bit CRS column layouts are built when first selected,
  // tuned or queried when the shared-offset layout was built (the default)
  bool layouts_deferred = false;
  int sho_u = 8;                      // entries (node blocks) per lane in flight: 8 or 9
  // node-block size: 1, or D when the pattern is made of D x D node blocks
  // (offsets and mask bits per block; entry (k, d) of a lane at
  // sho_off[s] + 64 (k D + d) + lane, the list at sho_tab[sho_off[s] / (64 D)])
  int sho_d = 1;
  int64_t sho_nnz = 0;
  int *sho_off = nullptr;             // nslice + 1
  int *sho_tab = nullptr;             // sho_nnz / 64
  unsigned *sho_mask = nullptr;       // n
  double *sho64 = nullptr;
  float *sho32 = nullptr;
  uint64_t sho64_ver = 0, sho32_ver = 0;
  // workspace of the multi-system engine, kept across admissions (a
  // hipFree per retired request would synchronise the whole device)
  double *bw64 = nullptr;
  size_t bw64_n = 0;
  float *bw32 = nullptr;
  size_t bw32_n = 0;
  int fmt64 = 0, fmt32 = 0;           // SpMV format per precision: 0 CSR stream, 1 SELL-64,
                                      // 2 SELL-64 with shared offsets
  uint64_t vals_ver = 1, vals32_ver = 1, sell64_ver = 0, sell32_ver = 0;
  double *vals = nullptr, *rhs = nullptr, *x = nullptr;
  float *vals32 = nullptr;
  double *dscale = nullptr;  // ScaleLinearSystem's D (n)
  bool has_values = false, has_x = false;
  ehip::SpmvPlan plan64, plan32;
  bool tuned64 = false, tuned32 = false;
  // Krylov workspace
  double *work = nullptr;  // n * nwork doubles
  int nwork = 0;
  float *work32 = nullptr;
  int nwork32 = 0;
  double *partials = nullptr;  // per-block partial sums (4 * maxgrid), FP64
  float *partials32 = nullptr; // per-block partial sums (4 * maxgrid), FP32
  ehip::KState *st = nullptr;    // FP64 solver scalars
  ehip::KState32 *st32 = nullptr;// FP32 inner-solver scalars
  ehip::KState *h_st = nullptr;  // pinned mirror
  ehip::KState32 *h_st32 = nullptr;
  double *hist = nullptr;        // device residual history (hist_cap)
  int hist_cap = 0;
  // row-owner gather maps (from ehip_system_from_mesh)
  int nelem = 0, npe = 0;
  int *elements = nullptr;        // nelem*npe, 0-based node ids (device)
  // Built row-major (rhs_start: each row's entries, in element order), then
  // made slice-major for the gather (slice_maps): entry k of row 64 s + l
  // at map_base[s] + 64 k + l, each slice padded to its longest row with
  // code -1; rhs_start is freed then.
  int *rhs_start = nullptr;       // n+1 (row-major build only)
  int *map_base = nullptr;        // nslice+1 (slice-major)
  int *rhs_contrib = nullptr;     // per entry: e*ld + the row's local index (-1: padding)
  void *row_off = nullptr;        // per entry, ld column positions in the row (uint8 or uint16)
  int row_off_bytes = 0;          // 1 or 2
  int64_t map_entries = 0;        // entries of rhs_contrib
  double *emat = nullptr;         // nelem*ld*ld element matrices (row-major, or packed upper triangles)
  double *evec = nullptr;         // nelem*ld element force vectors
  // assembly inputs kept on the device between calls (no allocation, free or
  // table upload per call once the first call has sized them)
  double *asm_xyz = nullptr;      // nnodes*3, valid after the first call with coordinates
  double *asm_in = nullptr;       // coef, conv, load (nelem each), velo (3*nelem)
  double *asm_tab = nullptr;      // basis, dLBasis, weights of element type asm_code
  int asm_code = 0;
  int *asm_bad = nullptr;         // degenerate Gauss-point metrics of the last call
  int asm_pending = 0;            // that count is not checked yet (asynchronous call)
  // Dirichlet rows and values of the last ehip_dirichlet call (device flags
  // and values per row; host copies to recognise the same call again)
  signed char *bc_flag = nullptr;
  double *bc_val = nullptr;
  std::vector<int> h_bc_rows;
  std::vector<double> h_bc_vals;
  int64_t bytes = 0;
  bool pattern_on_device = false; // rows/cols/diag were built on the device (not uploaded)
  // sampled SpMV timing inside the Krylov loop
  int prof_every = 0, prof_count = 0, prof_n = 0;
  hipEvent_t prof_ev[2 * 512] = {};
  // events for the next SpMV launch when it is a sample (hipExtLaunchKernelGGL
  // timestamps the kernel itself), else null
  hipEvent_t prof_start = nullptr, prof_stop = nullptr;
  // recorded blocks of 1, 2, 4, ..., 32 FP32 CG iterations (mixed form),
  // valid while cg32_key (the buffers and SpMV schedule they bake in) holds
  hipGraphExec_t cg32_graph[6] = {};
  uint64_t cg32_key = 0;
  // ILU(k) preconditioner (ilu.hip): pattern order (-1: none), the 0-based
  // device pattern (the matrix's own arrays when k = 0) with the position
  // of every matrix nonzero in it (k > 0), the factors, rows grouped by
  // forward / backward dependency level (host offsets per level), and the
  // recorded sweeps of up to two (source, destination) vector pairs
  int ilu_order = -1;
  int64_t ilu_nnz = 0;
  int *ilu_rows = nullptr, *ilu_cols = nullptr, *ilu_diag = nullptr, *ilu_amap = nullptr;
  double *ilu = nullptr;
  int *ilu_lord = nullptr, *ilu_uord = nullptr;
  double *ilu_y = nullptr;             // n: the forward sweep's result
  int *ilu_sync = nullptr;             // the sweeps' two tickets and error flag
  std::vector<int> ilu_loff, ilu_uoff;
  hipGraphExec_t ilu_graph[2] = {};
  const double *ilu_graph_src[2] = {};
  double *ilu_graph_dst[2] = {};
};

namespace ehip {

// allocation helpers that account for device bytes per system
int dmalloc(ehip_system *s, void **p, size_t bytes);
void dfree(ehip_system *s, void *p, size_t bytes);

int ensure_work(ehip_system *s, int nvec);
int ensure_work32(ehip_system *s, int nvec);
int ensure_state(ehip_system *s);

// Values (or the FP32 copy) changed: SELL copies must be refreshed before use.
inline void touch_vals(ehip_system *s, bool v64, bool v32) {
  if (v64) ++s->vals_ver;
  if (v32) ++s->vals32_ver;
}
// SELL copies of the current values (refreshed on the system's stream when stale)
const double *sell_values64(ehip_system *s);
const float *sell_values32(ehip_system *s);
const double *sho_values64(ehip_system *s);
const float *sho_values32(ehip_system *s);

// SELL-64 launch shape: nonzeros per lane in flight (kSellU, or 9 when that
// wastes fewer loads on the pattern's slice widths), and the grid (slice
// groups of four per workgroup step, at most 4096 workgroups -- measured
// 2-3% faster than 2048 for FP32 and FP64 at 128^3 -- a multiple of 8)
constexpr int kSellU = 8;
constexpr int kSell2U = 4;  // pairs per lane in flight (pair-packed FP64)
constexpr int kSellBU = 4;  // node blocks per lane in flight (node-blocked)
// Infinity Cache (MALL) of the MI355X: a matrix values copy smaller than it
// is read with cached loads and stays resident between SpMVs (the 230 MB
// FP32 copy at 128^3: 44.5 -> 35.9 us per SpMV); a larger one is streamed
// with non-temporal loads, which keeps it from evicting the vectors (the
// 460 MB FP64 copy: 106 us streamed, 112 us cached).
constexpr int64_t kMallBytes = 256ll << 20;
inline bool values_nt(int64_t bytes) { return bytes > kMallBytes; }
// Of a larger copy, the slices before sho_nt_from are read with cached loads
// -- about kMallWorkingBytes less the solver's vectors (six of n entries),
// proportionally to the slice count -- and the rest streamed.  Measured
// optimum of that budget: 128 MB for the FP64 copy at 128^3 (FP64 fast CG
// 59.9 -> 54.2 ms), 96 MB for FP64 and >= 160 MB for FP32 on 96^3 Hex8
// elasticity: each where cached values and vectors come to about 228 MB.
// EHIP_MALL_CACHED_MB=m (read at every launch) sets the budget to m MB for
// every copy, smaller ones included.
constexpr int64_t kMallWorkingBytes = 228ll << 20;
inline int64_t mall_cached_override() {
  const char *v = std::getenv("EHIP_MALL_CACHED_MB");
  return v && *v ? (int64_t)std::atoll(v) << 20 : (int64_t)-1;
}
// whether a copy of `bytes` takes the streamed-loads kernel
inline bool sho_streamed(int64_t bytes) { return values_nt(bytes) || mall_cached_override() >= 0; }
inline int sho_nt_from(int nslice, int64_t bytes, int64_t n, int elem) {
  const int64_t o = mall_cached_override();
  if (o < 0 && !values_nt(bytes)) return nslice;
  const int64_t c = o >= 0 ? o : kMallWorkingBytes - 6 * n * elem;
  if (c <= 0) return 0;
  if (c >= bytes) return nslice;
  return (int)((double)nslice * (double)c / (double)bytes);
}

inline int sell_grid(const ehip_system *s) {
  int g = (s->nslice + 3) / 4;
  if (g > kEpiStride) g = kEpiStride;
  if (g < 8) g = 8;
  return (g + 7) / 8 * 8;
}

// SpMV
// Chunk cut of the CSR stream schedule: the first row of every chunk of at
// most nb nonzeros (a longer row alone), then n.  rows is 1-based (any
// common offset works: only differences are used).  Writes entries
// [first, first + cap) of the list to out (may be null) and returns the
// length of the whole list.
int chunk_rows(const int *rows, int n, int nb, int *out, int first, int cap);
// Fills a plan's shape (block, nb, chunk count, XCD-aware grid).
void plan_shape(SpmvPlan &plan, int block, int nb, int nchunks);
int plan_spmv(ehip_system *s, SpmvPlan &plan, int block, int nb);
int spmv_f64(ehip_system *s, const double *x, double *y);
int spmv_f32(ehip_system *s, const float *x, float *y);

// scaling (device): D from the diagonal, values *= D_i D_j, b *= D
// bnorm_dev non-null (with rhs_scaling): ||b|| is written there and used on
// the device, and nothing waits; bnorm_out is then not set.
int scale_system(ehip_system *s, double *b, bool rhs_scaling, double *x, double *bnorm_out,
                 bool make_f32, double *bnorm_dev = nullptr);
// rhs_scale_dev non-null: ||b|| is read from the device instead of rhs_scale
int backscale_system(ehip_system *s, double *b, double *x, double rhs_scale, const double *rhs_scale_dev = nullptr);

int max_grid();

// CRS pattern and row-owner maps of a mesh built on the device
// (pattern_gpu.hip): fills s->rows/cols/diag/rhs_start/rhs_contrib/row_off
// on the device, s->n, s->nnz, the bandwidth fields and the 1-based host
// copies h_rows/h_cols/h_diag/h_perm; el0_dev: 0-based connectivity.
// Row-major gather maps (rhs_start, rhs_contrib, row_off) into the
// slice-major layout the gather reads (map_base, rhs_contrib, row_off).
int slice_maps(ehip_system *s);
int build_device_pattern(ehip_system *s, int nnodes, int nelem, int npe, const int *el0_dev, int nbnd, int npe_b,
                         const int *bel, int dofs, int optimize, int *max_len);

// Reports the degenerate Gauss-point metrics of an ehip_assemble call made
// without waiting (EHIP_EARG), once; waits for the system's stream.
int assembly_check(ehip_system *s);

// ILU(k) (ilu.hip): pattern and level schedule of order k (kept while the
// order stays), factorization of the current Values, and the LU solve
// dst = (LU)^-1 src enqueued on the system's stream.
int ilu_setup(ehip_system *s, int order);
int ilu_factor(ehip_system *s);
int ilu_apply(ehip_system *s, const double *src, double *dst);
int ilu_check(ehip_system *s);  // a sweep gave up waiting: EHIP_EHIP
void ilu_release(ehip_system *s);

// ehip_solve's body.  events = false times the solve with host clock
// readings instead of HIP events (for callers that create no runtime objects).
int solve_impl(ehip_system *s, const ehip_params *p, double *x, ehip_report *r, double *hist, bool events);

// SpMV launch sampling (ehip_params.profile_every)
void prof_begin(ehip_system *s);
void prof_end(ehip_system *s);
void prof_collect(ehip_system *s, double *mean_us, int *samples);

}  // namespace ehip
