// CSR SpMV core for CDNA4: LDS-staged products with one lane per row.
//
//   phase 1: the workgroup streams its chunk's Values and Cols with
//            coalesced non-temporal loads -- all of a lane's loads are issued
//            before any x gather, and all gathers before any product, so a
//            chunk costs two memory latencies, not one per element -- and
//            stores the rounded products Values(j)*x(Cols(j)) in LDS;
//   phase 2: one lane per row adds its products left to right starting from
//            zero -- the order and roundings of CRS_MatrixVectorProd
//            (fem/src/CRSMatrix.src:3565-3573), so FP64 results are
//            bit-identical to the reference's serial loop;
//   epilogue: a functor receives (row, sum) and fuses the Krylov vector
//            updates and dot-product partials of the calling kernel.  The
//            row operands it needs are loaded before phase 1 so their latency
//            hides behind the streaming.
#pragma once
#include <type_traits>
#include <utility>
#include "reduce.h"

namespace ehip {

template <typename T>
__device__ __forceinline__ T ntload(const T *p) {
  return __builtin_nontemporal_load(p);
}

// Column index sources.  raw(j) loads what is stored for nonzero j;
// col(raw, j, seg) decodes it, seg = seg(j0) of the chunk that starts at
// nonzero j0; at_any(j) loads and decodes in one step.
// Kernels take the three index arrays as separate __restrict__ parameters
// (so that uniform loads from them become scalar loads) and build the
// source they use with CS(cols32, cols16, seg_base).
struct ColsI32 {  // the CRS Cols array, 0-based int32
  using Raw = int;
  struct Seg {};
  const int *c;
  __device__ __forceinline__ ColsI32(const int *c32, const unsigned short *, const int *) : c(c32) {}
  __device__ __forceinline__ Seg seg(int) const { return Seg{}; }
  __device__ __forceinline__ Raw raw(int j) const { return ntload(c + j); }
  __device__ __forceinline__ int col(Raw v, int, const Seg &) const { return v; }
  __device__ __forceinline__ int at_any(int j) const { return ntload(c + j); }
};

// Compressed columns: nonzeros are cut into fixed segments of 2^kSegShift;
// column = base[segment] + a 16-bit offset.  Used when every segment's
// column span is below 65536 (any banded ordering of a 3-D mesh up to
// ~2^16 nodes of bandwidth), it moves 2 instead of 4 bytes per nonzero.
// A chunk of at most 2^kSegShift nonzeros touches at most two segments,
// whose bases are read once per chunk.
constexpr int kSegShift = 11;
struct ColsU16 {
  using Raw = int;  // the zero-extended offset: one register per nonzero, no packing
  struct Seg { int s0, b0, b1; };
  const unsigned short *c;
  const int *base;  // nseg + 2 entries
  __device__ __forceinline__ ColsU16(const int *, const unsigned short *c16, const int *sb) : c(c16), base(sb) {}
  __device__ __forceinline__ Seg seg(int j0) const {
    const int s0 = j0 >> kSegShift;
    return Seg{s0, base[s0], base[s0 + 1]};
  }
  __device__ __forceinline__ Raw raw(int j) const { return (int)ntload(c + j); }
  __device__ __forceinline__ int col(Raw v, int j, const Seg &g) const {
    return ((j >> kSegShift) != g.s0 ? g.b1 : g.b0) + v;
  }
  __device__ __forceinline__ int at_any(int j) const { return base[j >> kSegShift] + (int)ntload(c + j); }
};

// Workgroup barrier that orders LDS only: global loads still in flight (the
// next chunk's stream) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Index of the k-th nonzero lane t streams from the chunk [j0, j0+len):
// t + k*BS, clamped to the chunk's last nonzero (or to the one before the
// chunk when it is empty) so that every load is unconditional and in
// bounds; the matrix has at least one nonzero.
template <int BS>
__device__ __forceinline__ int lane_nz(int j0, int len, int k) {
  return min(j0 + (int)threadIdx.x + k * BS, max(j0 + len - 1, 0));
}

template <typename V, int BS, int ITEMS, class CS>
__device__ __forceinline__ void stream_chunk(const V *__restrict__ vals, const CS &cols, int j0, int len,
                                             V (&vv)[ITEMS], typename CS::Raw (&cr)[ITEMS]) {
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int j = lane_nz<BS>(j0, len, k);
    vv[k] = ntload(vals + j);
    cr[k] = cols.raw(j);
  }
}

// Epi must provide:
//   static constexpr int NV;                 number of reduced values (0..4)
//   struct Pre;  Pre load(int r);            row operands, loaded early
//   __device__ void init();                  read scalars once per workgroup
//   __device__ void row(int r, A s, const Pre &, A (&acc)[NV]);
//   __device__ void finish(A (&tot)[NV]);    run by the finalize kernel
//   A *partials; int stride; unsigned *ticket;
// and may provide `bool skip`, set by init() the same way in every
// workgroup: the workgroup then returns before any row (and publishes no
// partials).
template <class E, class = void>
struct EpiHasSkip : std::false_type {};
template <class E>
struct EpiHasSkip<E, std::void_t<decltype(std::declval<E &>().skip)>> : std::true_type {};
template <class E>
__device__ __forceinline__ bool epi_skipped(const E &e) {
  if constexpr (EpiHasSkip<E>::value) return e.skip;
  else return false;
}

// One chunk in flight: its bounds, the streamed values / column indices, and
// this lane's first row's bounds and epilogue operands.  All of it is loaded
// one chunk ahead, so nothing the row sums use is loaded (and waited for)
// after the next chunk's stream has been issued.
template <typename V, typename Raw, int ITEMS, class Pre>
struct ChunkBuf {
  V vv[ITEMS];
  Raw cr[ITEMS];
  int r0, r1, j0, len;
  int ra, rb;
  Pre pre;
};

// Software pipeline: the next chunk's value/index stream is issued right
// after this chunk's x gathers, into the other of two register buffers, so
// its memory latency overlaps the gathers, the products and the LDS row sums
// of this chunk.  (Issued any earlier, the in-order vmcnt wait for the
// gathers would also wait for it.)
template <typename V, typename A, int BS, int NB, class CS, class Epi>
__device__ __forceinline__ void spmv_stream_body(const int *__restrict__ rows,
                                                 const CS &cols,
                                                 const V *__restrict__ vals,
                                                 const A *__restrict__ x,
                                                 const int *__restrict__ chunk_row, int nchunks,
                                                 Epi &epi) {
  static_assert(NB % BS == 0, "chunk capacity must be a multiple of the workgroup size");
  static_assert(NB <= (1 << kSegShift), "a chunk may touch at most two column segments");
  constexpr int ITEMS = NB / BS;
  using Buf = ChunkBuf<V, typename CS::Raw, ITEMS, typename Epi::Pre>;
  __shared__ A prod[NB];
  constexpr int NV = Epi::NV;
  constexpr int NA = NV > 0 ? NV : 1;
  A acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = (A)0;
  epi.init();
  if (epi_skipped(epi)) return;
  const int tid = (int)threadIdx.x;
  const XcdWalk w(nchunks);

  // loads chunk c's bounds into b and issues its stream (one cache line when
  // c is past this workgroup's range -- b is then the current chunk `self`
  // again and never processed -- or when c is a long row)
  auto fill = [&](Buf &b, int c, int self) {
    const bool valid = c < w.last;
    const int cq = valid ? c : self;
    b.r0 = chunk_row[cq];
    b.r1 = chunk_row[cq + 1];
    b.j0 = rows[b.r0];
    b.len = rows[b.r1] - b.j0;
    stream_chunk<V, BS, ITEMS>(vals, cols, b.j0, (valid && b.len <= NB) ? b.len : 1, b.vv, b.cr);
    // this lane's first row, clamped into the chunk so that the loads are
    // unconditional (used only when r0 + tid < r1)
    const int rl = min(b.r0 + tid, b.r1 - 1);
    b.ra = rows[rl];
    b.rb = rows[rl + 1];
    b.pre = epi.load(rl);
  };

  // processes the chunk held in b (chunk c) and streams chunk `next` into nb
  auto process = [&](Buf &b, Buf &nb, int c, int next) {
    const int r0 = b.r0, r1 = b.r1, j0 = b.j0, len = b.len;
    const int rr = r0 + tid;
    const int ra = b.ra, rb = b.rb;
    const typename Epi::Pre pre = b.pre;
    if (len <= NB) {
      const typename CS::Seg g = cols.seg(j0);
      A xv[ITEMS];
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) xv[k] = x[cols.col(b.cr[k], lane_nz<BS>(j0, len, k), g)];
      fill(nb, next, c);
      // entries past the chunk are written too and never read
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) prod[tid + k * BS] = (A)b.vv[k] * xv[k];
      lds_barrier();
      // (the row sums stay inside this branch: a join with the long-row path
      // would make the compiler wait for the in-flight stream here)
      if (rr < r1) {
        A s = (A)0;
        for (int k = ra - j0; k < rb - j0; ++k) s = s + prod[k];
        epi.row(rr, s, pre, acc);
      }
      for (int r = rr + BS; r < r1; r += BS) {  // lanes holding more than one row
        const int a = rows[r] - j0, bb = rows[r + 1] - j0;
        A s = (A)0;
        for (int k = a; k < bb; ++k) s = s + prod[k];
        const typename Epi::Pre p2 = epi.load(r);
        epi.row(r, s, p2, acc);
      }
      lds_barrier();
    } else {
      // a single row longer than the chunk capacity: workgroup-parallel sum
      fill(nb, next, c);
      __shared__ A sh[16];
      A part = (A)0;
      for (int j = j0 + tid; j < j0 + len; j += BS) part += (A)ntload(vals + j) * x[cols.at_any(j)];
      const A s = block_sum(part, sh);
      if (tid == 0) epi.row(r0, s, pre, acc);
      __syncthreads();
    }
  };

  if (w.first < w.last) {
    Buf b0, b1;
    int c = w.first;
    fill(b0, c, c);
    for (;;) {
      process(b0, b1, c, c + w.step);
      c += w.step;
      if (c >= w.last) break;
      process(b1, b0, c, c + w.step);
      c += w.step;
      if (c >= w.last) break;
    }
  }
  if constexpr (NV > 0) block_publish<A, NV>(acc, epi.partials, epi.stride);
}

// ---------------------------------------------------------------- SELL-64
// Sliced ELLPACK with slices of 64 rows (one wavefront): inside slice s the
// k-th nonzero of slice row 64 s + l sits at position off[s] + 64 k + l;
// rows shorter than the slice's longest are padded (value 0, a valid
// column).  Slice row i is matrix row i, or rmap[i] when the rows were
// sorted by length inside windows to keep the padding small (SELL-C-sigma:
// patterns whose neighbouring rows differ in length, like Hex27 elements);
// rmap[i] >= n marks a padding lane.
// One lane owns one row and adds its products for k = 0, 1, ... in order --
// the CRS order, so FP64 results stay bit-identical to the reference loop --
// with no LDS and no barriers.  Consecutive lanes read consecutive positions
// of the value and index streams, and for mesh matrices (neighbouring rows
// have neighbouring columns) their x gathers fall in a few cache lines:
// the CSR-stream kernel's per-lane scattered gathers, bound by the L1 tag
// rate rather than HBM, become coalesced loads.
template <bool C16>
struct SellCols {
  const int *c32;
  const unsigned short *c16;
  const int *base;  // per-slice column base (16-bit form)
  __device__ __forceinline__ int col(int idx, int slice) const {
    if constexpr (C16) return base[slice] + (int)ntload(c16 + idx);
    else return ntload(c32 + idx);
  }
};

// U: nonzeros per lane loaded before any of them is used
// FP64 values stop at the row's own length, so the sum is the reference's to
// the bit whatever x holds.  FP32 copies (no bit-exact contract) sum the
// padding too: a padding entry is +0 times x at a real column, which leaves
// a finite sum unchanged, and the row-pointer loads are saved.
template <typename V, typename A, int BS, int U, class SC, class Epi>
__device__ __forceinline__ void spmv_sell_body(const int *__restrict__ rows, int n, const int *__restrict__ rmap,
                                               const SC &sc,
                                               const int *__restrict__ off, const V *__restrict__ vals,
                                               const A *__restrict__ x, int nslice, Epi &epi) {
  constexpr int WPB = BS / 64;  // slices per workgroup step
  constexpr bool kF32 = sizeof(V) == 4;
  constexpr int NV = Epi::NV;
  constexpr int NA = NV > 0 ? NV : 1;
  A acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = (A)0;
  epi.init();
  if (epi_skipped(epi)) return;
  const int lane = (int)threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const XcdWalk w((nslice + WPB - 1) / WPB);
  for (int gidx = w.first; gidx < w.last; gidx += w.step) {
    const int sl = gidx * WPB + wv;  // wave-uniform
    if (sl >= nslice) continue;
    const int r = rmap ? rmap[sl * 64 + lane] : sl * 64 + lane;  // >= n: a padding lane
    const int rl = min(r, n - 1);
    const typename Epi::Pre pre = epi.load(rl);
    const int len = kF32 || r >= n ? 0 : rows[rl + 1] - rows[rl];
    const int p0 = off[sl];
    const int wd = (off[sl + 1] - p0) >> 6;  // slice width
    A sum = (A)0;
    for (int k0 = 0; k0 < wd; k0 += U) {
      V vv[U];
      int cc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = p0 + min(k0 + u, wd - 1) * 64 + lane;  // clamped into the slice
        vv[u] = ntload(vals + idx);
        cc[u] = sc.col(idx, sl);
      }
      A xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = x[cc[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        // the product is rounded before it is added, as in the reference's
        // loop (no fused multiply-add)
#pragma clang fp contract(off)
        const A t = sum + (A)vv[u] * xv[u];
        if constexpr (kF32) sum = (k0 + u < wd) ? t : sum;  // only the clamped repeats are skipped
        else sum = (k0 + u < len) ? t : sum;
      }
    }
    if (r < n) epi.row(r, sum, pre, acc);
  }
  if constexpr (NV > 0) block_publish<A, NV>(acc, epi.partials, epi.stride);
}

// Pair-packed SELL-64 for 8-byte values (the FP64 copy): inside slice s the
// k-th nonzero of row 64 s + l sits at off2[s] + 128 (k/2) + 2 l + (k%2),
// each slice's width rounded up to even.  A lane loads its row's next two
// values as one 16-byte load and their two 16-bit column offsets as one
// 4-byte load -- half the load instructions of the unpacked layout, which
// for FP64 are the limit -- and still adds k = 0, 1, ... in order.
typedef double EhipDouble2 __attribute__((ext_vector_type(2)));
template <typename A, int BS, int U, class Epi>
__device__ __forceinline__ void spmv_sell2_body(const int *__restrict__ rows, int n, const int *__restrict__ rmap,
                                                const unsigned short *__restrict__ c16,
                                                const int *__restrict__ base, const int *__restrict__ off2,
                                                const double *__restrict__ vals, const A *__restrict__ x, int nslice,
                                                Epi &epi) {
  constexpr int WPB = BS / 64;
  constexpr int NV = Epi::NV;
  constexpr int NA = NV > 0 ? NV : 1;
  A acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = (A)0;
  epi.init();
  if (epi_skipped(epi)) return;
  const int lane = (int)threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const XcdWalk w((nslice + WPB - 1) / WPB);
  for (int gidx = w.first; gidx < w.last; gidx += w.step) {
    const int sl = gidx * WPB + wv;  // wave-uniform
    if (sl >= nslice) continue;
    const int r = rmap ? rmap[sl * 64 + lane] : sl * 64 + lane;  // >= n: a padding lane
    const int rl = min(r, n - 1);
    const typename Epi::Pre pre = epi.load(rl);
    const int len = r < n ? rows[rl + 1] - rows[rl] : 0;
    const int p0 = off2[sl];
    const int np = (off2[sl + 1] - p0) >> 7;  // pairs per row in this slice
    const int b = base[sl];
    A sum = (A)0;
    for (int k0 = 0; k0 < np; k0 += U) {
      EhipDouble2 vv[U];
      unsigned int cc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = p0 + min(k0 + u, np - 1) * 128 + 2 * lane;  // clamped into the slice
        vv[u] = ntload(reinterpret_cast<const EhipDouble2 *>(vals + idx));
        cc[u] = ntload(reinterpret_cast<const unsigned int *>(c16 + idx));
      }
      A xv[U][2];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xv[u][0] = x[b + (int)(cc[u] & 0xffffu)];
        xv[u][1] = x[b + (int)(cc[u] >> 16)];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma clang fp contract(off)
          const A t = sum + (A)vv[u][h] * xv[u][h];
          sum = (2 * (k0 + u) + h < len) ? t : sum;
        }
    }
    if (r < n) epi.row(r, sum, pre, acc);
  }
  if constexpr (NV > 0) block_publish<A, NV>(acc, epi.partials, epi.stride);
}

// SELL-64 over node blocks, for patterns whose unknowns are interleaved per
// node (D unknowns per node, every row made of whole D-column blocks
// node*D .. node*D+D-1, as CRS patterns of D-DOF equations are): values lie
// as in spmv_sell_body (entry k*D+d of slice row 64 s + l at
// off[s] + 64 (k D + d) + l), and one 16-bit node offset from nbase[s] per
// block, at off[s]/D + 64 k + l, replaces D column indices.  Summation order
// is the row's ascending column order, as in the other layouts.
template <typename V, typename A, int BS, int U, int D, class Epi>
__device__ __forceinline__ void spmv_sellb_body(const int *__restrict__ rows, int n, const int *__restrict__ rmap,
                                                const unsigned short *__restrict__ cb,
                                                const int *__restrict__ nbase, const int *__restrict__ off,
                                                const V *__restrict__ vals, const A *__restrict__ x, int nslice,
                                                Epi &epi) {
  constexpr int WPB = BS / 64;
  constexpr bool kF32 = sizeof(V) == 4;  // as in spmv_sell_body
  constexpr int NV = Epi::NV;
  constexpr int NA = NV > 0 ? NV : 1;
  A acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = (A)0;
  epi.init();
  if (epi_skipped(epi)) return;
  const int lane = (int)threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const XcdWalk w((nslice + WPB - 1) / WPB);
  for (int gidx = w.first; gidx < w.last; gidx += w.step) {
    const int sl = gidx * WPB + wv;  // wave-uniform
    if (sl >= nslice) continue;
    const int r = rmap ? rmap[sl * 64 + lane] : sl * 64 + lane;  // >= n: a padding lane
    const int rl = min(r, n - 1);
    const typename Epi::Pre pre = epi.load(rl);
    const int nb = kF32 || r >= n ? 0 : (rows[rl + 1] - rows[rl]) / D;  // blocks of this row
    const int p0 = off[sl];
    const int wb = (off[sl + 1] - p0) / (64 * D);  // slice width in blocks
    const int q0 = p0 / D;
    const int b = nbase[sl];
    A sum = (A)0;
    for (int k0 = 0; k0 < wb; k0 += U) {
      V vv[U][D];
      int cc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kk = min(k0 + u, wb - 1);  // clamped into the slice
        cc[u] = D * (b + (int)ntload(cb + q0 + kk * 64 + lane));
#pragma unroll
        for (int d = 0; d < D; ++d) vv[u][d] = ntload(vals + p0 + (kk * D + d) * 64 + lane);
      }
      A xv[U][D];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int d = 0; d < D; ++d) xv[u][d] = x[cc[u] + d];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int d = 0; d < D; ++d) {
#pragma clang fp contract(off)
          const A t = sum + (A)vv[u][d] * xv[u][d];
          if constexpr (kF32) sum = (k0 + u < wb) ? t : sum;
          else sum = (k0 + u < nb) ? t : sum;
        }
    }
    if (r < n) epi.row(r, sum, pre, acc);
  }
  if constexpr (NV > 0) block_publish<A, NV>(acc, epi.partials, epi.stride);
}

// SELL-64 with slice-shared column offsets, for stencil-like patterns
// (meshes of regular topology in natural node order): entry k of every lane
// of slice s multiplies x at row + tab[k], one offset list per slice (the
// sorted union of its rows' column offsets), an entry that is not one of
// the row's nonzeros holding value 0.  The column stream disappears -- a
// wave reads one offset per entry, a wave-uniform (scalar) load, instead of
// 64 indices -- and a wave's x loads are 64 consecutive elements.  Entries
// are added in ascending offset order, i.e. the row's CRS order.  FP64 adds
// only the row's own nonzeros (bit k of mask[row]), so results stay
// bit-identical to the reference loop whatever x holds; FP32 adds the
// absent entries too, as +0 times x at a column inside the matrix (clamped),
// which leaves a finite sum unchanged (the same results as the other
// layouts).  D > 1: rows made of whole D x D node blocks (D unknowns per
// node, interleaved); the list holds node offsets, block k of row r covers
// columns D (r / D + tab[k]) + d, d < D, its values at off[s] + 64 (k D + d)
// + lane, and mask bit k stands for the whole block.  NT: the values of
// slices from nt_from on are streamed with non-temporal loads and those
// before it read with cached loads; without NT every value load is cached.
// Cached values stay in the Infinity Cache from one SpMV to the next, so a
// values copy smaller than that cache is read cached whole, and a larger
// one cached up to a budget and streamed after it (sho_nt_from picks).
template <typename V, typename A, int BS, int U, int D, bool NT, class Epi>
__device__ __forceinline__ void spmv_sho_body(int n, const int *__restrict__ off, const int *__restrict__ tab,
                                              const unsigned *__restrict__ mask, const V *__restrict__ vals,
                                              const A *__restrict__ x, int nslice, int nt_from, Epi &epi) {
  constexpr int WPB = BS / 64;
  constexpr bool kF32 = sizeof(V) == 4;
  constexpr int NV = Epi::NV;
  constexpr int NA = NV > 0 ? NV : 1;
  A acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = (A)0;
  epi.init();
  if (epi_skipped(epi)) return;
  const int lane = (int)threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const XcdWalk w((nslice + WPB - 1) / WPB);
  // the wave's slices in [s0, s1), values read with streamed (true) or
  // cached loads; called once per kind, so each loop has registers of its own
  auto walk = [&](auto streamed, int s0, int s1) {
    constexpr bool kNt = decltype(streamed)::value;
    for (int gidx = w.first; gidx < w.last; gidx += w.step) {
      const int sl = gidx * WPB + wv;  // wave-uniform
      if (sl < s0 || sl >= s1) continue;
      const int r = sl * 64 + lane;  // >= n: a padding lane
      const int rl = min(r, n - 1);
      const typename Epi::Pre pre = epi.load(rl);
      const unsigned m = kF32 || r >= n ? 0u : mask[rl];
      const int node = rl / D;
      const int p0 = off[sl];
      const int wd = (off[sl + 1] - p0) / (64 * D);  // slice width in blocks
      const int *__restrict__ t = tab + p0 / (64 * D);
      A sum = (A)0;
      for (int k0 = 0; k0 < wd; k0 += U) {
        V vv[U][D];
        int cc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kk = min(k0 + u, wd - 1);  // clamped into the slice
          cc[u] = min(max(D * (node + t[kk]), 0), n - D);
#pragma unroll
          for (int d = 0; d < D; ++d)
            vv[u][d] = kNt ? ntload(vals + p0 + (kk * D + d) * 64 + lane) : vals[p0 + (kk * D + d) * 64 + lane];
        }
        A xv[U][D];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int d = 0; d < D; ++d) xv[u][d] = x[cc[u] + d];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + u;  // mask bits stop below wd <= 32
          const bool on = kF32 ? k < wd : (k < 32 && ((m >> (k & 31)) & 1u));
#pragma unroll
          for (int d = 0; d < D; ++d) {
#pragma clang fp contract(off)
            const A tv = sum + (A)vv[u][d] * xv[u][d];
            sum = on ? tv : sum;
          }
        }
      }
      if (r < n) epi.row(r, sum, pre, acc);
    }
  };
  if constexpr (NT) {
    walk(std::false_type{}, 0, nt_from);
    walk(std::true_type{}, nt_from, nslice);
  } else {
    walk(std::false_type{}, 0, nslice);
  }
  if constexpr (NV > 0) block_publish<A, NV>(acc, epi.partials, epi.stride);
}

}  // namespace ehip
