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
//
// Software pipeline per workgroup: the Values/column stream of chunk c+1 is
// issued before the row sums of chunk c, so its memory latency overlaps
// the LDS phase instead of following it.
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
  __shared__ A prod[NB];
  constexpr int NV = Epi::NV;
  constexpr int NA = NV > 0 ? NV : 1;
  A acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = (A)0;
  epi.init();
  const int tid = (int)threadIdx.x;
  XcdWalk w(nchunks);
  if (w.first < w.last) {
    V vv[ITEMS];
    typename CS::Raw cr[ITEMS];
    int c = w.first;
    int r0 = chunk_row[c], r1 = chunk_row[c + 1];
    int j0 = rows[r0];
    int len = rows[r1] - j0;
    stream_chunk<V, BS, ITEMS>(vals, cols, j0, len <= NB ? len : 1, vv, cr);
    while (c < w.last) {
      // the next chunk's bounds (the current chunk again after the last one)
      const int cn = c + w.step;
      const bool more = cn < w.last;
      const int cq = more ? cn : c;
      const int nr0 = chunk_row[cq], nr1 = chunk_row[cq + 1];
      const int nj0 = rows[nr0];
      const int nlen = rows[nr1] - nj0;
      // this lane's first row, clamped into the chunk so that its operand
      // and bound loads are unconditional (used only when rr < r1)
      const int rr = r0 + tid;
      const int rl = min(rr, r1 - 1);
      const typename Epi::Pre pre = epi.load(rl);
      const int ra = rows[rl], rb = rows[rl + 1];
      if (len <= NB) {
        const typename CS::Seg g = cols.seg(j0);
        A xv[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) xv[k] = x[cols.col(cr[k], lane_nz<BS>(j0, len, k), g)];
        // entries past the chunk are written too and never read
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) prod[tid + k * BS] = (A)vv[k] * xv[k];
      } else {
        // a single row longer than the chunk capacity: workgroup-parallel sum
        __shared__ A sh[16];
        A part = (A)0;
        for (int j = j0 + tid; j < j0 + len; j += BS) part += (A)ntload(vals + j) * x[cols.at_any(j)];
        const A s = block_sum(part, sh);
        if (tid == 0) epi.row(r0, s, pre, acc);
      }
      // the next chunk's stream, in flight during this chunk's row sums (one
      // cache line when there is no next chunk or it is a long row)
      stream_chunk<V, BS, ITEMS>(vals, cols, nj0, (more && nlen <= NB) ? nlen : 1, vv, cr);
      lds_barrier();
      if (len <= NB) {
        if (rr < r1) {
          A s = (A)0;
          for (int k = ra - j0; k < rb - j0; ++k) s = s + prod[k];
          epi.row(rr, s, pre, acc);
        }
        for (int r = rr + BS; r < r1; r += BS) {  // lanes holding more than one row
          const int a = rows[r] - j0, b = rows[r + 1] - j0;
          A s = (A)0;
          for (int k = a; k < b; ++k) s = s + prod[k];
          const typename Epi::Pre p2 = epi.load(r);
          epi.row(r, s, p2, acc);
        }
      }
      lds_barrier();
      c = cn;
      r0 = nr0;
      r1 = nr1;
      j0 = nj0;
      len = nlen;
    }
  }
  if constexpr (NV > 0) block_publish<A, NV>(acc, epi.partials, epi.stride);
}

}  // namespace ehip
