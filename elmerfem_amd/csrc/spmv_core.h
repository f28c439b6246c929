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

// Epi must provide:
//   static constexpr int NV;                 number of reduced values (0..4)
//   struct Pre;  Pre load(int r);            row operands, loaded early
//   __device__ void init();                  read scalars once per workgroup
//   __device__ void row(int r, A s, const Pre &, A (&acc)[NV]);
//   __device__ void finish(A (&tot)[NV]);    run by the finalize kernel
//   A *partials; int stride; unsigned *ticket;
template <typename V, typename A, int BS, int NB, class Epi>
__device__ __forceinline__ void spmv_stream_body(const int *__restrict__ rows,
                                                 const int *__restrict__ cols,
                                                 const V *__restrict__ vals,
                                                 const A *__restrict__ x,
                                                 const int *__restrict__ chunk_row, int nchunks,
                                                 Epi &epi) {
  static_assert(NB % BS == 0, "chunk capacity must be a multiple of the workgroup size");
  constexpr int ITEMS = NB / BS;
  __shared__ A prod[NB];
  constexpr int NV = Epi::NV;
  constexpr int NA = NV > 0 ? NV : 1;
  A acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = (A)0;
  epi.init();
  XcdWalk w(nchunks);
  for (int c = w.first; c < w.last; c += w.step) {
    const int r0 = chunk_row[c], r1 = chunk_row[c + 1];
    const int j0 = rows[r0], j1 = rows[r1];
    const int len = j1 - j0;
    if (len <= NB) {
      // row operands of this lane's first row
      const int rr = r0 + (int)threadIdx.x;
      typename Epi::Pre pre{};
      if (rr < r1) pre = epi.load(rr);
      V vv[ITEMS];
      int cc[ITEMS];
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int j = (int)threadIdx.x + k * BS;
        if (j < len) {
          vv[k] = ntload(vals + j0 + j);
          cc[k] = ntload(cols + j0 + j);
        }
      }
      A xv[ITEMS];
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int j = (int)threadIdx.x + k * BS;
        if (j < len) xv[k] = x[cc[k]];
      }
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int j = (int)threadIdx.x + k * BS;
        if (j < len) prod[j] = (A)vv[k] * xv[k];
      }
      __syncthreads();
      for (int r = rr; r < r1; r += BS) {
        const int a = rows[r] - j0, b = rows[r + 1] - j0;
        A s = (A)0;
        for (int k = a; k < b; ++k) s = s + prod[k];
        if (r == rr) {
          epi.row(r, s, pre, acc);
        } else {
          const typename Epi::Pre p2 = epi.load(r);
          epi.row(r, s, p2, acc);
        }
      }
      __syncthreads();
    } else {
      // a single row longer than the chunk capacity: workgroup-parallel sum
      __shared__ A sh[16];
      A part = (A)0;
      for (int j = j0 + threadIdx.x; j < j1; j += BS) part += (A)ntload(vals + j) * x[ntload(cols + j)];
      A s = block_sum(part, sh);
      if (threadIdx.x == 0) {
        const typename Epi::Pre p2 = epi.load(r0);
        epi.row(r0, s, p2, acc);
      }
      __syncthreads();
    }
  }
  if constexpr (NV > 0) block_publish<A, NV>(acc, epi.partials, epi.stride);
}

}  // namespace ehip
