// CSR SpMV core for CDNA4: LDS-staged products with one lane per row.
//
//   phase 1: the workgroup streams its chunk's Values and Cols with
//            coalesced non-temporal loads, gathers x, and stores the
//            rounded products Values(j)*x(Cols(j)) in LDS;
//   phase 2: one lane per row adds its products left to right starting from
//            zero -- the order and roundings of CRS_MatrixVectorProd
//            (fem/src/CRSMatrix.src:3565-3573), so FP64 results are
//            bit-identical to the reference's serial loop;
//   epilogue: a functor receives (row, sum) and fuses the Krylov vector
//            updates and dot-product partials of the calling kernel.
#pragma once
#include "reduce.h"

namespace ehip {

template <typename T>
__device__ __forceinline__ T ntload(const T *p) {
  return __builtin_nontemporal_load(p);
}

// Epi must provide:
//   static constexpr int NV;           number of reduced values (0..4)
//   __device__ void init();                 read scalars once per workgroup
//   __device__ void row(int r, A s, A (&acc)[NV]);
//   __device__ void finish(A (&tot)[NV]);   run by the last workgroup
//   A *partials; int stride; unsigned *ticket;
template <typename V, typename A, int BS, int NB, class Epi>
__device__ __forceinline__ void spmv_stream_body(const int *__restrict__ rows,
                                                 const int *__restrict__ cols,
                                                 const V *__restrict__ vals,
                                                 const A *__restrict__ x,
                                                 const int *__restrict__ chunk_row, int nchunks,
                                                 Epi &epi) {
  __shared__ A prod[NB];
  constexpr int NV = Epi::NV;
  A acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int k = 0; k < (NV > 0 ? NV : 1); ++k) acc[k] = (A)0;
  epi.init();
  XcdWalk w(nchunks);
  for (int c = w.first; c < w.last; c += w.step) {
    const int r0 = chunk_row[c], r1 = chunk_row[c + 1];
    const int j0 = rows[r0], j1 = rows[r1];
    const int len = j1 - j0;
    if (len <= NB) {
#pragma unroll 4
      for (int j = threadIdx.x; j < len; j += BS) {
        const V v = ntload(vals + j0 + j);
        const int col = ntload(cols + j0 + j);
        prod[j] = (A)v * x[col];
      }
      __syncthreads();
      for (int r = r0 + threadIdx.x; r < r1; r += BS) {
        const int a = rows[r] - j0, b = rows[r + 1] - j0;
        A s = (A)0;
        for (int k = a; k < b; ++k) s = s + prod[k];
        if constexpr (NV > 0) epi.row(r, s, acc);
        else { A dummy[1]; epi.row(r, s, dummy); }
      }
      __syncthreads();
    } else {
      // a single row longer than the chunk capacity: workgroup-parallel sum
      __shared__ A sh[16];
      A part = (A)0;
      for (int j = j0 + threadIdx.x; j < j1; j += BS) part += (A)ntload(vals + j) * x[ntload(cols + j)];
      A s = block_sum(part, sh);
      if (threadIdx.x == 0) {
        if constexpr (NV > 0) epi.row(r0, s, acc);
        else { A dummy[1]; epi.row(r0, s, dummy); }
      }
      __syncthreads();
    }
  }
  if constexpr (NV > 0) {
    A tot[NV];
    if (last_block_reduce<A, NV>(acc, epi.partials, epi.stride, epi.ticket, tot)) epi.finish(tot);
  }
}

}  // namespace ehip
