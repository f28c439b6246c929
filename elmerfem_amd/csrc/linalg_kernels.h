// Row/nonzero walkers and vector kernels shared by scaling, Dirichlet and
// the Krylov solvers.
#pragma once
#include "ehip_internal.h"
#include "reduce.h"

namespace ehip {

constexpr int kVecBlock = 256;

// The Krylov work vectors of one system (HUTI's work array columns).
template <typename T>
struct Vecs {
  T *r, *z, *p, *q, *t, *dg, *b, *x;
  T *rtld, *v, *sv, *t1v, *t2v;  // BiCGStab
};

// Grid for vector kernels: enough workgroups to fill 256 CUs, a multiple of 8.
// Vector kernels give each lane kVI elements per tile (coalesced, all loads
// of a tile issued before any arithmetic).  At most 4096 workgroups, so two
// partial sums per workgroup fit the 4 x 2048 partials buffer.
constexpr int kVI = 4;
inline int vec_grid(int n) {
  int g = (n + kVecBlock * kVI - 1) / (kVecBlock * kVI);
  if (g > 4096) g = 4096;
  if (g < 8) g = 8;
  return (g + 7) / 8 * 8;
}

// Grid for the CG vector steps that open with a reduction of the SpMV's
// partial sums (every workgroup reads all of them): capped so that those
// reads stay small beside the vector traffic.  Each workgroup walks several
// tiles, the next tile's loads in flight while the current one is computed.
// (1024 measured best of 256/512/1024/4096 on the 2.1M-row Hex8 system.)
constexpr int kFusedCap = 1024;
inline int fused_grid(int n) { return std::min(vec_grid(n), kFusedCap); }

// Calls f(i, j) for every stored nonzero j (0-based) of row i, walking the
// SpMV chunks so that loads of Values/Cols stay coalesced: each thread finds
// the row of its nonzero by binary search over the chunk's row pointers held
// in LDS.  NB must be >= the plan's chunk capacity.
template <int BS, int NB, class F>
__device__ __forceinline__ void for_each_nonzero(const int *__restrict__ rows,
                                                 const int *__restrict__ chunk_row, int nchunks,
                                                 F &f) {
  __shared__ int rp[NB + 1];
  XcdWalk w(nchunks);
  for (int c = w.first; c < w.last; c += w.step) {
    const int r0 = chunk_row[c], r1 = chunk_row[c + 1];
    const int nr = r1 - r0;
    const int j0 = rows[r0], j1 = rows[r1];
    if (nr <= NB) {
      for (int k = threadIdx.x; k <= nr; k += BS) rp[k] = rows[r0 + k];
      __syncthreads();
      for (int j = j0 + threadIdx.x; j < j1; j += BS) {
        int lo = 0, hi = nr - 1;  // last row with rp[row] <= j
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (rp[mid] <= j) lo = mid; else hi = mid - 1;
        }
        f(r0 + lo, j);
      }
      __syncthreads();
    } else {
      for (int r = r0; r < r1; ++r)
        for (int j = rows[r] + threadIdx.x; j < rows[r + 1]; j += BS) f(r, j);
    }
  }
}

template <class F>
__global__ __launch_bounds__(256) void k_nonzeros(const int *__restrict__ rows,
                                                  const int *__restrict__ chunk_row, int nchunks, F f) {
  for_each_nonzero<256, 4096>(rows, chunk_row, nchunks, f);
}

}  // namespace ehip
