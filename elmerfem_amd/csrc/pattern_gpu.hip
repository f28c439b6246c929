// Device build of the CRS topology and the row-owner assembly maps
// (CreateMatrix, fem/src/ElementUtils.src:698-935, with MakeListMatrix
// :279-620; the maps of DefaultUpdateEquations' glue, CRSMatrix.src:
// 507-590).  Integer work only, so every array equals the host build's
// (pattern.cpp) bit for bit:
//   node -> element incidence by counting, filling and a segmented sort
//   (ascending element order per node);
//   node adjacency as the sorted, de-duplicated union of the nodes of the
//   bulk and boundary elements touching each node;
//   OptimizeBandwidth on the host when asked (a sequential walk over the
//   adjacency, BandwidthOptimize.src:108-355), else the identity and the
//   half bandwidth by a device max-reduction;
//   CRS rows (scan of degree x dofs), columns (permuted neighbours x dofs,
//   sorted per row) and Diag;
//   per row its (element, local row) entries in element order and the
//   position in the row of each element column.
// Counting uses integer atomics, whose results do not depend on order;
// everything order-dependent goes through a sort.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "ehip_internal.h"
#include "pattern.h"

namespace ehip {

namespace {

__global__ void k_count(int64_t m, const int *__restrict__ el, int *cnt) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[el[k]], 1);
}

// inc[start[node] + (arrival order)] = k (k = element * npe + local index)
__global__ void k_fill_inc(int64_t m, const int *__restrict__ el, const int *__restrict__ start, int *fill, int *inc) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int nd = el[k];
    inc[start[nd] + atomicAdd(&fill[nd], 1)] = (int)k;
  }
}

// candidates per node: the npe (npe_b) nodes of every bulk (boundary) element touching it
__global__ void k_cand_count(int nn, const int *__restrict__ is, int npe, const int *__restrict__ bs, int npe_b,
                             int *cc) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x)
    cc[i] = (is[i + 1] - is[i]) * npe + (bs ? (bs[i + 1] - bs[i]) * npe_b : 0);
}

__global__ void k_cand_fill(int nn, const int *__restrict__ el, const int *__restrict__ is,
                            const int *__restrict__ inc, int npe, const int *__restrict__ bel,
                            const int *__restrict__ bs, const int *__restrict__ binc, int npe_b,
                            const int *__restrict__ cs, int *cand) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
    int p = cs[i];
    for (int q = is[i]; q < is[i + 1]; ++q) {
      const int e = inc[q] / npe;
      for (int b = 0; b < npe; ++b) cand[p++] = el[(int64_t)e * npe + b];
    }
    if (bs)
      for (int q = bs[i]; q < bs[i + 1]; ++q) {
        const int e = binc[q] / npe_b;
        for (int b = 0; b < npe_b; ++b) cand[p++] = bel[(int64_t)e * npe_b + b];
      }
  }
}

// distinct values of each sorted segment
__global__ void k_uniq_count(int nn, const int *__restrict__ cs, const int *__restrict__ cand, int *deg) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
    int d = 0, prev = -1;
    for (int q = cs[i]; q < cs[i + 1]; ++q) {
      const int v = cand[q];
      d += v != prev;
      prev = v;
    }
    deg[i] = d;
  }
}

__global__ void k_uniq_write(int nn, const int *__restrict__ cs, const int *__restrict__ cand,
                             const int *__restrict__ as, int *adj) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
    int p = as[i], prev = -1;
    for (int q = cs[i]; q < cs[i + 1]; ++q) {
      const int v = cand[q];
      if (v != prev) adj[p++] = v;
      prev = v;
    }
  }
}

// ComputeBandwidth (BandwidthOptimize.src:60-90) of the unpermuted adjacency
__global__ void k_half_bw(int nn, const int *__restrict__ as, const int *__restrict__ adj, int *hb) {
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x)
    for (int q = as[i]; q < as[i + 1]; ++q) m = max(m, abs(i - adj[q]));
  atomicMax(hb, m);
}

// rows r = dofs * perm0[node] + l hold deg(node) * dofs entries
__global__ void k_row_len(int nn, int dofs, const int *__restrict__ inv, const int *__restrict__ as, int *len) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nn; j += gridDim.x * blockDim.x) {
    const int node = inv[j];
    const int d = (as[node + 1] - as[node]) * dofs;
    for (int l = 0; l < dofs; ++l) len[(int64_t)dofs * j + l] = d;
  }
}

// columns of every row of node `node`, unsorted (a segmented sort follows)
__global__ void k_cols(int nn, int dofs, const int *__restrict__ inv, const int *__restrict__ perm0,
                       const int *__restrict__ as, const int *__restrict__ adj, const int *__restrict__ rows,
                       int *cols) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nn; j += gridDim.x * blockDim.x) {
    const int node = inv[j];
    for (int l = 0; l < dofs; ++l) {
      int p = rows[(int64_t)dofs * j + l];
      for (int q = as[node]; q < as[node + 1]; ++q) {
        const int pk = perm0[adj[q]];
        for (int d = 0; d < dofs; ++d) cols[p++] = dofs * pk + d;
      }
    }
  }
}

__device__ __forceinline__ int lower_pos(const int *c, int len, int v) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void k_diag(int n, const int *__restrict__ rows, const int *__restrict__ cols, int *diag) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    diag[r] = rows[r] + lower_pos(cols + rows[r], rows[r + 1] - rows[r], r);
}

// entries per row: the elements touching the row's node
__global__ void k_ent_count(int n, int dofs, const int *__restrict__ inv, const int *__restrict__ is, int *ec) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int node = inv[r / dofs];
    ec[r] = is[node + 1] - is[node];
  }
}

// row maps: per entry e*ld + li, and the position in the row of each element column
template <typename OffT>
__global__ void k_row_maps(int n, int dofs, int npe, const int *__restrict__ el, const int *__restrict__ inv,
                           const int *__restrict__ perm0, const int *__restrict__ is, const int *__restrict__ inc,
                           const int *__restrict__ rows, const int *__restrict__ cols, const int *__restrict__ rs,
                           int *rcontrib, OffT *roff) {
  const int ld = npe * dofs;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int node = inv[r / dofs], l = r % dofs;
    const int *c = cols + rows[r];
    const int len = rows[r + 1] - rows[r];
    int k = rs[r];
    for (int q = is[node]; q < is[node + 1]; ++q, ++k) {
      const int e = inc[q] / npe, la = inc[q] - e * npe;
      rcontrib[k] = e * ld + la * dofs + l;
      for (int bl = 0; bl < npe; ++bl) {
        const int pk = perm0[el[(int64_t)e * npe + bl]];
        for (int m = 0; m < dofs; ++m) roff[(int64_t)k * ld + bl * dofs + m] = (OffT)lower_pos(c, len, dofs * pk + m);
      }
    }
  }
}

__global__ void k_inv(int nn, const int *__restrict__ perm0, int *inv) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) inv[perm0[i]] = i;
}

__global__ void k_iota(int nn, int *p) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) p[i] = i;
}

// device buffers of one build, freed together
struct Dev {
  std::vector<void *> ptrs;
  ~Dev() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int get(T **p, int64_t count) {
    *p = nullptr;
    if (hipMalloc((void **)p, sizeof(T) * std::max<int64_t>(count, 1)) != hipSuccess) {
      set_error("pattern build: device allocation of %lld bytes failed", (long long)(sizeof(T) * count));
      return EHIP_ENOMEM;
    }
    ptrs.push_back(*p);
    return EHIP_OK;
  }
};

constexpr int kGrid = 4096, kBlock = 256;

// out[0..n] = exclusive prefix sums of in[0..n-1] (out[n] = total); int32 totals checked
static int scan(Dev &D, hipStream_t st, const int *in, int *out, int n, int64_t *total) {
  size_t tb = 0;
  EHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n + 1, st));
  void *tmp = nullptr;
  int rc;
  if ((rc = D.get((char **)&tmp, (int64_t)tb))) return rc;
  // out has n+1 entries; in[n] is read as the zero appended below
  EHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n + 1, st));
  int last = 0;
  EHIP_HIP_TRY(hipMemcpyAsync(&last, out + n, sizeof(int), hipMemcpyDeviceToHost, st));
  EHIP_HIP_TRY(hipStreamSynchronize(st));
  *total = last;
  return EHIP_OK;
}

// sorts every segment [off[i], off[i+1]) of keys ascending (in place through a copy)
static int seg_sort(Dev &D, hipStream_t st, int *keys, int64_t count, int nseg, const int *off) {
  if (count == 0) return EHIP_OK;
  int *alt = nullptr;
  int rc;
  if ((rc = D.get(&alt, count))) return rc;
  hipcub::DoubleBuffer<int> buf(keys, alt);
  size_t tb = 0;
  EHIP_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, buf, (int)count, nseg, off, off + 1, 0,
                                                          32, st));
  void *tmp = nullptr;
  if ((rc = D.get((char **)&tmp, (int64_t)tb))) return rc;
  EHIP_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortKeys(tmp, tb, buf, (int)count, nseg, off, off + 1, 0, 32, st));
  if (buf.Current() != keys)
    EHIP_HIP_TRY(hipMemcpyAsync(keys, buf.Current(), sizeof(int) * count, hipMemcpyDeviceToDevice, st));
  return EHIP_OK;
}

}  // namespace

int build_device_pattern(ehip_system *s, int nnodes, int nelem, int npe, const int *el0_dev, int nbnd, int npe_b,
                         const int *bel, int dofs, int optimize, int *max_len_out) {
  hipStream_t st = s->stream;
  Dev D;
  int rc;
  const int nn = nnodes;
  const int64_t m = (int64_t)nelem * npe, mb = (int64_t)nbnd * npe_b;
  if (m >= INT32_MAX || mb >= INT32_MAX) {
    set_error("pattern build: more than 2^31 element-node entries");
    return EHIP_ECAP;
  }
  // boundary elements to the device (0-based)
  int *bel0 = nullptr;
  if (nbnd > 0) {
    std::vector<int> h((size_t)mb);
    for (int64_t k = 0; k < mb; ++k) {
      if (bel[k] < 1 || bel[k] > nnodes) {
        set_error("ehip_system_from_mesh: boundary node id %d out of range", bel[k]);
        return EHIP_EARG;
      }
      h[k] = bel[k] - 1;
    }
    if ((rc = D.get(&bel0, mb))) return rc;
    EHIP_HIP_TRY(hipMemcpyAsync(bel0, h.data(), sizeof(int) * mb, hipMemcpyHostToDevice, st));
    EHIP_HIP_TRY(hipStreamSynchronize(st));
  }
  // incidence
  int *cnt, *is, *fill, *inc;
  if ((rc = D.get(&cnt, nn + 1)) || (rc = D.get(&is, nn + 1)) || (rc = D.get(&fill, nn)) || (rc = D.get(&inc, m)))
    return rc;
  EHIP_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(int) * (nn + 1), st));
  EHIP_HIP_TRY(hipMemsetAsync(fill, 0, sizeof(int) * nn, st));
  hipLaunchKernelGGL(k_count, dim3(kGrid), dim3(kBlock), 0, st, m, el0_dev, cnt);
  int64_t tot = 0;
  if ((rc = scan(D, st, cnt, is, nn, &tot))) return rc;
  hipLaunchKernelGGL(k_fill_inc, dim3(kGrid), dim3(kBlock), 0, st, m, el0_dev, is, fill, inc);
  if ((rc = seg_sort(D, st, inc, m, nn, is))) return rc;
  int *bs = nullptr, *binc = nullptr;
  if (nbnd > 0) {
    int *bcnt, *bfill;
    if ((rc = D.get(&bcnt, nn + 1)) || (rc = D.get(&bs, nn + 1)) || (rc = D.get(&bfill, nn)) ||
        (rc = D.get(&binc, mb)))
      return rc;
    EHIP_HIP_TRY(hipMemsetAsync(bcnt, 0, sizeof(int) * (nn + 1), st));
    EHIP_HIP_TRY(hipMemsetAsync(bfill, 0, sizeof(int) * nn, st));
    hipLaunchKernelGGL(k_count, dim3(kGrid), dim3(kBlock), 0, st, mb, bel0, bcnt);
    if ((rc = scan(D, st, bcnt, bs, nn, &tot))) return rc;
    hipLaunchKernelGGL(k_fill_inc, dim3(kGrid), dim3(kBlock), 0, st, mb, bel0, bs, bfill, binc);
  }
  // adjacency
  int *cc, *cs, *cand, *deg, *as, *adj;
  if ((rc = D.get(&cc, nn + 1)) || (rc = D.get(&cs, nn + 1)) || (rc = D.get(&deg, nn + 1)) || (rc = D.get(&as, nn + 1)))
    return rc;
  EHIP_HIP_TRY(hipMemsetAsync(cc + nn, 0, sizeof(int), st));
  EHIP_HIP_TRY(hipMemsetAsync(deg + nn, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_cand_count, dim3(kGrid), dim3(kBlock), 0, st, nn, is, npe, bs, npe_b, cc);
  int64_t ncand = 0;
  if ((rc = scan(D, st, cc, cs, nn, &ncand))) return rc;
  if ((rc = D.get(&cand, ncand))) return rc;
  hipLaunchKernelGGL(k_cand_fill, dim3(kGrid), dim3(kBlock), 0, st, nn, el0_dev, is, inc, npe, bel0, bs, binc,
                     npe_b, cs, cand);
  if ((rc = seg_sort(D, st, cand, ncand, nn, cs))) return rc;
  hipLaunchKernelGGL(k_uniq_count, dim3(kGrid), dim3(kBlock), 0, st, nn, cs, cand, deg);
  int64_t nadj = 0;
  if ((rc = scan(D, st, deg, as, nn, &nadj))) return rc;
  if ((rc = D.get(&adj, nadj))) return rc;
  hipLaunchKernelGGL(k_uniq_write, dim3(kGrid), dim3(kBlock), 0, st, nn, cs, cand, as, adj);
  // a node in no element has no row to hold its diagonal
  {
    std::vector<int> h((size_t)nn + 1);
    EHIP_HIP_TRY(hipMemcpyAsync(h.data(), as, sizeof(int) * (nn + 1), hipMemcpyDeviceToHost, st));
    EHIP_HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < nn; ++i)
      if (h[i + 1] == h[i]) {
        set_error("ehip_system_from_mesh: node %d is in no element", i + 1);
        return EHIP_EARG;
      }
  }
  // permutation (0-based position of each node) and half bandwidths
  int *perm0, *inv;
  if ((rc = D.get(&perm0, nn)) || (rc = D.get(&inv, nn))) return rc;
  std::vector<int> hperm((size_t)nn + 1);
  if (optimize) {
    std::vector<int64_t> hstart((size_t)nn + 2, 0);
    std::vector<int> hs((size_t)nn + 1), hadj((size_t)nadj);
    EHIP_HIP_TRY(hipMemcpyAsync(hs.data(), as, sizeof(int) * (nn + 1), hipMemcpyDeviceToHost, st));
    EHIP_HIP_TRY(hipMemcpyAsync(hadj.data(), adj, sizeof(int) * nadj, hipMemcpyDeviceToHost, st));
    EHIP_HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i <= nn; ++i) hstart[i + 1] = hs[i];
    for (int64_t q = 0; q < nadj; ++q) hadj[q] += 1;
    bandwidth_permutation(nn, hstart.data(), hadj.data(), hperm, s->bw_before, s->bw_after, s->bw_accepted);
    std::vector<int> p0(nn);
    for (int i = 0; i < nn; ++i) p0[i] = hperm[i + 1] - 1;
    EHIP_HIP_TRY(hipMemcpyAsync(perm0, p0.data(), sizeof(int) * nn, hipMemcpyHostToDevice, st));
    EHIP_HIP_TRY(hipStreamSynchronize(st));
  } else {
    int *hb;
    if ((rc = D.get(&hb, 1))) return rc;
    EHIP_HIP_TRY(hipMemsetAsync(hb, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_half_bw, dim3(kGrid), dim3(kBlock), 0, st, nn, as, adj, hb);
    hipLaunchKernelGGL(k_iota, dim3(kGrid), dim3(kBlock), 0, st, nn, perm0);
    int h = 0;
    EHIP_HIP_TRY(hipMemcpyAsync(&h, hb, sizeof(int), hipMemcpyDeviceToHost, st));
    EHIP_HIP_TRY(hipStreamSynchronize(st));
    s->bw_before = s->bw_after = h + 1;
    s->bw_accepted = 0;
    for (int i = 0; i <= nn; ++i) hperm[i] = i;
  }
  hipLaunchKernelGGL(k_inv, dim3(kGrid), dim3(kBlock), 0, st, nn, perm0, inv);
  // CRS
  const int n = nn * dofs;
  int *rlen;
  if ((rc = D.get(&rlen, n + 1))) return rc;
  EHIP_HIP_TRY(hipMemsetAsync(rlen + n, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_row_len, dim3(kGrid), dim3(kBlock), 0, st, nn, dofs, inv, as, rlen);
  if ((rc = dmalloc(s, (void **)&s->rows, sizeof(int) * (n + 1)))) return rc;
  int64_t nnz = 0;
  if ((rc = scan(D, st, rlen, s->rows, n, &nnz))) return rc;
  if (nnz >= INT32_MAX) {
    set_error("pattern build: nnz >= 2^31");
    return EHIP_ECAP;
  }
  if ((rc = dmalloc(s, (void **)&s->cols, sizeof(int) * std::max<int64_t>(nnz, 1))) ||
      (rc = dmalloc(s, (void **)&s->diag, sizeof(int) * n)))
    return rc;
  hipLaunchKernelGGL(k_cols, dim3(kGrid), dim3(kBlock), 0, st, nn, dofs, inv, perm0, as, adj, s->rows, s->cols);
  if (optimize && s->bw_accepted)
    if ((rc = seg_sort(D, st, s->cols, nnz, n, s->rows))) return rc;
  hipLaunchKernelGGL(k_diag, dim3(kGrid), dim3(kBlock), 0, st, n, s->rows, s->cols, s->diag);
  // row maps
  const int ld = npe * dofs;
  int *ec;
  if ((rc = D.get(&ec, n + 1))) return rc;
  EHIP_HIP_TRY(hipMemsetAsync(ec + n, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_ent_count, dim3(kGrid), dim3(kBlock), 0, st, n, dofs, inv, is, ec);
  if ((rc = dmalloc(s, (void **)&s->rhs_start, sizeof(int) * (n + 1)))) return rc;
  int64_t nent = 0;
  if ((rc = scan(D, st, ec, s->rhs_start, n, &nent))) return rc;
  if ((int64_t)nelem * ld * ld >= INT32_MAX) {
    set_error("pattern build: element blocks exceed int32");
    return EHIP_ECAP;
  }
  // longest row, for the width of the column positions
  std::vector<int> hrows((size_t)n + 1);
  EHIP_HIP_TRY(hipMemcpyAsync(hrows.data(), s->rows, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st));
  EHIP_HIP_TRY(hipStreamSynchronize(st));
  int max_len = 0;
  for (int r = 0; r < n; ++r) max_len = std::max(max_len, hrows[r + 1] - hrows[r]);
  if (max_len > 65536) {
    set_error("pattern build: a row is longer than 65536");
    return EHIP_ECAP;
  }
  s->row_off_bytes = max_len <= 256 ? 1 : 2;
  if ((rc = dmalloc(s, (void **)&s->rhs_contrib, sizeof(int) * std::max<int64_t>(nent, 1))) ||
      (rc = dmalloc(s, &s->row_off, (size_t)nent * ld * s->row_off_bytes)))
    return rc;
  if (s->row_off_bytes == 1)
    hipLaunchKernelGGL(k_row_maps<uint8_t>, dim3(kGrid), dim3(kBlock), 0, st, n, dofs, npe, el0_dev, inv, perm0, is,
                       inc, s->rows, s->cols, s->rhs_start, s->rhs_contrib, (uint8_t *)s->row_off);
  else
    hipLaunchKernelGGL(k_row_maps<uint16_t>, dim3(kGrid), dim3(kBlock), 0, st, n, dofs, npe, el0_dev, inv, perm0, is,
                       inc, s->rows, s->cols, s->rhs_start, s->rhs_contrib, (uint16_t *)s->row_off);
  EHIP_HIP_TRY(hipGetLastError());
  // host copies of the pattern (1-based, as the host build makes them)
  s->n = n;
  s->nnz = nnz;
  s->h_rows.resize((size_t)n + 1);
  s->h_cols.resize((size_t)nnz);
  s->h_diag.resize((size_t)n);
  EHIP_HIP_TRY(hipMemcpyAsync(s->h_cols.data(), s->cols, sizeof(int) * nnz, hipMemcpyDeviceToHost, st));
  EHIP_HIP_TRY(hipMemcpyAsync(s->h_diag.data(), s->diag, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  EHIP_HIP_TRY(hipStreamSynchronize(st));
  for (int r = 0; r <= n; ++r) s->h_rows[r] = hrows[r] + 1;
  for (int64_t j = 0; j < nnz; ++j) s->h_cols[j] += 1;
  for (int r = 0; r < n; ++r) s->h_diag[r] += 1;
  s->h_perm.assign(hperm.begin() + 1, hperm.end());
  if (max_len_out) *max_len_out = max_len;
  return EHIP_OK;
}

}  // namespace ehip
