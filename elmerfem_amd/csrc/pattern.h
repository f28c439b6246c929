// Host-side CRS topology and element-to-slot maps (plain C++, OpenMP).
#pragma once
#include <cstdint>
#include <vector>

namespace ehip {

struct HostPattern {
  int nnodes = 0, dofs = 1, n = 0;
  int64_t nnz = 0;
  std::vector<int> perm;               // 1-based, per node
  std::vector<int> rows, cols, diag;   // 1-based CRS
  int bw_before = 0, bw_after = 0, accepted = 0;
};

// The CreateMatrix chain (fem/src/ElementUtils.src:698-935): identity
// InitialPermutation, MakeListMatrix adjacency over bulk and boundary
// elements, OptimizeBandwidth (fem/src/BandwidthOptimize.src:108-355) when
// `optimize`, then sorted CRS rows with Diag.  Element arrays are 1-based.
int build_host_pattern(int nnodes, int nelem, int npe, const int *el, int nbnd, int npe_b, const int *bel,
                       int dofs, int optimize, HostPattern &out);

// OptimizeBandwidth (BandwidthOptimize.src:108-355) over a given node
// adjacency (node i, 1-based, at idx[start[i] .. start[i+1]), sorted, self
// included): perm[1..n] (resized to n+1), the half bandwidths + 1 before
// and after, and whether the new ordering was accepted (else identity).
void bandwidth_permutation(int n, const int64_t *start, const int *idx, std::vector<int> &perm, int &before,
                           int &after, int &accepted);

// Row-owner gather maps for deterministic assembly.  For every matrix row
// the elements touching it in ascending element order -- the order in which
// CRS_GlueLocalMatrix adds their blocks (CRSMatrix.src:507-590) -- each as
// e*ld + i (i the row's local index in the element, ld = npe*dofs), and for
// every such entry the position within the row of each of the element's ld
// columns (row_off[k*ld + j]).  max_len: the longest row.
int build_row_maps(const HostPattern &P, int nelem, int npe, const int *el, std::vector<int> &row_start,
                   std::vector<int> &row_contrib, std::vector<uint16_t> &row_off, int *max_len);

// Slice-shared column offsets of a pattern cut into slices of 64 rows: for
// slice s, the sorted union of (column - row) over its rows at
// tab[s * max_len ..] (len[s] entries), and for every row the bit mask of
// the union entries that are its nonzeros (bit k: entry k).  rows1 is the
// 1-based row pointer, c0 the 0-based columns, each row sorted.  With D > 1
// (rows made of whole D-column node blocks, D unknowns per node) the
// offsets and mask bits are per block: column / D - row / D.  Returns false
// as soon as one slice's union exceeds max_len (<= 32) offsets.
bool slice_offsets(int n, const int *rows1, const int *c0, int max_len, std::vector<int> &len,
                   std::vector<int> &tab, std::vector<unsigned> &mask, int D = 1);

}  // namespace ehip
