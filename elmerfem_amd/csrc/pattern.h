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

// Gather maps for deterministic assembly.  For every CRS slot the list of
// (element, local entry) contributions in ascending element order -- the
// order in which CRS_GlueLocalMatrix adds them (CRSMatrix.src:507-590) --
// encoded as e*ld*ld + (i*ld + j) with ld = npe*dofs and the local matrix
// stored row-major.  Likewise for RHS rows: e*ld + i.
int build_slot_maps(const HostPattern &P, int nelem, int npe, const int *el, std::vector<int> &slot_start,
                    std::vector<int> &slot_contrib, std::vector<int> &rhs_start, std::vector<int> &rhs_contrib);

}  // namespace ehip
