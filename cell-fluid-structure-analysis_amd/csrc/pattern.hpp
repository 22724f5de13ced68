// Sparsity patterns, node-to-tet incidence and tet colouring for the P2/P1
// operators.  Built once on the host when a mesh is attached; the values are
// assembled on the GPU into these fixed patterns.
//
// Pattern semantics follow Eigen's setFromTriplets as used by the reference
// (src/FEM/NavierStokes.cpp:123-129,220-221,287-288,351-352): the pattern is
// the union of all (row, col) pairs an element touches, explicit zeros are
// kept, and column indices are sorted ascending within each row.  M, K and the
// momentum matrix share the 10-node clique pattern; L, Gx, Gy, Gz share the
// 4-vertex clique pattern.
#pragma once

#include <cstdint>
#include <vector>

namespace cfsa {

struct CsrPattern {
  int64_t n = 0;
  std::vector<int32_t> rowptr;  // n + 1
  std::vector<int32_t> col;     // nnz, ascending per row
  std::vector<int32_t> diag;    // slot of (i, i) or -1
  int64_t nnz() const { return rowptr.empty() ? 0 : rowptr.back(); }
};

struct Incidence {
  std::vector<int32_t> ptr;  // n_nodes + 1
  std::vector<int32_t> tet;  // ascending tet ids per node
};

// nodes_per_tet = 4 (vertices only) or 10 (vertices, then nv + edge node ids).
// elem_nodes is nt x nodes_per_tet.
Incidence build_incidence(const std::vector<int32_t> &elem_nodes, int nodes_per_tet,
                          int64_t n_nodes);
// Rows 0..n-1 use inc_ptr[0..n] / inc_tet (a prefix of a larger incidence is
// fine: the P1 pattern reuses the vertex rows of the P2 incidence).
CsrPattern build_pattern(const std::vector<int32_t> &elem_nodes, int nodes_per_tet,
                         const int32_t *inc_ptr, const int32_t *inc_tet, int64_t n);

// Greedy first-fit colouring in tet order; two tets conflict when they share
// a vertex (which also covers sharing an edge node).  Returns the colour of
// every tet; order[] lists tets grouped by colour, ascending tet id inside a
// colour; offsets has ncolours + 1 entries.
struct Colouring {
  std::vector<int32_t> colour;
  std::vector<int32_t> order;
  std::vector<int32_t> offsets;
  int ncolours() const { return (int)offsets.size() - 1; }
};
Colouring colour_tets(const int32_t *tets, int64_t nt, int64_t nv);

// Sliced-ELL copy of a CSR pattern for the Krylov SpMV: rows are grouped in
// slices of kSellSlice (one wavefront, one row per lane); inside a window of
// kSellWindow consecutive rows, rows are stably ordered by decreasing length
// so the lanes of a slice have nearly equal lengths.  Entry j of the row at
// position p = s * kSellSlice + l is stored at sptr[s] + j * kSellSlice + l,
// in the CSR order of that row (ascending columns), so a lane that walks
// j = 0, 1, ... sums its row in the same order as the CSR product.
constexpr int kSellSlice = 64;
constexpr int kSellWindow = 4096;  // multiple of kSellSlice
struct SellPattern {
  int64_t n = 0;
  int nslices = 0;
  std::vector<int32_t> sptr;  // nslices + 1 offsets of the slices' entries
  std::vector<int32_t> rlen;  // per position (nslices * kSellSlice), 0 past n
  std::vector<int32_t> perm;  // per position: CSR row (0 past n)
  std::vector<int32_t> col;   // per stored entry (0 in padding)
  std::vector<int32_t> src;   // per stored entry: CSR slot, -1 in padding
  int64_t stored() const { return sptr.empty() ? 0 : sptr.back(); }
};
SellPattern build_sell(const CsrPattern &P);

// CSR-stream windows: consecutive rows packed greedily so that each window
// holds at most max_nnz nonzeros and at most max_rows rows; returns the
// nwin + 1 window start rows.
std::vector<int32_t> stream_windows(const CsrPattern &P, int max_nnz, int max_rows);

}  // namespace cfsa
