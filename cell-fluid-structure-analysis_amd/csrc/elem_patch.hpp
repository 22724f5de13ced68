// Element-patch form of the momentum operator for the FP32 inner solves.
//
// The momentum operator is a sum of 10 x 10 element matrices
// (NavierStokes.cpp:135-225 stiffness, :227-310 mass, :460-495 lhs):
//   lhs_e = (rho V / dt) C + (V / mu) B Phi Phi^T B^T,
// C the reference's constant P2 mass pattern (vertex-vertex 1/10 diagonal and
// 1/20 off it, vertex-edge 1/30, edge-edge 1/15 and 1/60), Phi the 4 x 3
// vertex gradients and B = [I; E] (an edge node's gradient is the sum of its
// two vertices' gradients).  y = lhs x can therefore be applied tet by tet
// from the volume and three gradients: per right-hand side about 90 flops and
// 10 gathered nodes per tet, instead of 28.5 stored nonzeros per row.
//
// Rows are cut into patches that are compact in space (consecutive rows in the
// Morton order of the node positions).  A patch owns its rows and visits every
// tet that touches one of them (tets on a patch border are visited by each
// patch they touch); a block computes its tets' contributions into LDS slots,
// one slot per (owned row, incident tet), and then sums each owned row's
// slots in a fixed order, so the product is deterministic and needs neither
// atomics nor a second pass.
#pragma once

#include <cstdint>
#include <vector>

namespace cfsa {

constexpr int kElemSlots = 4096;  // contribution slots of a patch (12 B of LDS each)
constexpr int kElemNodes = 2048;  // nodes of a patch's tets (local ids)
constexpr int kElemRows = 1024;   // rows a patch owns
constexpr int kElemStat = 8;      // 32-bit words per visit: local ids, slots, the tet
constexpr int kElemGeo = 8;       // floats per tet: rho V / dt, (V / mu) grad phi_k . grad phi_l

struct ElemPatchHost {
  int64_t npatch = 0, nvisit = 0;
  std::vector<int32_t> node_off;  // npatch + 1
  std::vector<int32_t> nodes;     // per patch: owned rows first, then the other nodes
  std::vector<int32_t> row_off;   // npatch + 1: offsets of the owned rows
  std::vector<int32_t> rows;      // global id of each owned row (row_off order)
  std::vector<uint16_t> starts;   // per patch (owned rows + 1) slot starts, at row_off[p] + p
  std::vector<int32_t> visit_off; // npatch + 1
  std::vector<int32_t> vtet;      // tet of each visit (mesh numbering)
  std::vector<int32_t> tet_order; // patch numbering of the tets: tet j is tet_order[j]
  // kElemStat words per visit, two 128-bit little-endian words: A (words 0-3)
  // holds the 10 local node ids (11 bits each, id l at bit 11 l) and bits 0-17
  // of the tet in patch numbering (tet_order) at bit 110; B (words 4-7) holds
  // the LDS slot of each node the patch owns (12 bits at bit 12 l: the owned
  // row's first slot plus the tet's rank among the row's tets in ascending
  // mesh order; 0 for a node the patch does not own) and bits 18-25 of the
  // tet at bit 120
  std::vector<uint32_t> vstat;
  int max_rows = 0, max_nodes = 0, max_slots = 0;
};

// nu nodes (the nv vertices, then the edge nodes), fewer than 2^26 tets,
// tets nt x 4, tet edge
// nodes nt x 6 (edge ids 0..nu-nv-1, local edges 01 02 03 12 13 23), vertex
// coordinates nv x 3.  Throws if a node lies in more than 254 tets or a single
// row does not fit a patch.
ElemPatchHost build_elem_patches(int64_t nu, int64_t nv, int64_t nt, const int32_t *tets,
                                 const int32_t *ten, const double *xyz);

}  // namespace cfsa
