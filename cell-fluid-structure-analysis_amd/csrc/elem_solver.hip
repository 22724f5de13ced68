// Element patches of the momentum operator on the device (elem_spmv): host
// set-up once per mesh topology (elem_patch.cpp), upload, and the device
// descriptor the FP32 SpMV launcher reads.
#include <chrono>

#include "elem_patch.hpp"
#include "solver.hpp"

namespace cfsa {

void NSSolver::ensure_elem() {
  if (elem_built_) return;
  const auto t0 = std::chrono::steady_clock::now();
  const ElemPatchHost E =
      build_elem_patches(nu, nv, nt, h_tets.data(), h_ten.data(), h_coords.data());
  if (E.max_rows > kElemRows || E.max_nodes > kElemNodes || E.max_slots > kElemSlots)
    throw std::logic_error("element patches exceed the kernel's LDS tables");
  if (E.npatch <= 0 || (int64_t)E.rows.size() != nu || (int64_t)E.tet_order.size() != nt)
    throw std::logic_error("element patches do not own every row once");
  d_enode_off.upload(E.node_off.data(), E.node_off.size());
  // on the device the nodes are numbered in patch order: row j of that order
  // is node rows[j] (the owned rows of patch p are rows row_off[p] ..)
  {
    std::vector<int32_t> pos((size_t)nu);
    for (size_t j = 0; j < E.rows.size(); ++j) pos[(size_t)E.rows[j]] = (int32_t)j;
    std::vector<int32_t> nodes(E.nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) nodes[i] = pos[(size_t)E.nodes[i]];
    d_enodes.upload(nodes.data(), nodes.size());
  }
  d_erow_off.upload(E.row_off.data(), E.row_off.size());
  d_erows.upload(E.rows.data(), E.rows.size());
  d_estarts.upload(E.starts.data(), E.starts.size());
  d_evisit_off.upload(E.visit_off.data(), E.visit_off.size());
  d_etorder.upload(E.tet_order.data(), E.tet_order.size());
  // the device keeps the 8 words of a visit as two planes of 4
  {
    std::vector<uint32_t> planes(E.vstat.size());
    const size_t nvis = (size_t)E.nvisit;
    for (size_t v = 0; v < nvis; ++v)
      for (int k = 0; k < kElemStat; ++k)
        planes[(size_t)(k / 4) * nvis * 4 + v * 4 + (size_t)(k % 4)] = E.vstat[v * kElemStat + k];
    d_evstat.upload(planes.data(), planes.size());
  }
  d_evgeo.alloc((size_t)nt * kElemGeo);
  d_epen.alloc((size_t)nu);
  d_einvd32.alloc((size_t)nu);
  elem_dev_ = ElemPatchDev{};
  elem_dev_.npatch = (int)E.npatch;
  elem_dev_.nvisit = E.nvisit;
  elem_dev_.ntet = nt;
  elem_dev_.node_off = d_enode_off.p;
  elem_dev_.nodes = d_enodes.p;
  elem_dev_.row_off = d_erow_off.p;
  elem_dev_.rows = d_erows.p;
  elem_dev_.starts = d_estarts.p;
  elem_dev_.visit_off = d_evisit_off.p;
  elem_dev_.vstat = d_evstat.p;
  elem_dev_.vgeo = d_evgeo.p;
  elem_dev_.pen = d_epen.p;
  elem_nvisit_ = E.nvisit;
  elem_nrows_ = nu;
  elem_built_ = true;
  elem_fresh_ = false;
  clear_graphs();  // captured chunks may hold the CSR operator
  elem_setup_ms_ =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace cfsa
