// Host set-up of the element patches (see elem_patch.hpp).
#include "elem_patch.hpp"

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>

namespace cfsa {

static_assert(kElemNodes <= 2048, "local node ids are 11 bits");
static_assert(kElemSlots <= 4096, "slots are 12 bits");

namespace {

// 21 bits of v spread to every third bit
uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

// bits b .. b + n - 1 of the 128-bit little-endian word w[0..3] (n <= 32, a
// field ending in word 3 when it starts there): or-in the low n bits of v,
// read, clear
void put_bits(uint32_t *w, int b, int n, uint32_t v) {
  const int k = b >> 5, o = b & 31;
  const uint64_t f = ((uint64_t)v & ((1ull << n) - 1)) << o;
  w[k] |= (uint32_t)f;
  if (k < 3) w[k + 1] |= (uint32_t)(f >> 32);
}
uint32_t get_bits(const uint32_t *w, int b, int n) {
  const int k = b >> 5, o = b & 31;
  const uint64_t v = (uint64_t)w[k] | (k < 3 ? (uint64_t)w[k + 1] << 32 : 0);
  return (uint32_t)((v >> o) & ((1ull << n) - 1));
}
void clear_bits(uint32_t *w, int b, int n) {
  const int k = b >> 5, o = b & 31;
  const uint64_t f = ((1ull << n) - 1) << o;
  w[k] &= ~(uint32_t)f;
  if (k < 3) w[k + 1] &= ~(uint32_t)(f >> 32);
}

}  // namespace

ElemPatchHost build_elem_patches(int64_t nu, int64_t nv, int64_t nt, const int32_t *tets,
                                 const int32_t *ten, const double *xyz) {
  if (nu <= 0 || nv <= 0 || nt <= 0 || nu < nv) throw std::invalid_argument("elem patches: sizes");
  if (nt >= (int64_t)1 << 26) throw std::runtime_error("elem patches: more than 2^26 tets");
  auto node = [&](int64_t t, int l) -> int64_t {
    return l < 4 ? (int64_t)tets[t * 4 + l] : nv + (int64_t)ten[t * 6 + (l - 4)];
  };
  // node -> incident tets in ascending tet order; rank of (t, l) in its node's list
  std::vector<int64_t> inc_off((size_t)nu + 1, 0);
  for (int64_t t = 0; t < nt; ++t)
    for (int l = 0; l < 10; ++l) {
      const int64_t j = node(t, l);
      if (j < 0 || j >= nu) throw std::invalid_argument("elem patches: node id out of range");
      inc_off[(size_t)j + 1]++;
    }
  for (int64_t j = 0; j < nu; ++j) {
    if (inc_off[(size_t)j + 1] > 254)
      throw std::runtime_error("elem patches: a node lies in more than 254 tets");
    inc_off[(size_t)j + 1] += inc_off[(size_t)j];
  }
  std::vector<int32_t> inc_t((size_t)inc_off[(size_t)nu]);
  std::vector<uint8_t> rank((size_t)nt * 10);
  {
    std::vector<int64_t> pos(inc_off.begin(), inc_off.end() - 1);
    for (int64_t t = 0; t < nt; ++t)
      for (int l = 0; l < 10; ++l) {
        const int64_t j = node(t, l);
        rank[(size_t)t * 10 + l] = (uint8_t)(pos[(size_t)j] - inc_off[(size_t)j]);
        inc_t[(size_t)pos[(size_t)j]++] = (int32_t)t;
      }
  }
  // Morton order of the node positions (edge nodes at their midpoints)
  std::vector<double> P((size_t)nu * 3);
  for (int64_t i = 0; i < nv * 3; ++i) P[(size_t)i] = xyz[i];
  for (int64_t t = 0; t < nt; ++t) {
    static const int ev[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int k = 0; k < 6; ++k) {
      const int64_t e = nv + ten[t * 6 + k];
      const int64_t a = tets[t * 4 + ev[k][0]], b = tets[t * 4 + ev[k][1]];
      for (int c = 0; c < 3; ++c) P[(size_t)e * 3 + c] = 0.5 * (xyz[a * 3 + c] + xyz[b * 3 + c]);
    }
  }
  double lo[3], hi[3];
  for (int c = 0; c < 3; ++c) {
    lo[c] = hi[c] = P[(size_t)c];
    for (int64_t i = 0; i < nu; ++i) {
      lo[c] = std::min(lo[c], P[(size_t)i * 3 + c]);
      hi[c] = std::max(hi[c], P[(size_t)i * 3 + c]);
    }
  }
  std::vector<uint64_t> code((size_t)nu);
  for (int64_t i = 0; i < nu; ++i) {
    uint64_t m = 0;
    for (int c = 0; c < 3; ++c) {
      const double w = hi[c] > lo[c] ? (P[(size_t)i * 3 + c] - lo[c]) / (hi[c] - lo[c]) : 0.0;
      const uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, w * 2097151.0));
      m |= spread3(q) << c;
    }
    code[(size_t)i] = m;
  }
  std::vector<int32_t> order((size_t)nu);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return code[(size_t)a] < code[(size_t)b]; });

  // Morton rank of each node: the patches take consecutive ranks, so a node's
  // rank is its index in the patch order the device numbers the nodes in
  std::vector<int32_t> mrank((size_t)nu);
  for (int64_t o = 0; o < nu; ++o) mrank[(size_t)order[(size_t)o]] = (int32_t)o;
  // tets are numbered in the order patches first visit them (the device keeps
  // per-tet geometry in that order, so a patch's tets are mostly contiguous)
  std::vector<int32_t> tnew((size_t)nt, -1);
  ElemPatchHost E;
  E.tet_order.reserve((size_t)nt);
  E.node_off.push_back(0);
  E.row_off.push_back(0);
  E.visit_off.push_back(0);
  // per-patch marks (patch index + 1) and a trial stamp for counting
  std::vector<int64_t> tmark((size_t)nt, 0), nmark((size_t)nu, 0), ntrial((size_t)nu, 0);
  std::vector<int32_t> loc((size_t)nu, -1);
  int64_t stamp = 0;
  std::vector<int32_t> prow, ptets, pother;  // the open patch
  int slots = 0, nnodes = 0;
  int64_t pid = 1;
  auto close_patch = [&]() {
    // node table: owned rows, then the other nodes (put in device order once
    // every patch is built)
    const int nr = (int)prow.size();
    // owned rows by descending slot count (a vertex lies in up to ~24 tets, an
    // edge node in 4-6), so the threads of one wavefront sum rows of equal length
    std::stable_sort(prow.begin(), prow.end(), [&](int32_t a, int32_t b) {
      return inc_off[(size_t)a + 1] - inc_off[(size_t)a] > inc_off[(size_t)b + 1] - inc_off[(size_t)b];
    });
    std::sort(pother.begin(), pother.end(),
              [&](int32_t a, int32_t b) { return mrank[(size_t)a] < mrank[(size_t)b]; });
    for (int i = 0; i < nr; ++i) loc[(size_t)prow[(size_t)i]] = i;
    int k = nr;
    for (int32_t j : pother) loc[(size_t)j] = k++;
    for (int32_t j : prow) E.nodes.push_back(j);
    for (int32_t j : pother) E.nodes.push_back(j);
    E.node_off.push_back((int32_t)E.nodes.size());
    for (int32_t j : prow) E.rows.push_back(j);
    E.row_off.push_back((int32_t)E.rows.size());
    std::vector<int> pst((size_t)nr + 1, 0);
    for (int i = 0; i < nr; ++i)
      pst[(size_t)i + 1] =
          pst[(size_t)i] + (int)(inc_off[(size_t)prow[(size_t)i] + 1] - inc_off[(size_t)prow[(size_t)i]]);
    const int s = pst[(size_t)nr];
    for (int i = 0; i <= nr; ++i) E.starts.push_back((uint16_t)pst[(size_t)i]);
    std::sort(ptets.begin(), ptets.end());
    for (int32_t t : ptets)
      if (tnew[(size_t)t] < 0) {
        tnew[(size_t)t] = (int32_t)E.tet_order.size();
        E.tet_order.push_back(t);
      }
    std::sort(ptets.begin(), ptets.end(),
              [&](int32_t a, int32_t b) { return tnew[(size_t)a] < tnew[(size_t)b]; });
    for (int32_t t : ptets) {
      E.vtet.push_back(t);
      // words 0-3: the 10 local ids (11 bits at bit 11 l), tet bits 0-17 at
      // bit 110; words 4-7: the 10 LDS slots (12 bits at bit 12 l; 0 for a
      // node the patch does not own), tet bits 18-25 at bit 120
      uint32_t w[kElemStat] = {0, 0, 0, 0, 0, 0, 0, 0};
      const uint32_t tp = (uint32_t)tnew[(size_t)t];
      for (int l = 0; l < 10; ++l) {
        const int64_t j = node(t, l);
        const int lid = loc[(size_t)j];
        put_bits(w, 11 * l, 11, (uint32_t)lid);
        if (lid < nr) put_bits(w + 4, 12 * l, 12, (uint32_t)(pst[(size_t)lid] + rank[(size_t)t * 10 + l]));
      }
      put_bits(w, 110, 18, tp & 0x3ffffu);
      put_bits(w + 4, 120, 8, tp >> 18);
      E.vstat.insert(E.vstat.end(), w, w + kElemStat);
    }
    E.visit_off.push_back((int32_t)E.vtet.size());
    E.max_rows = std::max(E.max_rows, nr);
    E.max_nodes = std::max(E.max_nodes, nnodes);
    E.max_slots = std::max(E.max_slots, s);
    for (int32_t j : prow) loc[(size_t)j] = -1;
    for (int32_t j : pother) loc[(size_t)j] = -1;
    prow.clear();
    ptets.clear();
    pother.clear();
    slots = nnodes = 0;
    ++pid;
  };
  for (int64_t o = 0; o < nu; ++o) {
    const int32_t r = order[(size_t)o];
    const int add_slots = (int)(inc_off[(size_t)r + 1] - inc_off[(size_t)r]);
    // nodes this row would add: those of its tets that are new to the patch
    // (counted exactly only when the cheap bound, 10 per new tet, is too big)
    int add_nodes = 1;
    for (int64_t q = inc_off[(size_t)r]; q < inc_off[(size_t)r + 1]; ++q)
      if (tmark[(size_t)inc_t[(size_t)q]] != pid) add_nodes += 10;
    if (nnodes + add_nodes > kElemNodes) {
      ++stamp;
      add_nodes = 0;
      for (int64_t q = inc_off[(size_t)r]; q < inc_off[(size_t)r + 1]; ++q) {
        const int32_t t = inc_t[(size_t)q];
        if (tmark[(size_t)t] == pid) continue;
        for (int l = 0; l < 10; ++l) {
          const int64_t j = node(t, l);
          if (nmark[(size_t)j] != pid && ntrial[(size_t)j] != stamp) {
            ntrial[(size_t)j] = stamp;
            ++add_nodes;
          }
        }
      }
      if (nmark[(size_t)r] != pid && ntrial[(size_t)r] != stamp) ++add_nodes;
    }
    if (!prow.empty() && ((int)prow.size() + 1 > kElemRows || slots + add_slots > kElemSlots ||
                          nnodes + add_nodes > kElemNodes)) {
      close_patch();
      --o;  // the row opens the next patch
      continue;
    }
    if (add_slots > kElemSlots || add_nodes > kElemNodes)
      throw std::runtime_error("elem patches: one row exceeds a patch");
    prow.push_back(r);
    slots += add_slots;
    if (nmark[(size_t)r] != pid) {
      nmark[(size_t)r] = pid;
      ++nnodes;
    } else {
      // a node already in the patch as a neighbour becomes an owned row
      pother.erase(std::find(pother.begin(), pother.end(), r));
    }
    for (int64_t q = inc_off[(size_t)r]; q < inc_off[(size_t)r + 1]; ++q) {
      const int32_t t = inc_t[(size_t)q];
      if (tmark[(size_t)t] == pid) continue;
      tmark[(size_t)t] = pid;
      ptets.push_back(t);
      for (int l = 0; l < 10; ++l) {
        const int64_t j = node(t, l);
        if (nmark[(size_t)j] == pid) continue;
        nmark[(size_t)j] = pid;
        ++nnodes;
        if (j != r) pother.push_back((int32_t)j);
      }
    }
  }
  if (!prow.empty()) close_patch();
  E.npatch = (int64_t)E.visit_off.size() - 1;
  E.nvisit = (int64_t)E.vtet.size();
  // The other nodes of each table were ordered by Morton rank; now that every
  // patch has ordered its owned rows, put them in the device order (position
  // in rows), so the gathers that stage them read runs of neighbouring
  // patches' rows, and renumber the local ids of the patch's visits to match.
  {
    std::vector<int32_t> pos((size_t)nu);
    for (size_t j = 0; j < E.rows.size(); ++j) pos[(size_t)E.rows[j]] = (int32_t)j;
    std::vector<int32_t> idx, remap;
    for (int64_t p = 0; p < E.npatch; ++p) {
      const int b = E.node_off[(size_t)p], e = E.node_off[(size_t)p + 1];
      const int nr = E.row_off[(size_t)p + 1] - E.row_off[(size_t)p];
      const int no = e - b - nr;
      idx.resize((size_t)no);
      std::iota(idx.begin(), idx.end(), 0);
      std::sort(idx.begin(), idx.end(), [&](int32_t u, int32_t v) {
        return pos[(size_t)E.nodes[(size_t)(b + nr + u)]] < pos[(size_t)E.nodes[(size_t)(b + nr + v)]];
      });
      remap.resize((size_t)(nr + no));
      for (int i = 0; i < nr; ++i) remap[(size_t)i] = i;
      std::vector<int32_t> moved((size_t)no);
      for (int k = 0; k < no; ++k) {
        moved[(size_t)k] = E.nodes[(size_t)(b + nr + idx[(size_t)k])];
        remap[(size_t)(nr + idx[(size_t)k])] = nr + k;
      }
      std::copy(moved.begin(), moved.end(), E.nodes.begin() + b + nr);
      for (int v = E.visit_off[(size_t)p]; v < E.visit_off[(size_t)p + 1]; ++v) {
        uint32_t *w = E.vstat.data() + (size_t)v * kElemStat;
        for (int l = 0; l < 10; ++l) {
          const uint32_t lid = get_bits(w, 11 * l, 11);
          clear_bits(w, 11 * l, 11);
          put_bits(w, 11 * l, 11, (uint32_t)remap[lid]);
        }
      }
    }
  }
  return E;
}

}  // namespace cfsa
