#include "pattern.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "parallel.hpp"

namespace cfsa {

std::vector<int32_t> stream_windows(const CsrPattern &P, int max_nnz, int max_rows) {
  std::vector<int32_t> w{0};
  const int64_t n = P.n;
  int64_t r = 0;
  while (r < n) {
    const int32_t k0 = P.rowptr[(size_t)r];
    int64_t e = r;
    while (e < n && e - r < max_rows && P.rowptr[(size_t)e + 1] - k0 <= max_nnz) ++e;
    if (e == r) throw std::runtime_error("CSR row longer than a stream window");
    w.push_back((int32_t)e);
    r = e;
  }
  return w;
}

SellPattern build_sell(const CsrPattern &P) {
  SellPattern S;
  const int64_t n = P.n;
  S.n = n;
  const int64_t ns = (n + kSellSlice - 1) / kSellSlice;
  if (ns > INT32_MAX) throw std::runtime_error("sliced-ELL: too many slices");
  S.nslices = (int)ns;
  const int64_t npos = ns * kSellSlice;
  S.rlen.assign((size_t)npos, 0);
  S.perm.assign((size_t)npos, 0);
  std::vector<int32_t> idx;
  for (int64_t w0 = 0; w0 < n; w0 += kSellWindow) {
    const int64_t w1 = std::min(n, w0 + kSellWindow);
    idx.resize((size_t)(w1 - w0));
    for (int64_t i = w0; i < w1; ++i) idx[(size_t)(i - w0)] = (int32_t)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
      return P.rowptr[(size_t)a + 1] - P.rowptr[(size_t)a] >
             P.rowptr[(size_t)b + 1] - P.rowptr[(size_t)b];
    });
    for (int64_t k = 0; k < w1 - w0; ++k) {
      const int32_t r = idx[(size_t)k];
      S.perm[(size_t)(w0 + k)] = r;
      S.rlen[(size_t)(w0 + k)] = P.rowptr[(size_t)r + 1] - P.rowptr[(size_t)r];
    }
  }
  S.sptr.assign((size_t)ns + 1, 0);
  int64_t total = 0;
  for (int64_t s = 0; s < ns; ++s) {
    int32_t mx = 0;
    for (int l = 0; l < kSellSlice; ++l)
      mx = std::max(mx, S.rlen[(size_t)(s * kSellSlice + l)]);
    total += (int64_t)mx * kSellSlice;
    if (total > INT32_MAX) throw std::runtime_error("sliced-ELL exceeds int32 offsets");
    S.sptr[(size_t)s + 1] = (int32_t)total;
  }
  S.col.assign((size_t)total, 0);
  S.src.assign((size_t)total, -1);
  parallel_chunks(ns, [&](int, int64_t a, int64_t b) {
    for (int64_t s = a; s < b; ++s)
      for (int l = 0; l < kSellSlice; ++l) {
        const int64_t p = s * kSellSlice + l;
        if (p >= n) continue;
        const int32_t r = S.perm[(size_t)p];
        const int32_t k0 = P.rowptr[(size_t)r];
        for (int32_t j = 0; j < S.rlen[(size_t)p]; ++j) {
          const size_t e = (size_t)S.sptr[(size_t)s] + (size_t)j * kSellSlice + (size_t)l;
          S.col[e] = P.col[(size_t)(k0 + j)];
          S.src[e] = k0 + j;
        }
      }
  });
  return S;
}

Incidence build_incidence(const std::vector<int32_t> &elem_nodes, int npt,
                          int64_t n_nodes) {
  const int64_t nt = (int64_t)elem_nodes.size() / npt;
  if (nt * npt > INT32_MAX) throw std::runtime_error("incidence exceeds int32");
  Incidence inc;
  inc.ptr.assign((size_t)n_nodes + 1, 0);
  for (int64_t i = 0; i < nt * npt; ++i) inc.ptr[(size_t)elem_nodes[(size_t)i] + 1]++;
  for (int64_t i = 0; i < n_nodes; ++i) inc.ptr[(size_t)i + 1] += inc.ptr[(size_t)i];
  inc.tet.resize((size_t)nt * npt);
  std::vector<int32_t> fill(inc.ptr.begin(), inc.ptr.end() - 1);
  // Scanning tets in order leaves every list ascending.
  for (int64_t t = 0; t < nt; ++t)
    for (int a = 0; a < npt; ++a) {
      const int32_t node = elem_nodes[(size_t)t * npt + a];
      inc.tet[(size_t)fill[(size_t)node]++] = (int32_t)t;
    }
  return inc;
}

CsrPattern build_pattern(const std::vector<int32_t> &elem_nodes, int npt,
                         const int32_t *inc_ptr, const int32_t *inc_tet, int64_t n) {
  CsrPattern P;
  P.n = n;
  const int T = host_threads();
  std::vector<std::vector<int32_t>> cols((size_t)T);
  std::vector<std::vector<int32_t>> cnt((size_t)T);
  std::vector<int64_t> row0((size_t)T, 0), row1((size_t)T, 0);
  parallel_chunks(n, [&](int t, int64_t a, int64_t b) {
    row0[(size_t)t] = a;
    row1[(size_t)t] = b;
    std::vector<int32_t> scratch;
    auto &cv = cols[(size_t)t];
    auto &cn = cnt[(size_t)t];
    cn.reserve((size_t)(b - a));
    for (int64_t r = a; r < b; ++r) {
      scratch.clear();
      for (int32_t q = inc_ptr[r]; q < inc_ptr[r + 1]; ++q) {
        const int32_t *en = &elem_nodes[(size_t)inc_tet[q] * npt];
        scratch.insert(scratch.end(), en, en + npt);
      }
      std::sort(scratch.begin(), scratch.end());
      scratch.erase(std::unique(scratch.begin(), scratch.end()), scratch.end());
      cn.push_back((int32_t)scratch.size());
      cv.insert(cv.end(), scratch.begin(), scratch.end());
    }
  });
  P.rowptr.assign((size_t)n + 1, 0);
  int64_t total = 0;
  for (int t = 0; t < T; ++t) {
    int64_t r = row0[(size_t)t];
    for (int32_t c : cnt[(size_t)t]) {
      total += c;
      if (total > INT32_MAX) throw std::runtime_error("pattern nnz exceeds int32");
      P.rowptr[(size_t)++r] = (int32_t)total;
    }
  }
  P.col.resize((size_t)total);
  for (int t = 0; t < T; ++t) {
    if (cols[(size_t)t].empty()) continue;
    std::memcpy(&P.col[(size_t)P.rowptr[(size_t)row0[(size_t)t]]], cols[(size_t)t].data(),
                cols[(size_t)t].size() * sizeof(int32_t));
    std::vector<int32_t>().swap(cols[(size_t)t]);
  }
  P.diag.assign((size_t)n, -1);
  parallel_chunks(n, [&](int, int64_t a, int64_t b) {
    for (int64_t r = a; r < b; ++r) {
      const int32_t *lo = P.col.data() + P.rowptr[(size_t)r];
      const int32_t *hi = P.col.data() + P.rowptr[(size_t)r + 1];
      const int32_t *it = std::lower_bound(lo, hi, (int32_t)r);
      if (it != hi && *it == (int32_t)r) P.diag[(size_t)r] = (int32_t)(it - P.col.data());
    }
  });
  return P;
}

Colouring colour_tets(const int32_t *tets, int64_t nt, int64_t nv) {
  constexpr int W = 4;  // up to 256 colours
  std::vector<uint64_t> used((size_t)nv * W, 0);
  Colouring C;
  C.colour.resize((size_t)nt);
  int maxc = -1;
  for (int64_t t = 0; t < nt; ++t) {
    uint64_t forb[W] = {0, 0, 0, 0};
    for (int a = 0; a < 4; ++a) {
      const uint64_t *u = &used[(size_t)tets[(size_t)t * 4 + a] * W];
      for (int w = 0; w < W; ++w) forb[w] |= u[w];
    }
    int c = -1;
    for (int w = 0; w < W && c < 0; ++w)
      if (~forb[w]) c = w * 64 + __builtin_ctzll(~forb[w]);
    if (c < 0) throw std::runtime_error("tet colouring needs more than 256 colours");
    C.colour[(size_t)t] = c;
    maxc = std::max(maxc, c);
    for (int a = 0; a < 4; ++a)
      used[(size_t)tets[(size_t)t * 4 + a] * W + c / 64] |= (1ULL << (c % 64));
  }
  const int nc = maxc + 1;
  C.offsets.assign((size_t)nc + 1, 0);
  for (int64_t t = 0; t < nt; ++t) C.offsets[(size_t)C.colour[(size_t)t] + 1]++;
  for (int c = 0; c < nc; ++c) C.offsets[(size_t)c + 1] += C.offsets[(size_t)c];
  C.order.resize((size_t)nt);
  std::vector<int32_t> fill(C.offsets.begin(), C.offsets.end() - 1);
  for (int64_t t = 0; t < nt; ++t) C.order[(size_t)fill[(size_t)C.colour[(size_t)t]]++] = (int32_t)t;
  return C;
}

}  // namespace cfsa
