// layout.cpp -- device data layout of the neighbor lists (host side).
//
// Internal node order.  Nodes are renumbered by the Morton (Z-order) index of
// their mesh cell, so that 256 consecutive internal nodes -- one workgroup of
// the step kernel -- cover a compact patch of the specimen (a 16x16 patch of a
// uniform 2D grid).  The rows of the neighbor lists keep the reference's order
// (ascending reference id, src/geometry/neighbor.cpp:69-89) so per-node sums
// and bond-bit positions are unchanged; only node storage is permuted.
//
// Slice layout (SELL-64).  64 consecutive internal nodes form a slice (one
// wavefront).  Slot k of lane l of slice s is entry sl_ptr[s] + 64 k + l;
// slots past a row's length repeat the node itself.
//
// Halo.  For each block of 256 internal nodes (128 or 512 when the staged
// kernel's tile is tuned) the sorted set of internal ids its rows reference
// (plus its own nodes) is the block's halo; the step kernel stages the halo's
// positions, displacements and volumes in LDS and reads neighbors through
// 16-bit halo-local indices (colL).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <thread>
#include <vector>

#include "layout.h"

namespace nlm {
namespace {

uint64_t spread2(uint64_t v) {  // 21 bits -> every 2nd bit
  v &= 0x1fffffull;
  v = (v | (v << 16)) & 0x0000ffff0000ffffull;
  v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}
uint64_t spread3(uint64_t v) {  // 21 bits -> every 3rd bit
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// Runs body(lo, hi) over [0, count) split into contiguous chunks, one per
// host thread (at most 16).
template <class F>
void parallel_chunks(int64_t count, F body) {
  int64_t nt = int64_t(std::thread::hardware_concurrency());
  nt = std::max<int64_t>(1, std::min<int64_t>(nt, 16));
  if (count < 4096) nt = 1;
  const int64_t chunk = (count + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int64_t t = 0; t < nt; t++) {
    const int64_t lo = t * chunk, hi = std::min(count, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=, &body] { body(lo, hi); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

void build_layout(int dim, int64_t n, const double* x, double cell,
                  const int64_t* rp_ref, const int32_t* col_ref,
                  HostLayout& L) {
  // --- permutation ------------------------------------------------------
  double lo[3] = {1e300, 1e300, 1e300};
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], x[3 * i + d]);
  std::vector<uint64_t> key(n);
  for (int64_t i = 0; i < n; i++) {
    uint64_t c[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
      double q = std::floor((x[3 * i + d] - lo[d]) / cell + 0.5);
      c[d] = q < 0 ? 0 : (q > 2097151.0 ? 2097151ull : uint64_t(q));
    }
    key[i] = dim == 3 ? (spread3(c[0]) | (spread3(c[1]) << 1) |
                         (spread3(c[2]) << 2))
                      : (spread2(c[0]) | (spread2(c[1]) << 1));
  }
  L.perm.resize(n);
  std::iota(L.perm.begin(), L.perm.end(), 0);
  std::stable_sort(L.perm.begin(), L.perm.end(), [&](int32_t a, int32_t b) {
    return key[a] < key[b];
  });
  L.inv.resize(n);
  for (int64_t a = 0; a < n; a++) L.inv[L.perm[a]] = int32_t(a);

  // --- internal CSR (rows in reference order) ------------------------------
  L.row_ptr.assign(n + 1, 0);
  for (int64_t a = 0; a < n; a++) {
    int64_t i = L.perm[a];
    L.row_ptr[a + 1] = L.row_ptr[a] + (rp_ref[i + 1] - rp_ref[i]);
  }
  const int64_t nnz = L.row_ptr[n];
  L.col.resize(nnz);
  L.ref_off.resize(n);
  parallel_chunks(n, [&](int64_t a0, int64_t a1) {
    for (int64_t a = a0; a < a1; a++) {
      int64_t i = L.perm[a];
      L.ref_off[a] = rp_ref[i];
      int64_t o = L.row_ptr[a];
      for (int64_t k = rp_ref[i]; k < rp_ref[i + 1]; k++)
        L.col[o++] = L.inv[col_ref[k]];
    }
  });

  // --- slices ----------------------------------------------------------------
  const int64_t nslices = (n + 63) / 64;
  L.nslices = nslices;
  L.sl_ptr.assign(nslices + 1, 0);
  L.rowlen.assign(n, 0);
  for (int64_t s = 0; s < nslices; s++) {
    int64_t w = 0;
    for (int64_t a = s * 64; a < std::min(n, s * 64 + 64); a++) {
      int64_t len = L.row_ptr[a + 1] - L.row_ptr[a];
      L.rowlen[a] = uint16_t(len);
      w = std::max(w, len);
    }
    L.sl_ptr[s + 1] = L.sl_ptr[s] + 64 * w;
  }
  const int64_t nent = L.sl_ptr[nslices];
  L.colS.resize(nent);
  parallel_chunks(nslices, [&](int64_t s0, int64_t s1) {
    for (int64_t s = s0; s < s1; s++) {
      int64_t w = (L.sl_ptr[s + 1] - L.sl_ptr[s]) / 64;
      for (int64_t k = 0; k < w; k++) {
        for (int l = 0; l < 64; l++) {
          int64_t a = s * 64 + l;
          int64_t len = a < n ? L.row_ptr[a + 1] - L.row_ptr[a] : 0;
          int32_t self = a < n ? int32_t(a) : 0;
          L.colS[L.sl_ptr[s] + 64 * k + l] =
              k < len ? L.col[L.row_ptr[a] + k] : self;
        }
      }
    }
  });

  // --- halos of 256-node blocks ----------------------------------------------
  L.nblocks = (nslices + 3) / 4;
  build_halos(n, nslices, L.sl_ptr.data(), L.colS.data(), 4, L.halo_ptr,
              L.halo, L.colL, L.halo_max);
}

void build_halos(int64_t n, int64_t nslices, const int64_t* sl_ptr,
                 const int32_t* colS, int bs, std::vector<int64_t>& halo_ptr,
                 std::vector<int32_t>& halo, std::vector<uint16_t>& colL,
                 int64_t& halo_max) {
  const int64_t nblocks = (nslices + bs - 1) / bs;
  const int64_t nent = sl_ptr[nslices];
  halo_ptr.assign(nblocks + 1, 0);
  colL.resize(nent);
  std::vector<std::vector<int32_t>> hb(nblocks);
  parallel_chunks(nblocks, [&](int64_t b0, int64_t b1) {
    std::vector<int32_t> h;
    for (int64_t b = b0; b < b1; b++) {
      int64_t s0 = int64_t(bs) * b, s1 = std::min(nslices, s0 + bs);
      int64_t e0 = sl_ptr[s0], e1 = sl_ptr[s1];
      h.assign(colS + e0, colS + e1);
      for (int64_t a = s0 * 64; a < std::min(n, s1 * 64); a++)
        h.push_back(int32_t(a));
      std::sort(h.begin(), h.end());
      h.erase(std::unique(h.begin(), h.end()), h.end());
      for (int64_t e = e0; e < e1; e++) {
        auto it = std::lower_bound(h.begin(), h.end(), colS[e]);
        colL[e] = uint16_t(std::min<int64_t>(it - h.begin(), 65535));
      }
      hb[b] = h;
    }
  });
  halo_max = 0;
  for (int64_t b = 0; b < nblocks; b++) {
    halo_max = std::max<int64_t>(halo_max, int64_t(hb[b].size()));
    halo_ptr[b + 1] = halo_ptr[b] + int64_t(hb[b].size());
  }
  halo.resize(halo_ptr[nblocks]);
  parallel_chunks(nblocks, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; b++)
      std::copy(hb[b].begin(), hb[b].end(), halo.begin() + halo_ptr[b]);
  });
}

}  // namespace nlm
