// layout.h -- host construction of the device neighbor-list layout.
#pragma once

#include <cstdint>
#include <vector>

namespace nlm {

struct HostLayout {
  std::vector<int32_t> perm;     // internal -> reference id
  std::vector<int32_t> inv;      // reference -> internal id
  std::vector<int64_t> row_ptr;  // internal CSR, rows in reference order
  std::vector<int32_t> col;      // internal ids
  std::vector<int64_t> ref_off;  // per internal node: its row start in the
                                 // reference-order CSR
  int64_t nslices = 0;
  std::vector<int64_t> sl_ptr;   // slice starts (entries, multiples of 64)
  std::vector<int32_t> colS;     // slice entries, internal ids
  std::vector<uint16_t> rowlen;
  int64_t nblocks = 0;           // blocks of 4 slices (256 nodes)
  std::vector<int64_t> halo_ptr;
  std::vector<int32_t> halo;     // per block: sorted internal ids
  std::vector<uint16_t> colL;    // slice entries, halo-local indices
  int64_t halo_max = 0;
};

// Builds the layout for n nodes at x (AoS Point3) with the reference-order
// CSR (rp_ref, col_ref); cell is the Morton cell edge (the mesh size).
void build_layout(int dim, int64_t n, const double* x, double cell,
                  const int64_t* rp_ref, const int32_t* col_ref,
                  HostLayout& L);

// Halos of blocks of bs consecutive slices (64 bs nodes, one workgroup of
// the staged step kernel): each block's sorted internal ids referenced by its
// rows plus its own nodes (halo_ptr / halo), the halo-local index of every
// slice entry (colL, clamped to 65535) and the largest halo.
void build_halos(int64_t n, int64_t nslices, const int64_t* sl_ptr,
                 const int32_t* colS, int bs, std::vector<int64_t>& halo_ptr,
                 std::vector<int32_t>& halo, std::vector<uint16_t>& colL,
                 int64_t& halo_max);

}  // namespace nlm
