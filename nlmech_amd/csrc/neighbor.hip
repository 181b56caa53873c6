// neighbor.hip -- GPU cell-list neighbor search.
//
// Replaces geometry::Neighbor::Neighbor (src/geometry/neighbor.cpp:24-92).
// The reference tests all N^2 pairs; here the candidates of node i come from
// the 3^dim cells around it on a uniform grid whose edge exceeds the largest
// accepted distance, and the acceptance predicate is the reference's own
// (definitelyLessThan(xi.dist(xj), horizon + 1e-10), neighbor.cpp:81-83,
// evaluated in the same double operation order).  Each row is then sorted so
// the lists equal the brute-force ones (ascending j).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "device_common.h"
#include "nlm_internal.h"

namespace nlm {
namespace {

constexpr int kMaxRow = 1024;  // longest neighbor row the builder accepts

struct Grid {
  double lo[3];
  double cs;
  int64_t nc[3];
};

__device__ __forceinline__ void cell_of(const Grid& g, double px, double py,
                                        double pz, int64_t* c) {
  double p[3] = {px, py, pz};
  for (int d = 0; d < 3; d++) {
    int64_t k = int64_t((p[d] - g.lo[d]) / g.cs);
    k = k < 0 ? 0 : k;
    c[d] = k >= g.nc[d] ? g.nc[d] - 1 : k;
  }
}

__global__ void k_cell_count(int64_t n, const double* __restrict__ x, Grid g,
                             int32_t* __restrict__ cell_id,
                             int32_t* __restrict__ count) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c[3];
  cell_of(g, x[3 * i], x[3 * i + 1], x[3 * i + 2], c);
  int64_t cid = (c[2] * g.nc[1] + c[1]) * g.nc[0] + c[0];
  cell_id[i] = int32_t(cid);
  atomicAdd(&count[cid], 1);
}

__global__ void k_cell_fill(int64_t n, const int32_t* __restrict__ cell_id,
                            const int64_t* __restrict__ cell_start,
                            int32_t* __restrict__ fill,
                            int32_t* __restrict__ members) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t cid = cell_id[i];
  int32_t pos = atomicAdd(&fill[cid], 1);
  members[cell_start[cid] + pos] = int32_t(i);
}

// Exclusive scan of n int32 counts into n+1 int64 offsets, one workgroup of
// 1024 threads walking the array in chunks (setup path only).
__global__ __launch_bounds__(1024) void k_scan(int64_t n,
                                               const int32_t* __restrict__ in,
                                               int64_t* __restrict__ out) {
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    int64_t idx = base + threadIdx.x;
    int64_t v = idx < n ? int64_t(in[idx]) : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int64_t t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (idx < n) out[idx + 1] = carry + buf[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = 0;
}

// One wavefront (= one workgroup of 64 lanes) per node.  COUNT: row length.
// FILL: rows written sorted ascending (bitonic sort in LDS).
template <bool FILL>
__global__ __launch_bounds__(64) void k_nbr_rows(
    int64_t n, int dim, const double* __restrict__ x, Grid g, double horizon,
    const int64_t* __restrict__ cell_start,
    const int32_t* __restrict__ members, int32_t* __restrict__ rowlen,
    const int64_t* __restrict__ row_ptr, int32_t* __restrict__ col,
    int32_t* __restrict__ overflow) {
  __shared__ int32_t row[kMaxRow];
  int64_t i = blockIdx.x;
  if (i >= n) return;
  int lane = threadIdx.x;
  double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  int64_t c[3];
  cell_of(g, xi0, xi1, xi2, c);
  const double lim = horizon + 1.0E-10;
  int count = 0;
  int zr = dim == 3 ? 1 : 0;
  int yr = dim >= 2 ? 1 : 0;
  for (int64_t dz = -zr; dz <= zr; dz++)
    for (int64_t dy = -yr; dy <= yr; dy++)
      for (int64_t dx = -1; dx <= 1; dx++) {
        int64_t cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= g.nc[0] || cy >= g.nc[1] ||
            cz >= g.nc[2])
          continue;
        int64_t cc = (cz * g.nc[1] + cy) * g.nc[0] + cx;
        int64_t s = cell_start[cc], e = cell_start[cc + 1];
        for (int64_t b = s; b < e; b += 64) {
          int64_t idx = b + lane;
          bool ok = false;
          int32_t j = -1;
          if (idx < e) {
            j = members[idx];
            if (j != i) {
              // Point3::dist, src/util/point.h:165-169
              double a0 = xi0 - x[3 * int64_t(j)];
              double a1 = xi1 - x[3 * int64_t(j) + 1];
              double a2 = xi2 - x[3 * int64_t(j) + 2];
              double d = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
              ok = def_less(d, lim);
            }
          }
          uint64_t m = __ballot(ok);
          if (FILL) {
            int pos = count + __popcll(m & ((1ull << lane) - 1ull));
            if (ok && pos < kMaxRow) row[pos] = j;
          }
          count += __popcll(m);
        }
      }
  if (!FILL) {
    if (lane == 0) {
      rowlen[i] = count;
      if (count > kMaxRow) atomicAdd(overflow, 1);
    }
    return;
  }
  if (count > kMaxRow) return;  // reported by the COUNT pass
  int P = 64;
  while (P < count) P <<= 1;
  for (int t = count + lane; t < P; t += 64) row[t] = 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int t = lane; t < P; t += 64) {
        int ixj = t ^ jj;
        if (ixj > t) {
          int32_t a = row[t], b = row[ixj];
          bool up = (t & k) == 0;
          if ((a > b) == up) {
            row[t] = b;
            row[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  int64_t base = row_ptr[i];
  for (int t = lane; t < count; t += 64) col[base + t] = row[t];
}

}  // namespace

}  // namespace nlm

using namespace nlm;

extern "C" int nlm_neighbor_build(nlm_ctx* c, int dim, int64_t n,
                                  const double* x, double horizon,
                                  int64_t* nnz_out) {
  if (!c) return NLM_ERR_INVALID;
  if (n <= 0 || !x || dim < 1 || dim > 3 || !(horizon > 0.))
    return fail(c, NLM_ERR_INVALID, "nlm_neighbor_build: bad arguments");
  if (n >= (int64_t(1) << 31))
    return fail(c, NLM_ERR_INVALID, "nlm_neighbor_build: n >= 2^31");
  c->nbr.ready = false;
  // Uniform grid: edge cs exceeds every accepted distance (< horizon + 1e-10)
  // with a relative margin, so accepted pairs lie in adjacent cells.
  Grid g;
  double hi[3];
  for (int d = 0; d < 3; d++) {
    g.lo[d] = 1e300;
    hi[d] = -1e300;
  }
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      g.lo[d] = std::min(g.lo[d], x[3 * i + d]);
      hi[d] = std::max(hi[d], x[3 * i + d]);
    }
  g.cs = (horizon + 1.0E-10) * (1.0 + 1.0E-6);
  int64_t ncell = 1;
  for (int d = 0; d < 3; d++) {
    g.nc[d] = int64_t((hi[d] - g.lo[d]) / g.cs) + 1;
    ncell *= g.nc[d];
  }
  if (ncell >= (int64_t(1) << 31))
    return fail(c, NLM_ERR_INVALID, "nlm_neighbor_build: too many cells");
  hipStream_t st = c->stream;
  int rc;
  DevBuf dx, cell_id, count, cell_start, fill, members, rowlen, overflow;
  if ((rc = alloc(c, dx, sizeof(double) * 3 * n))) return rc;
  if ((rc = alloc(c, cell_id, sizeof(int32_t) * n))) return rc;
  if ((rc = alloc(c, count, sizeof(int32_t) * ncell))) return rc;
  if ((rc = alloc(c, cell_start, sizeof(int64_t) * (ncell + 1)))) return rc;
  if ((rc = alloc(c, fill, sizeof(int32_t) * ncell))) return rc;
  if ((rc = alloc(c, members, sizeof(int32_t) * n))) return rc;
  if ((rc = alloc(c, rowlen, sizeof(int32_t) * n))) return rc;
  if ((rc = alloc(c, overflow, sizeof(int32_t)))) return rc;
  if ((rc = alloc(c, c->nbr.row_ptr, sizeof(int64_t) * (n + 1)))) return rc;
  NLM_HIP(c, hipMemcpyAsync(dx.p, x, sizeof(double) * 3 * n,
                            hipMemcpyHostToDevice, st));
  NLM_HIP(c, hipMemsetAsync(count.p, 0, count.bytes, st));
  NLM_HIP(c, hipMemsetAsync(fill.p, 0, fill.bytes, st));
  NLM_HIP(c, hipMemsetAsync(overflow.p, 0, overflow.bytes, st));
  unsigned gb = unsigned((n + 255) / 256);
  hipLaunchKernelGGL(k_cell_count, dim3(gb), dim3(256), 0, st, n,
                     dx.as<double>(), g, cell_id.as<int32_t>(),
                     count.as<int32_t>());
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, ncell,
                     count.as<int32_t>(), cell_start.as<int64_t>());
  hipLaunchKernelGGL(k_cell_fill, dim3(gb), dim3(256), 0, st, n,
                     cell_id.as<int32_t>(), cell_start.as<int64_t>(),
                     fill.as<int32_t>(), members.as<int32_t>());
  hipLaunchKernelGGL(k_nbr_rows<false>, dim3(unsigned(n)), dim3(64), 0, st, n,
                     dim, dx.as<double>(), g, horizon,
                     cell_start.as<int64_t>(), members.as<int32_t>(),
                     rowlen.as<int32_t>(), (const int64_t*)nullptr,
                     (int32_t*)nullptr, overflow.as<int32_t>());
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, n,
                     rowlen.as<int32_t>(), c->nbr.row_ptr.as<int64_t>());
  NLM_HIP(c, hipGetLastError());
  int64_t nnz = 0;
  int32_t ovf = 0;
  NLM_HIP(c, hipMemcpyAsync(&nnz, c->nbr.row_ptr.as<int64_t>() + n,
                            sizeof(int64_t), hipMemcpyDeviceToHost, st));
  NLM_HIP(c, hipMemcpyAsync(&ovf, overflow.p, sizeof(int32_t),
                            hipMemcpyDeviceToHost, st));
  NLM_HIP(c, hipStreamSynchronize(st));
  if (ovf)
    return fail(c, NLM_ERR_INVALID,
                "nlm_neighbor_build: a node has more than 1024 neighbors");
  if ((rc = alloc(c, c->nbr.col, sizeof(int32_t) * (nnz + 1)))) return rc;
  hipLaunchKernelGGL(k_nbr_rows<true>, dim3(unsigned(n)), dim3(64), 0, st, n,
                     dim, dx.as<double>(), g, horizon,
                     cell_start.as<int64_t>(), members.as<int32_t>(),
                     rowlen.as<int32_t>(), c->nbr.row_ptr.as<int64_t>(),
                     c->nbr.col.as<int32_t>(), overflow.as<int32_t>());
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(st));
  c->nbr.n = n;
  c->nbr.nnz = nnz;
  c->nbr.ready = true;
  if (nnz_out) *nnz_out = nnz;
  return NLM_OK;
}

extern "C" int nlm_neighbor_get(nlm_ctx* c, int64_t* row_ptr, int32_t* col) {
  if (!c) return NLM_ERR_INVALID;
  if (!c->nbr.ready)
    return fail(c, NLM_ERR_STATE, "nlm_neighbor_get: no neighbor lists built");
  if (row_ptr)
    NLM_HIP(c, hipMemcpyAsync(row_ptr, c->nbr.row_ptr.p,
                              sizeof(int64_t) * (c->nbr.n + 1),
                              hipMemcpyDeviceToHost, c->stream));
  if (col && c->nbr.nnz > 0)
    NLM_HIP(c, hipMemcpyAsync(col, c->nbr.col.p, sizeof(int32_t) * c->nbr.nnz,
                              hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}
