// gnn.hip -- graph network that predicts the sparse approximate-inverse
// preconditioner of the quasi-static tangent, with its training.
//
// Graph: the neighbor graph in the slice layout (an edge is a slice entry
// (i, slot k), j = col; padding entries are masked).  Layer products are
// dense [rows x K] x [K x 64] matrix products on the matrix cores
// (gfx950's v_mfma_f32_32x32x16_bf16: BF16 inputs, FP32 accumulation, K = 16
// per instruction); one wavefront computes a 32-row x 64-column tile.  Weight
// gradients X^T dY reduce over the rows with the same instruction on
// LDS-staged tiles, per workgroup, and a second kernel adds the workgroups'
// partials in a fixed order.
//
// Lane layout of v_mfma_f32_32x32x16_bf16 (64 lanes): A[i][k] from lane
// i + 32 (k / 8), element k % 8; B[k][j] from lane j + 32 (k / 8), element
// k % 8; D[i][j] in lane j + 32 ((i / 4) % 2), element (i % 4) + 4 (i / 8).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fd_common.h"
#include "nlm_internal.h"

namespace nlm {
namespace {

constexpr int kL = NLM_GNN_LAYERS;
constexpr int kH = NLM_GNN_WIDTH;
constexpr int kFn = 8;  // node input features
constexpr int kFe = 8;  // edge input features

// FP32 parameter offsets.
constexpr int64_t kOffWn = 0, kOffBn = 512, kOffWe = 576, kOffBe = 1088;
constexpr int64_t kOffLayer = 1152, kLayerSize = 20608;
constexpr int64_t kLWa = 0, kLWb = 4096, kLWc = 8192, kLBm = 12288,
                  kLWd = 12352, kLWf = 16448, kLBu = 20544;
constexpr int64_t kOffWoe = kOffLayer + kL * kLayerSize;  // [64][4]
constexpr int64_t kOffBoe = kOffWoe + 256;
constexpr int64_t kOffWon = kOffBoe + 4;
constexpr int64_t kOffBon = kOffWon + 256;
constexpr int64_t kNumParams = kOffBon + 4;

// BF16 copies: forward (transposed, [64][K]) and backward ([K][64]).
constexpr int64_t kFwdWn = 0, kFwdWe = 512, kFwdLayer = 1024,
                  kFwdLayerSize = 20480;
constexpr int64_t kFWa = 0, kFWb = 4096, kFWc = 8192, kFWu = 12288;
constexpr int64_t kFwdSize = kFwdLayer + kL * kFwdLayerSize;
constexpr int64_t kBwdLayerSize = 20480;
constexpr int64_t kBWc = 0, kBWd = 4096, kBWf = 8192, kBWab = 12288;
constexpr int64_t kBwdSize = kL * kBwdLayerSize;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  uint32_t u = uint32_t(b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ bool fixed_dof(uint32_t fl, int d) {
  return ((fl >> (kFlagUbcShift + 4 * d)) & 0xfu) != 0u;
}

// ---- matrix products --------------------------------------------------------

struct MmArgs {
  int64_t M;
  int K;                       // 8, 64 or 128 (two 64-wide halves)
  const uint16_t* X1;          // [M][min(K, 64)]
  const uint16_t* X2;          // [M][64] second half when K == 128
  const uint16_t* Wt;          // [64][K]
  const float* bias;           // [64] or null
  const float* P;              // gather-added rows P[prow[r]] (or null)
  const float* Q;              // gather-added rows Q[qrow[r]] (or null)
  const int32_t* prow;
  const int32_t* qrow;
  const uint8_t* valid;        // row mask (or null)
  int relu;
  int accumulate;              // Yf += result
  uint16_t* Ybf;               // [M][64] (or null)
  float* Yf;                   // [M][64] (or null)
};

__device__ __forceinline__ v4s load4(const uint16_t* p) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  v4s r;
  r[0] = short(v.x & 0xffffu);
  r[1] = short(v.x >> 16);
  r[2] = short(v.y & 0xffffu);
  r[3] = short(v.y >> 16);
  return r;
}

// Eight consecutive BF16 values (16-byte aligned).
__device__ __forceinline__ v8bf load8(const uint16_t* p) {
  return __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(p));
}
__device__ __forceinline__ v8bf zero8() {
  return __builtin_bit_cast(v8bf, make_uint4(0u, 0u, 0u, 0u));
}

// Y = act(X W + bias + P[prow] + Q[qrow]) masked; one wavefront per 32 rows.
__global__ __launch_bounds__(256) void k_mm(MmArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t r0 = (int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 32;
  if (r0 >= a.M) return;
  v16f acc0, acc1;
  for (int t = 0; t < 16; t++) {
    acc0[t] = 0.f;
    acc1[t] = 0.f;
  }
  const int64_t arow = r0 + (lane & 31);
  const bool arow_ok = arow < a.M;
  const int kq = 8 * (lane >> 5);
  const int ld1 = a.K < 64 ? a.K : 64;
  // K = 8: the upper half of the 16-deep step is zero
  for (int k0 = 0; k0 < a.K; k0 += 16) {
    const int kk = k0 + kq;
    const bool kin = kk < a.K;
    v8bf av = zero8(), b0 = zero8(), b1 = zero8();
    if (arow_ok && kin) {
      if (kk < 64)
        av = load8(a.X1 + arow * ld1 + kk);
      else
        av = load8(a.X2 + arow * 64 + (kk - 64));
    }
    if (kin) {
      b0 = load8(a.Wt + int64_t(lane & 31) * a.K + kk);
      b1 = load8(a.Wt + int64_t(32 + (lane & 31)) * a.K + kk);
    }
    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b0, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b1, acc1, 0, 0, 0);
  }
  for (int t = 0; t < 16; t++) {
    const int64_t row = r0 + 8 * (t >> 2) + 4 * (lane >> 5) + (t & 3);
    if (row >= a.M) continue;
    const bool ok = a.valid ? a.valid[row] != 0 : true;
    const int64_t pr = a.P ? int64_t(a.prow ? a.prow[row] : row) : 0;
    const int64_t qr = a.Q ? int64_t(a.qrow[row]) : 0;
    for (int half = 0; half < 2; half++) {
      const int col = (lane & 31) + 32 * half;
      float y = half ? acc1[t] : acc0[t];
      if (a.bias) y += a.bias[col];
      if (a.P) y += a.P[pr * 64 + col];
      if (a.Q) y += a.Q[qr * 64 + col];
      if (a.relu) y = fmaxf(y, 0.f);
      if (!ok) y = 0.f;
      const int64_t o = row * 64 + col;
      if (a.Ybf) a.Ybf[o] = f2bf(y);
      if (a.Yf) a.Yf[o] = a.accumulate ? a.Yf[o] + y : y;
    }
  }
}

// Partial X^T dY ([K][64]) of a workgroup over row chunks of 64 (grid-stride),
// K = 8, 64 or 128 (halves X1, X2); dY is [M][64].
constexpr int kTnRows = 64;
__global__ __launch_bounds__(256) void k_mm_tn(int64_t M, int K,
                                               const uint16_t* __restrict__ X1,
                                               const uint16_t* __restrict__ X2,
                                               const uint16_t* __restrict__ dY,
                                               float* __restrict__ part) {
  __shared__ uint16_t sX[kTnRows * 128];
  __shared__ uint16_t sD[kTnRows * 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ktiles = (K + 31) / 32;      // output row tiles of 32
  const int ntiles = ktiles * 2;         // x 2 column tiles
  const int ld1 = K < 64 ? K : 64;
  v16f acc[2];
  for (int u = 0; u < 2; u++)
    for (int t = 0; t < 16; t++) acc[u][t] = 0.f;
  const int64_t nchunk = (M + kTnRows - 1) / kTnRows;
  for (int64_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    const int64_t base = ch * kTnRows;
    // stage X (K columns, zero padded to 32 for K = 8) and dY
    for (int idx = threadIdx.x; idx < kTnRows * K; idx += 256) {
      const int r = idx / K, k = idx % K;
      const int64_t row = base + r;
      uint16_t v = 0;
      if (row < M)
        v = k < 64 ? X1[row * ld1 + k] : X2[row * 64 + (k - 64)];
      sX[r * 128 + k] = v;
    }
    if (K < 32)
      for (int idx = threadIdx.x; idx < kTnRows * 32; idx += 256) {
        const int r = idx / 32, k = idx % 32;
        if (k >= K) sX[r * 128 + k] = 0;
      }
    for (int idx = threadIdx.x; idx < kTnRows * 64; idx += 256) {
      const int r = idx / 64, c = idx % 64;
      const int64_t row = base + r;
      sD[r * 64 + c] = row < M ? dY[row * 64 + c] : uint16_t(0);
    }
    __syncthreads();
    for (int u = 0; u < 2; u++) {
      const int tile = wave + 4 * u;
      if (tile >= ntiles) continue;
      const int kt = tile >> 1, nt = tile & 1;
      for (int kd = 0; kd < kTnRows; kd += 16) {
        uint16_t av[8], bv[8];
        for (int e = 0; e < 8; e++) {
          const int r = kd + 8 * (lane >> 5) + e;
          av[e] = sX[r * 128 + kt * 32 + (lane & 31)];
          bv[e] = sD[r * 64 + nt * 32 + (lane & 31)];
        }
        acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            __builtin_bit_cast(v8bf, av), __builtin_bit_cast(v8bf, bv), acc[u],
            0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* out = part + int64_t(blockIdx.x) * K * 64;
  for (int u = 0; u < 2; u++) {
    const int tile = wave + 4 * u;
    if (tile >= ntiles) continue;
    const int kt = tile >> 1, nt = tile & 1;
    for (int t = 0; t < 16; t++) {
      const int i = kt * 32 + 8 * (t >> 2) + 4 * (lane >> 5) + (t & 3);
      const int j = nt * 32 + (lane & 31);
      if (i < K) out[i * 64 + j] = acc[u][t];
    }
  }
}

// grad[off + t] += sum_b part[b][t], t < size (fixed order).
// A workgroup owns 16 entries; its 16 thread groups sum every 16th part
// and the groups' sums are added in group order.
constexpr int kRedCols = 16;
__global__ __launch_bounds__(256) void k_reduce_add(
    const float* __restrict__ part, int nparts, int size,
    float* __restrict__ grad) {
  __shared__ float red[16][kRedCols + 1];
  const int tc = threadIdx.x % kRedCols, gq = threadIdx.x / kRedCols;
  const int t = blockIdx.x * kRedCols + tc;
  float s = 0.f;
  if (t < size)
    for (int b = gq; b < nparts; b += 16) s += part[int64_t(b) * size + t];
  red[gq][tc] = s;
  __syncthreads();
  if (gq == 0 && t < size) {
    float r = 0.f;
    for (int k = 0; k < 16; k++) r += red[k][tc];
    grad[t] += r;
  }
}

// Column sums of dY [M][64] (BF16 or FP32) into part[block][64].
template <class T>
__global__ __launch_bounds__(256) void k_colsum(int64_t M,
                                                const T* __restrict__ dY,
                                                float* __restrict__ part) {
  __shared__ float red[256];
  const int c = threadIdx.x & 63, rq = threadIdx.x >> 6;
  float s = 0.f;
  for (int64_t r = int64_t(blockIdx.x) * 4 + rq; r < M; r += int64_t(gridDim.x) * 4) {
    if constexpr (sizeof(T) == 2)
      s += bf2f(dY[r * 64 + c]);
    else
      s += dY[r * 64 + c];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < 64)
    part[int64_t(blockIdx.x) * 64 + c] =
        red[c] + red[64 + c] + red[128 + c] + red[192 + c];
}

// ---- graph kernels ---------------------------------------------------------

struct GraphArgs {
  int64_t n, nslices;
  const int64_t* sl_ptr;
  const int32_t* col;
  const uint16_t* rowlen;
  const uint32_t* flags;
  const double2* x;
  const double4* blk;   // tangent blocks
  const double4* diag;
  double horizon;
};

__device__ __forceinline__ bool slice_lane(int64_t nslices, int64_t& s,
                                           int& lane) {
  s = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  lane = threadIdx.x & 63;
  return s < nslices;
}

// Reverse slice entry of every bond: rev[e(i, k)] = e(j, k') with j the
// neighbor in slot k of node i and i the neighbor in slot k' of node j
// (a scan of row j: rows keep the reference's order, not the internal one).
// A missing reverse bond sets *bad.
__global__ __launch_bounds__(256) void k_gnn_rev(
    int64_t n, int64_t nslices, const int64_t* __restrict__ sl_ptr,
    const int32_t* __restrict__ col, const uint16_t* __restrict__ rowlen,
    int32_t* __restrict__ rev, int* __restrict__ bad) {
  int64_t sl;
  int lane;
  if (!slice_lane(nslices, sl, lane)) return;
  const int64_t i = sl * 64 + lane;
  if (i >= n) return;
  const int64_t cb = sl_ptr[sl];
  const int len = rowlen[i];
  for (int k = 0; k < len; k++) {
    const int64_t e = cb + 64 * int64_t(k) + lane;
    const int32_t j = col[e];
    const int64_t cj = sl_ptr[j >> 6] + (j & 63);
    const int lj = rowlen[j];
    int kk = 0;
    while (kk < lj && col[cj + 64 * int64_t(kk)] != int32_t(i)) kk++;
    const bool ok = kk < lj;
    rev[e] = ok ? int32_t(cj + 64 * int64_t(kk)) : 0;
    if (!ok) bad[0] = 1;
  }
}

// Node scale s_i = (A_ii[0][0] + A_ii[1][1]) / 2, 1/sqrt(s_i), degree.
__global__ __launch_bounds__(256) void k_gnn_scale(GraphArgs g,
                                                   double* __restrict__ sc,
                                                   float* __restrict__ isq,
                                                   float* __restrict__ deg) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= g.n) return;
  const double4 D = g.diag[i];
  double s = 0.5 * (D.x + D.w);
  if (!(s > 0.)) s = 1.;
  sc[i] = s;
  isq[i] = float(1. / sqrt(s));
  deg[i] = float(g.rowlen[i]);
}

// Input features of nodes and edges, edge row index and mask.
__global__ __launch_bounds__(256) void k_gnn_feat(
    GraphArgs g, const double* __restrict__ sc, uint16_t* __restrict__ xn,
    uint16_t* __restrict__ xe, int32_t* __restrict__ erow,
    uint8_t* __restrict__ valid) {
  int64_t s;
  int lane;
  if (!slice_lane(g.nslices, s, lane)) return;
  const int64_t i = s * 64 + lane;
  const bool ok = i < g.n;
  const int64_t cb = g.sl_ptr[s];
  const int width = int((g.sl_ptr[s + 1] - cb) >> 6);
  const int len = ok ? int(g.rowlen[i]) : 0;
  const double si = ok ? sc[i] : 1.;
  const double2 xi = ok ? g.x[i] : make_double2(0., 0.);
  for (int k = 0; k < width; k++) {
    const int64_t e = cb + int64_t(k) * 64 + lane;
    float f[kFe] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool v = k < len;
    if (v) {
      const int32_t j = g.col[e];
      const double4 B = g.blk[e];
      const double w = 1. / sqrt(si * sc[j]);
      const double2 xj = g.x[j];
      const double dx = (xj.x - xi.x) / g.horizon, dy = (xj.y - xi.y) / g.horizon;
      f[0] = float(B.x * w);
      f[1] = float(B.y * w);
      f[2] = float(B.z * w);
      f[3] = float(B.w * w);
      f[4] = float(dx);
      f[5] = float(dy);
      f[6] = float(sqrt(dx * dx + dy * dy));
      f[7] = 1.f;
    }
    for (int t = 0; t < kFe; t++) xe[e * kFe + t] = f2bf(f[t]);
    erow[e] = ok ? int32_t(i) : 0;
    valid[e] = v ? 1 : 0;
  }
  if (!ok) return;
  const double4 D = g.diag[i];
  const uint32_t fl = g.flags[i];
  float f[kFn] = {float(D.x / si), float(D.y / si), float(D.z / si),
                  float(D.w / si), fixed_dof(fl, 0) ? 1.f : 0.f,
                  fixed_dof(fl, 1) ? 1.f : 0.f, float(len) / 32.f, 1.f};
  for (int t = 0; t < kFn; t++) xn[i * kFn + t] = f2bf(f[t]);
}

// agg_i = mean over the valid slots of m (BF16 [E][64]).
__global__ __launch_bounds__(256) void k_gnn_agg(GraphArgs g,
                                                 const uint16_t* __restrict__ m,
                                                 uint16_t* __restrict__ agg) {
  int64_t s;
  int lane;
  if (!slice_lane(g.nslices, s, lane)) return;
  const int64_t i = s * 64 + lane;
  if (i >= g.n) return;
  const int64_t cb = g.sl_ptr[s];
  const int len = int(g.rowlen[i]);
  float acc[64];
  for (int c = 0; c < 64; c++) acc[c] = 0.f;
  for (int k = 0; k < len; k++) {
    const uint16_t* row = m + (cb + int64_t(k) * 64 + lane) * 64;
    for (int c = 0; c < 64; c += 4) {
      const v4s v = load4(row + c);
      for (int t = 0; t < 4; t++) acc[c + t] += bf2f(uint16_t(v[t]));
    }
  }
  const float inv = len > 0 ? 1.f / float(len) : 0.f;
  for (int c = 0; c < 64; c++) agg[i * 64 + c] = f2bf(acc[c] * inv);
}

struct DecArgs {
  GraphArgs g;
  const float* w;        // FP32 master weights
  const uint16_t* gL;    // [E][64]
  const uint16_t* hL;    // [N][64]
  const float* isq;
  float4* mblk;
  float4* mdiag;
};

// 2x2 mask of the preconditioner block (i rows, j columns): prescribed dofs
// have neither rows nor columns.
__device__ __forceinline__ void block_mask(uint32_t fi, uint32_t fj,
                                           float* mk) {
  const bool ri0 = !fixed_dof(fi, 0), ri1 = !fixed_dof(fi, 1);
  const bool cj0 = !fixed_dof(fj, 0), cj1 = !fixed_dof(fj, 1);
  mk[0] = ri0 && cj0;
  mk[1] = ri0 && cj1;
  mk[2] = ri1 && cj0;
  mk[3] = ri1 && cj1;
}

// Decoders: edge block Me = g_L Woe + boe, node block Mn = h_L Won + bon + I,
// scaled by 1/sqrt(s_i s_j) (1/s_i) and masked.
__global__ __launch_bounds__(256) void k_gnn_decode(DecArgs a) {
  int64_t s;
  int lane;
  if (!slice_lane(a.g.nslices, s, lane)) return;
  const int64_t i = s * 64 + lane;
  const bool ok = i < a.g.n;
  const int64_t cb = a.g.sl_ptr[s];
  const int width = int((a.g.sl_ptr[s + 1] - cb) >> 6);
  const int len = ok ? int(a.g.rowlen[i]) : 0;
  const float* Woe = a.w + kOffWoe;
  const float* boe = a.w + kOffBoe;
  const uint32_t fi = ok ? a.g.flags[i] : 0u;
  const float isqi = ok ? a.isq[i] : 0.f;
  for (int k = 0; k < width; k++) {
    const int64_t e = cb + int64_t(k) * 64 + lane;
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k < len) {
      const int32_t j = a.g.col[e];
      float o[4] = {boe[0], boe[1], boe[2], boe[3]};
      const uint16_t* row = a.gL + e * 64;
      for (int c = 0; c < 64; c++) {
        const float v = bf2f(row[c]);
        for (int t = 0; t < 4; t++) o[t] += v * Woe[c * 4 + t];
      }
      float mk[4];
      block_mask(fi, a.g.flags[j], mk);
      const double w = double(isqi) * double(a.isq[j]);
      out = make_float4(float(o[0] * mk[0] * w), float(o[1] * mk[1] * w),
                        float(o[2] * mk[2] * w), float(o[3] * mk[3] * w));
    }
    a.mblk[e] = out;
  }
  if (!ok) return;
  const float* Won = a.w + kOffWon;
  const float* bon = a.w + kOffBon;
  float o[4] = {bon[0] + 1.f, bon[1], bon[2], bon[3] + 1.f};
  const uint16_t* row = a.hL + i * 64;
  for (int c = 0; c < 64; c++) {
    const float v = bf2f(row[c]);
    for (int t = 0; t < 4; t++) o[t] += v * Won[c * 4 + t];
  }
  float mk[4];
  block_mask(fi, fi, mk);
  const double w = double(isqi) * double(isqi);
  a.mdiag[i] = make_float4(float(o[0] * mk[0] * w), float(o[1] * mk[1] * w),
                           float(o[2] * mk[2] * w), float(o[3] * mk[3] * w));
}

// ---- loss ------------------------------------------------------------------

// z_k = M r_k for K vectors (double2 per node, K blocks of n).
__global__ __launch_bounds__(256) void k_gnn_spmm(
    GraphArgs g, const float4* __restrict__ mblk,
    const float4* __restrict__ mdiag, int K, const double2* __restrict__ r,
    double2* __restrict__ z) {
  int64_t s;
  int lane;
  if (!slice_lane(g.nslices, s, lane)) return;
  const int64_t i = s * 64 + lane;
  if (i >= g.n) return;
  const int64_t cb = g.sl_ptr[s];
  const int len = int(g.rowlen[i]);
  for (int kv = 0; kv < K; kv++) {
    const double2* rv = r + int64_t(kv) * g.n;
    const float4 D = mdiag[i];
    const double2 ri = rv[i];
    double ax = D.x * ri.x + D.y * ri.y, ay = D.z * ri.x + D.w * ri.y;
    for (int k = 0; k < len; k++) {
      const int64_t e = cb + int64_t(k) * 64 + lane;
      const float4 B = mblk[e];
      const double2 rj = rv[g.col[e]];
      ax += B.x * rj.x + B.y * rj.y;
      ay += B.z * rj.x + B.w * rj.y;
    }
    z[int64_t(kv) * g.n + i] = make_double2(ax, ay);
  }
}

// w_k = z_k - e_k (in place in z); loss partial sum_k c_k |w_k|^2 per block.
__global__ __launch_bounds__(256) void k_gnn_resid(
    int64_t n, int K, const double* __restrict__ ck, double2* __restrict__ z,
    const double2* __restrict__ e, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double acc = 0.;
  if (i < n)
    for (int kv = 0; kv < K; kv++) {
      const int64_t o = int64_t(kv) * n + i;
      double2 w = z[o];
      const double2 ev = e[o];
      w.x -= ev.x;
      w.y -= ev.y;
      z[o] = w;
      acc += ck[kv] * (w.x * w.x + w.y * w.y);
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dL/dMe and dL/dMn: 2 sum_k c_k w_k,i r_k,j^T, masked and scaled like the
// decoder's output.
__global__ __launch_bounds__(256) void k_gnn_dm(
    GraphArgs g, int K, const double* __restrict__ ck,
    const double2* __restrict__ w, const double2* __restrict__ r,
    const float* __restrict__ isq, float* __restrict__ dme,
    float* __restrict__ dmn) {
  int64_t s;
  int lane;
  if (!slice_lane(g.nslices, s, lane)) return;
  const int64_t i = s * 64 + lane;
  const bool ok = i < g.n;
  const int64_t cb = g.sl_ptr[s];
  const int width = int((g.sl_ptr[s + 1] - cb) >> 6);
  const int len = ok ? int(g.rowlen[i]) : 0;
  const uint32_t fi = ok ? g.flags[i] : 0u;
  const float isqi = ok ? isq[i] : 0.f;
  for (int k = 0; k < width; k++) {
    const int64_t e = cb + int64_t(k) * 64 + lane;
    float o[4] = {0.f, 0.f, 0.f, 0.f};
    if (k < len) {
      const int32_t j = g.col[e];
      double d[4] = {0., 0., 0., 0.};
      for (int kv = 0; kv < K; kv++) {
        const double2 wi = w[int64_t(kv) * g.n + i];
        const double2 rj = r[int64_t(kv) * g.n + j];
        const double c2 = 2. * ck[kv];
        d[0] += c2 * wi.x * rj.x;
        d[1] += c2 * wi.x * rj.y;
        d[2] += c2 * wi.y * rj.x;
        d[3] += c2 * wi.y * rj.y;
      }
      float mk[4];
      block_mask(fi, g.flags[j], mk);
      const double sc = double(isqi) * double(isq[j]);
      for (int t = 0; t < 4; t++) o[t] = float(d[t] * sc * mk[t]);
    }
    for (int t = 0; t < 4; t++) dme[e * 4 + t] = o[t];
  }
  if (!ok) return;
  double d[4] = {0., 0., 0., 0.};
  for (int kv = 0; kv < K; kv++) {
    const double2 wi = w[int64_t(kv) * g.n + i];
    const double2 ri = r[int64_t(kv) * g.n + i];
    const double c2 = 2. * ck[kv];
    d[0] += c2 * wi.x * ri.x;
    d[1] += c2 * wi.x * ri.y;
    d[2] += c2 * wi.y * ri.x;
    d[3] += c2 * wi.y * ri.y;
  }
  float mk[4];
  block_mask(fi, fi, mk);
  const double sc = double(isqi) * double(isqi);
  for (int t = 0; t < 4; t++) dmn[i * 4 + t] = float(d[t] * sc * mk[t]);
}

// ---- backward ----------------------------------------------------------------

// Decoder backward: dX[r][c] = sum_t dM[r][t] W[c][t] over rows x 64.
template <bool OUT_F32>
__global__ __launch_bounds__(256) void k_dec_dx(int64_t M,
                                                const float* __restrict__ dM,
                                                const float* __restrict__ W,
                                                uint16_t* __restrict__ dX16,
                                                float* __restrict__ dX32) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= M * 64) return;
  const int64_t r = t >> 6;
  const int c = int(t & 63);
  const float* d = dM + r * 4;
  const float* wr = W + c * 4;
  const float v = d[0] * wr[0] + d[1] * wr[1] + d[2] * wr[2] + d[3] * wr[3];
  if (OUT_F32)
    dX32[t] = v;
  else
    dX16[t] = f2bf(v);
}

// Decoder weight gradient of a workgroup (rows b, b + G, ...): thread (c, t)
// accumulates sum_r X[r][c] dM[r][t]; part[b][c*4 + t], bias part[b][256+t].
__global__ __launch_bounds__(256) void k_dec_dw(int64_t M,
                                                const float* __restrict__ dM,
                                                const uint16_t* __restrict__ X,
                                                float* __restrict__ part) {
  const int c = threadIdx.x >> 2, t = threadIdx.x & 3;
  float acc = 0.f, accb = 0.f;
  for (int64_t r = blockIdx.x; r < M; r += gridDim.x) {
    const float d = dM[r * 4 + t];
    acc += bf2f(X[r * 64 + c]) * d;
    accb += d;
  }
  float* out = part + int64_t(blockIdx.x) * 260;
  out[c * 4 + t] = acc;
  if (c == 0) out[256 + t] = accb;
}

// out = d * (act > 0) (BF16), d FP32 or BF16.
template <class T>
__global__ __launch_bounds__(256) void k_relu_mask(int64_t cnt,
                                                   const T* __restrict__ d,
                                                   const uint16_t* __restrict__ act,
                                                   uint16_t* __restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= cnt) return;
  float v;
  if constexpr (sizeof(T) == 2)
    v = bf2f(d[t]);
  else
    v = d[t];
  out[t] = bf2f(act[t]) > 0.f ? f2bf(v) : uint16_t(0);
}

// dpre_m[e] = (dg[e] + dagg[row(e)] / deg[row(e)]) * (m[e] > 0) * valid[e].
__global__ __launch_bounds__(256) void k_edge_dm(
    int64_t ne, const uint16_t* __restrict__ dg, const float* __restrict__ dagg,
    const int32_t* __restrict__ erow, const float* __restrict__ deg,
    const uint8_t* __restrict__ valid, const uint16_t* __restrict__ m,
    uint16_t* __restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= ne * 64) return;
  const int64_t e = t >> 6;
  const int c = int(t & 63);
  float v = 0.f;
  if (valid[e] && bf2f(m[t]) > 0.f) {
    const int32_t i = erow[e];
    v = bf2f(dg[t]) + dagg[int64_t(i) * 64 + c] / deg[i];
  }
  out[t] = f2bf(v);
}

// dP_i = sum_k dpre[e(i,k)], dQ_i = sum_k dpre[rev(e(i,k))] (FP32 and BF16).
__global__ __launch_bounds__(256) void k_dpq(GraphArgs g,
                                             const uint16_t* __restrict__ dpre,
                                             const int32_t* __restrict__ rev,
                                             float* __restrict__ dP,
                                             float* __restrict__ dQ,
                                             uint16_t* __restrict__ dP16,
                                             uint16_t* __restrict__ dQ16) {
  int64_t s;
  int lane;
  if (!slice_lane(g.nslices, s, lane)) return;
  const int64_t i = s * 64 + lane;
  if (i >= g.n) return;
  const int64_t cb = g.sl_ptr[s];
  const int len = int(g.rowlen[i]);
  for (int c0 = 0; c0 < 64; c0 += 16) {
    float ap[16], aq[16];
    for (int c = 0; c < 16; c++) ap[c] = aq[c] = 0.f;
    for (int k = 0; k < len; k++) {
      const int64_t e = cb + int64_t(k) * 64 + lane;
      const int64_t er = rev[e];
      for (int c = 0; c < 16; c += 4) {
        const v4s a = load4(dpre + e * 64 + c0 + c);
        const v4s b = load4(dpre + er * 64 + c0 + c);
        for (int t = 0; t < 4; t++) {
          ap[c + t] += bf2f(uint16_t(a[t]));
          aq[c + t] += bf2f(uint16_t(b[t]));
        }
      }
    }
    for (int c = 0; c < 16; c++) {
      dP[i * 64 + c0 + c] = ap[c];
      dQ[i * 64 + c0 + c] = aq[c];
      dP16[i * 64 + c0 + c] = f2bf(ap[c]);
      dQ16[i * 64 + c0 + c] = f2bf(aq[c]);
    }
  }
}

__global__ void k_to_bf16_copies(const float* __restrict__ w,
                                 uint16_t* __restrict__ fwd,
                                 uint16_t* __restrict__ bwd) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t < kFwdSize) {
    float v;
    if (t < kFwdLayer) {
      // Wn_t / We_t: [64][8] from W[8][64]
      const int64_t base = t < kFwdWe ? kOffWn : kOffWe;
      const int64_t o = t < kFwdWe ? t : t - kFwdWe;
      const int n = int(o / 8), k = int(o % 8);
      v = w[base + k * 64 + n];
    } else {
      const int64_t lo = t - kFwdLayer;
      const int l = int(lo / kFwdLayerSize);
      const int64_t o = lo % kFwdLayerSize;
      const int64_t L = kOffLayer + l * kLayerSize;
      if (o < kFWu) {
        const int64_t mat = o / 4096, q = o % 4096;
        const int n = int(q / 64), k = int(q % 64);
        const int64_t src = mat == 0 ? kLWa : (mat == 1 ? kLWb : kLWc);
        v = w[L + src + k * 64 + n];
      } else {
        const int64_t q = o - kFWu;  // [64][128]
        const int n = int(q / 128), k = int(q % 128);
        v = k < 64 ? w[L + kLWd + k * 64 + n] : w[L + kLWf + (k - 64) * 64 + n];
      }
    }
    fwd[t] = f2bf(v);
  }
  if (t < kBwdSize) {
    const int l = int(t / kBwdLayerSize);
    const int64_t o = t % kBwdLayerSize;
    const int64_t L = kOffLayer + l * kLayerSize;
    float v;
    if (o < kBWab) {
      const int64_t mat = o / 4096, q = o % 4096;
      const int64_t src = mat == 0 ? kLWc : (mat == 1 ? kLWd : kLWf);
      v = w[L + src + q];  // row-major [64][64] as stored
    } else {
      const int64_t q = o - kBWab;  // [64][128]: row n = [Wa[n][:], Wb[n][:]]
      const int n = int(q / 128), k = int(q % 128);
      v = k < 64 ? w[L + kLWa + n * 64 + k] : w[L + kLWb + n * 64 + (k - 64)];
    }
    bwd[t] = f2bf(v);
  }
}

__global__ void k_adam(int64_t np, float* __restrict__ w,
                       float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, float lr, float c1, float c2) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= np) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gr = grad[t];
  const float mt = b1 * m[t] + (1.f - b1) * gr;
  const float vt = b2 * v[t] + (1.f - b2) * gr * gr;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  grad[t] = 0.f;
}

// ---- host side -------------------------------------------------------------

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Initial weights: uniform(-a, a) with a = sqrt(6 / (fan_in + fan_out)),
// biases 0, decoders scaled by 0.1 (the network starts near point Jacobi).
std::vector<float> init_weights(uint64_t seed) {
  std::vector<float> w(kNumParams, 0.f);
  uint64_t st = seed;
  auto fill = [&](int64_t off, int fan_in, int fan_out, int cnt, float scale) {
    const double a = scale * std::sqrt(6.0 / double(fan_in + fan_out));
    for (int t = 0; t < cnt; t++) {
      const double u = double(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
      w[off + t] = float((2.0 * u - 1.0) * a);
    }
  };
  fill(kOffWn, kFn, kH, kFn * kH, 1.f);
  fill(kOffWe, kFe, kH, kFe * kH, 1.f);
  for (int l = 0; l < kL; l++) {
    const int64_t L = kOffLayer + l * kLayerSize;
    fill(L + kLWa, kH, kH, 4096, 1.f);
    fill(L + kLWb, kH, kH, 4096, 1.f);
    fill(L + kLWc, kH, kH, 4096, 1.f);
    fill(L + kLWd, 2 * kH, kH, 4096, 1.f);
    fill(L + kLWf, 2 * kH, kH, 4096, 1.f);
  }
  fill(kOffWoe, kH, 4, 256, 0.1f);
  fill(kOffWon, kH, 4, 256, 0.1f);
  return w;
}

GraphArgs graph_args(nlm_ctx* c) {
  FdState& s = c->fd;
  GraphArgs g;
  g.n = s.n;
  g.nslices = s.sell.nslices;
  g.sl_ptr = s.sell.sl_ptr.as<int64_t>();
  g.col = s.sell.col.as<int32_t>();
  g.rowlen = s.sell.rowlen.as<uint16_t>();
  g.flags = s.flags.as<uint32_t>();
  g.x = s.x.as<double2>();
  g.blk = c->qs.blk.as<double4>();
  g.diag = c->qs.diag.as<double4>();
  g.horizon = s.mat.horizon;
  return g;
}

int gnn_check(nlm_ctx* c, const char* fn) {
  if (!c) return NLM_ERR_INVALID;
  if (!c->fd.ready)
    return fail(c, NLM_ERR_STATE, std::string(fn) + ": nlm_fd_setup not done");
  if (c->fd.dim != 2 || c->fd.material != NLM_MATERIAL_RNP_BOND)
    return fail(c, NLM_ERR_INVALID,
                std::string(fn) + ": 2D bond-based material only");
  if (!c->gnn.ready)
    return fail(c, NLM_ERR_STATE, std::string(fn) + ": no network attached");
  return NLM_OK;
}

// Activation buffers and the reverse-entry map for the context's graph.
int gnn_size(nlm_ctx* c) {
  FdState& s = c->fd;
  GnnState& G = c->gnn;
  const int64_t n = s.n, ne = s.sell.nentries;
  if (G.n == n && G.ne == ne) return NLM_OK;
  int rc;
  auto A = [&](DevBuf& b, size_t bytes) { return alloc(c, b, bytes + 256); };
  if ((rc = A(G.xn, n * kFn * 2)) || (rc = A(G.xe, ne * kFe * 2)) ||
      (rc = A(G.h, (kL + 1) * n * 64 * 2)) ||
      (rc = A(G.g, (kL + 1) * ne * 64 * 2)) ||
      (rc = A(G.agg, kL * n * 64 * 2)) || (rc = A(G.P, n * 64 * 4)) ||
      (rc = A(G.Q, n * 64 * 4)) || (rc = A(G.scale, n * 8)) ||
      (rc = A(G.isq, n * 4)) || (rc = A(G.deg, n * 4)) ||
      (rc = A(G.erow, ne * 4)) || (rc = A(G.rev, ne * 4)) ||
      (rc = A(G.valid, ne)) || (rc = A(G.dh, n * 64 * 4)) ||
      (rc = A(G.dh16, n * 64 * 2)) || (rc = A(G.dg, ne * 64 * 2)) ||
      (rc = A(G.dg2, ne * 64 * 2)) || (rc = A(G.dpre, ne * 64 * 2)) ||
      (rc = A(G.dP, n * 64 * 4)) || (rc = A(G.dQ, n * 64 * 4)) ||
      (rc = A(G.dPQ16, 2 * n * 64 * 2)) || (rc = A(G.dagg, n * 64 * 4)) ||
      (rc = A(G.dme, ne * 4 * 4)) || (rc = A(G.dmn, n * 4 * 4)) ||
      (rc = A(G.lpart, ((n + 255) / 256) * 8)))
    return rc;
  // reverse entries: the slice entry of bond (j, i) for entry (i, k)
  DevBuf bad;
  if ((rc = alloc(c, bad, 256))) return rc;
  NLM_HIP(c, hipMemsetAsync(bad.p, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_gnn_rev, dim3(unsigned((s.sell.nslices + 3) / 4)),
                     dim3(256), 0, c->stream, n, s.sell.nslices,
                     s.sell.sl_ptr.as<int64_t>(), s.sell.col.as<int32_t>(),
                     s.sell.rowlen.as<uint16_t>(), G.rev.as<int32_t>(),
                     bad.as<int>());
  NLM_HIP(c, hipGetLastError());
  int hbad = 0;
  NLM_HIP(c, hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost,
                            c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  if (hbad)
    return fail(c, NLM_ERR_INVALID, "nlm_gnn: neighbor lists are not symmetric");
  G.n = n;
  G.ne = ne;
  return NLM_OK;
}

unsigned mm_grid(int64_t M) { return unsigned((M + 127) / 128); }

void mm(nlm_ctx* c, const MmArgs& a) {
  if (a.M <= 0) return;
  hipLaunchKernelGGL(k_mm, dim3(mm_grid(a.M)), dim3(256), 0, c->stream, a);
}

MmArgs mm_args(int64_t M, int K, const uint16_t* X1, const uint16_t* X2,
               const uint16_t* Wt) {
  MmArgs a;
  std::memset(&a, 0, sizeof(a));
  a.M = M;
  a.K = K;
  a.X1 = X1;
  a.X2 = X2;
  a.Wt = Wt;
  return a;
}

// grad[off .. off + K*64) += X^T dY over M rows.
int mm_tn(nlm_ctx* c, int64_t M, int K, const uint16_t* X1, const uint16_t* X2,
          const uint16_t* dY, int64_t off) {
  GnnState& G = c->gnn;
  const int64_t nchunk = (M + kTnRows - 1) / kTnRows;
  const int nb = int(std::min<int64_t>(nchunk, 512));
  const size_t need = size_t(nb) * K * 64 * 4 + 256;
  if (!G.part.p || G.part.bytes < need) {
    int rc = alloc(c, G.part, need);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_mm_tn, dim3(nb), dim3(256), 0, c->stream, M, K, X1, X2,
                     dY, G.part.as<float>());
  hipLaunchKernelGGL(k_reduce_add,
                     dim3(unsigned((K * 64 + kRedCols - 1) / kRedCols)),
                     dim3(256), 0, c->stream, G.part.as<float>(), nb, K * 64,
                     G.grad.as<float>() + off);
  return NLM_OK;
}

template <class T>
int colsum(nlm_ctx* c, int64_t M, const T* dY, int64_t off) {
  GnnState& G = c->gnn;
  const int nb = int(std::min<int64_t>((M + 3) / 4, 512));
  const size_t need = size_t(nb) * 64 * 4 + 256;
  if (!G.part.p || G.part.bytes < need) {
    int rc = alloc(c, G.part, need);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_colsum<T>, dim3(nb), dim3(256), 0, c->stream, M, dY,
                     G.part.as<float>());
  hipLaunchKernelGGL(k_reduce_add, dim3(64 / kRedCols), dim3(256), 0,
                     c->stream, G.part.as<float>(), nb, 64,
                     G.grad.as<float>() + off);
  return NLM_OK;
}

uint16_t* hbuf(GnnState& G, int l) {
  return G.h.as<uint16_t>() + int64_t(l) * G.n * 64;
}
uint16_t* gbuf(GnnState& G, int l) {
  return G.g.as<uint16_t>() + int64_t(l) * G.ne * 64;
}
uint16_t* aggbuf(GnnState& G, int l) {
  return G.agg.as<uint16_t>() + int64_t(l) * G.n * 64;
}

// Forward pass on the assembled tangent; writes the preconditioner blocks.
int gnn_forward(nlm_ctx* c) {
  FdState& s = c->fd;
  QsState& q = c->qs;
  GnnState& G = c->gnn;
  int rc;
  if ((rc = gnn_size(c))) return rc;
  const int64_t n = s.n, ne = s.sell.nentries;
  if (!q.mblk.p || q.mblk.bytes < sizeof(float4) * ne) {
    if ((rc = alloc(c, q.mblk, sizeof(float4) * ne + 256))) return rc;
    if ((rc = alloc(c, q.mdiag, sizeof(float4) * n + 256))) return rc;
  }
  const GraphArgs ga = graph_args(c);
  const unsigned gn = unsigned((n + 255) / 256);
  const unsigned gs = unsigned((s.sell.nslices + 3) / 4);
  hipLaunchKernelGGL(k_gnn_scale, dim3(gn), dim3(256), 0, c->stream, ga,
                     G.scale.as<double>(), G.isq.as<float>(), G.deg.as<float>());
  hipLaunchKernelGGL(k_gnn_feat, dim3(gs), dim3(256), 0, c->stream, ga,
                     G.scale.as<double>(), G.xn.as<uint16_t>(),
                     G.xe.as<uint16_t>(), G.erow.as<int32_t>(),
                     G.valid.as<uint8_t>());
  const float* w = G.w.as<float>();
  const uint16_t* wf = G.wfwd.as<uint16_t>();
  // encoders
  MmArgs a = mm_args(n, kFn, G.xn.as<uint16_t>(), nullptr, wf + kFwdWn);
  a.bias = w + kOffBn;
  a.relu = 1;
  a.Ybf = hbuf(G, 0);
  mm(c, a);
  a = mm_args(ne, kFe, G.xe.as<uint16_t>(), nullptr, wf + kFwdWe);
  a.bias = w + kOffBe;
  a.relu = 1;
  a.valid = G.valid.as<uint8_t>();
  a.Ybf = gbuf(G, 0);
  mm(c, a);
  for (int l = 0; l < kL; l++) {
    const uint16_t* WL = wf + kFwdLayer + l * kFwdLayerSize;
    const float* wl = w + kOffLayer + l * kLayerSize;
    a = mm_args(n, 64, hbuf(G, l), nullptr, WL + kFWa);
    a.Yf = G.P.as<float>();
    mm(c, a);
    a = mm_args(n, 64, hbuf(G, l), nullptr, WL + kFWb);
    a.Yf = G.Q.as<float>();
    mm(c, a);
    a = mm_args(ne, 64, gbuf(G, l), nullptr, WL + kFWc);
    a.bias = wl + kLBm;
    a.P = G.P.as<float>();
    a.prow = G.erow.as<int32_t>();
    a.Q = G.Q.as<float>();
    a.qrow = s.sell.col.as<int32_t>();
    a.valid = G.valid.as<uint8_t>();
    a.relu = 1;
    a.Ybf = gbuf(G, l + 1);
    mm(c, a);
    hipLaunchKernelGGL(k_gnn_agg, dim3(gs), dim3(256), 0, c->stream, ga,
                       gbuf(G, l + 1), aggbuf(G, l));
    a = mm_args(n, 128, hbuf(G, l), aggbuf(G, l), WL + kFWu);
    a.bias = wl + kLBu;
    a.relu = 1;
    a.Ybf = hbuf(G, l + 1);
    mm(c, a);
  }
  DecArgs d;
  d.g = ga;
  d.w = w;
  d.gL = gbuf(G, kL);
  d.hL = hbuf(G, kL);
  d.isq = G.isq.as<float>();
  d.mblk = q.mblk.as<float4>();
  d.mdiag = q.mdiag.as<float4>();
  hipLaunchKernelGGL(k_gnn_decode, dim3(gs), dim3(256), 0, c->stream, d);
  NLM_HIP(c, hipGetLastError());
  q.has_m = true;
  q.mhalf_ok = false;
  return NLM_OK;
}

// Backward pass from dme / dmn; adds every parameter gradient to G.grad.
int gnn_backward(nlm_ctx* c) {
  FdState& s = c->fd;
  GnnState& G = c->gnn;
  const int64_t n = s.n, ne = s.sell.nentries;
  const GraphArgs ga = graph_args(c);
  const unsigned gs = unsigned((s.sell.nslices + 3) / 4);
  const float* w = G.w.as<float>();
  const uint16_t* wb = G.wbwd.as<uint16_t>();
  float* grad = G.grad.as<float>();
  int rc;
  // decoders: dg_L (BF16) and dh (FP32), weight and bias gradients
  {
    const int nbw = 1024;
    if (!G.part.p || G.part.bytes < size_t(nbw) * 260 * 4 + 256)
      if ((rc = alloc(c, G.part, size_t(nbw) * 260 * 4 + 256))) return rc;
    hipLaunchKernelGGL(k_dec_dx<false>, dim3(unsigned((ne * 64 + 255) / 256)),
                       dim3(256), 0, c->stream, ne, G.dme.as<float>(),
                       w + kOffWoe, G.dg.as<uint16_t>(), nullptr);
    const int nb1 = int(std::min<int64_t>(ne, nbw));
    hipLaunchKernelGGL(k_dec_dw, dim3(nb1), dim3(256), 0, c->stream, ne,
                       G.dme.as<float>(), gbuf(G, kL), G.part.as<float>());
    hipLaunchKernelGGL(k_reduce_add, dim3((260 + kRedCols - 1) / kRedCols),
                       dim3(256), 0, c->stream, G.part.as<float>(), nb1, 260, grad + kOffWoe);
    hipLaunchKernelGGL(k_dec_dx<true>, dim3(unsigned((n * 64 + 255) / 256)),
                       dim3(256), 0, c->stream, n, G.dmn.as<float>(),
                       w + kOffWon, nullptr, G.dh.as<float>());
    const int nb2 = int(std::min<int64_t>(n, nbw));
    hipLaunchKernelGGL(k_dec_dw, dim3(nb2), dim3(256), 0, c->stream, n,
                       G.dmn.as<float>(), hbuf(G, kL), G.part.as<float>());
    hipLaunchKernelGGL(k_reduce_add, dim3((260 + kRedCols - 1) / kRedCols),
                       dim3(256), 0, c->stream, G.part.as<float>(), nb2, 260, grad + kOffWon);
  }
  uint16_t* dpre_n = G.dh16.as<uint16_t>();
  uint16_t* dP16 = G.dPQ16.as<uint16_t>();
  uint16_t* dQ16 = dP16 + n * 64;
  const unsigned gN64 = unsigned((n * 64 + 255) / 256);
  const unsigned gE64 = unsigned((ne * 64 + 255) / 256);
  for (int l = kL - 1; l >= 0; l--) {
    const int64_t L = kOffLayer + l * kLayerSize;
    const uint16_t* WB = wb + l * kBwdLayerSize;
    // node update: h_{l+1} = relu([h_l, agg_l] [Wd; Wf] + bu)
    hipLaunchKernelGGL(k_relu_mask<float>, dim3(gN64), dim3(256), 0, c->stream,
                       n * 64, G.dh.as<float>(), hbuf(G, l + 1), dpre_n);
    if ((rc = mm_tn(c, n, 128, hbuf(G, l), aggbuf(G, l), dpre_n, L + kLWd)))
      return rc;
    if ((rc = colsum(c, n, dpre_n, L + kLBu))) return rc;
    MmArgs a = mm_args(n, 64, dpre_n, nullptr, WB + kBWd);
    a.Yf = G.dh.as<float>();  // dh_l (first part)
    mm(c, a);
    a = mm_args(n, 64, dpre_n, nullptr, WB + kBWf);
    a.Yf = G.dagg.as<float>();
    mm(c, a);
    // messages: m = relu(g_l Wc + P_i + Q_j + bm), g_{l+1} = m
    hipLaunchKernelGGL(k_edge_dm, dim3(gE64), dim3(256), 0, c->stream, ne,
                       G.dg.as<uint16_t>(), G.dagg.as<float>(),
                       G.erow.as<int32_t>(), G.deg.as<float>(),
                       G.valid.as<uint8_t>(), gbuf(G, l + 1),
                       G.dpre.as<uint16_t>());
    if ((rc = mm_tn(c, ne, 64, gbuf(G, l), nullptr, G.dpre.as<uint16_t>(),
                    L + kLWc)))
      return rc;
    if ((rc = colsum(c, ne, G.dpre.as<uint16_t>(), L + kLBm))) return rc;
    a = mm_args(ne, 64, G.dpre.as<uint16_t>(), nullptr, WB + kBWc);
    a.Ybf = G.dg2.as<uint16_t>();  // d g_l
    mm(c, a);
    hipLaunchKernelGGL(k_dpq, dim3(gs), dim3(256), 0, c->stream, ga,
                       G.dpre.as<uint16_t>(), G.rev.as<int32_t>(),
                       G.dP.as<float>(), G.dQ.as<float>(), dP16, dQ16);
    if ((rc = mm_tn(c, n, 64, hbuf(G, l), nullptr, dP16, L + kLWa))) return rc;
    if ((rc = mm_tn(c, n, 64, hbuf(G, l), nullptr, dQ16, L + kLWb))) return rc;
    a = mm_args(n, 128, dP16, dQ16, WB + kBWab);
    a.Yf = G.dh.as<float>();
    a.accumulate = 1;
    mm(c, a);
    std::swap(G.dg.p, G.dg2.p);
  }
  // encoders
  hipLaunchKernelGGL(k_relu_mask<float>, dim3(gN64), dim3(256), 0, c->stream,
                     n * 64, G.dh.as<float>(), hbuf(G, 0), dpre_n);
  if ((rc = mm_tn(c, n, kFn, G.xn.as<uint16_t>(), nullptr, dpre_n, kOffWn)))
    return rc;
  if ((rc = colsum(c, n, dpre_n, kOffBn))) return rc;
  hipLaunchKernelGGL(k_relu_mask<uint16_t>, dim3(gE64), dim3(256), 0,
                     c->stream, ne * 64, G.dg.as<uint16_t>(), gbuf(G, 0),
                     G.dpre.as<uint16_t>());
  if ((rc = mm_tn(c, ne, kFe, G.xe.as<uint16_t>(), nullptr,
                  G.dpre.as<uint16_t>(), kOffWe)))
    return rc;
  if ((rc = colsum(c, ne, G.dpre.as<uint16_t>(), kOffBe))) return rc;
  NLM_HIP(c, hipGetLastError());
  return NLM_OK;
}

int refresh_bf16(nlm_ctx* c) {
  GnnState& G = c->gnn;
  const int64_t cnt = std::max(kFwdSize, kBwdSize);
  hipLaunchKernelGGL(k_to_bf16_copies, dim3(unsigned((cnt + 255) / 256)),
                     dim3(256), 0, c->stream, G.w.as<float>(),
                     G.wfwd.as<uint16_t>(), G.wbwd.as<uint16_t>());
  NLM_HIP(c, hipGetLastError());
  return NLM_OK;
}

}  // namespace

int gnn_predict_precond(nlm_ctx* c) {
  if (!c->gnn.ready) return fail(c, NLM_ERR_STATE, "no graph network attached");
  return gnn_forward(c);
}

}  // namespace nlm

using namespace nlm;

extern "C" int64_t nlm_gnn_num_params(void) { return kNumParams; }

extern "C" int nlm_gnn_create(nlm_ctx* c, uint64_t seed) {
  if (!c) return NLM_ERR_INVALID;
  if (!c->fd.ready)
    return fail(c, NLM_ERR_STATE, "nlm_gnn_create: nlm_fd_setup not done");
  if (c->fd.dim != 2 || c->fd.material != NLM_MATERIAL_RNP_BOND)
    return fail(c, NLM_ERR_INVALID, "nlm_gnn_create: 2D bond-based only");
  GnnState& G = c->gnn;
  int rc;
  const size_t pb = sizeof(float) * kNumParams + 256;
  if ((rc = alloc(c, G.w, pb)) || (rc = alloc(c, G.grad, pb)) ||
      (rc = alloc(c, G.am, pb)) || (rc = alloc(c, G.av, pb)) ||
      (rc = alloc(c, G.wfwd, sizeof(uint16_t) * kFwdSize + 256)) ||
      (rc = alloc(c, G.wbwd, sizeof(uint16_t) * kBwdSize + 256)))
    return rc;
  G.np = kNumParams;
  std::vector<float> w = init_weights(seed);
  NLM_HIP(c, hipMemcpy(G.w.p, w.data(), sizeof(float) * kNumParams,
                       hipMemcpyHostToDevice));
  NLM_HIP(c, hipMemset(G.grad.p, 0, pb));
  NLM_HIP(c, hipMemset(G.am.p, 0, pb));
  NLM_HIP(c, hipMemset(G.av.p, 0, pb));
  G.adam_t = 0;
  G.n = G.ne = -1;
  if ((rc = refresh_bf16(c))) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  G.ready = true;
  return NLM_OK;
}

extern "C" int nlm_gnn_get_weights(nlm_ctx* c, float* w) {
  int rc = gnn_check(c, "nlm_gnn_get_weights");
  if (rc) return rc;
  if (!w) return fail(c, NLM_ERR_INVALID, "nlm_gnn_get_weights: null");
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  NLM_HIP(c, hipMemcpy(w, c->gnn.w.p, sizeof(float) * kNumParams,
                       hipMemcpyDeviceToHost));
  return NLM_OK;
}

extern "C" int nlm_gnn_set_weights(nlm_ctx* c, const float* w) {
  int rc = gnn_check(c, "nlm_gnn_set_weights");
  if (rc) return rc;
  if (!w) return fail(c, NLM_ERR_INVALID, "nlm_gnn_set_weights: null");
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  NLM_HIP(c, hipMemcpy(c->gnn.w.p, w, sizeof(float) * kNumParams,
                       hipMemcpyHostToDevice));
  if ((rc = refresh_bf16(c))) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_gnn_predict(nlm_ctx* c) {
  int rc = gnn_check(c, "nlm_gnn_predict");
  if (rc) return rc;
  if (!c->qs.assembled)
    return fail(c, NLM_ERR_STATE, "nlm_gnn_predict: no tangent assembled");
  if ((rc = gnn_forward(c))) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_gnn_loss_grad(nlm_ctx* c, int K, const double* r,
                                 const double* e, double* loss, float* grad) {
  int rc = gnn_check(c, "nlm_gnn_loss_grad");
  if (rc) return rc;
  if (K < 1 || !r || !e)
    return fail(c, NLM_ERR_INVALID, "nlm_gnn_loss_grad: bad arguments");
  if (!c->qs.assembled)
    return fail(c, NLM_ERR_STATE, "nlm_gnn_loss_grad: no tangent assembled");
  FdState& s = c->fd;
  GnnState& G = c->gnn;
  const int64_t n = s.n;
  if ((rc = gnn_forward(c))) return rc;
  const size_t vb = sizeof(double2) * n * K + 256;
  if (!G.vr.p || G.vr.bytes < vb) {
    if ((rc = alloc(c, G.vr, vb)) || (rc = alloc(c, G.ve, vb)) ||
        (rc = alloc(c, G.vz, vb)))
      return rc;
  }
  // vectors to internal order; weights c_k = 1 / (K |e_k|^2)
  std::vector<double2> hr(size_t(n) * K), he(size_t(n) * K);
  std::vector<double> ck(K + 1, 0.);
  for (int k = 0; k < K; k++) {
    double ee = 0.;
    for (int64_t a = 0; a < n; a++) {
      const int64_t i = s.perm_host[a];
      const double* rv = r + (int64_t(k) * n + i) * 3;
      const double* ev = e + (int64_t(k) * n + i) * 3;
      hr[int64_t(k) * n + a] = make_double2(rv[0], rv[1]);
      he[int64_t(k) * n + a] = make_double2(ev[0], ev[1]);
      ee += ev[0] * ev[0] + ev[1] * ev[1];
    }
    ck[k] = ee > 0. ? 1. / (double(K) * ee) : 0.;
  }
  DevBuf dck;
  if ((rc = alloc(c, dck, sizeof(double) * (K + 1)))) return rc;
  NLM_HIP(c, hipMemcpyAsync(G.vr.p, hr.data(), sizeof(double2) * n * K,
                            hipMemcpyHostToDevice, c->stream));
  NLM_HIP(c, hipMemcpyAsync(G.ve.p, he.data(), sizeof(double2) * n * K,
                            hipMemcpyHostToDevice, c->stream));
  NLM_HIP(c, hipMemcpyAsync(dck.p, ck.data(), sizeof(double) * (K + 1),
                            hipMemcpyHostToDevice, c->stream));
  const GraphArgs ga = graph_args(c);
  const unsigned gs = unsigned((s.sell.nslices + 3) / 4);
  const unsigned gn = unsigned((n + 255) / 256);
  hipLaunchKernelGGL(k_gnn_spmm, dim3(gs), dim3(256), 0, c->stream, ga,
                     c->qs.mblk.as<float4>(), c->qs.mdiag.as<float4>(), K,
                     G.vr.as<double2>(), G.vz.as<double2>());
  hipLaunchKernelGGL(k_gnn_resid, dim3(gn), dim3(256), 0, c->stream, n, K,
                     dck.as<double>(), G.vz.as<double2>(), G.ve.as<double2>(),
                     G.lpart.as<double>());
  hipLaunchKernelGGL(k_gnn_dm, dim3(gs), dim3(256), 0, c->stream, ga, K,
                     dck.as<double>(), G.vz.as<double2>(), G.vr.as<double2>(),
                     G.isq.as<float>(), G.dme.as<float>(), G.dmn.as<float>());
  if ((rc = gnn_backward(c))) return rc;
  std::vector<double> lp(gn);
  NLM_HIP(c, hipMemcpyAsync(lp.data(), G.lpart.p, sizeof(double) * gn,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  double L = 0.;
  for (double v : lp) L += v;
  if (loss) *loss = L;
  if (grad)
    NLM_HIP(c, hipMemcpy(grad, G.grad.p, sizeof(float) * kNumParams,
                         hipMemcpyDeviceToHost));
  return NLM_OK;
}

extern "C" int nlm_gnn_zero_grad(nlm_ctx* c) {
  int rc = gnn_check(c, "nlm_gnn_zero_grad");
  if (rc) return rc;
  GnnState& G = c->gnn;
  NLM_HIP(c, hipMemsetAsync(G.grad.p, 0, sizeof(float) * kNumParams,
                            c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_gnn_adam_step(nlm_ctx* c, double lr) {
  int rc = gnn_check(c, "nlm_gnn_adam_step");
  if (rc) return rc;
  GnnState& G = c->gnn;
  G.adam_t++;
  const float c1 = float(1. - std::pow(0.9, double(G.adam_t)));
  const float c2 = float(1. - std::pow(0.999, double(G.adam_t)));
  hipLaunchKernelGGL(k_adam, dim3(unsigned((kNumParams + 255) / 256)),
                     dim3(256), 0, c->stream, int64_t(kNumParams),
                     G.w.as<float>(), G.grad.as<float>(), G.am.as<float>(),
                     G.av.as<float>(), float(lr), c1, c2);
  if ((rc = refresh_bf16(c))) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}
