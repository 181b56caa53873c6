// qs.hip -- quasi-static Newton solver of the bond-based material on the GPU
// (model::QuasiStaticModel<T>, src/model/quasistatic/QuasiStaticModel.cpp).
//
// Tangent.  The bond force of RNPBond::getBondEF (rnpBond.cpp:216-301) is
// F_ij = g(S) e with e = dx/|dx| fixed and S = e.du/|dx|, so
// dF_ij/du_j = k_ij e e^T = -dF_ij/du_i with k_ij = g'(S)/|dx|.  The solver
// works with A = -df_int/du: per slice entry (i, j) the block -k_ij e e^T, per
// node the block sum_j k_ij e e^T.  The reference assembles a central-
// difference Jacobian that keeps only same-component entries (:363-454);
// NLM_QS_TANGENT_COMPONENT keeps the same entries of the exact derivative.
//
// Krylov solve.  BiCGSTAB on A_ff (prescribed dofs held at zero in every
// vector, their rows masked in every product).  The scalars live on the
// device: each kernel that needs a dot product reduces the 256-node blocks'
// partial sums itself (every workgroup in the same order, so all agree), and
// workgroup 0 records the scalar for later kernels.  Convergence sets a flag
// that turns the remaining launches into no-ops, so the host checks it only
// every few iterations.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bond_rnp.h"
#include "device_common.h"
#include "fd_common.h"
#include "nlm_internal.h"

namespace nlm {
namespace {

struct QsArgs {
  int64_t n;
  int64_t nslices;
  int64_t nb;
  const double2* __restrict__ x;
  const double2* __restrict__ u;
  const double* __restrict__ vol;
  const uint32_t* __restrict__ flags;
  const int64_t* __restrict__ sl_ptr;
  const int32_t* __restrict__ col;
  const uint16_t* __restrict__ rowlen;
  const unsigned long long* __restrict__ bits;
  RnpDev m;
};

__device__ __forceinline__ bool dof_fixed(uint32_t fl, int d) {
  return ((fl >> (kFlagUbcShift + 4 * d)) & 0xfu) != 0u;
}

// Sum of nb partials, the same order in every workgroup.
__device__ double block_sum(const double* __restrict__ P, int64_t nb) {
  __shared__ double red[256];
  double s = 0.;
  for (int64_t k = threadIdx.x; k < nb; k += 256) s += P[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// Workgroup sum of one value per thread, stored to P[b] by thread 0.
__device__ void block_partial(double v, double* __restrict__ P, int64_t b) {
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) P[b] = red[0];
  __syncthreads();
}

// Workgroup sums of N values per thread, every thread receiving the N sums:
// a butterfly within each wavefront, then the four wavefronts' sums in a
// fixed order -- one barrier for all N, the same order in every workgroup.
template <int N>
__device__ __forceinline__ void block_reduce(double (&v)[N]) {
  __shared__ double red[N][4];
  for (int off = 32; off > 0; off >>= 1)
    for (int q = 0; q < N; q++) v[q] += __shfl_xor(v[q], off);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < N; q++) red[q][w] = v[q];
  __syncthreads();
  for (int q = 0; q < N; q++)
    v[q] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  __syncthreads();  // red is reused by the next call
}

// Sums of N partial arrays of nb values each (their loads in flight together).
template <int N>
__device__ __forceinline__ void block_sums(const double* const (&P)[N],
                                           int64_t nb, double (&out)[N]) {
  for (int q = 0; q < N; q++) out[q] = 0.;
  for (int64_t k = threadIdx.x; k < nb; k += 256)
    for (int q = 0; q < N; q++) out[q] += P[q][k];
  block_reduce<N>(out);
}

// Workgroup sums of N values per thread, stored to P[q][b] by thread 0.
template <int N>
__device__ __forceinline__ void block_partials(double (&v)[N],
                                               double* const (&P)[N],
                                               int64_t b) {
  block_reduce<N>(v);
  if (threadIdx.x == 0)
    for (int q = 0; q < N; q++) P[q][b] = v[q];
}

// Tangent blocks of every slice entry and the diagonal blocks.
__global__ __launch_bounds__(256) void k_qs_tangent(
    QsArgs a, int component, double4* __restrict__ blk,
    double4* __restrict__ diag, double2* __restrict__ dinv,
    double* __restrict__ bg) {
  const int64_t nblk = (a.nslices * 64 + 255) / 256;
  const int64_t lb = xcd_block(nblk);
  if (lb >= nblk) return;
  const int lane = threadIdx.x & 63;
  const int64_t s = lb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (s >= a.nslices) return;
  const int64_t i = s * 64 + lane;
  const bool valid = i < a.n;
  const int64_t cb = a.sl_ptr[s];
  const int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  const int64_t bb = cb >> 6;
  const int len = valid ? int(a.rowlen[i]) : 0;
  double2 xi = valid ? a.x[i] : make_double2(0., 0.);
  double2 ui = valid ? a.u[i] : make_double2(0., 0.);
  const uint32_t fli = valid ? a.flags[i] : 0u;
  const bool int_i = !(fli & kFlagNonInterior);
  const RnpDev& m = a.m;
  double4 D = make_double4(0., 0., 0., 0.);
  for (int k = 0; k < width; k++) {
    const int64_t e = cb + int64_t(k) * 64 + lane;
    const unsigned long long w = a.bits[bb + k];
    double4 B = make_double4(0., 0., 0., 0.);
    double gb = 0.;
    if (k < len) {
      const int32_t j = a.col[e];
      const double2 xj = a.x[j], uj = a.u[j];
      const bool brk = int_i && !(a.flags[j] & kFlagNonInterior);
      const bool fs = (w >> lane) & 1ull;
      const double dx0 = xj.x - xi.x, dx1 = xj.y - xi.y;
      const double du0 = uj.x - ui.x, du1 = uj.y - ui.y;
      const double dd = dx0 * dx0 + dx1 * dx1;
      const double r = sqrt(dd);
      const double S = (dx0 * du0 + dx1 * du1) / dd;
      double volj = a.vol[j];
      if (def_greater(r, m.check_low)) volj *= (m.check_up - r) / m.h;
      const double infl = influence_exact(m, r);
      const double c = infl * 4. * m.C * m.beta / m.inv_factor * volj;
      double kk;
      if (brk) {
        const double q = m.beta * r * S * S;
        kk = fs ? 0. : c * exp(-q) * (1. - 2. * q) / r;
      } else {
        kk = c / r;
      }
      const double e0 = dx0 / r, e1 = dx1 / r;
      // the block is rank one (or diagonal): -B = gb dx dx^T, dx the
      // reference bond vector
      gb = -(kk / dd);
      B.x = kk * e0 * e0;
      B.y = component ? 0. : kk * e0 * e1;
      B.z = component ? 0. : kk * e1 * e0;
      B.w = kk * e1 * e1;
      D.x += B.x;
      D.y += B.y;
      D.z += B.z;
      D.w += B.w;
    }
    blk[e] = make_double4(-B.x, -B.y, -B.z, -B.w);
    bg[e] = gb;
  }
  if (!valid) return;
  diag[i] = D;
  dinv[i] = make_double2(!dof_fixed(fli, 0) && D.x != 0. ? 1. / D.x : 0.,
                         !dof_fixed(fli, 1) && D.w != 0. ? 1. / D.w : 0.);
}

// Finite-difference tangent (assembly_jacobian_matrix_part,
// QuasiStaticModel.cpp:363-454).  u <- u0 +/- eps on component r of the
// nodes of color cl.
__global__ __launch_bounds__(256) void k_fd_perturb(
    int64_t n, const double2* __restrict__ u0, const int* __restrict__ color,
    int cl, int r, double eps, double2* __restrict__ u) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  double2 v = u0[i];
  if (color[i] == cl) {
    if (r == 0)
      v.x = v.x + eps;
    else
      v.y = v.y + eps;
  }
  u[i] = v;
}

// A_ij (component r, r) = -(f+_i - f-_i)_r / (2 eps) for every slice entry
// whose neighbor j has color cl, and A_ii when i has color cl.  Every node
// within the force's reach of row i (two horizons) carries a distinct color,
// so f+_i - f-_i holds the perturbation of j alone.
__global__ __launch_bounds__(256) void k_fd_collect(
    QsArgs a, const double2* __restrict__ fp, const double2* __restrict__ fm,
    const int* __restrict__ color, int cl, int r, double eps,
    double4* __restrict__ blk, double4* __restrict__ diag) {
  const int64_t nblk = (a.nslices * 64 + 255) / 256;
  const int64_t lb = xcd_block(nblk);
  if (lb >= nblk) return;
  const int lane = threadIdx.x & 63;
  const int64_t s = lb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (s >= a.nslices) return;
  const int64_t i = s * 64 + lane;
  if (i >= a.n) return;
  const int64_t cb = a.sl_ptr[s];
  const int len = int(a.rowlen[i]);
  const double2 dp = fp[i], dm = fm[i];
  const double df = r == 0 ? dp.x - dm.x : dp.y - dm.y;
  const double val = -(df / (2. * eps));
  for (int k = 0; k < len; k++) {
    const int64_t e = cb + int64_t(k) * 64 + lane;
    if (color[a.col[e]] != cl) continue;
    if (r == 0)
      blk[e].x = val;
    else
      blk[e].w = val;
  }
  if (color[i] == cl) {
    if (r == 0)
      diag[i].x = val;
    else
      diag[i].w = val;
  }
}

__global__ __launch_bounds__(256) void k_fd_dinv(
    int64_t n, const uint32_t* __restrict__ flags,
    const double4* __restrict__ diag, double2* __restrict__ dinv) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double4 D = diag[i];
  const uint32_t fl = flags[i];
  dinv[i] = make_double2(!dof_fixed(fl, 0) && D.x != 0. ? 1. / D.x : 0.,
                         !dof_fixed(fl, 1) && D.w != 0. ? 1. / D.w : 0.);
}

// y = (D x_i + sum_j B_ij x_j) on free dofs, 0 on prescribed dofs.
// DOT 1: partial w.y into P0; DOT 2: partials w.y into P0 and y.y into P1.
template <int DOT, class BT>
__global__ __launch_bounds__(256) void k_qs_spmv(
    QsArgs a, const BT* __restrict__ blk, const BT* __restrict__ dg,
    const double2* __restrict__ xv, double2* __restrict__ y,
    const double2* __restrict__ wv, double* __restrict__ P0,
    double* __restrict__ P1, const int* __restrict__ ctl,
    const double* __restrict__ zc) {
  if (ctl && ctl[0]) return;
  const int64_t nblk = (a.nslices * 64 + 255) / 256;
  const int64_t lb = xcd_block(nblk);
  if (lb >= nblk) return;
  const int lane = threadIdx.x & 63;
  const int64_t s = lb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool slice_ok = s < a.nslices;
  const int64_t i = s * 64 + lane;
  const bool valid = slice_ok && i < a.n;
  double p0 = 0., p1 = 0.;
  if (slice_ok) {
    const int64_t cb = a.sl_ptr[s];
    const int width = int((a.sl_ptr[s + 1] - cb) >> 6);
    double2 acc = make_double2(0., 0.);
    for (int k = 0; k < width; k++) {
      const int64_t e = cb + int64_t(k) * 64 + lane;
      const BT B = blk[e];
      const double2 xj = xv[a.col[e]];
      acc.x += double(B.x) * xj.x + double(B.y) * xj.y;
      acc.y += double(B.z) * xj.x + double(B.w) * xj.y;
    }
    if (valid) {
      const BT D = dg[i];
      const double2 xi = xv[i];
      acc.x += double(D.x) * xi.x + double(D.y) * xi.y;
      acc.y += double(D.z) * xi.x + double(D.w) * xi.y;
      if (zc) {
        acc.x += zc[2 * (i >> 8)];
        acc.y += zc[2 * (i >> 8) + 1];
      }
      const uint32_t fl = a.flags[i];
      if (dof_fixed(fl, 0)) acc.x = 0.;
      if (dof_fixed(fl, 1)) acc.y = 0.;
      y[i] = acc;
      if (DOT >= 1) {
        const double2 w = wv[i];
        p0 = w.x * acc.x + w.y * acc.y;
      }
      if (DOT == 2) p1 = acc.x * acc.x + acc.y * acc.y;
    }
  }
  if (DOT >= 1) block_partial(p0, P0, lb);
  if (DOT == 2) block_partial(p1, P1, lb);
}

// Same product with one workgroup per slice: wavefront q of the four takes
// slots q, q+4, ..., their sums are added in a fixed order by wavefront 0.
// Partials (DOT) are per slice.  Four times the parallelism of k_qs_spmv, for
// the latency-bound Krylov products of small systems.
// Coarse correction of aggregate b: rows 2b and 2b+1 of the dense m x m
// inverse times the restricted vector rc, summed by the whole 256-thread
// workgroup in a fixed order (every workgroup of the aggregate obtains the
// same value).
__device__ double2 coarse_rows(const double* __restrict__ Ainv, int m,
                               const double* __restrict__ rc, int64_t b) {
  __shared__ double2 cr[256];
  const double* r0 = Ainv + 2 * b * int64_t(m);
  const double* r1 = r0 + m;
  double s0 = 0., s1 = 0.;
  for (int k = threadIdx.x; k < m; k += 256) {
    const double v = rc[k];
    s0 += r0[k] * v;
    s1 += r1[k] * v;
  }
  cr[threadIdx.x] = make_double2(s0, s1);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      cr[threadIdx.x].x += cr[threadIdx.x + w].x;
      cr[threadIdx.x].y += cr[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  const double2 z = cr[0];
  __syncthreads();
  return z;
}

// BF16 copy of the preconditioner's bond blocks (8 bytes per slot): the
// rounding (relative 2^-9) perturbs the preconditioner only, never the
// operator.
struct __align__(8) Bf4 {
  uint16_t x, y, z, w;
};
__device__ __forceinline__ float bf16_to_f(uint16_t b) {
  return __uint_as_float(uint32_t(b) << 16);
}
__device__ __forceinline__ uint16_t f_to_bf16(float f) {  // nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__device__ __forceinline__ double4 d4(const double4& b) { return b; }
__device__ __forceinline__ double4 d4(const float4& b) {
  return make_double4(b.x, b.y, b.z, b.w);
}
__device__ __forceinline__ double4 d4(const Bf4& b) {
  return make_double4(bf16_to_f(b.x), bf16_to_f(b.y), bf16_to_f(b.z),
                      bf16_to_f(b.w));
}

__global__ __launch_bounds__(256) void k_blocks_bf16(
    int64_t nent, const float4* __restrict__ src, Bf4* __restrict__ dst) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= nent) return;
  const float4 b = src[e];
  Bf4 o;
  o.x = f_to_bf16(b.x);
  o.y = f_to_bf16(b.y);
  o.z = f_to_bf16(b.z);
  o.w = f_to_bf16(b.w);
  dst[e] = o;
}

// y = A x over the slices (one workgroup per slice, wavefront q taking
// slots q, q+4, ...), A from its blocks (BS 0), or from one scalar per bond
// (the analytic bond tangent: off-diagonal block -B_ij = g_ij dx dx^T, BS 1;
// the component tangent diag(g dx_0^2, g dx_1^2), BS 2) with dx = x_j - x_i
// formed from the positions -- 12 bytes per slot instead of 36.  DOT 1:
// partial (w, y) per slice; DOT 2: also (y, y).
template <int DOT, class BT, int BS = 0, class DT = BT>
__global__ __launch_bounds__(256) void k_qs_spmv4(
    QsArgs a, const BT* __restrict__ blk, const DT* __restrict__ dg,
    const double2* __restrict__ xv, double2* __restrict__ y,
    const double2* __restrict__ wv, double* __restrict__ P0,
    double* __restrict__ P1, const int* __restrict__ ctl,
    const double* __restrict__ Ainv, int m, const double* __restrict__ rc,
    const double* __restrict__ bg) {
  // the control word is read with the first loads and tested after the
  // products (a finished solve only skips the stores)
  const int done = ctl ? ctl[0] : 0;
  const int64_t s = xcd_block(a.nslices);
  if (s >= a.nslices) return;
  __shared__ double2 red[3][64];
  // two-level: the slice's aggregate (4 slices) adds its coarse correction
  double2 zc = make_double2(0., 0.);
  if (Ainv) zc = coarse_rows(Ainv, m, rc, s >> 2);
  const int lane = threadIdx.x & 63;
  const int wq = threadIdx.x >> 6;
  const int64_t i = s * 64 + lane;
  const bool valid = i < a.n;
  const int64_t cb = a.sl_ptr[s];
  const int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  double2 acc = make_double2(0., 0.);
  if (BS == 0) {
    for (int k = wq; k < width; k += 4) {
      const int64_t e = cb + int64_t(k) * 64 + lane;
      const double4 B = d4(blk[e]);
      const double2 xj = xv[a.col[e]];
      acc.x += B.x * xj.x + B.y * xj.y;
      acc.y += B.z * xj.x + B.w * xj.y;
    }
  } else {
    // padding slots hold g = 0 and a valid column
    const double2 pi = valid ? a.x[i] : make_double2(0., 0.);
    auto slot = [&](double gb, double2 pj, double2 xj) {
      const double d0 = pj.x - pi.x, d1 = pj.y - pi.y;
      if (BS == 1) {
        const double w = gb * (d0 * xj.x + d1 * xj.y);
        acc.x += w * d0;
        acc.y += w * d1;
      } else {
        acc.x += gb * d0 * d0 * xj.x;
        acc.y += gb * d1 * d1 * xj.y;
      }
    };
    // two slots per trip, both slots' loads in flight together
    int k = wq;
    for (; k + 4 < width; k += 8) {
      const int64_t e0 = cb + int64_t(k) * 64 + lane, e1 = e0 + 256;
      const double g0 = bg[e0], g1 = bg[e1];
      const int32_t j0 = a.col[e0], j1 = a.col[e1];
      const double2 p0 = a.x[j0], x0 = xv[j0], p1 = a.x[j1], x1 = xv[j1];
      slot(g0, p0, x0);
      slot(g1, p1, x1);
    }
    if (k < width) {
      const int64_t e = cb + int64_t(k) * 64 + lane;
      const int32_t j = a.col[e];
      slot(bg[e], a.x[j], xv[j]);
    }
  }
  if (wq > 0) red[wq - 1][lane] = acc;
  __syncthreads();
  if (wq != 0 || done) return;
  acc.x += red[0][lane].x + red[1][lane].x + red[2][lane].x;
  acc.y += red[0][lane].y + red[1][lane].y + red[2][lane].y;
  double p0 = 0., p1 = 0.;
  if (valid) {
    const double4 D = d4(dg[i]);
    const double2 xi = xv[i];
    acc.x += D.x * xi.x + D.y * xi.y;
    acc.y += D.z * xi.x + D.w * xi.y;
    acc.x += zc.x;
    acc.y += zc.y;
    const uint32_t fl = a.flags[i];
    if (dof_fixed(fl, 0)) acc.x = 0.;
    if (dof_fixed(fl, 1)) acc.y = 0.;
    y[i] = acc;
    if (DOT >= 1) {
      const double2 w = wv[i];
      p0 = w.x * acc.x + w.y * acc.y;
    }
    if (DOT == 2) p1 = acc.x * acc.x + acc.y * acc.y;
  }
  if (DOT >= 1) {
    for (int off = 32; off > 0; off >>= 1) {
      p0 += __shfl_xor(p0, off);
      if (DOT == 2) p1 += __shfl_xor(p1, off);
    }
    if (lane == 0) {
      P0[s] = p0;
      if (DOT == 2) P1[s] = p1;
    }
  }
}

// Scalars of iteration it: rho[it] = (rh, r_it), alpha[it], omega[it],
// rr[it] = (r_it, r_it), at offsets 8 + {0, 1, 2, 3} * L + it.
// The latest rho, alpha and omega are also kept in two slots each, by the
// parity of their iteration, so a step can load both slots together with
// its control words and pick one once they arrive (no dependent load).
struct Scal {
  double* base;
  int L;
  __device__ double& tol2() const { return base[0]; }
  __device__ double& rho(int it) const { return base[8 + it]; }
  __device__ double& alpha(int it) const { return base[8 + L + it]; }
  __device__ double& omega(int it) const { return base[8 + 2 * L + it]; }
  __device__ double& rr(int it) const { return base[8 + 3 * L + it]; }
  __device__ double& rho_p(int par) const { return base[2 + par]; }
  __device__ double& alpha_p(int par) const { return base[4 + par]; }
  __device__ double& omega_p(int par) const { return base[6 + par]; }
};

struct VecArgs {
  int64_t n, nb;
  int64_t nbs;  // partials of the Krylov products (P2, P3): one per slice
  const uint32_t* __restrict__ flags;
  double2 *r, *rh, *p, *v, *s, *t, *x, *ph, *sh;
  const double2* __restrict__ dinv;
  double *P0, *P1, *P2, *P3;
  double *rc;  // restriction of p (then s): per 256-node aggregate, 2 dofs
  int* ctl;     // [0] done (1 converged, 2 breakdown), [1] its iteration,
                // [2] first iteration of the current batch
  int max_it;
  Scal sc;
};

// Restriction R^T v of the workgroup's 256 nodes (one aggregate of the
// coarse level): rc[2 b + d] = sum over the aggregate of v_d.
__device__ void block_restrict(double vx, double vy, double* __restrict__ rc,
                               int64_t b) {
  block_partial(vx, rc, 2 * b);
  block_partial(vy, rc, 2 * b + 1);
}

// r = rh = b (masked), p = v = x = 0; partials (rh, r) and (r, r).
__global__ __launch_bounds__(256) void k_bicg_init(VecArgs a,
                                                   const double2* __restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double p0 = 0.;
  if (i < a.n) {
    double2 bi = b[i];
    const uint32_t fl = a.flags[i];
    if (dof_fixed(fl, 0)) bi.x = 0.;
    if (dof_fixed(fl, 1)) bi.y = 0.;
    const double2 z = make_double2(0., 0.);
    a.r[i] = bi;
    a.rh[i] = bi;
    a.p[i] = z;
    a.v[i] = z;
    a.x[i] = z;
    a.ph[i] = z;
    a.sh[i] = z;
    p0 = bi.x * bi.x + bi.y * bi.y;
  }
  {
    double v[2] = {p0, p0};
    double* const P[2] = {a.P0, a.P1};
    block_partials<2>(v, P, blockIdx.x);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctl[0] = 0;
    a.ctl[1] = 0;
    a.ctl[2] = 1;
  }
}

// Moves the batch's first iteration on by n (one thread).
__global__ void k_iter_advance(int* __restrict__ ctl, int n) {
  if (threadIdx.x == 0) ctl[2] += n;
}

// Iteration it, step 1: convergence test on r_{it-1}; p = r + beta (p -
// omega v); ph = M p for the diagonal preconditioners (PRE 0 none, 1 Jacobi;
// PRE 2 leaves ph to a block product).  Every load of the step -- control
// words, scalar slots, the node's vectors, the partials -- is issued before
// the first wait.
template <int PRE, bool RESTRICT = false>
__global__ __launch_bounds__(256) void k_bicg_p(VecArgs a, int k) {
  const int done = a.ctl[0];
  const int it = a.ctl[2] + k;
  const double rho2[2] = {a.sc.rho_p(0), a.sc.rho_p(1)};
  const double al1[2] = {a.sc.alpha_p(0), a.sc.alpha_p(1)};
  const double om1[2] = {a.sc.omega_p(0), a.sc.omega_p(1)};
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const bool own = i < a.n;
  const double2 z = make_double2(0., 0.);
  const double2 r = own ? a.r[i] : z, p = own ? a.p[i] : z,
                v = own ? a.v[i] : z;
  const double2 di = (PRE == 1 && own) ? a.dinv[i] : z;
  double red[2];
  {
    const double* const P[2] = {a.P0, a.P1};
    block_sums<2>(P, a.nb, red);
  }
  if (done || it > a.max_it) return;
  const double rho_i = red[0], rr = red[1];
  const bool conv = !(rr > a.sc.tol2());
  const bool brk = !conv && rho_i == 0.;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc.rho(it - 1) = rho_i;
    a.sc.rho_p((it - 1) & 1) = rho_i;
    a.sc.rr(it - 1) = rr;
    if (conv || brk) {
      a.ctl[0] = conv ? 1 : 2;
      a.ctl[1] = it - 1;
    }
  }
  if (conv || brk) return;
  double beta = 0., om = 0.;
  if (it > 1) {
    om = om1[(it - 1) & 1];
    beta = (rho_i / rho2[it & 1]) * (al1[(it - 1) & 1] / om);
  }
  double2 pn = make_double2(0., 0.);
  if (own) {
    pn.x = r.x + beta * (p.x - om * v.x);
    pn.y = r.y + beta * (p.y - om * v.y);
    a.p[i] = pn;
    if (PRE == 0) a.ph[i] = pn;
    if (PRE == 1) a.ph[i] = make_double2(di.x * pn.x, di.y * pn.y);
  }
  if (RESTRICT) block_restrict(pn.x, pn.y, a.rc, blockIdx.x);
}

// Step 2 (after v = A ph with partial (rh, v) in P2): alpha, s = r - alpha v.
template <int PRE, bool RESTRICT = false>
__global__ __launch_bounds__(256) void k_bicg_s(VecArgs a, int k) {
  const int done = a.ctl[0];
  const int it = a.ctl[2] + k;
  const double rho1[2] = {a.sc.rho_p(0), a.sc.rho_p(1)};
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const bool own = i < a.n;
  const double2 z = make_double2(0., 0.);
  const double2 r = own ? a.r[i] : z, v = own ? a.v[i] : z;
  const double2 di = (PRE == 1 && own) ? a.dinv[i] : z;
  double rv1[1];
  {
    const double* const P[1] = {a.P2};
    block_sums<1>(P, a.nbs, rv1);
  }
  if (done || it > a.max_it) return;
  const double rv = rv1[0];
  const double rho_i = rho1[(it - 1) & 1];
  const bool brk = rv == 0.;
  const double alpha = brk ? 0. : rho_i / rv;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc.alpha(it) = alpha;
    a.sc.alpha_p(it & 1) = alpha;
    if (brk) {
      a.ctl[0] = 2;
      a.ctl[1] = it - 1;
    }
  }
  if (brk) return;
  double2 sn = make_double2(0., 0.);
  if (own) {
    sn = make_double2(r.x - alpha * v.x, r.y - alpha * v.y);
    a.s[i] = sn;
    if (PRE == 0) a.sh[i] = sn;
    if (PRE == 1) a.sh[i] = make_double2(di.x * sn.x, di.y * sn.y);
  }
  if (RESTRICT) block_restrict(sn.x, sn.y, a.rc, blockIdx.x);
}

// Step 2 with the preconditioner's bond blocks, one level: alpha, s = r -
// alpha v and sh = M s in one launch.  Each row forms its neighbors' s_j =
// r_j - alpha v_j itself (r and v are not written by this step, so every
// row reads the same operands the owner does); the owner stores its s_i.
// One workgroup per slice, wavefront q taking slots q, q+4, ...
__global__ __launch_bounds__(256) void k_bicg_s_blocks(
    VecArgs a, QsArgs q, const Bf4* __restrict__ mb,
    const float4* __restrict__ md, int k) {
  const int done = a.ctl[0];
  const int it = a.ctl[2] + k;
  const double rho1[2] = {a.sc.rho_p(0), a.sc.rho_p(1)};
  const int64_t sl = xcd_block(q.nslices);
  if (sl >= q.nslices) return;  // whole workgroups
  double rv1[1];
  {
    const double* const P[1] = {a.P2};
    block_sums<1>(P, a.nbs, rv1);
  }
  if (done || it > a.max_it) return;
  const double rv = rv1[0];
  const bool brk = rv == 0.;
  const double alpha = brk ? 0. : rho1[(it - 1) & 1] / rv;
  if (sl == 0 && threadIdx.x == 0) {
    a.sc.alpha(it) = alpha;
    a.sc.alpha_p(it & 1) = alpha;
    if (brk) {
      a.ctl[0] = 2;
      a.ctl[1] = it - 1;
    }
  }
  if (brk) return;
  __shared__ double2 red[3][64];
  const int lane = threadIdx.x & 63;
  const int wq = threadIdx.x >> 6;
  const int64_t i = sl * 64 + lane;
  const bool valid = i < a.n;
  const int64_t cb = q.sl_ptr[sl];
  const int width = int((q.sl_ptr[sl + 1] - cb) >> 6);
  double2 acc = make_double2(0., 0.);
  for (int kk = wq; kk < width; kk += 4) {
    const int64_t e = cb + int64_t(kk) * 64 + lane;
    const double4 B = d4(mb[e]);
    const int32_t j = q.col[e];
    const double2 rj = a.r[j], vj = a.v[j];
    const double2 sj = make_double2(rj.x - alpha * vj.x, rj.y - alpha * vj.y);
    acc.x += B.x * sj.x + B.y * sj.y;
    acc.y += B.z * sj.x + B.w * sj.y;
  }
  if (wq > 0) red[wq - 1][lane] = acc;
  __syncthreads();
  if (wq != 0 || !valid) return;
  acc.x += red[0][lane].x + red[1][lane].x + red[2][lane].x;
  acc.y += red[0][lane].y + red[1][lane].y + red[2][lane].y;
  const double2 r = a.r[i], v = a.v[i];
  const double2 si = make_double2(r.x - alpha * v.x, r.y - alpha * v.y);
  a.s[i] = si;
  const double4 D = d4(md[i]);
  acc.x += D.x * si.x + D.y * si.y;
  acc.y += D.z * si.x + D.w * si.y;
  const uint32_t fl = a.flags[i];
  if (dof_fixed(fl, 0)) acc.x = 0.;
  if (dof_fixed(fl, 1)) acc.y = 0.;
  a.sh[i] = acc;
}

// ---- coarse level ------------------------------------------------------------
// Aggregates are the layout's 256-node blocks (compact Morton patches); the
// coarse space is piecewise constant per aggregate and dof.  A_c = R^T A R
// on the free dofs, inverted densely once per refresh (Gauss-Jordan sweeps,
// the matrix being symmetric positive definite up to the volume asymmetry of
// the bond-based tangent); the preconditioner is S + R A_c^-1 R^T.

// zc = Ainv rc (dense m x m, one wavefront per row).
__global__ __launch_bounds__(256) void k_coarse_mv(const double* __restrict__ Ainv,
                                                   int m,
                                                   const double* __restrict__ rc,
                                                   double* __restrict__ zc,
                                                   const int* __restrict__ ctl) {
  if (ctl && ctl[0]) return;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= m) return;
  const double* ar = Ainv + int64_t(row) * m;
  double s = 0.;
  for (int k = lane; k < m; k += 64) s += ar[k] * rc[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) zc[row] = s;
}

// dst = dinv * src + zc[aggregate] on free dofs (point-Jacobi smoother plus
// the prolongated coarse correction).
__global__ __launch_bounds__(256) void k_smooth_jacobi(
    VecArgs a, const double2* __restrict__ src, double2* __restrict__ dst,
    const double* __restrict__ Ainv, int m) {
  if (a.ctl && a.ctl[0]) return;
  // the workgroup's 256 nodes are one aggregate
  double2 zc = make_double2(0., 0.);
  if (Ainv) zc = coarse_rows(Ainv, m, a.rc, blockIdx.x);
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  const double2 di = a.dinv[i], v = src[i];
  const uint32_t fl = a.flags[i];
  double2 o = make_double2(di.x * v.x + zc.x, di.y * v.y + zc.y);
  if (dof_fixed(fl, 0)) o.x = 0.;
  if (dof_fixed(fl, 1)) o.y = 0.;
  dst[i] = o;
}

// Coarse matrix rows of aggregate b: A_c[(b, d)][(J, e)] for the aggregates J
// adjacent to b (adj list), free rows and columns only; dense row-major.
__global__ __launch_bounds__(256) void k_coarse_assemble(
    QsArgs a, const double4* __restrict__ blk, const double4* __restrict__ dg,
    const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj,
    int m, double* __restrict__ Ac) {
  __shared__ double red[4][256];
  const int64_t b = blockIdx.x;
  const int64_t i = b * 256 + threadIdx.x;
  const bool valid = i < a.n;
  const int64_t s = i >> 6;
  const int lane = int(i & 63);
  const uint32_t fi = valid ? a.flags[i] : 0u;
  const double ri0 = dof_fixed(fi, 0) ? 0. : 1., ri1 = dof_fixed(fi, 1) ? 0. : 1.;
  const int64_t cb = valid ? a.sl_ptr[s] : 0;
  const int len = valid ? int(a.rowlen[i]) : 0;
  for (int64_t q = adj_ptr[b]; q < adj_ptr[b + 1]; q++) {
    const int32_t J = adj[q];
    double acc[4] = {0., 0., 0., 0.};
    for (int k = 0; k < len; k++) {
      const int64_t e = cb + int64_t(k) * 64 + lane;
      const int32_t j = a.col[e];
      if ((j >> 8) != J) continue;
      const uint32_t fj = a.flags[j];
      const double cj0 = dof_fixed(fj, 0) ? 0. : 1.;
      const double cj1 = dof_fixed(fj, 1) ? 0. : 1.;
      const double4 B = blk[e];
      acc[0] += B.x * ri0 * cj0;
      acc[1] += B.y * ri0 * cj1;
      acc[2] += B.z * ri1 * cj0;
      acc[3] += B.w * ri1 * cj1;
    }
    if (valid && J == b) {
      const double4 D = dg[i];
      acc[0] += D.x * ri0 * ri0;
      acc[1] += D.y * ri0 * ri1;
      acc[2] += D.z * ri1 * ri0;
      acc[3] += D.w * ri1 * ri1;
    }
    for (int t = 0; t < 4; t++) red[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w)
        for (int t = 0; t < 4; t++)
          red[t][threadIdx.x] += red[t][threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      Ac[(2 * b) * m + 2 * J] = red[0][0];
      Ac[(2 * b) * m + 2 * J + 1] = red[1][0];
      Ac[(2 * b + 1) * m + 2 * J] = red[2][0];
      Ac[(2 * b + 1) * m + 2 * J + 1] = red[3][0];
    }
    __syncthreads();
  }
}

__global__ void k_fill(double* __restrict__ p, int64_t cnt, double v) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t < cnt) p[t] = v;
}

// rc = R^T v (free dofs) per aggregate.
__global__ __launch_bounds__(256) void k_restrict(VecArgs a,
                                                  const double2* __restrict__ v) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double2 w = make_double2(0., 0.);
  if (i < a.n) {
    w = v[i];
    const uint32_t fl = a.flags[i];
    if (dof_fixed(fl, 0)) w.x = 0.;
    if (dof_fixed(fl, 1)) w.y = 0.;
  }
  block_restrict(w.x, w.y, a.rc, blockIdx.x);
}

// z = R zc on free dofs.
__global__ __launch_bounds__(256) void k_prolong(VecArgs a,
                                                 const double* __restrict__ zc,
                                                 double2* __restrict__ z) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t fl = a.flags[i];
  const int64_t b = i >> 8;
  z[i] = make_double2(dof_fixed(fl, 0) ? 0. : zc[2 * b],
                      dof_fixed(fl, 1) ? 0. : zc[2 * b + 1]);
}

// Coarse dofs without a free fine dof get an identity row and column.
__global__ void k_coarse_fixdiag(double* __restrict__ Ac, int m) {
  const int d = blockIdx.x * 256 + threadIdx.x;
  if (d < m && Ac[int64_t(d) * m + d] == 0.) Ac[int64_t(d) * m + d] = 1.;
}

// One Gauss-Jordan sweep on pivot k, A -> B (after all m sweeps, A^-1).
__global__ void k_gj_sweep(const double* __restrict__ A, double* __restrict__ B,
                           int m, int k) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= int64_t(m) * m) return;
  const int i = int(t / m), j = int(t % m);
  const double p = A[int64_t(k) * m + k];
  double v;
  if (i == k && j == k)
    v = 1. / p;
  else if (i == k)
    v = A[int64_t(k) * m + j] / p;
  else if (j == k)
    v = -A[int64_t(i) * m + k] / p;
  else
    v = A[t] - A[int64_t(i) * m + k] * A[int64_t(k) * m + j] / p;
  B[t] = v;
}

// Step 3 (after t = A sh with partials (t, s) in P2 and (t, t) in P3):
// omega, x += alpha ph + omega sh, r = s - omega t; partials (rh, r), (r, r).
__global__ __launch_bounds__(256) void k_bicg_x(VecArgs a, int k) {
  const int done = a.ctl[0];
  const int it = a.ctl[2] + k;
  const double al[2] = {a.sc.alpha_p(0), a.sc.alpha_p(1)};
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const bool own = i < a.n;
  const double2 z = make_double2(0., 0.);
  const double2 ph = own ? a.ph[i] : z, sh = own ? a.sh[i] : z,
                s = own ? a.s[i] : z, t = own ? a.t[i] : z,
                x0 = own ? a.x[i] : z, rh = own ? a.rh[i] : z;
  double red[2];
  {
    const double* const P[2] = {a.P2, a.P3};
    block_sums<2>(P, a.nbs, red);
  }
  if (done || it > a.max_it) return;
  const double ts = red[0], tt = red[1];
  const double omega = tt > 0. ? ts / tt : 0.;
  const double alpha = al[it & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc.omega(it) = omega;
    a.sc.omega_p(it & 1) = omega;
  }
  double p0 = 0., p1 = 0.;
  if (own) {
    double2 x = x0;
    x.x += alpha * ph.x + omega * sh.x;
    x.y += alpha * ph.y + omega * sh.y;
    a.x[i] = x;
    const double2 r = make_double2(s.x - omega * t.x, s.y - omega * t.y);
    a.r[i] = r;
    p0 = rh.x * r.x + rh.y * r.y;
    p1 = r.x * r.x + r.y * r.y;
  }
  double v[2] = {p0, p1};
  double* const P[2] = {a.P0, a.P1};
  block_partials<2>(v, P, blockIdx.x);
}

// Newton helpers ------------------------------------------------------------

// R = f on free dofs (f = f_int + b), 0 on prescribed; partial (R, R).
__global__ __launch_bounds__(256) void k_qs_residual(
    int64_t n, const uint32_t* __restrict__ flags,
    const double2* __restrict__ f, double2* __restrict__ R,
    double* __restrict__ P) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double p = 0.;
  if (i < n) {
    double2 v = f[i];
    const uint32_t fl = flags[i];
    if (dof_fixed(fl, 0)) v.x = 0.;
    if (dof_fixed(fl, 1)) v.y = 0.;
    R[i] = v;
    p = v.x * v.x + v.y * v.y;
  }
  block_partial(p, P, blockIdx.x);
}

// dup = ULoading(t) - u on prescribed dofs, 0 on free dofs.
__global__ __launch_bounds__(256) void k_qs_dup(
    int64_t n, const uint32_t* __restrict__ flags,
    const double2* __restrict__ x, const double2* __restrict__ u,
    const BcTable* __restrict__ bc, const double* __restrict__ node_rate,
    double t, double2* __restrict__ dup) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t fl = flags[i];
  const double2 xi = x[i], ui = u[i];
  double d[2] = {0., 0.};
  const double uu[2] = {ui.x, ui.y};
  for (int c = 0; c < 2; c++) {
    const uint32_t id = (fl >> (kFlagUbcShift + 4 * c)) & 0xfu;
    if (id) {
      double du, dv;
      if (id == kNodeListBc) {
        du = node_rate[3 * i + c] * t;
      } else {
        ubc_eval(bc->u[id - 1], xi.x, xi.y, 0., t, du, dv);
      }
      d[c] = du - uu[c];
    }
  }
  dup[i] = make_double2(d[0], d[1]);
}

// b = R - y on free dofs; partial (b, b).
__global__ __launch_bounds__(256) void k_qs_sub(int64_t n,
                                                const uint32_t* __restrict__ flags,
                                                double2* __restrict__ R,
                                                const double2* __restrict__ y,
                                                double* __restrict__ P) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double p = 0.;
  if (i < n) {
    double2 v = R[i];
    const double2 w = y[i];
    v.x -= w.x;
    v.y -= w.y;
    const uint32_t fl = flags[i];
    if (dof_fixed(fl, 0)) v.x = 0.;
    if (dof_fixed(fl, 1)) v.y = 0.;
    R[i] = v;
    p = v.x * v.x + v.y * v.y;
  }
  block_partial(p, P, blockIdx.x);
}

// u += delta on free dofs; u += dup on prescribed dofs (dup may be null:
// prescribed dofs unchanged).
__global__ __launch_bounds__(256) void k_qs_update_u(
    int64_t n, const uint32_t* __restrict__ flags, double2* __restrict__ u,
    const double2* __restrict__ delta, const double2* __restrict__ dup) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t fl = flags[i];
  double2 ui = u[i];
  const double2 d = delta[i];
  const double2 p = dup ? dup[i] : make_double2(0., 0.);
  ui.x += dof_fixed(fl, 0) ? p.x : d.x;
  ui.y += dof_fixed(fl, 1) ? p.y : d.y;
  u[i] = ui;
}

// ---- host side ------------------------------------------------------------

QsArgs qs_args(nlm_ctx* c) {
  FdState& s = c->fd;
  QsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = s.n;
  a.nslices = s.sell.nslices;
  a.nb = (s.n + 255) / 256;
  a.x = s.x.as<double2>();
  a.u = s.u[s.cur].as<double2>();
  a.vol = s.vol.as<double>();
  a.flags = s.flags.as<uint32_t>();
  a.sl_ptr = s.sell.sl_ptr.as<int64_t>();
  a.col = s.sell.col.as<int32_t>();
  a.rowlen = s.sell.rowlen.as<uint16_t>();
  a.bits = s.sell.bits.as<unsigned long long>();
  a.m = s.rnp;
  return a;
}

int qs_check(nlm_ctx* c, const char* fn) {
  if (!c) return NLM_ERR_INVALID;
  if (!c->fd.ready)
    return fail(c, NLM_ERR_STATE, std::string(fn) + ": nlm_fd_setup not done");
  if (c->fd.dim != 2)
    return fail(c, NLM_ERR_INVALID, std::string(fn) + ": 2D models only");
  if (c->fd.material == NLM_MATERIAL_ELASTIC_STATE &&
      c->fd.es_par.critical_stretch > 0.)
    return fail(c, NLM_ERR_INVALID,
                std::string(fn) + ": the state-based material must not break "
                                  "bonds in a quasi-static solve");
  return NLM_OK;
}

// Checks that the tangent kind fits the material: the analytic tangents
// are the bond-based force's, the finite-difference tangent needs a force
// without bond breaking (the state-based material here).
int tangent_check(nlm_ctx* c, int tangent, const char* fn) {
  const bool es = c->fd.material == NLM_MATERIAL_ELASTIC_STATE;
  if (tangent == NLM_QS_TANGENT_FD && !es)
    return fail(c, NLM_ERR_INVALID,
                std::string(fn) + ": the finite-difference tangent needs the "
                                  "state-based material");
  if (tangent != NLM_QS_TANGENT_FD && es)
    return fail(c, NLM_ERR_INVALID,
                std::string(fn) + ": the state-based material needs "
                                  "NLM_QS_TANGENT_FD");
  if (tangent != NLM_QS_TANGENT_FULL && tangent != NLM_QS_TANGENT_COMPONENT &&
      tangent != NLM_QS_TANGENT_FD)
    return fail(c, NLM_ERR_INVALID, std::string(fn) + ": unknown tangent");
  return NLM_OK;
}

int qs_alloc(nlm_ctx* c, int max_iter) {
  FdState& s = c->fd;
  QsState& q = c->qs;
  const int64_t n = s.n;
  const size_t nent = size_t(s.sell.nentries);
  int rc;
  const bool grow_vec = !q.blk.p || q.blk.bytes < nent * sizeof(double4);
  if ((grow_vec || max_iter + 2 > q.max_iter_alloc) && q.bicg_graph) {
    // buffers the captured batch references are about to be replaced: the
    // executable graph goes first, so no live graph ever refers to freed
    // memory
    (void)hipGraphExecDestroy(q.bicg_graph);
    q.bicg_graph = nullptr;
    q.bicg_key.clear();
  }
  if (grow_vec) {
    if ((rc = alloc(c, q.blk, nent * sizeof(double4) + 256))) return rc;
    if ((rc = alloc(c, q.bg, nent * sizeof(double) + 256))) return rc;
    if ((rc = alloc(c, q.diag, n * sizeof(double4) + 256))) return rc;
    if ((rc = alloc(c, q.dinv, n * sizeof(double2) + 256))) return rc;
    DevBuf* vec[] = {&q.r, &q.rh, &q.p, &q.v, &q.s, &q.t, &q.x,
                     &q.ph, &q.sh, &q.b, &q.y, &q.dup, &q.u0};
    for (DevBuf* d : vec)
      if ((rc = alloc(c, *d, n * sizeof(double2) + 256))) return rc;
    q.nb = (n + 255) / 256;
    if ((rc = alloc(c, q.part, (2 * q.nb + 2 * s.sell.nslices) *
                                   sizeof(double) + 256)))
      return rc;
    if ((rc = alloc(c, q.ctl, 64))) return rc;
    if (!q.e0) NLM_HIP(c, hipEventCreate(&q.e0));
    if (!q.e1) NLM_HIP(c, hipEventCreate(&q.e1));
    if (!q.hctl)
      NLM_HIP(c, hipHostMalloc(reinterpret_cast<void**>(&q.hctl),
                               4 * sizeof(int), hipHostMallocDefault));
  }
  if (max_iter + 2 > q.max_iter_alloc) {
    q.max_iter_alloc = max_iter + 2;
    if ((rc = alloc(c, q.scal, (8 + 4 * size_t(q.max_iter_alloc)) *
                                   sizeof(double))))
      return rc;
  }
  return NLM_OK;
}

VecArgs vec_args(nlm_ctx* c) {
  QsState& q = c->qs;
  VecArgs a;
  std::memset(&a, 0, sizeof(a));  // padding too: the graph key compares bytes
  a.n = c->fd.n;
  a.nb = q.nb;
  a.flags = c->fd.flags.as<uint32_t>();
  a.r = q.r.as<double2>();
  a.rh = q.rh.as<double2>();
  a.p = q.p.as<double2>();
  a.v = q.v.as<double2>();
  a.s = q.s.as<double2>();
  a.t = q.t.as<double2>();
  a.x = q.x.as<double2>();
  a.ph = q.ph.as<double2>();
  a.sh = q.sh.as<double2>();
  a.dinv = q.dinv.as<double2>();
  double* P = q.part.as<double>();
  a.nbs = c->fd.sell.nslices;
  a.P0 = P;
  a.P1 = P + q.nb;
  a.P2 = P + 2 * q.nb;
  a.P3 = P + 2 * q.nb + a.nbs;
  a.rc = nullptr;
  a.ctl = q.ctl.as<int>();
  a.sc.base = q.scal.as<double>();
  a.sc.L = q.max_iter_alloc;
  return a;
}

// BiCGSTAB iterations per host check (and per graph replay).
constexpr int kBicgBatch = 8;

// Everything a captured BiCGSTAB batch depends on.
struct BicgKey {
  VecArgs a;
  QsArgs qa;
  int precond;
  int coarse;
  int m;
  unsigned g, gmv;
  const double* cinv;
  const void *blk, *diag, *mblk, *mdiag, *zcp, *zcs;
};

unsigned node_grid(const FdState& s) { return unsigned((s.n + 255) / 256); }
unsigned slice_grid(const FdState& s) {
  return grid_xcd(force_blocks(s) * 256, 256);
}

// Host sum of one partial slot (nb doubles).
int partial_sum(nlm_ctx* c, const double* P, double* out) {
  QsState& q = c->qs;
  std::vector<double> h(q.nb);
  NLM_HIP(c, hipMemcpyAsync(h.data(), P, sizeof(double) * q.nb,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  double s = 0.;
  for (double v : h) s += v;
  *out = s;
  return NLM_OK;
}

// Greedy coloring of the nodes (internal order): two nodes closer than
// 4 horizons (plus the neighbor test's tolerance) get different colors, so
// no force f_i, which reads the displacements within two horizons of i,
// sees two perturbed nodes.
int fd_colors(nlm_ctx* c) {
  FdState& s = c->fd;
  QsState& q = c->qs;
  if (q.ncolors > 0) return NLM_OK;
  const int64_t n = s.n;
  std::vector<double2> x(n);
  NLM_HIP(c, hipMemcpy(x.data(), s.x.p, sizeof(double2) * n,
                       hipMemcpyDeviceToHost));
  const double horizon = s.material == NLM_MATERIAL_ELASTIC_STATE
                             ? s.es_par.horizon
                             : s.mat.horizon;
  const double R = 4. * (horizon + 1e-10) * (1. + 2e-5);
  const double R2 = R * R;
  double lo0 = x[0].x, lo1 = x[0].y, hi0 = lo0, hi1 = lo1;
  for (const auto& p : x) {
    lo0 = std::min(lo0, p.x);
    lo1 = std::min(lo1, p.y);
    hi0 = std::max(hi0, p.x);
    hi1 = std::max(hi1, p.y);
  }
  const int64_t nc0 = int64_t((hi0 - lo0) / R) + 1;
  const int64_t nc1 = int64_t((hi1 - lo1) / R) + 1;
  auto cell = [&](const double2& p, int64_t* a, int64_t* b) {
    *a = std::min(nc0 - 1, int64_t((p.x - lo0) / R));
    *b = std::min(nc1 - 1, int64_t((p.y - lo1) / R));
  };
  std::vector<int64_t> head(nc0 * nc1, -1), next(n, -1);
  for (int64_t i = 0; i < n; i++) {
    int64_t a, b;
    cell(x[i], &a, &b);
    next[i] = head[a * nc1 + b];
    head[a * nc1 + b] = i;
  }
  std::vector<int> color(n, -1);
  std::vector<int64_t> seen;  // seen[c] == i: color c taken near node i
  int ncol = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t a, b;
    cell(x[i], &a, &b);
    for (int64_t da = -1; da <= 1; da++)
      for (int64_t db = -1; db <= 1; db++) {
        const int64_t ca = a + da, cb = b + db;
        if (ca < 0 || cb < 0 || ca >= nc0 || cb >= nc1) continue;
        for (int64_t j = head[ca * nc1 + cb]; j >= 0; j = next[j]) {
          if (color[j] < 0) continue;
          const double d0 = x[j].x - x[i].x, d1 = x[j].y - x[i].y;
          if (d0 * d0 + d1 * d1 <= R2) seen[color[j]] = i;
        }
      }
    int cl = 0;
    while (cl < ncol && seen[cl] == i) cl++;
    if (cl == ncol) {
      ncol++;
      seen.push_back(-1);
    }
    color[i] = cl;
  }
  int rc;
  if ((rc = alloc(c, q.color, sizeof(int) * n))) return rc;
  NLM_HIP(c, hipMemcpy(q.color.p, color.data(), sizeof(int) * n,
                       hipMemcpyHostToDevice));
  q.ncolors = ncol;
  return NLM_OK;
}

// The reference's Jacobian by central differences of the force, colored:
// 2 * dim force evaluations per color.  u is restored and the force of the
// current state recomputed at the end.
int assemble_fd(nlm_ctx* c) {
  FdState& s = c->fd;
  QsState& q = c->qs;
  int rc;
  if ((rc = fd_colors(c))) return rc;
  const double h = s.material == NLM_MATERIAL_ELASTIC_STATE ? s.es_par.h
                                                             : s.mat.h;
  const double eps = q.fd_perturbation * h;
  if (!(eps > 0.))
    return fail(c, NLM_ERR_INVALID, "finite-difference tangent: eps <= 0");
  const size_t nent = size_t(s.sell.nentries);
  NLM_HIP(c, hipMemsetAsync(q.blk.p, 0, nent * sizeof(double4), c->stream));
  NLM_HIP(c, hipMemsetAsync(q.diag.p, 0, s.n * sizeof(double4), c->stream));
  double2* u = s.u[s.cur].as<double2>();
  NLM_HIP(c, hipMemcpyAsync(q.u0.p, u, sizeof(double2) * s.n,
                            hipMemcpyDeviceToDevice, c->stream));
  const unsigned g = node_grid(s);
  const double t = s.time;
  for (int cl = 0; cl < q.ncolors; cl++)
    for (int r = 0; r < 2; r++) {
      hipLaunchKernelGGL(k_fd_perturb, dim3(g), dim3(256), 0, c->stream, s.n,
                         q.u0.as<double2>(), q.color.as<int>(), cl, r, eps, u);
      fd_launch_force(c, false, t);
      NLM_HIP(c, hipMemcpyAsync(q.y.p, s.f.p, sizeof(double2) * s.n,
                                hipMemcpyDeviceToDevice, c->stream));
      hipLaunchKernelGGL(k_fd_perturb, dim3(g), dim3(256), 0, c->stream, s.n,
                         q.u0.as<double2>(), q.color.as<int>(), cl, r, -eps,
                         u);
      fd_launch_force(c, false, t);
      hipLaunchKernelGGL(k_fd_collect, dim3(slice_grid(s)), dim3(256), 0,
                         c->stream, qs_args(c), q.y.as<double2>(),
                         s.f.as<double2>(), q.color.as<int>(), cl, r, eps,
                         q.blk.as<double4>(), q.diag.as<double4>());
    }
  NLM_HIP(c, hipMemcpyAsync(u, q.u0.p, sizeof(double2) * s.n,
                            hipMemcpyDeviceToDevice, c->stream));
  fd_launch_force(c, false, t);
  hipLaunchKernelGGL(k_fd_dinv, dim3(g), dim3(256), 0, c->stream, s.n,
                     s.flags.as<uint32_t>(), q.diag.as<double4>(),
                     q.dinv.as<double2>());
  NLM_HIP(c, hipGetLastError());
  return NLM_OK;
}

int assemble(nlm_ctx* c, int tangent) {
  QsState& q = c->qs;
  int rc;
  if ((rc = qs_alloc(c, 8))) return rc;
  q.bs_mode = 0;
  if (tangent == NLM_QS_TANGENT_FD) {
    if ((rc = assemble_fd(c))) return rc;
    q.assembled = true;
    q.tangent = tangent;
    return NLM_OK;
  }
  hipLaunchKernelGGL(k_qs_tangent, dim3(slice_grid(c->fd)), dim3(256), 0,
                     c->stream, qs_args(c), tangent, q.blk.as<double4>(),
                     q.diag.as<double4>(), q.dinv.as<double2>(),
                     q.bg.as<double>());
  NLM_HIP(c, hipGetLastError());
  q.assembled = true;
  q.tangent = tangent;
  q.bs_mode = tangent == NLM_QS_TANGENT_COMPONENT ? 2 : 1;
  return NLM_OK;
}

template <int DOT, class BT = double4, class DT = BT>
void spmv(nlm_ctx* c, const DevBuf& blk, const DevBuf& dg, const double2* x,
          double2* y, const double2* w, double* P0, double* P1,
          const int* ctl, const double* rc = nullptr) {
  const unsigned g = unsigned((c->fd.sell.nslices + 7) / 8 * 8);
  const QsState& q = c->qs;
  // the tangent's products from its bond scalars where it has them
  const int bs = (&blk == &q.blk && q.bond_scalar) ? q.bs_mode : 0;
  const double* Ai = rc ? q.cinv : nullptr;
  const double* bg = q.bg.as<double>();
  if (bs == 1)
    hipLaunchKernelGGL((k_qs_spmv4<DOT, BT, 1, DT>), dim3(g), dim3(256), 0,
                       c->stream, qs_args(c), blk.as<BT>(), dg.as<DT>(), x, y,
                       w, P0, P1, ctl, Ai, q.m, rc, bg);
  else if (bs == 2)
    hipLaunchKernelGGL((k_qs_spmv4<DOT, BT, 2, DT>), dim3(g), dim3(256), 0,
                       c->stream, qs_args(c), blk.as<BT>(), dg.as<DT>(), x, y,
                       w, P0, P1, ctl, Ai, q.m, rc, bg);
  else
    hipLaunchKernelGGL((k_qs_spmv4<DOT, BT, 0, DT>), dim3(g), dim3(256), 0,
                       c->stream, qs_args(c), blk.as<BT>(), dg.as<DT>(), x, y,
                       w, P0, P1, ctl, Ai, q.m, rc, nullptr);
}
// The preconditioner's bond blocks are read from their BF16 copy, its node
// blocks in FP32 (rounding perturbs the preconditioner only, never the
// operator).  rc non-null: add the coarse correction of the restricted
// vector rc.
void mapply(nlm_ctx* c, const double2* x, double2* y, const int* ctl,
            const double* rc) {
  spmv<0, Bf4, float4>(c, c->qs.mhalf, c->qs.mdiag, x, y, nullptr, nullptr,
                       nullptr, ctl, rc);
}

// The BF16 copy of the preconditioner's bond blocks, made once after they
// change.
int precond_bf16(nlm_ctx* c) {
  QsState& q = c->qs;
  if (q.mhalf_ok) return NLM_OK;
  const int64_t nent = c->fd.sell.nentries;
  int rc;
  if (!q.mhalf.p || q.mhalf.bytes < sizeof(Bf4) * nent) {
    if (q.bicg_graph) {  // it may reference the buffer being replaced
      (void)hipGraphExecDestroy(q.bicg_graph);
      q.bicg_graph = nullptr;
      q.bicg_key.clear();
    }
    if ((rc = alloc(c, q.mhalf, sizeof(Bf4) * nent + 256))) return rc;
  }
  hipLaunchKernelGGL(k_blocks_bf16, dim3(unsigned((nent + 255) / 256)),
                     dim3(256), 0, c->stream, nent, q.mblk.as<float4>(),
                     q.mhalf.as<Bf4>());
  NLM_HIP(c, hipGetLastError());
  q.mhalf_ok = true;
  return NLM_OK;
}

// Coarse operator of the assembled tangent: adjacency of the aggregates
// (once per layout), A_c assembly and its inverse in q.cinv.
int coarse_setup(nlm_ctx* c) {
  FdState& s = c->fd;
  QsState& q = c->qs;
  int rc;
  const int64_t nagg = q.nb;
  const int m = int(2 * nagg);
  if (m > 8192)
    return fail(c, NLM_ERR_INVALID,
                "two-level preconditioner: more than 4096 aggregates");
  if (!q.adj_ptr.p) {
    std::vector<int64_t> slp(s.sell.nslices + 1);
    std::vector<int32_t> col(s.sell.nentries);
    std::vector<uint16_t> rl(s.n);
    NLM_HIP(c, hipStreamSynchronize(c->stream));
    NLM_HIP(c, hipMemcpy(slp.data(), s.sell.sl_ptr.p,
                         sizeof(int64_t) * slp.size(), hipMemcpyDeviceToHost));
    NLM_HIP(c, hipMemcpy(col.data(), s.sell.col.p,
                         sizeof(int32_t) * col.size(), hipMemcpyDeviceToHost));
    NLM_HIP(c, hipMemcpy(rl.data(), s.sell.rowlen.p, sizeof(uint16_t) * s.n,
                         hipMemcpyDeviceToHost));
    std::vector<int64_t> ap(nagg + 1, 0);
    std::vector<int32_t> adj;
    for (int64_t b = 0; b < nagg; b++) {
      std::vector<int32_t> J{int32_t(b)};
      for (int64_t i = b * 256; i < std::min(s.n, b * 256 + 256); i++)
        for (int k = 0; k < rl[i]; k++)
          J.push_back(col[slp[i / 64] + 64 * k + i % 64] >> 8);
      std::sort(J.begin(), J.end());
      J.erase(std::unique(J.begin(), J.end()), J.end());
      adj.insert(adj.end(), J.begin(), J.end());
      ap[b + 1] = int64_t(adj.size());
    }
    if ((rc = alloc(c, q.adj_ptr, sizeof(int64_t) * ap.size()))) return rc;
    if ((rc = alloc(c, q.adj, sizeof(int32_t) * adj.size()))) return rc;
    NLM_HIP(c, hipMemcpy(q.adj_ptr.p, ap.data(), sizeof(int64_t) * ap.size(),
                         hipMemcpyHostToDevice));
    NLM_HIP(c, hipMemcpy(q.adj.p, adj.data(), sizeof(int32_t) * adj.size(),
                         hipMemcpyHostToDevice));
    const size_t mb = sizeof(double) * size_t(m) * m + 256;
    if ((rc = alloc(c, q.cA, mb)) || (rc = alloc(c, q.cB, mb)) ||
        (rc = alloc(c, q.rc, sizeof(double) * (m + 64))) ||
        (rc = alloc(c, q.zcp, sizeof(double) * (m + 64))) ||
        (rc = alloc(c, q.zcs, sizeof(double) * (m + 64))))
      return rc;
  }
  const int64_t mm2 = int64_t(m) * m;
  const unsigned gm = unsigned((mm2 + 255) / 256);
  hipLaunchKernelGGL(k_fill, dim3(gm), dim3(256), 0, c->stream,
                     q.cA.as<double>(), mm2, 0.);
  hipLaunchKernelGGL(k_coarse_assemble, dim3(unsigned(nagg)), dim3(256), 0,
                     c->stream, qs_args(c), q.blk.as<double4>(),
                     q.diag.as<double4>(), q.adj_ptr.as<int64_t>(),
                     q.adj.as<int32_t>(), m, q.cA.as<double>());
  hipLaunchKernelGGL(k_coarse_fixdiag, dim3(unsigned((m + 255) / 256)),
                     dim3(256), 0, c->stream, q.cA.as<double>(), m);
  double* A = q.cA.as<double>();
  double* B = q.cB.as<double>();
  for (int k = 0; k < m; k++) {
    hipLaunchKernelGGL(k_gj_sweep, dim3(gm), dim3(256), 0, c->stream, A, B, m,
                       k);
    std::swap(A, B);
  }
  q.cinv = A;
  q.m = m;
  NLM_HIP(c, hipGetLastError());
  return NLM_OK;
}

// BiCGSTAB on A_ff x = b (q.b -> q.x).  Returns iterations, relative
// residual and whether it converged.
int bicgstab(nlm_ctx* c, int precond, double tol, int max_iter, int* iters,
             double* relres, bool* ok) {
  FdState& s = c->fd;
  QsState& q = c->qs;
  int rc;
  if ((rc = qs_alloc(c, max_iter))) return rc;
  const bool coarse = (precond & NLM_QS_PRECOND_COARSE) != 0;
  precond &= ~NLM_QS_PRECOND_COARSE;
  if (precond == NLM_QS_PRECOND_BLOCKS && !q.has_m)
    return fail(c, NLM_ERR_STATE, "BiCGSTAB: no preconditioner blocks set");
  if (precond == NLM_QS_PRECOND_BLOCKS && (rc = precond_bf16(c))) return rc;
  if (coarse && (!q.cinv || precond == NLM_QS_PRECOND_NONE))
    return fail(c, NLM_ERR_STATE, "BiCGSTAB: no coarse operator set up");
  VecArgs a = vec_args(c);
  a.rc = coarse ? q.rc.as<double>() : nullptr;
  const unsigned gmv = unsigned((q.m + 3) / 4);
  const unsigned g = node_grid(s);
  const unsigned gsl = unsigned((s.sell.nslices + 7) / 8 * 8);
  hipLaunchKernelGGL(k_bicg_init, dim3(g), dim3(256), 0, c->stream, a,
                     q.b.as<double2>());
  NLM_HIP(c, hipGetLastError());
  double bb = 0.;
  if ((rc = partial_sum(c, a.P1, &bb))) return rc;
  const double tol2 = tol * tol * bb;
  NLM_HIP(c, hipMemcpyAsync(q.scal.p, &tol2, sizeof(double),
                            hipMemcpyHostToDevice, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  if (bb == 0.) {
    *iters = 0;
    *relres = 0.;
    *ok = true;
    return NLM_OK;
  }
  const double2* rh = q.rh.as<double2>();
  a.max_it = max_iter;
  int ctl[2] = {0, 0};
  const int batch = kBicgBatch;
  // one batch: iterations ctl[2] .. ctl[2] + batch - 1 (no-ops past
  // max_iter or after convergence), then the batch start moves on
  auto launch_batch = [&]() {
    for (int k = 0; k < batch && coarse; k++) {
      // two-level: smoother + R A_c^-1 R^T, for p and then for s
      hipLaunchKernelGGL((k_bicg_p<2, true>), dim3(g), dim3(256), 0, c->stream,
                         a, k);
      if (precond == NLM_QS_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_smooth_jacobi, dim3(g), dim3(256), 0, c->stream,
                           a, a.p, a.ph, q.cinv, q.m);
      else
        mapply(c, a.p, a.ph, a.ctl, a.rc);
      spmv<1>(c, q.blk, q.diag, a.ph, a.v, rh, a.P2, nullptr, a.ctl);
      hipLaunchKernelGGL((k_bicg_s<2, true>), dim3(g), dim3(256), 0, c->stream,
                         a, k);
      if (precond == NLM_QS_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_smooth_jacobi, dim3(g), dim3(256), 0, c->stream,
                           a, a.s, a.sh, q.cinv, q.m);
      else
        mapply(c, a.s, a.sh, a.ctl, a.rc);
      spmv<2>(c, q.blk, q.diag, a.sh, a.t, a.s, a.P2, a.P3, a.ctl);
      hipLaunchKernelGGL(k_bicg_x, dim3(g), dim3(256), 0, c->stream, a, k);
    }
    for (int k = 0; k < batch && !coarse; k++) {
      if (precond == NLM_QS_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_bicg_p<1>, dim3(g), dim3(256), 0, c->stream, a,
                           k);
      else if (precond == NLM_QS_PRECOND_NONE)
        hipLaunchKernelGGL(k_bicg_p<0>, dim3(g), dim3(256), 0, c->stream, a,
                           k);
      else {
        hipLaunchKernelGGL(k_bicg_p<2>, dim3(g), dim3(256), 0, c->stream, a,
                           k);
        mapply(c, a.p, a.ph, a.ctl, nullptr);
      }
      spmv<1>(c, q.blk, q.diag, a.ph, a.v, rh, a.P2, nullptr, a.ctl);
      if (precond == NLM_QS_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_bicg_s<1>, dim3(g), dim3(256), 0, c->stream, a,
                           k);
      else if (precond == NLM_QS_PRECOND_NONE)
        hipLaunchKernelGGL(k_bicg_s<0>, dim3(g), dim3(256), 0, c->stream, a,
                           k);
      else
        hipLaunchKernelGGL(k_bicg_s_blocks, dim3(gsl), dim3(256), 0,
                           c->stream, a, qs_args(c), q.mhalf.as<Bf4>(),
                           q.mdiag.as<float4>(), k);
      spmv<2>(c, q.blk, q.diag, a.sh, a.t, a.s, a.P2, a.P3, a.ctl);
      hipLaunchKernelGGL(k_bicg_x, dim3(g), dim3(256), 0, c->stream, a, k);
    }
    hipLaunchKernelGGL(k_iter_advance, dim3(1), dim3(64), 0, c->stream, a.ctl,
                       batch);
  };
  hipGraphExec_t exec = nullptr;
  if (q.use_graph) {
    // the captured batch depends on every kernel argument: recapture when
    // any of them changed since the last capture
    BicgKey key;
    std::memset(&key, 0, sizeof(key));
    std::memcpy(&key.a, &a, sizeof(a));
    const QsArgs qa = qs_args(c);
    std::memcpy(&key.qa, &qa, sizeof(qa));
    key.precond = precond;
    key.coarse = coarse;
    key.m = q.m;
    key.cinv = q.cinv;
    key.blk = q.blk.p;
    key.diag = q.diag.p;
    key.mblk = q.mhalf.p;
    key.mdiag = q.mdiag.p;
    key.zcp = q.zcp.p;
    key.zcs = q.zcs.p;
    key.g = g;
    key.gmv = gmv;
    const unsigned char* kb = reinterpret_cast<const unsigned char*>(&key);
    if (!q.bicg_graph || q.bicg_key.size() != sizeof(key) ||
        std::memcmp(kb, q.bicg_key.data(), sizeof(key)) != 0) {
      if (q.bicg_graph) (void)hipGraphExecDestroy(q.bicg_graph);
      q.bicg_graph = nullptr;
      hipGraph_t gr = nullptr;
      NLM_HIP(c, hipStreamBeginCapture(c->stream,
                                       hipStreamCaptureModeThreadLocal));
      launch_batch();
      NLM_HIP(c, hipStreamEndCapture(c->stream, &gr));
      hipError_t e = hipGraphInstantiate(&q.bicg_graph, gr, nullptr, nullptr,
                                         0);
      (void)hipGraphDestroy(gr);
      if (e != hipSuccess) {
        q.bicg_graph = nullptr;
        return fail(c, NLM_ERR_HIP, "BiCGSTAB: graph instantiation failed");
      }
      q.bicg_key.assign(kb, kb + sizeof(key));
    }
    exec = q.bicg_graph;
  }
  // One batch stays queued ahead of the host's check of the previous one,
  // so the device does not idle through the round trip; a batch queued
  // after the finish (convergence, breakdown or the iteration cap) runs as
  // no-ops.  Each batch's control words land in their own pinned slot.
  const int nbatch = (max_iter + batch - 1) / batch;
  hipEvent_t ev[2] = {q.e0, q.e1};
  auto queue = [&](int b) -> int {
    if (exec)
      NLM_HIP(c, hipGraphLaunch(exec, c->stream));
    else
      launch_batch();
    NLM_HIP(c, hipGetLastError());
    NLM_HIP(c, hipMemcpyAsync(q.hctl + 2 * (b & 1), q.ctl.p, 2 * sizeof(int),
                              hipMemcpyDeviceToHost, c->stream));
    NLM_HIP(c, hipEventRecord(ev[b & 1], c->stream));
    return NLM_OK;
  };
  if ((rc = queue(0))) return rc;
  for (int b = 0; b < nbatch; b++) {
    if (b + 1 < nbatch && (rc = queue(b + 1))) return rc;
    NLM_HIP(c, hipEventSynchronize(ev[b & 1]));
    ctl[0] = q.hctl[2 * (b & 1)];
    ctl[1] = q.hctl[2 * (b & 1) + 1];
    if (ctl[0]) break;
  }
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  double rr = 0.;
  int done_it = max_iter;
  if (ctl[0]) {
    done_it = ctl[1];
    const size_t off = 8 + 3 * size_t(q.max_iter_alloc) + size_t(done_it);
    NLM_HIP(c, hipMemcpy(&rr, q.scal.as<double>() + off, sizeof(double),
                         hipMemcpyDeviceToHost));
  } else {
    if ((rc = partial_sum(c, a.P1, &rr))) return rc;
  }
  *iters = done_it;
  *relres = std::sqrt(rr / bb);
  *ok = ctl[0] == 1;
  return NLM_OK;
}

// ||R||_2 of the residual of the current force (q.b <- R).
int residual(nlm_ctx* c, double* norm) {
  FdState& s = c->fd;
  QsState& q = c->qs;
  hipLaunchKernelGGL(k_qs_residual, dim3(node_grid(s)), dim3(256), 0,
                     c->stream, s.n, s.flags.as<uint32_t>(), s.f.as<double2>(),
                     q.b.as<double2>(), q.part.as<double>());
  NLM_HIP(c, hipGetLastError());
  double rr = 0.;
  int rc = partial_sum(c, q.part.as<double>(), &rr);
  *norm = std::sqrt(rr);
  return rc;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now() - t0)
      .count();
}

}  // namespace

// Preconditioner construction by the graph network (gnn.hip); returns
// NLM_ERR_STATE when no network is attached to the context.
int gnn_predict_precond(nlm_ctx* c);

}  // namespace nlm

using namespace nlm;

extern "C" int nlm_qs_set_graph(nlm_ctx* c, int on) {
  if (!c) return NLM_ERR_INVALID;
  c->qs.use_graph = on != 0;
  return NLM_OK;
}

extern "C" int nlm_qs_set_launch(nlm_ctx* c, int graph, int bond_scalar) {
  if (!c) return NLM_ERR_INVALID;
  c->qs.use_graph = graph != 0;
  c->qs.bond_scalar = bond_scalar != 0;
  return NLM_OK;
}

extern "C" int nlm_qs_assemble(nlm_ctx* c, int tangent) {
  int rc = qs_check(c, "nlm_qs_assemble");
  if (rc) return rc;
  if ((rc = tangent_check(c, tangent, "nlm_qs_assemble"))) return rc;
  if ((rc = assemble(c, tangent))) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

namespace {
// Copies slice-layout blocks and per-node blocks to the reference layout.
template <class BT>
int blocks_to_host(nlm_ctx* c, const DevBuf& blk, const DevBuf& dg,
                   double* blocks, double* diag) {
  FdState& s = c->fd;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  const int64_t nent = s.sell.nentries;
  std::vector<BT> hb(nent), hd(s.n);
  std::vector<int64_t> slp(s.sell.nslices + 1);
  std::vector<int64_t> roff(s.n);
  std::vector<uint16_t> rl(s.n);
  NLM_HIP(c, hipMemcpy(hb.data(), blk.p, sizeof(BT) * nent,
                       hipMemcpyDeviceToHost));
  NLM_HIP(c, hipMemcpy(hd.data(), dg.p, sizeof(BT) * s.n,
                       hipMemcpyDeviceToHost));
  NLM_HIP(c, hipMemcpy(slp.data(), s.sell.sl_ptr.p,
                       sizeof(int64_t) * slp.size(), hipMemcpyDeviceToHost));
  NLM_HIP(c, hipMemcpy(roff.data(), s.ref_off.p, sizeof(int64_t) * s.n,
                       hipMemcpyDeviceToHost));
  NLM_HIP(c, hipMemcpy(rl.data(), s.sell.rowlen.p, sizeof(uint16_t) * s.n,
                       hipMemcpyDeviceToHost));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t sl = a / 64, lane = a % 64;
    if (blocks)
      for (int k = 0; k < rl[a]; k++) {
        const BT B = hb[slp[sl] + 64 * k + lane];
        double* o = blocks + 4 * (roff[a] + k);
        o[0] = B.x;
        o[1] = B.y;
        o[2] = B.z;
        o[3] = B.w;
      }
    if (diag) {
      const BT D = hd[a];
      double* o = diag + 4 * int64_t(s.perm_host[a]);
      o[0] = D.x;
      o[1] = D.y;
      o[2] = D.z;
      o[3] = D.w;
    }
  }
  return NLM_OK;
}
}  // namespace

extern "C" int nlm_qs_get_tangent(nlm_ctx* c, double* blocks, double* diag) {
  int rc = qs_check(c, "nlm_qs_get_tangent");
  if (rc) return rc;
  if (!c->qs.assembled)
    return fail(c, NLM_ERR_STATE, "nlm_qs_get_tangent: nothing assembled");
  return blocks_to_host<double4>(c, c->qs.blk, c->qs.diag, blocks, diag);
}

extern "C" int nlm_qs_get_precond(nlm_ctx* c, double* blocks, double* diag) {
  int rc = qs_check(c, "nlm_qs_get_precond");
  if (rc) return rc;
  if (!c->qs.has_m)
    return fail(c, NLM_ERR_STATE, "nlm_qs_get_precond: no blocks");
  return blocks_to_host<float4>(c, c->qs.mblk, c->qs.mdiag, blocks, diag);
}

extern "C" int nlm_qs_set_precond(nlm_ctx* c, const double* blocks,
                                  const double* diag) {
  int rc = qs_check(c, "nlm_qs_set_precond");
  if (rc) return rc;
  if (!blocks || !diag)
    return fail(c, NLM_ERR_INVALID, "nlm_qs_set_precond: null blocks");
  FdState& s = c->fd;
  QsState& q = c->qs;
  if ((rc = qs_alloc(c, 8))) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  const int64_t nent = s.sell.nentries;
  if (!q.mblk.p || q.mblk.bytes < sizeof(float4) * nent) {
    if ((rc = alloc(c, q.mblk, sizeof(float4) * nent + 256))) return rc;
    if ((rc = alloc(c, q.mdiag, sizeof(float4) * s.n + 256))) return rc;
  }
  std::vector<float4> hb(nent, make_float4(0.f, 0.f, 0.f, 0.f)), hd(s.n);
  std::vector<int64_t> slp(s.sell.nslices + 1), roff(s.n);
  std::vector<uint16_t> rl(s.n);
  NLM_HIP(c, hipMemcpy(slp.data(), s.sell.sl_ptr.p,
                       sizeof(int64_t) * slp.size(), hipMemcpyDeviceToHost));
  NLM_HIP(c, hipMemcpy(roff.data(), s.ref_off.p, sizeof(int64_t) * s.n,
                       hipMemcpyDeviceToHost));
  NLM_HIP(c, hipMemcpy(rl.data(), s.sell.rowlen.p, sizeof(uint16_t) * s.n,
                       hipMemcpyDeviceToHost));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t sl = a / 64, lane = a % 64;
    for (int k = 0; k < rl[a]; k++) {
      const double* b = blocks + 4 * (roff[a] + k);
      hb[slp[sl] + 64 * k + lane] =
          make_float4(float(b[0]), float(b[1]), float(b[2]), float(b[3]));
    }
    const double* d = diag + 4 * int64_t(s.perm_host[a]);
    hd[a] = make_float4(float(d[0]), float(d[1]), float(d[2]), float(d[3]));
  }
  NLM_HIP(c, hipMemcpy(q.mblk.p, hb.data(), sizeof(float4) * nent,
                       hipMemcpyHostToDevice));
  NLM_HIP(c, hipMemcpy(q.mdiag.p, hd.data(), sizeof(float4) * s.n,
                       hipMemcpyHostToDevice));
  q.has_m = true;
  q.mhalf_ok = false;
  return NLM_OK;
}

namespace {
// AoS double[3n] (reference order) <-> double2 per internal node.
int aos_to_dev(nlm_ctx* c, const double* src, DevBuf& dst) {
  FdState& s = c->fd;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double2> h(s.n);
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    h[a] = make_double2(src[3 * i], src[3 * i + 1]);
  }
  NLM_HIP(c, hipMemcpy(dst.p, h.data(), sizeof(double2) * s.n,
                       hipMemcpyHostToDevice));
  return NLM_OK;
}
int dev_to_aos(nlm_ctx* c, const DevBuf& src, double* dst) {
  FdState& s = c->fd;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double2> h(s.n);
  NLM_HIP(c, hipMemcpy(h.data(), src.p, sizeof(double2) * s.n,
                       hipMemcpyDeviceToHost));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    dst[3 * i] = h[a].x;
    dst[3 * i + 1] = h[a].y;
    dst[3 * i + 2] = 0.;
  }
  return NLM_OK;
}
}  // namespace

extern "C" int nlm_qs_apply(nlm_ctx* c, const double* x, double* y) {
  int rc = qs_check(c, "nlm_qs_apply");
  if (rc) return rc;
  QsState& q = c->qs;
  if (!q.assembled)
    return fail(c, NLM_ERR_STATE, "nlm_qs_apply: nothing assembled");
  if (!x || !y) return fail(c, NLM_ERR_INVALID, "nlm_qs_apply: null vector");
  if ((rc = aos_to_dev(c, x, q.p))) return rc;
  // prescribed entries of x are ignored: zero them through the init kernel's
  // masking (b -> r)
  VecArgs a = vec_args(c);
  NLM_HIP(c, hipMemcpyAsync(q.b.p, q.p.p, sizeof(double2) * c->fd.n,
                            hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_bicg_init, dim3(node_grid(c->fd)), dim3(256), 0,
                     c->stream, a, q.b.as<double2>());
  spmv<0>(c, q.blk, q.diag, a.r, q.y.as<double2>(), nullptr, nullptr, nullptr,
          nullptr);
  NLM_HIP(c, hipGetLastError());
  return dev_to_aos(c, q.y, y);
}

extern "C" int nlm_qs_coarse_apply(nlm_ctx* c, const double* r, double* z) {
  int rc = qs_check(c, "nlm_qs_coarse_apply");
  if (rc) return rc;
  QsState& q = c->qs;
  if (!q.assembled)
    return fail(c, NLM_ERR_STATE, "nlm_qs_coarse_apply: nothing assembled");
  if (!r || !z) return fail(c, NLM_ERR_INVALID, "nlm_qs_coarse_apply: null");
  if ((rc = coarse_setup(c))) return rc;
  if ((rc = aos_to_dev(c, r, q.b))) return rc;
  VecArgs a = vec_args(c);
  a.rc = q.rc.as<double>();
  const unsigned g = node_grid(c->fd);
  hipLaunchKernelGGL(k_restrict, dim3(g), dim3(256), 0, c->stream, a,
                     q.b.as<double2>());
  hipLaunchKernelGGL(k_coarse_mv, dim3(unsigned((q.m + 3) / 4)), dim3(256), 0,
                     c->stream, q.cinv, q.m, a.rc, q.zcp.as<double>(), nullptr);
  hipLaunchKernelGGL(k_prolong, dim3(g), dim3(256), 0, c->stream, a,
                     q.zcp.as<double>(), q.y.as<double2>());
  NLM_HIP(c, hipGetLastError());
  return dev_to_aos(c, q.y, z);
}

extern "C" int nlm_qs_linear_solve(nlm_ctx* c, const double* b, double* x,
                                   int precond, double tol, int max_iter,
                                   int* iters, double* relres) {
  int rc = qs_check(c, "nlm_qs_linear_solve");
  if (rc) return rc;
  QsState& q = c->qs;
  if (!q.assembled)
    return fail(c, NLM_ERR_STATE, "nlm_qs_linear_solve: nothing assembled");
  if (!b || !x || max_iter < 1 || !(tol > 0.))
    return fail(c, NLM_ERR_INVALID, "nlm_qs_linear_solve: bad arguments");
  const int base = precond & ~NLM_QS_PRECOND_COARSE;
  if (base < NLM_QS_PRECOND_NONE || base > NLM_QS_PRECOND_BLOCKS ||
      (precond != base && base == NLM_QS_PRECOND_NONE))
    return fail(c, NLM_ERR_INVALID, "nlm_qs_linear_solve: unknown precond");
  if ((rc = qs_alloc(c, max_iter))) return rc;
  if (precond != base && (rc = coarse_setup(c))) return rc;
  if ((rc = aos_to_dev(c, b, q.b))) return rc;
  int it = 0;
  double rr = 0.;
  bool ok = false;
  if ((rc = bicgstab(c, precond, tol, max_iter, &it, &rr, &ok))) return rc;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return dev_to_aos(c, q.x, x);
}

extern "C" int nlm_qs_solve(nlm_ctx* c, const nlm_qs_options* o,
                            nlm_qs_stats* st) {
  int rc = qs_check(c, "nlm_qs_solve");
  if (rc) return rc;
  if (!o || o->nsteps < 0 || !(o->dt > 0.) || o->max_newton < 0 ||
      o->lin_max_iter < 1 || !(o->lin_tol > 0.))
    return fail(c, NLM_ERR_INVALID, "nlm_qs_solve: bad options");
  if ((rc = tangent_check(c, o->tangent, "nlm_qs_solve"))) return rc;
  if (o->perturbation > 0.) c->qs.fd_perturbation = o->perturbation;
  const int pbase = o->precond & ~NLM_QS_PRECOND_COARSE;
  const bool pcoarse = o->precond != pbase;
  if (pbase < NLM_QS_PRECOND_NONE || pbase > NLM_QS_PRECOND_BLOCKS ||
      (pcoarse && pbase == NLM_QS_PRECOND_NONE))
    return fail(c, NLM_ERR_INVALID, "nlm_qs_solve: unknown precond");
  FdState& s = c->fd;
  QsState& q = c->qs;
  if ((rc = qs_alloc(c, o->lin_max_iter))) return rc;
  nlm_qs_stats stats;
  std::memset(&stats, 0, sizeof(stats));
  const auto t_all = std::chrono::steady_clock::now();
  const unsigned g = node_grid(s);
  auto u_cur = [&]() { return s.u[s.cur].as<double2>(); };
  auto precond_setup = [&]() -> int {
    if (pbase != NLM_QS_PRECOND_BLOCKS && !pcoarse) return NLM_OK;
    const auto t0 = std::chrono::steady_clock::now();
    int r = NLM_OK;
    if (pbase == NLM_QS_PRECOND_BLOCKS) {
      r = gnn_predict_precond(c);
      if (r == NLM_ERR_STATE && q.has_m) r = NLM_OK;  // blocks set by hand
    }
    if (r == NLM_OK && pcoarse) r = coarse_setup(c);
    NLM_HIP(c, hipStreamSynchronize(c->stream));
    stats.ms_precond += ms_since(t0);
    return r;
  };
  auto linear = [&](int* it_out) -> int {
    const auto t0 = std::chrono::steady_clock::now();
    int it = 0;
    double rr = 0.;
    bool ok = false;
    int r = bicgstab(c, o->precond, o->lin_tol, o->lin_max_iter, &it, &rr,
                     &ok);
    stats.ms_linear += ms_since(t0);
    stats.linear_iters += it;
    if (!ok) stats.linear_failures++;
    *it_out = it;
    return r;
  };
  auto tangent = [&]() -> int {
    const auto t0 = std::chrono::steady_clock::now();
    int r = assemble(c, o->tangent);
    NLM_HIP(c, hipStreamSynchronize(c->stream));
    stats.ms_tangent += ms_since(t0);
    return r;
  };
  bool have_precond = false;
  for (int step = 0; step < o->nsteps; step++) {
    // QuasiStaticModel::solver: d_time = d_n * dt (QuasiStaticModel.cpp:539)
    const double t = double(s.step + 1) * o->dt;
    double rnorm = 0., r0 = 0.;
    int newton = 0, lit = 0;
    if (o->predictor) {
      // linearized first iteration at the previous equilibrium
      hipLaunchKernelGGL(k_qs_dup, dim3(g), dim3(256), 0, c->stream, s.n,
                         s.flags.as<uint32_t>(), s.x.as<double2>(), u_cur(),
                         s.bctab.as<BcTable>(), s.node_rate.as<double>(), t,
                         q.dup.as<double2>());
      fd_launch_force(c, false, t);
      if ((rc = residual(c, &rnorm))) return rc;
      if ((rc = tangent())) return rc;
      spmv<0>(c, q.blk, q.diag, q.dup.as<double2>(), q.y.as<double2>(),
              nullptr, nullptr, nullptr, nullptr);
      hipLaunchKernelGGL(k_qs_sub, dim3(g), dim3(256), 0, c->stream, s.n,
                         s.flags.as<uint32_t>(), q.b.as<double2>(),
                         q.y.as<double2>(), q.part.as<double>());
      double bb = 0.;
      if ((rc = partial_sum(c, q.part.as<double>(), &bb))) return rc;
      r0 = std::sqrt(bb);
      if (!(o->gnn_refresh == 2 && have_precond)) {
        if ((rc = precond_setup())) return rc;
        have_precond = true;
      }
      if ((rc = linear(&lit))) return rc;
      hipLaunchKernelGGL(k_qs_update_u, dim3(g), dim3(256), 0, c->stream, s.n,
                         s.flags.as<uint32_t>(), u_cur(), q.x.as<double2>(),
                         q.dup.as<double2>());
      newton = 1;
      fd_launch_force(c, false, t);
      if ((rc = residual(c, &rnorm))) return rc;
    } else {
      // ULoading at t on the prescribed dofs, then the residual
      hipLaunchKernelGGL(k_qs_dup, dim3(g), dim3(256), 0, c->stream, s.n,
                         s.flags.as<uint32_t>(), s.x.as<double2>(), u_cur(),
                         s.bctab.as<BcTable>(), s.node_rate.as<double>(), t,
                         q.dup.as<double2>());
      NLM_HIP(c, hipMemsetAsync(q.x.p, 0, sizeof(double2) * s.n, c->stream));
      hipLaunchKernelGGL(k_qs_update_u, dim3(g), dim3(256), 0, c->stream, s.n,
                         s.flags.as<uint32_t>(), u_cur(), q.x.as<double2>(),
                         q.dup.as<double2>());
      fd_launch_force(c, false, t);
      if ((rc = residual(c, &rnorm))) return rc;
      r0 = rnorm;
    }
    const double target = std::max(o->tol_abs, o->tol_rel * r0);
    bool refreshed = o->predictor || (o->gnn_refresh == 2 && have_precond);
    while (rnorm > target && newton < o->max_newton) {
      if ((rc = tangent())) return rc;
      if (o->gnn_refresh == 1 || !refreshed) {
        if ((rc = precond_setup())) return rc;
        refreshed = true;
        have_precond = true;
      }
      if ((rc = linear(&lit))) return rc;
      hipLaunchKernelGGL(k_qs_update_u, dim3(g), dim3(256), 0, c->stream, s.n,
                         s.flags.as<uint32_t>(), u_cur(), q.x.as<double2>(),
                         nullptr);
      newton++;
      fd_launch_force(c, false, t);
      if ((rc = residual(c, &rnorm))) return rc;
    }
    s.time = t;
    s.step++;
    stats.load_steps++;
    stats.newton_iters += newton;
    stats.residual = rnorm;
    stats.residual0 = r0;
  }
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  stats.ms_total = ms_since(t_all);
  if (st) *st = stats;
  return NLM_OK;
}
