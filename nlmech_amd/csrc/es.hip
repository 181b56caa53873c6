// es.hip -- explicit dynamics of the linear state-based material
// (material::pd::ElasticState) on the GPU.
//
// A force evaluation is two kernels over the neighbor slices (one lane per
// node, one wavefront per 64-node slice, rows in list order):
//   k_es_dil    extensions, critical-stretch break test with the bond-bit
//               update, dilatation theta_i (StateBasedHelperFunctions::
//               dilatation, stateBasedHelperFunctions.cpp:26-83) and the
//               node's force coefficients;
//   k_es_force  the ElasticState bond force of both directions of every bond
//               (ElasticState.cpp:160-271) gathered at node i as
//               QuasiStaticModel::computeForces scatters it
//               (QuasiStaticModel.cpp:169-239), plus the force BCs, fused
//               with the central-difference update and the displacement BCs
//               at t+dt (FDModel::integrateCD, fDModel.cpp:377-443).
// The second kernel reads the coefficients of the neighbors, so the two are
// separate launches.  The weighted volume m_i (volumeCorrection.cpp:51-67) is
// computed once at setup.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "bond_es.h"
#include "fd_common.h"
#include "nlm_internal.h"

namespace nlm {
namespace {

template <int DIM>
struct EsArgs {
  FdArgs<DIM> f;
  const double* __restrict__ mw;
  double* __restrict__ theta;
  double* __restrict__ P;
  double* __restrict__ Q;
  EsDev m;
};

// One halo node staged in LDS by the tiled passes (3D): reference position,
// displacement, volume and the force coefficient (pass 2) in one 64-byte
// record.
struct alignas(16) EsRec {
  double x0, x1, x2;
  double u0, u1, u2;
  double vol, P;
};

// Stages the halo of layout block lb (every node its four slices reference)
// in LDS; WITH_P also copies the force coefficient of pass 1.
template <bool WITH_P>
__device__ __forceinline__ void es_stage_halo(const FdArgs<3>& a,
                                              const double* __restrict__ P,
                                              int64_t lb, EsRec* smem) {
  const int64_t hb = a.halo_ptr[lb];
  const int hn = int(a.halo_ptr[lb + 1] - hb);
  for (int t = threadIdx.x; t < hn; t += int(blockDim.x)) {
    const uint32_t g = uint32_t(a.halo[hb + t]);
    const double4 xx = ld32(a.x, g);
    const double4 uu = ld32(a.u, g);
    EsRec r;
    r.x0 = xx.x;
    r.x1 = xx.y;
    r.x2 = xx.z;
    r.u0 = uu.x;
    r.u1 = uu.y;
    r.u2 = uu.z;
    r.vol = ld32(a.vol, g);
    r.P = WITH_P ? ld32(P, g) : 0.;
    smem[t] = r;
  }
  __syncthreads();
}

__device__ __forceinline__ const EsRec& es_rec(const EsRec* smem, uint32_t j) {
  return *reinterpret_cast<const EsRec*>(reinterpret_cast<const char*>(smem) +
                                         j * uint32_t(sizeof(EsRec)));
}

// The fast passes stage the halo as two arrays: FP64 position and
// displacement (48 bytes a node) and the volume and force coefficient (in
// FP32 in the FP32-force mode, which rounds both anyway).  The 12-dword
// stride of the first spreads a wavefront's reads over four times as many
// LDS bank groups as a 16-dword record, at no more LDS.
struct alignas(16) EsXU {
  double x[3];
  double u[3];
};
template <class R>
struct EsVP {
  R vol;
  R P;
};
template <int MODE>
using EsVPOf =
    EsVP<typename std::conditional<MODE == kModeFp32, float, double>::type>;

template <bool WITH_P, class VP>
__device__ __forceinline__ void es_stage_halo_split(
    const FdArgs<3>& a, const double* __restrict__ P, int64_t lb, EsXU* xu,
    VP* vp) {
  const int64_t hb = a.halo_ptr[lb];
  const int hn = int(a.halo_ptr[lb + 1] - hb);
  for (int t = threadIdx.x; t < hn; t += int(blockDim.x)) {
    const uint32_t g = uint32_t(a.halo[hb + t]);
    const double4 xx = ld32(a.x, g);
    const double4 uu = ld32(a.u, g);
    EsXU r;
    r.x[0] = xx.x;
    r.x[1] = xx.y;
    r.x[2] = xx.z;
    r.u[0] = uu.x;
    r.u[1] = uu.y;
    r.u[2] = uu.z;
    xu[t] = r;
    VP q;
    q.vol = decltype(q.vol)(ld32(a.vol, g));
    q.P = WITH_P ? decltype(q.P)(ld32(P, g)) : decltype(q.P)(0);
    vp[t] = q;
  }
  __syncthreads();
}

// The FP32-force passes stage each halo node as its displacement in FP64
// (24 bytes) and its position as a fixed-point offset from the block's
// origin (three int32 in units of the quantum q, a power of two) with its
// FP32 volume (16 bytes); the force pass adds the force coefficient (4
// bytes).  The bond arithmetic runs in units of q: the position difference
// is an exact integer difference rounded once to FP32, the displacement
// difference (the displacements scaled by 1/q, exact) is formed in FP64 and
// rounded once -- both cancellation-free, a relative error of one FP32
// rounding -- and a block's halo takes 40 / 44 bytes a node instead of 56,
// so three blocks fit a CU's LDS instead of two.
struct EsU64 {
  double u[3];
};
struct alignas(16) EsXV32 {
  int32_t x[3];
  float vol;
};

// fixed-point offset of a coordinate from the origin, in units of q
__device__ __forceinline__ int32_t es_fix(double x, double o, double inv_q) {
  return int32_t(__double2int_rn((x - o) * inv_q));
}

template <bool WITH_P>
__device__ __forceinline__ void es_stage_halo_r32(
    const FdArgs<3>& a, const double* __restrict__ P, int64_t lb,
    const double4& org, const EsDev& m, EsU64* su, EsXV32* sx, float* sp) {
  const int64_t hb = a.halo_ptr[lb];
  const int hn = int(a.halo_ptr[lb + 1] - hb);
  for (int t = threadIdx.x; t < hn; t += int(blockDim.x)) {
    const uint32_t g = uint32_t(a.halo[hb + t]);
    const double4 xx = ld32(a.x, g);
    const double4 uu = ld32(a.u, g);
    EsU64 ru;
    ru.u[0] = uu.x * m.inv_q;
    ru.u[1] = uu.y * m.inv_q;
    ru.u[2] = uu.z * m.inv_q;
    su[t] = ru;
    EsXV32 rx;
    rx.x[0] = es_fix(xx.x, org.x, m.inv_q);
    rx.x[1] = es_fix(xx.y, org.y, m.inv_q);
    rx.x[2] = es_fix(xx.z, org.z, m.inv_q);
    rx.vol = float(ld32(a.vol, g));
    sx[t] = rx;
    // force coefficient in units of the bond arithmetic (times q)
    if (WITH_P) sp[t] = float(ld32(P, g) * m.q);
  }
  __syncthreads();
}

// m_i = sum_j |X| |X| c_ij V_j in list order (weightedVolume, :51-67), over
// the internal CSR whose rows keep the reference order.
template <int DIM>
__global__ __launch_bounds__(256) void k_es_mw(FdArgs<DIM> a,
                                               const int64_t* __restrict__ rp,
                                               const int32_t* __restrict__ col,
                                               EsDev m,
                                               double* __restrict__ mw) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  double xi[3];
  load3<DIM>(a.x[i], xi);
  double tmp = 0.;
  for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
    const int32_t j = col[k];
    double xj[3];
    load3<DIM>(a.x[j], xj);
    double X0 = xj[0] - xi[0], X1 = xj[1] - xi[1], X2 = xj[2] - xi[2];
    double lx = sqrt(X0 * X0 + X1 * X1 + X2 * X2);
    double c = es_corr_exact(m, lx);
    tmp += (lx * lx * c * a.vol[j]);
  }
  mw[i] = tmp;
}

// Pass 1: break test, dilatation, force coefficients.
// TILED (3D): the block's halo is staged in LDS and neighbors are read
// through 16-bit halo-local indices.
template <int DIM, int MODE, int U, bool TILED>
__global__ __launch_bounds__(TILED ? 512 : 256) void k_es_dil(
    EsArgs<DIM> e, const unsigned long long* __restrict__ bits_ro) {
  static_assert(!TILED || DIM == 3, "the LDS-staged passes are 3D");
  const FdArgs<DIM>& a = e.f;
  // staged passes: the workgroup's nodes (its size: 128, 256 or 512, the
  // tuned tile) are one block of the layout's halos
  const int bn = TILED ? int(blockDim.x) : 256;
  const int64_t nblk = (a.nslices * 64 + bn - 1) / bn;
  const int64_t lb = xcd_block(nblk);
  if (lb >= nblk) return;
  extern __shared__ EsRec es_smem[];
  // fast staged passes: split arrays (FP64 mode: FP64 x, u; volume.
  // FP32-force mode: FP64 u; FP32 block-relative x and volume)
  constexpr bool R32 = TILED && MODE == kModeFp32;
  constexpr bool SPLIT = TILED && MODE == kModeFp64;
  typedef EsVPOf<MODE> VP;
  EsXU* const es_xu = reinterpret_cast<EsXU*>(es_smem);
  VP* const es_vp = reinterpret_cast<VP*>(es_xu + a.halo_cap);
  EsU64* const es_u64 = reinterpret_cast<EsU64*>(es_smem);
  EsXV32* const es_xv = reinterpret_cast<EsXV32*>(es_u64 + a.halo_cap);
  double4 org = make_double4(0., 0., 0., 0.);
  if constexpr (R32) {
    org = reinterpret_cast<const double4*>(a.blk_org)[lb];
    es_stage_halo_r32<false>(a, e.P, lb, org, e.m, es_u64, es_xv, nullptr);
  } else if constexpr (SPLIT) {
    es_stage_halo_split<false>(a, e.P, lb, es_xu, es_vp);
  } else if constexpr (TILED) {
    es_stage_halo<false>(a, e.P, lb, es_smem);
  }
  const int lane = threadIdx.x & 63;
  const int64_t s =
      lb * (bn >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (s >= a.nslices) return;  // whole wavefront
  const int64_t i = s * 64 + lane;
  const bool valid = i < a.n;
  const int64_t cb = a.sl_ptr[s];
  const int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  const int64_t bb = cb >> 6;
  const int len = valid ? int(a.rowlen[i]) : 0;
  double xi[3] = {0., 0., 0.}, ui[3] = {0., 0., 0.};
  double mwi = 1.;
  if (valid) {
    load3<DIM>(a.x[i], xi);
    load3<DIM>(a.u[i], ui);
    mwi = e.mw[i];
  }
  // the node's fixed-point position and scaled displacement (FP32-force
  // staged pass)
  int32_t xqi[3] = {0, 0, 0};
  double uqi[3] = {0., 0., 0.};
  if constexpr (R32) {
    xqi[0] = es_fix(xi[0], org.x, e.m.inv_q);
    xqi[1] = es_fix(xi[1], org.y, e.m.inv_q);
    xqi[2] = es_fix(xi[2], org.z, e.m.inv_q);
    for (int d = 0; d < 3; d++) uqi[d] = ui[d] * e.m.inv_q;
  }
  const int64_t hb = TILED ? a.halo_ptr[lb] : 0;
  // (3/m_i) in 3D, (2/m_i) factor2D in 2D, leading factors of the reference
  // product (stateBasedHelperFunctions.cpp:58-71)
  const double c3 = DIM == 3 ? 3. / mwi : (2. / mwi) * e.m.f2d;
  const float c3f = float(c3);
  const int32_t* __restrict__ colS = a.col + cb;
  const uint16_t* __restrict__ colL = a.colL + cb;
  const unsigned long long* __restrict__ bits_in = bits_ro + bb;
  const int wlast = width > 0 ? width - 1 : 0;
  auto slot_index = [&](int k) -> uint32_t {
    const uint32_t en = uint32_t(64 * (k < wlast ? k : wlast) + lane);
    if constexpr (TILED) return uint32_t(ld32(colL, en));
    else return uint32_t(ld32(colS, en));
  };
  double th = 0.;
  // fast modes: the slot-pair sums accumulate without the node's factor c3
  // (times q^2 in units of q), which multiplies their total once at the end
  double thf = 0.;
  const double sc3 = R32 ? c3 * e.m.q * e.m.q : c3;
  uint32_t jn[U];
#pragma unroll
  for (int q = 0; q < U; ++q) jn[q] = slot_index(q);
  for (int k0 = 0; k0 < width; k0 += U) {
    uint32_t jc[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      jc[q] = jn[q];
      jn[q] = slot_index(k0 + U + q);
    }
    double xj[U][3], uj[U][3], vj[U];
    float dxf[U][3], duf[U][3];
    unsigned long long w[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      if constexpr (R32) {
        const EsU64& ru = es_u64[jc[q]];
        const EsXV32 rx = es_xv[jc[q]];
        for (int d = 0; d < 3; d++) {
          dxf[q][d] = float(rx.x[d] - xqi[d]);
          duf[q][d] = float(ru.u[d] - uqi[d]);
        }
        vj[q] = double(rx.vol);
      } else if constexpr (SPLIT) {
        const EsXU& r = es_xu[jc[q]];
        for (int d = 0; d < 3; d++) {
          xj[q][d] = r.x[d];
          uj[q][d] = r.u[d];
        }
        vj[q] = double(es_vp[jc[q]].vol);
      } else if constexpr (TILED) {
        const EsRec& r = es_rec(es_smem, jc[q]);
        xj[q][0] = r.x0;
        xj[q][1] = r.x1;
        xj[q][2] = r.x2;
        uj[q][0] = r.u0;
        uj[q][1] = r.u1;
        uj[q][2] = r.u2;
        vj[q] = r.vol;
      } else {
        load3<DIM>(ld32(a.x, jc[q]), xj[q]);
        load3<DIM>(ld32(a.u, jc[q]), uj[q]);
        vj[q] = ld32(a.vol, jc[q]);
      }
      const int kq = k0 + q;
      w[q] = bits_in[kq < wlast ? kq : wlast];
    }
    if constexpr (MODE != kModeExact) {
      // Fast modes: branch-free common path; decisions are votes on single
      // comparisons (their lane masks) combined by scalar mask arithmetic;
      // bonds inside a guard band (break test, volume-correction edge) are
      // evaluated exactly after the chunk; dilatation terms summed over slot
      // pairs (2m, 2m+1) -- in FP32 in the FP32-force mode -- each pair
      // added to the FP64 sum.
      typedef typename std::conditional<MODE == kModeFp32, float, double>::type
          Real;
      constexpr bool F32 = MODE == kModeFp32;
      auto fmaR = [](Real x, Real y, Real z) -> Real {
        if constexpr (F32) return __builtin_fmaf(x, y, z);
        else return __builtin_fma(x, y, z);
      };
      Real pf[(U + 1) / 2];
      unsigned long long nb[U], band[U], ex[U];
      unsigned long long ex_any = 0ull;
      // (the staged FP32-force pass works in units of q: lengths of the
      // edge test and the dilatation factor scaled accordingly)
      const Real vtol = R32   ? Real(e.m.vex_tol_qf)
                        : F32 ? Real(e.m.vex_tol_f)
                              : Real(e.m.vex_tol);
      const Real lowb = R32   ? Real(e.m.low_band_qf)
                        : F32 ? Real(e.m.low_band_f)
                              : Real(e.m.low_band);
      const Real blo = F32 ? Real(e.m.brk_lo_f) : Real(e.m.brk_lo);
      const Real cup = R32   ? Real(e.m.check_up_qf)
                       : F32 ? Real(e.m.check_up_f)
                             : Real(e.m.check_up);
      const Real rch = R32   ? Real(e.m.rcp_h_qf)
                       : F32 ? Real(e.m.rcp_h_f)
                             : Real(e.m.rcp_h);
      const Real uprch = Real(e.m.up_rcp_h_qf);
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const unsigned long long act = __builtin_amdgcn_ballot_w64(k0 + q < len);
        Real dx[3], du[3];
        for (int d = 0; d < 3; d++) {
          if constexpr (R32) {
            dx[d] = dxf[q][d];
            du[d] = duf[q][d];
          } else {
            dx[d] = d < DIM ? Real(xj[q][d] - xi[d]) : Real(0);
            du[d] = d < DIM ? Real(uj[q][d] - ui[d]) : Real(0);
          }
        }
        Real dd = fmaR(dx[1], dx[1], dx[0] * dx[0]);
        Real g = fmaR(du[1], fmaR(Real(2), dx[1], du[1]),
                      du[0] * fmaR(Real(2), dx[0], du[0]));
        if (DIM == 3) {
          dd = fmaR(dx[2], dx[2], dd);
          g = fmaR(du[2], fmaR(Real(2), dx[2], du[2]), g);
        }
        Real rx, ly, ext;
        const Real yy = dd + g;
        if constexpr (F32) {
          rx = __builtin_amdgcn_rsqf(dd);
          ly = yy * __builtin_amdgcn_rsqf(yy);
        } else {
          rx = rsqrt_f64(dd);
          ly = yy * rsqrt_f64(yy);
        }
        const Real lx = dd * rx;
        if constexpr (F32) ext = g * __builtin_amdgcn_rcpf(ly + lx);
        else ext = g / (ly + lx);
        const Real st = ext * rx;
        const unsigned long long v_vex =
            __builtin_amdgcn_ballot_w64(fabs(lx - lowb) <= vtol);
        // one vote on the guard band's lower bound: a bond that may break
        // (inside the band, or beyond it) is decided, and its term taken, by
        // the exact evaluation after the chunk, which also reports the new
        // broken bits
        const unsigned long long v_lo = __builtin_amdgcn_ballot_w64(st >= blo);
        nb[q] = 0ull;
        band[q] = act & ~w[q] & v_lo;
        ex[q] = (v_vex & act) | band[q];
        const unsigned long long keep = act & ~w[q] & ~ex[q];
        // volume-correction factor (check_up - |X|)/h clamped to [0, 1]
        // (in FP32 one multiply-add with the clamp as its output modifier)
        Real c;
        if constexpr (F32)
          c = __builtin_amdgcn_fmed3f(__builtin_fmaf(lx, -rch, uprch), 0.f,
                                      1.f);
        else
          c = fmin((cup - lx) * rch, Real(1));
        // (the odd slot's term fused into its pair: fma(l e c, V_j, T_2m))
        Real lec = lx * ext * c;
        lec = __builtin_amdgcn_inverse_ballot_w64(keep) ? lec : Real(0);
        if (q % 2 == 0) pf[q / 2] = lec * Real(vj[q]);
        else pf[q / 2] = fmaR(lec, Real(vj[q]), pf[q / 2]);
        ex_any |= ex[q];
      }
      if (ex_any != 0ull) {  // rare: whole wavefronts skip it
#pragma unroll
        for (int q = 0; q < U; ++q) {
          if (ex[q] == 0ull) continue;
          bool fsn = false;
          if (__builtin_amdgcn_inverse_ballot_w64(ex[q])) {
            const bool fs0 = __builtin_amdgcn_inverse_ballot_w64(w[q]);
            bool fs = fs0;
            double xs[3], us[3], vs;
            if constexpr (R32) {
              // FP64 values of the neighbor from its halo entry
              const uint32_t g = uint32_t(a.halo[hb + jc[q]]);
              load3<DIM>(ld32(a.x, g), xs);
              load3<DIM>(ld32(a.u, g), us);
              vs = ld32(a.vol, g);
            } else {
              for (int d = 0; d < 3; d++) {
                xs[d] = xj[q][d];
                us[d] = uj[q][d];
              }
              vs = vj[q];
            }
            const double d = es_dil_exact(e.m, xi, ui, xs, us, vs, c3, fs);
            fsn = fs && !fs0;
            if (!fs) pf[q / 2] += Real(d / sc3);
          }
          // the guard band's bonds were decided exactly here
          nb[q] |= band[q] & __builtin_amdgcn_ballot_w64(fsn);
        }
      }
#pragma unroll
      for (int mm = 0; mm < (U + 1) / 2; ++mm) thf += double(pf[mm]);
      unsigned long long nb_any = 0ull;
#pragma unroll
      for (int q = 0; q < U; ++q) nb_any |= nb[q];
      if (nb_any != 0ull) {
#pragma unroll
        for (int q = 0; q < U; ++q)
          if (nb[q] != 0ull && k0 + q < width)
            a.bits[bb + k0 + q] = w[q] | nb[q];
      }
      continue;
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const bool act = k0 + q < len;
      const bool fs0 = __builtin_amdgcn_inverse_ballot_w64(w[q]);
      bool fs = fs0;
      double d;
      if constexpr (MODE == kModeExact) {
        d = es_dil_exact(e.m, xi, ui, xj[q], uj[q], vj[q], c3, fs);
      } else {
        bool exn;
        if constexpr (MODE == kModeFp64)
          d = es_dil_fp64<DIM>(e.m, xi, ui, xj[q], uj[q], vj[q], c3, fs0, fs,
                               exn);
        else
          d = es_dil_fp32<DIM>(e.m, xi, ui, xj[q], uj[q], vj[q], c3f, fs0,
                               fs, exn);
        exn = exn && act;
        if (__ballot(exn) != 0ull) {
          if (exn) {
            fs = fs0;
            d = es_dil_exact(e.m, xi, ui, xj[q], uj[q], vj[q], c3, fs);
          }
        }
      }
      if (act && !fs) th += d;
      const bool newly = act && fs && !fs0;
      if (k0 + q < width) {
        const unsigned long long nb = __ballot(newly);
        if (nb != 0ull && lane == 0) a.bits[bb + k0 + q] = w[q] | nb;
      }
    }
  }
  th += sc3 * thf;
  if (!valid) return;
  e.theta[i] = th;
  if constexpr (DIM == 3) {
    // EXACT keeps alpha_s = (9/m) K (ElasticState.cpp:230-233); the fast
    // modes fold t_ij = alpha_s theta_i |X| / 3 into P_i |X|.
    e.P[i] = MODE == kModeExact ? (9. / mwi) * e.m.K : 3. * e.m.K * th / mwi;
  } else if (e.m.one_d) {
    // 1D: t_ij = (E/m_i) e_ij (ElasticState.cpp:187-193)
    e.P[i] = 0.;
    e.Q[i] = e.m.E / mwi;
  } else {
    const double as = (9. / mwi) * e.m.Kc;
    const double ad = (8. / mwi) * e.m.mu;
    const double C = 2. * e.m.f2d * as - (3. - 2. * e.m.f2d) * ad;
    // fast modes: t_ij = theta_i (C/9 - alpha_d/3) |X| + alpha_d e_ij
    e.P[i] = MODE == kModeExact ? C : th * (C / 9. - ad / 3.);
    e.Q[i] = ad;
  }
}

// Pass 2: bond forces of both directions gathered at node i, force BCs,
// central-difference update.
template <int DIM, int MODE, int KIND, bool BASE_F, int U, bool TILED>
__global__ __launch_bounds__(TILED ? 512 : 256) void k_es_force(
    EsArgs<DIM> e, const unsigned long long* __restrict__ bits_ro,
    double t_force, double t_next) {
  if (e.f.tdev) {
    t_force = e.f.tdev[0];
    t_next = e.f.tdev[1];
  }
  static_assert(!TILED || (DIM == 3 && MODE != kModeExact),
                "the LDS-staged force pass is 3D, fast modes");
  const FdArgs<DIM>& a = e.f;
  const int bn = TILED ? int(blockDim.x) : 256;
  const int64_t nblk = (a.nslices * 64 + bn - 1) / bn;
  const int64_t lb = xcd_block(nblk);
  if (lb >= nblk) return;
  extern __shared__ EsRec es_smem[];
  // fast staged passes: split arrays (FP64 mode: FP64 x, u; volume and
  // force coefficient.  FP32-force mode: FP64 u; FP32 block-relative x and
  // volume; FP32 force coefficient)
  constexpr bool R32 = TILED && MODE == kModeFp32;
  constexpr bool SPLIT = TILED && MODE == kModeFp64;
  typedef EsVPOf<MODE> VP;
  EsXU* const es_xu = reinterpret_cast<EsXU*>(es_smem);
  VP* const es_vp = reinterpret_cast<VP*>(es_xu + a.halo_cap);
  EsU64* const es_u64 = reinterpret_cast<EsU64*>(es_smem);
  EsXV32* const es_xv = reinterpret_cast<EsXV32*>(es_u64 + a.halo_cap);
  float* const es_p32 = reinterpret_cast<float*>(es_xv + a.halo_cap);
  double4 org = make_double4(0., 0., 0., 0.);
  if constexpr (R32) {
    org = reinterpret_cast<const double4*>(a.blk_org)[lb];
    es_stage_halo_r32<true>(a, e.P, lb, org, e.m, es_u64, es_xv, es_p32);
  } else if constexpr (SPLIT) {
    es_stage_halo_split<true>(a, e.P, lb, es_xu, es_vp);
  } else if constexpr (TILED) {
    es_stage_halo<true>(a, e.P, lb, es_smem);
  }
  const int lane = threadIdx.x & 63;
  const int64_t s =
      lb * (bn >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (s >= a.nslices) return;
  const int64_t i = s * 64 + lane;
  const bool valid = i < a.n;
  const int64_t cb = a.sl_ptr[s];
  const int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  const int64_t bb = cb >> 6;
  const int len = valid ? int(a.rowlen[i]) : 0;
  constexpr bool kExact = MODE == kModeExact;
  constexpr bool kNeedQ = DIM == 2;
  double xi[3] = {0., 0., 0.}, ui[3] = {0., 0., 0.};
  double Pi = 0., Qi = 0., thi = 0.;
  uint32_t fl = 0;
  if (valid) {
    load3<DIM>(a.x[i], xi);
    load3<DIM>(a.u[i], ui);
    Pi = e.P[i];
    if (kNeedQ) Qi = e.Q[i];
    // (1D: no dilatation term, t = Q e)
    if (kExact) thi = e.m.one_d ? 0. : e.theta[i];
    fl = a.flags[i];
  }
  // FP32-force staged pass: force coefficient times q (the bond arithmetic
  // runs in units of q), fixed-point position, scaled displacement
  const float Pif = float(R32 ? Pi * e.m.q : Pi), Qif = float(Qi);
  int32_t xqi[3] = {0, 0, 0};
  double uqi[3] = {0., 0., 0.};
  if constexpr (R32) {
    xqi[0] = es_fix(xi[0], org.x, e.m.inv_q);
    xqi[1] = es_fix(xi[1], org.y, e.m.inv_q);
    xqi[2] = es_fix(xi[2], org.z, e.m.inv_q);
    for (int d = 0; d < 3; d++) uqi[d] = ui[d] * e.m.inv_q;
  }
  const int64_t hb = TILED ? a.halo_ptr[lb] : 0;
  const int32_t* __restrict__ colS = a.col + cb;
  const uint16_t* __restrict__ colL = a.colL + cb;
  const unsigned long long* __restrict__ bits_in = bits_ro + bb;
  const int wlast = width > 0 ? width - 1 : 0;
  auto slot_index = [&](int k) -> uint32_t {
    const uint32_t en = uint32_t(64 * (k < wlast ? k : wlast) + lane);
    if constexpr (TILED) return uint32_t(ld32(colL, en));
    else return uint32_t(ld32(colS, en));
  };
  double acc[3] = {0., 0., 0.}, oth[3] = {0., 0., 0.};
  uint32_t jn[U];
#pragma unroll
  for (int q = 0; q < U; ++q) jn[q] = slot_index(q);
  for (int k0 = 0; k0 < width; k0 += U) {
    uint32_t jc[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      jc[q] = jn[q];
      jn[q] = slot_index(k0 + U + q);
    }
    double xj[U][3], uj[U][3], vj[U], Pj[U], Qj[U], thj[U];
    float dxf[U][3], duf[U][3];
    unsigned long long w[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t j = jc[q];
      if constexpr (R32) {
        const EsU64& ru = es_u64[j];
        const EsXV32 rx = es_xv[j];
        for (int d = 0; d < 3; d++) {
          dxf[q][d] = float(rx.x[d] - xqi[d]);
          duf[q][d] = float(ru.u[d] - uqi[d]);
        }
        vj[q] = double(rx.vol);
        Pj[q] = double(es_p32[j]);
        Qj[q] = 0.;
        thj[q] = 0.;
      } else if constexpr (SPLIT) {
        const EsXU& r = es_xu[j];
        for (int d = 0; d < 3; d++) {
          xj[q][d] = r.x[d];
          uj[q][d] = r.u[d];
        }
        const VP vpj = es_vp[j];
        vj[q] = double(vpj.vol);
        Pj[q] = double(vpj.P);
        Qj[q] = 0.;
        thj[q] = 0.;
      } else if constexpr (TILED) {
        const EsRec& r = es_rec(es_smem, j);
        xj[q][0] = r.x0;
        xj[q][1] = r.x1;
        xj[q][2] = r.x2;
        uj[q][0] = r.u0;
        uj[q][1] = r.u1;
        uj[q][2] = r.u2;
        vj[q] = r.vol;
        Pj[q] = r.P;
        Qj[q] = 0.;
        thj[q] = 0.;
      } else {
        load3<DIM>(ld32(a.x, j), xj[q]);
        load3<DIM>(ld32(a.u, j), uj[q]);
        vj[q] = ld32(a.vol, j);
        Pj[q] = ld32(e.P, j);
        Qj[q] = kNeedQ ? ld32(e.Q, j) : 0.;
        thj[q] = kExact && !e.m.one_d ? ld32(e.theta, j) : 0.;
      }
      const int kq = k0 + q;
      w[q] = bits_in[kq < wlast ? kq : wlast];
    }
    float pf[(U + 1) / 2][3];  // FP32-force: slot-pair sums
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const bool act = k0 + q < len;
      const bool fs0 = __builtin_amdgcn_inverse_ballot_w64(w[q]);
      const bool keep = act && !fs0;
      if constexpr (kExact) {
        double own[3], ot[3];
        es_force_exact<DIM>(e.m, xi, ui, xj[q], uj[q], vj[q], Pi, Qi, thi,
                            Pj[q], Qj[q], thj[q], own, ot);
        for (int d = 0; d < DIM; d++) {
          if (keep) {
            acc[d] += own[d];
            oth[d] += ot[d];
          }
        }
      } else if constexpr (MODE == kModeFp64) {
        double dx[3], du[3], Y[3];
        for (int d = 0; d < 3; d++) {
          dx[d] = d < DIM ? xj[q][d] - xi[d] : 0.;
          du[d] = d < DIM ? uj[q][d] - ui[d] : 0.;
          Y[d] = dx[d] + du[d];
        }
        double yy = __builtin_fma(Y[1], Y[1], Y[0] * Y[0]);
        double dd = __builtin_fma(dx[1], dx[1], dx[0] * dx[0]);
        if (DIM == 3) {
          yy = __builtin_fma(Y[2], Y[2], yy);
          dd = __builtin_fma(dx[2], dx[2], dd);
        }
        const double ry = rsqrt_f64(yy);
        const double rx = rsqrt_f64(dd);
        const double lx = dd * rx;
        const bool vex = keep && fabs(lx - e.m.low_band) <=
                                     kEsGuard64 * e.m.low_band;
        double c = lx > e.m.low_band ? (e.m.check_up - lx) * e.m.rcp_h : 1.;
        if (__ballot(vex) != 0ull) {
          if (vex) c = es_corr_exact(e.m, sqrt(dx[0] * dx[0] + dx[1] * dx[1] +
                                               dx[2] * dx[2]));
        }
        double t = (Pi + Pj[q]) * lx;
        if (DIM == 2) {
          double g = __builtin_fma(du[1], __builtin_fma(2., dx[1], du[1]),
                                   du[0] * __builtin_fma(2., dx[0], du[0]));
          double ext = g / (yy * ry + lx);
          t = __builtin_fma(Qi + Qj[q], ext, t);
        }
        double sr = t * c * vj[q] * ry;
        sr = keep ? sr : 0.;
        for (int d = 0; d < DIM; d++) acc[d] = __builtin_fma(sr, Y[d], acc[d]);
      } else {
        // FP32-force: the volume-correction edge test is a vote on a single
        // comparison (its lane mask), the forces are summed in FP32 over
        // slot pairs (2m, 2m+1) and each pair is added to the FP64 sum
        const unsigned long long live =
            __builtin_amdgcn_ballot_w64(k0 + q < len) & ~w[q];
        float dx[3], du[3], Y[3];
        for (int d = 0; d < 3; d++) {
          if constexpr (R32) {
            dx[d] = dxf[q][d];
            du[d] = duf[q][d];
          } else {
            dx[d] = d < DIM ? float(xj[q][d] - xi[d]) : 0.f;
            du[d] = d < DIM ? float(uj[q][d] - ui[d]) : 0.f;
          }
          Y[d] = dx[d] + du[d];
        }
        float yy = __builtin_fmaf(Y[1], Y[1], Y[0] * Y[0]);
        float dd = __builtin_fmaf(dx[1], dx[1], dx[0] * dx[0]);
        if (DIM == 3) {
          yy = __builtin_fmaf(Y[2], Y[2], yy);
          dd = __builtin_fmaf(dx[2], dx[2], dd);
        }
        const float ry = __builtin_amdgcn_rsqf(yy);
        const float rx = __builtin_amdgcn_rsqf(dd);
        const float lx = dd * rx;
        const float lowb = R32 ? e.m.low_band_qf : e.m.low_band_f;
        const float vtol = R32 ? e.m.vex_tol_qf : 3e-5f * e.m.low_band_f;
        const float rch = R32 ? e.m.rcp_h_qf : e.m.rcp_h_f;
        const unsigned long long vexm =
            __builtin_amdgcn_ballot_w64(fabsf(lx - lowb) <= vtol) & live;
        // (check_up - |X|)/h clamped to [0, 1]: one multiply-add with the
        // clamp as its output modifier
        float c = __builtin_amdgcn_fmed3f(
            __builtin_fmaf(lx, -rch, e.m.up_rcp_h_qf), 0.f, 1.f);
        if (vexm != 0ull) {  // rare: lengths at the correction's edge
          if (__builtin_amdgcn_inverse_ballot_w64(vexm)) {
            double xs[3];
            if constexpr (R32) {
              // FP64 position of the neighbor from its halo entry
              load3<DIM>(ld32(a.x, uint32_t(a.halo[hb + jc[q]])), xs);
            } else {
              for (int d = 0; d < 3; d++) xs[d] = xj[q][d];
            }
            double X0 = xs[0] - xi[0], X1 = xs[1] - xi[1];
            double X2 = DIM == 3 ? xs[2] - xi[2] : 0.;
            c = float(es_corr_exact(e.m, sqrt(X0 * X0 + X1 * X1 + X2 * X2)));
          }
        }
        float t = (Pif + float(Pj[q])) * lx;
        if (DIM == 2) {
          float g = __builtin_fmaf(du[1], __builtin_fmaf(2.f, dx[1], du[1]),
                                   du[0] * __builtin_fmaf(2.f, dx[0], du[0]));
          float ext = g * __builtin_amdgcn_rcpf(yy * ry + lx);
          t = __builtin_fmaf(Qif + float(Qj[q]), ext, t);
        }
        float sr = t * c * float(vj[q]) * ry;
        sr = __builtin_amdgcn_inverse_ballot_w64(live) ? sr : 0.f;
        // the odd slot's force fused into its pair: fma(s, Y, F_2m)
        for (int d = 0; d < DIM; d++) {
          if (q % 2 == 0) pf[q / 2][d] = sr * Y[d];
          else pf[q / 2][d] = __builtin_fmaf(sr, Y[d], pf[q / 2][d]);
        }
      }
    }
    if constexpr (MODE == kModeFp32) {
#pragma unroll
      for (int mm = 0; mm < (U + 1) / 2; ++mm)
        for (int d = 0; d < DIM; d++) acc[d] += double(pf[mm][d]);
    }
  }
  if (!valid) return;
  double f[3] = {0., 0., 0.};
  if (BASE_F) {
    load3<DIM>(a.f_in[i], f);
  } else if (a.has_fbc) {
    fext_node<DIM>(a.bc, fl, xi, t_force, f);
  }
  for (int d = 0; d < DIM; d++) f[d] += kExact ? acc[d] - oth[d] : acc[d];
  node_tail<DIM, KIND>(a, i, fl, xi, ui, f, t_next);
}

// ---- host side ------------------------------------------------------------

template <int DIM>
EsArgs<DIM> make_es_args(nlm_ctx* c) {
  FdState& s = c->fd;
  EsArgs<DIM> e;
  e.f = make_args<DIM>(c);
  e.mw = s.es_mw.as<double>();
  e.theta = s.es_theta.as<double>();
  e.P = s.es_P.as<double>();
  e.Q = s.es_Q.as<double>();
  e.m = s.es;
  return e;
}

// LDS bytes of a tiled launch: one record per halo entry (the fast passes:
// the two split arrays; the exact pass 1 keeps one record).
size_t es_record_bytes(int precision) {
  const size_t fast = precision == kModeFp32
                          ? sizeof(EsU64) + sizeof(EsXV32) + sizeof(float)
                          : sizeof(EsXU) + sizeof(EsVPOf<kModeFp64>);
  return precision == kModeExact ? sizeof(EsRec) : fast;
}
size_t es_lds_bytes(const FdState& s) {
  return size_t(s.halo_max + 1) * es_record_bytes(s.precision);
}

// Raises a kernel's dynamic-LDS limit to what the launch needs.  The grant
// is kept per kernel (keyed by its address: every template instance has its
// own attribute) and per device.
void allow_lds(const void* kernel, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> granted;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  size_t& g = granted[{dev, kernel}];
  if (bytes > g) {
    (void)hipFuncSetAttribute(kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(bytes));
    g = bytes;
  }
}

template <int DIM, int MODE>
void launch_dil(nlm_ctx* c, const EsArgs<DIM>& e) {
  unsigned g = grid_xcd(force_blocks(c->fd) * 256, 256);
  if constexpr (DIM == 3) {
    if (c->fd.es_tiled && c->fd.use_tiled) {
      // one workgroup per tile of the layout's halos
      const int bn = c->fd.tile;
      const unsigned gt = grid_xcd(c->fd.sell.nslices * 64, bn);
      const size_t lds = es_lds_bytes(c->fd);
      allow_lds(reinterpret_cast<const void*>(k_es_dil<DIM, MODE, 4, true>),
                lds);
      hipLaunchKernelGGL((k_es_dil<DIM, MODE, 4, true>), dim3(gt), dim3(bn),
                         lds, c->stream, e, e.f.bits);
      return;
    }
  }
  hipLaunchKernelGGL((k_es_dil<DIM, MODE, 4, false>), dim3(g), dim3(256), 0,
                     c->stream, e, e.f.bits);
}

template <int DIM, int MODE, int KIND, bool BASE_F, int U>
void launch_force_k(nlm_ctx* c, const EsArgs<DIM>& e, double tf, double tn) {
  unsigned g = grid_xcd(force_blocks(c->fd) * 256, 256);
  if constexpr (DIM == 3 && MODE != kModeExact) {
    if (c->fd.es_tiled && c->fd.use_tiled) {
      const int bn = c->fd.tile;
      const unsigned gt = grid_xcd(c->fd.sell.nslices * 64, bn);
      const size_t lds = es_lds_bytes(c->fd);
      allow_lds(reinterpret_cast<const void*>(
                    k_es_force<DIM, MODE, KIND, BASE_F, U, true>),
                lds);
      hipLaunchKernelGGL((k_es_force<DIM, MODE, KIND, BASE_F, U, true>),
                         dim3(gt), dim3(bn), lds, c->stream, e, e.f.bits, tf,
                         tn);
      return;
    }
  }
  hipLaunchKernelGGL((k_es_force<DIM, MODE, KIND, BASE_F, U, false>), dim3(g),
                     dim3(256), 0, c->stream, e, e.f.bits, tf, tn);
}

template <int DIM, int MODE>
void launch_pair(nlm_ctx* c, int kind, bool base_f, double tf, double tn,
                 int U, const double* tdev = nullptr) {
  EsArgs<DIM> e = make_es_args<DIM>(c);
  e.f.tdev = tdev;
  launch_dil<DIM, MODE>(c, e);
  if (kind == KIND_FORCE) {
    if (base_f)
      launch_force_k<DIM, MODE, KIND_FORCE, true, 4>(c, e, tf, tn);
    else
      launch_force_k<DIM, MODE, KIND_FORCE, false, 4>(c, e, tf, tn);
    return;
  }
  if constexpr (MODE != kModeExact) {
    if (U == 2) return launch_force_k<DIM, MODE, KIND_CD, false, 2>(c, e, tf, tn);
    if (U == 8) return launch_force_k<DIM, MODE, KIND_CD, false, 8>(c, e, tf, tn);
  }
  launch_force_k<DIM, MODE, KIND_CD, false, 4>(c, e, tf, tn);
}

template <int DIM>
void launch_dim(nlm_ctx* c, int kind, bool base_f, double tf, double tn,
                int U, const double* tdev) {
  switch (c->fd.precision) {
    case kModeExact:
      launch_pair<DIM, kModeExact>(c, kind, base_f, tf, tn, U, tdev);
      break;
    case kModeFp64:
      launch_pair<DIM, kModeFp64>(c, kind, base_f, tf, tn, U, tdev);
      break;
    default:
      launch_pair<DIM, kModeFp32>(c, kind, base_f, tf, tn, U, tdev);
      break;
  }
}

void launch_any(nlm_ctx* c, int kind, bool base_f, double tf, double tn,
                int U, const double* tdev = nullptr) {
  if (c->fd.dim == 2)
    launch_dim<2>(c, kind, base_f, tf, tn, U, tdev);
  else
    launch_dim<3>(c, kind, base_f, tf, tn, U, tdev);
}

}  // namespace

// Device arrays and the weighted volume; called by nlm_fd_setup after the
// node arrays and slices are in place.
int es_prepare(nlm_ctx* c) {
  FdState& s = c->fd;
  int rc;
  const size_t nb = sizeof(double) * size_t(s.n) + 256;
  if ((rc = alloc(c, s.es_mw, nb))) return rc;
  if ((rc = alloc(c, s.es_theta, nb))) return rc;
  if ((rc = alloc(c, s.es_P, nb))) return rc;
  if ((rc = alloc(c, s.es_Q, nb))) return rc;
  NLM_HIP(c, hipMemsetAsync(s.es_theta.p, 0, nb, c->stream));
  NLM_HIP(c, hipMemsetAsync(s.es_P.p, 0, nb, c->stream));
  NLM_HIP(c, hipMemsetAsync(s.es_Q.p, 0, nb, c->stream));
  // LDS-staged passes: 3D, when one block's halo fits the device's LDS
  int lds_max = 0;
  NLM_HIP(c, hipDeviceGetAttribute(&lds_max,
                                   hipDeviceAttributeMaxSharedMemoryPerBlock,
                                   c->device));
  s.es_tiled = s.dim == 3 && s.halo_max > 0 &&
               es_lds_bytes(s) <= size_t(lds_max);
  const unsigned g = unsigned((s.n + 255) / 256);
  if (s.dim == 2)
    hipLaunchKernelGGL(k_es_mw<2>, dim3(g), dim3(256), 0, c->stream,
                       make_args<2>(c), s.row_ptr.as<int64_t>(),
                       s.col.as<int32_t>(), s.es, s.es_mw.as<double>());
  else
    hipLaunchKernelGGL(k_es_mw<3>, dim3(g), dim3(256), 0, c->stream,
                       make_args<3>(c), s.row_ptr.as<int64_t>(),
                       s.col.as<int32_t>(), s.es, s.es_mw.as<double>());
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

size_t es_halo_lds_bytes(int precision, int64_t halo_max) {
  return size_t(halo_max + 1) * es_record_bytes(precision);
}

void es_launch_force(nlm_ctx* c, bool base_f, double t) {
  launch_any(c, KIND_FORCE, base_f, t, t, 4);
}

void es_launch_fused(nlm_ctx* c, double tf, double tn) {
  launch_any(c, KIND_CD, false, tf, tn, c->fd.unroll);
}

void es_launch_fused_dev(nlm_ctx* c, const double* tdev) {
  launch_any(c, KIND_CD, false, 0., 0., c->fd.unroll, tdev);
}

// Times the fused pass pair for each chunk width of the force pass on the
// current state; the trials write the spare u/v buffers, the node
// coefficient arrays (rewritten by the next pass 1) and a scratch copy of the
// bond bits, so the model state is unchanged.
int es_autotune(nlm_ctx* c, int* chosen, double* best_us, double* us3) {
  FdState& s = c->fd;
  DevBuf bits_copy;
  int rc;
  if ((rc = alloc(c, bits_copy, s.sell.bits.bytes))) return rc;
  NLM_HIP(c, hipMemcpyAsync(bits_copy.p, s.sell.bits.p, s.sell.bits.bytes,
                            hipMemcpyDeviceToDevice, c->stream));
  DevBuf th_save;
  if ((rc = alloc(c, th_save, s.es_theta.bytes))) return rc;
  NLM_HIP(c, hipMemcpyAsync(th_save.p, s.es_theta.p, s.es_theta.bytes,
                            hipMemcpyDeviceToDevice, c->stream));
  // the trials run on the bits_copy bits
  std::swap(s.sell.bits.p, bits_copy.p);
  const int cands[3] = {2, 4, 8};
  double best = 1e300;
  int pick = 4;
  hipEvent_t e0, e1;
  NLM_HIP(c, hipEventCreate(&e0));
  NLM_HIP(c, hipEventCreate(&e1));
  const int reps = 4;
  for (int k = 0; k < 3; k++) {
    const int U = cands[k];
    launch_any(c, KIND_CD, false, s.time, s.time + s.dt, U);
    NLM_HIP(c, hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++)
      launch_any(c, KIND_CD, false, s.time, s.time + s.dt, U);
    NLM_HIP(c, hipEventRecord(e1, c->stream));
    NLM_HIP(c, hipEventSynchronize(e1));
    float ms = 0.f;
    NLM_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    const double us = 1e3 * double(ms) / reps;
    if (us3) us3[k] = us;
    if (us < best) {
      best = us;
      pick = U;
    }
  }
  std::swap(s.sell.bits.p, bits_copy.p);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // restore pass 1's outputs of the current state
  NLM_HIP(c, hipMemcpyAsync(s.es_theta.p, th_save.p, s.es_theta.bytes,
                            hipMemcpyDeviceToDevice, c->stream));
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  s.unroll = pick;
  if (chosen) *chosen = pick;
  if (best_us) *best_us = best;
  return NLM_OK;
}

// ---- output fields of the quasi-static model ------------------------------

namespace {

// Output fields of QuasiStaticModel::computeForces(true)
// (QuasiStaticModel.cpp:206-229) at the current state, one thread per node,
// rows in list order: the strain energy of every unbroken bond
// (ElasticState::getBondEF's second value, ElasticState.cpp:250-262) added
// in FP32 to energy[i], the strain tensor (getStrain, :336-345, with
// deformation_gradient :304-334 and K_shape_tensor :287-302, whose
// volume-correction index stays at the row's first bond) and the stress
// tensor (getStress :432-471 with K_modulus_tensor :359-430 and dirac_delta
// :347-357), written as (00, 11, 22, 12, 02, 01).  The stress is summed in
// the factorized form of the reference's double loop over (j, k):
//   sum_jk K(j,k) (eps X_k) X_j^T c_k V_k c_j V_j
//     = lead w^2 (sum_k (X_k^T eps X_k) c_k V_k) (sum_j X_j X_j^T c_j V_j)
//     + sum_j a_d w (M_j^T eps X_j) M_j X_j^T c_j^3 V_j,
// since dirac_delta(X_k - X_j) is nonzero (c_k / V_k) only for k = j.
// edim is the model's dimension (1, 2 or 3).
template <int DIM>
__global__ __launch_bounds__(256) void k_es_fields(
    FdArgs<DIM> a, const int64_t* __restrict__ rp,
    const int32_t* __restrict__ col, EsDev m, const double* __restrict__ mw,
    int edim, float* __restrict__ energy, double* __restrict__ strain6,
    double* __restrict__ stress6) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  const int64_t sl = i >> 6;
  const int lane = int(i & 63);
  const int64_t bb = a.sl_ptr[sl] >> 6;
  auto broken = [&](int64_t k) {
    return ((a.bits[bb + (k - rp[i])] >> lane) & 1ull) != 0ull;
  };
  double xi[3], ui[3];
  load3<DIM>(a.x[i], xi);
  load3<DIM>(a.u[i], ui);
  const double mwi = mw[i];
  const double w = 1.;
  const double c3 = edim == 3 ? 3. / mwi
                    : edim == 2 ? (2. / mwi) * m.f2d
                                : 1. / mwi;
  // dilatation of the current state (the exact pass-1 expression)
  double th = 0.;
  for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
    if (broken(k)) continue;
    const int32_t j = col[k];
    double xj[3], uj[3];
    load3<DIM>(a.x[j], xj);
    load3<DIM>(a.u[j], uj);
    bool fs = true;  // no break test here
    th += es_dil_exact(m, xi, ui, xj, uj, a.vol[j], c3, fs);
  }
  // material constants of getBondEF (energy) and K_modulus_tensor (stress)
  double alpha = 0., as_e = 0., ad_e = 0., lead = 0., ad_s = 0.;
  if (edim == 1) {
    alpha = m.E / mwi;
  } else if (edim == 2) {
    as_e = (9. / mwi) * m.Kc;
    ad_e = (8. / mwi) * m.mu;
    // K_modulus_tensor takes plane stress and nu from K and mu
    const double Nu = (3. * m.K - 2. * m.mu) / (2. * (3. * m.K + m.mu));
    const double as_s =
        (9. / mwi) * (m.K + pow((Nu + 1.) / (2. * Nu - 1.), 2) * m.mu / 9.);
    ad_s = (8. / mwi) * m.mu;
    const double asb = (2. * m.f2d * as_s - (3. - 2. * m.f2d) * ad_s) / 3.;
    lead = (asb - ad_s) / mwi;
  } else {
    as_e = (9. / mwi) * m.K;
    ad_e = (15. / mwi) * m.mu;
    ad_s = ad_e;
    lead = (as_e - ad_e) / mwi;
  }
  float en = energy ? energy[i] : 0.f;
  double K[3][3] = {}, T[3][3] = {}, XX[3][3] = {};
  const int64_t k0 = rp[i];
  double c0 = 1.;
  if (rp[i + 1] > k0) {
    double x0[3];
    load3<DIM>(a.x[col[k0]], x0);
    const double X0[3] = {x0[0] - xi[0], x0[1] - xi[1], x0[2] - xi[2]};
    c0 = es_corr_exact(m, sqrt(X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]));
  }
  for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
    const int32_t j = col[k];
    double xj[3], uj[3];
    load3<DIM>(a.x[j], xj);
    load3<DIM>(a.u[j], uj);
    double Y[3], lx, ly, ext;
    es_geom_exact(xi, ui, xj, uj, Y, lx, ly, ext);
    const double X[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
    const double c = es_corr_exact(m, lx);
    const double vj = a.vol[j];
    if (energy && !broken(k)) {
      double eb;
      if (edim == 1) {
        eb = 0.5 * alpha * w * ext * ext * c * vj;
      } else {
        const double e_s = th * lx / 3.;
        const double e_d = edim == 2 ? ext - e_s : 0.;
        eb = 0.5 * w * (as_e * e_s * e_s + ad_e * e_d * e_d) * c * vj;
      }
      en += float(eb);
    }
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) {
        K[p][q] += X[p] * X[q] * w * c0 * vj;
        T[p][q] += Y[p] * X[q] * w * c * vj;
        XX[p][q] += X[p] * X[q] * c * vj;
      }
  }
  if (energy) energy[i] = en;
  if (!strain6 && !stress6) return;
  // (K^-1)^T on the edim x edim block
  double Si[3][3] = {};
  if (edim == 1) {
    Si[0][0] = 1. / K[0][0];
  } else if (edim == 2) {
    const double det = K[0][0] * K[1][1] - K[0][1] * K[1][0];
    const double inv[2][2] = {{K[1][1] / det, -K[0][1] / det},
                              {-K[1][0] / det, K[0][0] / det}};
    for (int p = 0; p < 2; p++)
      for (int q = 0; q < 2; q++) Si[p][q] = inv[q][p];
  } else {
    const double c00 = K[1][1] * K[2][2] - K[1][2] * K[2][1];
    const double c01 = K[1][2] * K[2][0] - K[1][0] * K[2][2];
    const double c02 = K[1][0] * K[2][1] - K[1][1] * K[2][0];
    const double det = K[0][0] * c00 + K[0][1] * c01 + K[0][2] * c02;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[1][0] = c01 / det;
    inv[2][0] = c02 / det;
    inv[0][1] = (K[0][2] * K[2][1] - K[0][1] * K[2][2]) / det;
    inv[1][1] = (K[0][0] * K[2][2] - K[0][2] * K[2][0]) / det;
    inv[2][1] = (K[0][1] * K[2][0] - K[0][0] * K[2][1]) / det;
    inv[0][2] = (K[0][1] * K[1][2] - K[0][2] * K[1][1]) / det;
    inv[1][2] = (K[0][2] * K[1][0] - K[0][0] * K[1][2]) / det;
    inv[2][2] = (K[0][0] * K[1][1] - K[0][1] * K[1][0]) / det;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) Si[p][q] = inv[q][p];
  }
  double E[3][3];
  {
    double F[3][3] = {};
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++)
        for (int r = 0; r < 3; r++) F[p][q] += T[p][r] * Si[r][q];
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++)
        E[p][q] = (F[p][q] + F[q][p]) * 0.5 - ((p == q && p < edim) ? 1. : 0.);
  }
  auto put = [&](double* o, const double (&A)[3][3]) {
    o[6 * i + 0] = A[0][0];
    o[6 * i + 1] = A[1][1];
    o[6 * i + 2] = A[2][2];
    o[6 * i + 3] = A[1][2];
    o[6 * i + 4] = A[0][2];
    o[6 * i + 5] = A[0][1];
  };
  if (strain6) put(strain6, E);
  if (!stress6) return;
  // second sweep: sum_k (X_k^T eps X_k) c_k V_k and the k = j terms
  double qsum = 0.;
  double S[3][3] = {};
  const double ad_w = (edim == 1 ? alpha : ad_s) * w;
  for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
    const int32_t j = col[k];
    double xj[3];
    load3<DIM>(a.x[j], xj);
    const double X[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
    const double lx = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    const double c = es_corr_exact(m, lx);
    const double vj = a.vol[j];
    double eX[3];
    for (int p = 0; p < 3; p++)
      eX[p] = E[p][0] * X[0] + E[p][1] * X[1] + E[p][2] * X[2];
    qsum += (X[0] * eX[0] + X[1] * eX[1] + X[2] * eX[2]) * c * vj;
    const double M[3] = {X[0] / lx, X[1] / lx, X[2] / lx};
    const double me = M[0] * eX[0] + M[1] * eX[1] + M[2] * eX[2];
    const double coef = ad_w * me * c * c * c * vj;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) S[p][q] += coef * M[p] * X[q];
  }
  if (edim > 1) {
    const double t1 = lead * w * w * qsum;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) S[p][q] += t1 * XX[p][q];
  }
  put(stress6, S);
}

}  // namespace

}  // namespace nlm

extern "C" int nlm_fd_es_fields(nlm_ctx* c, float* strain_energy,
                                double* strain, double* stress) {
  using namespace nlm;
  if (!c) return NLM_ERR_INVALID;
  FdState& s = c->fd;
  if (!s.ready) return fail(c, NLM_ERR_STATE, "nlm_fd_es_fields: no setup");
  if (s.material != NLM_MATERIAL_ELASTIC_STATE)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_es_fields: state-based material only");
  int rc;
  DevBuf de, dst, dss;
  if ((rc = alloc(c, de, sizeof(float) * size_t(s.n)))) return rc;
  if ((rc = alloc(c, dst, sizeof(double) * 6 * size_t(s.n)))) return rc;
  if ((rc = alloc(c, dss, sizeof(double) * 6 * size_t(s.n)))) return rc;
  std::vector<float> he(size_t(s.n), 0.f);
  if (strain_energy) {
    for (int64_t a = 0; a < s.n; a++) he[a] = strain_energy[s.perm_host[a]];
    NLM_HIP(c, hipMemcpyAsync(de.p, he.data(), sizeof(float) * s.n,
                              hipMemcpyHostToDevice, c->stream));
  }
  const unsigned g = unsigned((s.n + 255) / 256);
  float* pe = strain_energy ? de.as<float>() : nullptr;
  double* pst = strain ? dst.as<double>() : nullptr;
  double* pss = stress ? dss.as<double>() : nullptr;
  if (s.dim == 2)
    hipLaunchKernelGGL(k_es_fields<2>, dim3(g), dim3(256), 0, c->stream,
                       make_args<2>(c), s.row_ptr.as<int64_t>(),
                       s.col.as<int32_t>(), s.es, s.es_mw.as<double>(),
                       s.udim, pe, pst, pss);
  else
    hipLaunchKernelGGL(k_es_fields<3>, dim3(g), dim3(256), 0, c->stream,
                       make_args<3>(c), s.row_ptr.as<int64_t>(),
                       s.col.as<int32_t>(), s.es, s.es_mw.as<double>(),
                       s.udim, pe, pst, pss);
  NLM_HIP(c, hipGetLastError());
  std::vector<double> hs(6 * size_t(s.n)), hq(6 * size_t(s.n));
  if (strain_energy)
    NLM_HIP(c, hipMemcpyAsync(he.data(), de.p, sizeof(float) * s.n,
                              hipMemcpyDeviceToHost, c->stream));
  if (strain)
    NLM_HIP(c, hipMemcpyAsync(hs.data(), dst.p, sizeof(double) * 6 * s.n,
                              hipMemcpyDeviceToHost, c->stream));
  if (stress)
    NLM_HIP(c, hipMemcpyAsync(hq.data(), dss.p, sizeof(double) * 6 * s.n,
                              hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    if (strain_energy) strain_energy[i] = he[a];
    for (int k = 0; k < 6; k++) {
      if (strain) strain[6 * i + k] = hs[6 * a + k];
      if (stress) stress[6 * i + k] = hq[6 * a + k];
    }
  }
  return NLM_OK;
}

extern "C" int nlm_fd_get_node_fields(nlm_ctx* c, double* theta, double* mw) {
  using namespace nlm;
  if (!c) return NLM_ERR_INVALID;
  FdState& s = c->fd;
  if (!s.ready)
    return fail(c, NLM_ERR_STATE, "nlm_fd_get_node_fields: no setup");
  if (s.material != NLM_MATERIAL_ELASTIC_STATE)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_get_node_fields: state-based material only");
  std::vector<double> a(s.n), b(s.n);
  NLM_HIP(c, hipMemcpyAsync(a.data(), s.es_theta.p, sizeof(double) * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipMemcpyAsync(b.data(), s.es_mw.p, sizeof(double) * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t k = 0; k < s.n; k++) {
    const int64_t i = s.perm_host[k];
    if (theta) theta[i] = a[k];
    if (mw) mw[i] = b[k];
  }
  return NLM_OK;
}
