// bond_rnp.h -- per-bond RNPBond force and break test on the device.
//
// Follows RNPBond::getBondEF (src/material/pd/rnpBond.cpp:216-301).  Three
// arithmetic variants share one break decision:
//
//  EXACT  the reference's operation order: 3-term Point3 dot products with a
//         zero z in 2D, IEEE division and square root, no fused multiply-add
//         (the library is compiled with -ffp-contract=off).
//  FP64   the stretch S, the reference length r and the break test as EXACT
//         decides them; the force magnitude and direction from reciprocal
//         square roots and fused multiply-adds.
//  FP32F  position and displacement differences formed in FP64, then bond
//         length, stretch, force magnitude and direction in FP32; forces
//         accumulated in FP64 by the caller.
//
// Break decision.  The reference breaks a bond when
//   definitelyGreaterThan(|S|, k * (rbar / sqrt(r)))   (rnpBond.cpp:263-266)
// i.e. when rho = |S| / T exceeds 1/(1-1e-5) (T the threshold), up to the
// rounding of that expression.  FP64/FP32F evaluate rho^2 = S^2 r / (k rbar)^2
// with cheap arithmetic and decide directly when rho^2 lies outside a guard
// band around 1/(1-1e-5)^2 that is far wider than their rounding error; inside
// the band the EXACT expression is evaluated.  The decision therefore equals
// the reference's on the same displacement field.  The volume-correction test
// definitelyGreaterThan(r, horizon - h/2) (rnpBond.cpp:255-256) is handled the
// same way.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "nlm_internal.h"

namespace nlm {

constexpr int kModeExact = NLM_PREC_FP64_EXACT;
constexpr int kModeFp64 = NLM_PREC_FP64;
constexpr int kModeFp32 = NLM_PREC_FP32_FORCE;

// 1/(1-1e-5)^2: square of the ratio |S|/T at which definitelyGreaterThan
// switches (for |S| > T > 0).
constexpr double kBreakRho2 = 1.0 / ((1.0 - 1e-5) * (1.0 - 1e-5));
// Relative half-width of the guard band (well above FP32 rounding of rho^2).
constexpr double kGuard64 = 1e-9;
constexpr float kGuard32 = 1e-4f;

// Influence function J(r/horizon), src/material/pd/influenceFn.h:104-106,
// LinearInfluenceFn, GaussianInfluenceFn.
__device__ __forceinline__ double influence_exact(const RnpDev& m, double r) {
  if (m.inf_type == 0) return m.inf_a;
  double q = r / m.horizon;
  if (m.inf_type == 1) return m.inf_a + m.inf_b * q;
  return m.inf_a * exp(-q * q / m.inf_b);
}

// Exact restatement of getBondEF on explicit operands.  dx = xj - xi,
// du = uj - ui per component; for the broken-bond contact force the positions
// are needed as well.  Returns the force density contribution in F[0..2];
// energy only when WANT_E.
template <int DIM, bool WANT_E>
__device__ __forceinline__ void bond_exact(const RnpDev& m, const double* xi,
                                           const double* ui, const double* xj,
                                           const double* uj, double voli,
                                           double volj, bool brk, bool& fs,
                                           double* F, double* E) {
  double dx0 = xj[0] - xi[0], dx1 = xj[1] - xi[1];
  double dx2 = DIM == 3 ? xj[2] - xi[2] : 0.0;
  double du0 = uj[0] - ui[0], du1 = uj[1] - ui[1];
  double du2 = DIM == 3 ? uj[2] - ui[2] : 0.0;
  // xj.dist(xi) (point.h:165-169) and dx.dot(dx) are the same sum
  double dd = dx0 * dx0 + dx1 * dx1;
  dd = dd + dx2 * dx2;
  double rji = sqrt(dd);
  double num = dx0 * du0 + dx1 * du1;
  num = num + dx2 * du2;
  double S = num / dd;                                  // getS :303-306
  if (def_greater(rji, m.check_low)) volj *= (m.check_up - rji) / m.h;
  double infl = influence_exact(m, rji);
  F[0] = F[1] = F[2] = 0.;
  if (WANT_E) *E = 0.;
  if (brk) {
    if (m.irreversible && !fs &&
        def_greater(fabs(S), m.factor_sc * (m.rbar / sqrt(rji))))
      fs = true;
    if (!fs) {
      double ex = exp(-m.beta * rji * S * S);
      if (WANT_E)
        *E = (infl * m.C * (1. - exp(-m.beta * rji * S * S)) / m.inv_factor) *
             volj;
      double s = infl * 4. * S * m.C * m.beta * ex / m.inv_factor;
      s = s * volj;
      F[0] = s * (dx0 / rji);                           // eij :330-333
      F[1] = s * (dx1 / rji);
      if (DIM == 3) F[2] = s * (dx2 / rji);
      return;
    }
    if (WANT_E) *E = infl * m.C / m.inv_factor * volj;
    if (m.contact) {                                    // :286-291
      double y0 = (xj[0] + uj[0]) - (xi[0] + ui[0]);
      double y1 = (xj[1] + uj[1]) - (xi[1] + ui[1]);
      double y2 = DIM == 3 ? (xj[2] + uj[2]) - (xi[2] + ui[2]) : 0.0;
      double R = y0 * y0 + y1 * y1;
      R = sqrt(R + y2 * y2);
      double sf = m.kn * (voli * volj / (voli + volj)) * (m.rc - R) / R;
      if (sf < 0.) sf = 0.;
      F[0] = -sf * y0;
      F[1] = -sf * y1;
      if (DIM == 3) F[2] = -sf * y2;
    }
    return;
  }
  if (WANT_E)
    *E = (infl * m.C * m.beta * rji * S * S / m.inv_factor) * volj;
  double s = infl * 4. * S * m.C * m.beta / m.inv_factor;
  s = s * volj;
  F[0] = s * (dx0 / rji);
  F[1] = s * (dx1 / rji);
  if (DIM == 3) F[2] = s * (dx2 / rji);
}

// Exact break decision (no side effects).
template <int DIM>
__device__ __forceinline__ bool break_exact(const RnpDev& m, const double* xi,
                                            const double* ui,
                                            const double* xj,
                                            const double* uj) {
  double dx0 = xj[0] - xi[0], dx1 = xj[1] - xi[1];
  double dx2 = DIM == 3 ? xj[2] - xi[2] : 0.0;
  double du0 = uj[0] - ui[0], du1 = uj[1] - ui[1];
  double du2 = DIM == 3 ? uj[2] - ui[2] : 0.0;
  double dd = dx0 * dx0 + dx1 * dx1;
  dd = dd + dx2 * dx2;
  double rji = sqrt(dd);
  double num = dx0 * du0 + dx1 * du1;
  num = num + dx2 * du2;
  double S = num / dd;
  return def_greater(fabs(S), m.factor_sc * (m.rbar / sqrt(rji)));
}

// Exact volume-correction factor decision.
template <int DIM>
__device__ __forceinline__ double corrected_volj_exact(const RnpDev& m,
                                                       const double* xi,
                                                       const double* xj,
                                                       double volj) {
  double dx0 = xj[0] - xi[0], dx1 = xj[1] - xi[1];
  double dx2 = DIM == 3 ? xj[2] - xi[2] : 0.0;
  double dd = dx0 * dx0 + dx1 * dx1;
  dd = dd + dx2 * dx2;
  double rji = sqrt(dd);
  if (def_greater(rji, m.check_low)) volj *= (m.check_up - rji) / m.h;
  return volj;
}

// Reciprocal square root to full double precision: hardware estimate plus two
// Newton-Raphson steps.
__device__ __forceinline__ double rsqrt_f64(double a) {
  double y = __builtin_amdgcn_rsq(a);
  double h = a * y;
  double e = __builtin_fma(-h, y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  h = a * y;
  e = __builtin_fma(-h, y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}

// Influence function value for the fast paths (INF: 0 constant, 1 linear,
// 2 gaussian; r the reference bond length).
template <int INF>
__device__ __forceinline__ double influence_f64(const RnpDev& m, double r) {
  if constexpr (INF == 0) return m.inf_a;
  if constexpr (INF == 1) return __builtin_fma(m.inf_b, r * m.rcp_horizon, m.inf_a);
  double q = r * m.rcp_horizon;
  return m.inf_a * exp(-(q * q) * m.rcp_inf_b);
}
template <int INF>
__device__ __forceinline__ float influence_f32(const RnpDev& m, float r) {
  if constexpr (INF == 0) return m.inf_a_f;
  if constexpr (INF == 1) return __builtin_fmaf(m.inf_b_f, r * m.rcp_horizon_f, m.inf_a_f);
  float q = r * m.rcp_horizon_f;
  return m.inf_a_f * __expf(-(q * q) * m.rcp_inf_b_f);
}

// Contact force of a broken bond (rnpBond.cpp:285-291), FP64.
template <int DIM>
__device__ __forceinline__ void contact_f64(const RnpDev& m, const double* xi,
                                            const double* ui,
                                            const double* xj,
                                            const double* uj, double voli,
                                            double volj, double* F) {
  double y0 = (xj[0] + uj[0]) - (xi[0] + ui[0]);
  double y1 = (xj[1] + uj[1]) - (xi[1] + ui[1]);
  double y2 = DIM == 3 ? (xj[2] + uj[2]) - (xi[2] + ui[2]) : 0.0;
  double R2 = __builtin_fma(y1, y1, y0 * y0);
  if (DIM == 3) R2 = __builtin_fma(y2, y2, R2);
  double Rinv = rsqrt_f64(R2);
  double R = R2 * Rinv;
  double sf = m.kn * (voli * volj / (voli + volj)) * (m.rc - R) * Rinv;
  if (sf < 0.) sf = 0.;
  F[0] = -sf * y0;
  F[1] = -sf * y1;
  if (DIM == 3) F[2] = -sf * y2;
}

// FP64 fast variant.
template <int DIM, int INF>
__device__ __forceinline__ void bond_fp64(const RnpDev& m, const double* xi,
                                          const double* ui, const double* xj,
                                          const double* uj, double voli,
                                          double volj, bool brk, bool& fs,
                                          double* F) {
  double dx0 = xj[0] - xi[0], dx1 = xj[1] - xi[1];
  double dx2 = DIM == 3 ? xj[2] - xi[2] : 0.0;
  double du0 = uj[0] - ui[0], du1 = uj[1] - ui[1];
  double du2 = DIM == 3 ? uj[2] - ui[2] : 0.0;
  double dd = __builtin_fma(dx1, dx1, dx0 * dx0);
  double num = __builtin_fma(dx1, du1, dx0 * du0);
  if (DIM == 3) {
    dd = __builtin_fma(dx2, dx2, dd);
    num = __builtin_fma(dx2, du2, num);
  }
  double rinv = rsqrt_f64(dd);
  double r = dd * rinv;
  double S = num * (rinv * rinv);
  // volume correction, exact decision inside the guard band
  double band = r - m.low_band;
  if (fabs(band) <= 1e-9 * m.low_band)
    volj = corrected_volj_exact<DIM>(m, xi, xj, volj);
  else if (band > 0.)
    volj *= (m.check_up - r) * m.rcp_h;
  double infl = influence_f64<INF>(m, r);
  F[0] = F[1] = F[2] = 0.;
  if (brk) {
    if (m.irreversible && !fs) {
      double rho2 = (S * S) * r * m.rcp_krbar_sq;
      double d = rho2 - kBreakRho2;
      if (fabs(d) <= kGuard64 * kBreakRho2)
        fs = break_exact<DIM>(m, xi, ui, xj, uj);
      else
        fs = d > 0.;
    }
    if (!fs) {
      double ex = exp(-m.beta * r * S * S);
      double sr = (infl * m.c4) * S * ex * volj * rinv;
      F[0] = sr * dx0;
      F[1] = sr * dx1;
      if (DIM == 3) F[2] = sr * dx2;
      return;
    }
    if (m.contact) contact_f64<DIM>(m, xi, ui, xj, uj, voli, volj, F);
    return;
  }
  double sr = (infl * m.c4) * S * volj * rinv;
  F[0] = sr * dx0;
  F[1] = sr * dx1;
  if (DIM == 3) F[2] = sr * dx2;
}

// FP32-force variant: displacement and position differences in FP64 (no
// cancellation: u_j - u_i is formed before rounding), then bond geometry,
// stretch and force in FP32; the break decision is exact (guard band).
template <int DIM, int INF>
__device__ __forceinline__ void bond_fp32(const RnpDev& m, const double* xi,
                                          const double* ui, const double* xj,
                                          const double* uj, double voli,
                                          double volj, bool brk, bool& fs,
                                          double* F) {
  float dx0 = float(xj[0] - xi[0]), dx1 = float(xj[1] - xi[1]);
  float dx2 = DIM == 3 ? float(xj[2] - xi[2]) : 0.f;
  float du0 = float(uj[0] - ui[0]), du1 = float(uj[1] - ui[1]);
  float du2 = DIM == 3 ? float(uj[2] - ui[2]) : 0.f;
  float dd = __builtin_fmaf(dx1, dx1, dx0 * dx0);
  float num = __builtin_fmaf(dx1, du1, dx0 * du0);
  if (DIM == 3) {
    dd = __builtin_fmaf(dx2, dx2, dd);
    num = __builtin_fmaf(dx2, du2, num);
  }
  float rinv = __builtin_amdgcn_rsqf(dd);
  float r = dd * rinv;
  float S = num * (rinv * rinv);
  float band = r - m.low_band_f;
  float vj = float(volj);
  if (fabsf(band) <= 1e-5f * m.low_band_f)
    vj = float(corrected_volj_exact<DIM>(m, xi, xj, volj));
  else if (band > 0.f)
    vj *= (m.check_up_f - r) * m.rcp_h_f;
  float infl = influence_f32<INF>(m, r);
  F[0] = F[1] = F[2] = 0.;
  if (brk) {
    if (m.irreversible && !fs) {
      float rho2 = (S * S) * r * m.rcp_krbar_sq_f;
      float d = rho2 - float(kBreakRho2);
      if (fabsf(d) <= kGuard32 * float(kBreakRho2))
        fs = break_exact<DIM>(m, xi, ui, xj, uj);
      else
        fs = d > 0.f;
    }
    if (!fs) {
      float ex = __expf(-m.beta_f * r * S * S);
      float sr = (infl * m.c4f) * S * ex * vj * rinv;
      F[0] = double(sr * dx0);
      F[1] = double(sr * dx1);
      if (DIM == 3) F[2] = double(sr * dx2);
      return;
    }
    if (m.contact) contact_f64<DIM>(m, xi, ui, xj, uj, voli, volj, F);
    return;
  }
  float sr = (infl * m.c4f) * S * vj * rinv;
  F[0] = double(sr * dx0);
  F[1] = double(sr * dx1);
  if (DIM == 3) F[2] = double(sr * dx2);
}

// Wavefront masks (bit = lane) of one neighbor slot for the fast variants.
struct SlotMasks {
  unsigned long long act;  // the slot lies within the lane's row
  unsigned long long brk;  // both ends breakable (no-fail regions)
  unsigned long long w;    // broken before this step
  unsigned long long vex;  // length within the volume-correction edge band
};

// Branch-free common path of the fast variants.  Computes the force of a
// bond whose two boundary decisions (volume correction, break test) are
// clear of their guard bands; the decisions are wavefront votes on single
// comparisons (a vote on a comparison is its lane mask), combined by scalar
// mask arithmetic.  Outputs: nb, the bonds that break now outside the break
// guard band; band, the breakable bonds inside it; ex, the active bonds the
// caller must recompute with bond_fp64 / bond_fp32 (a decision inside a
// band, or a contact force), whose force here is 0.  Broken bonds and
// inactive slots contribute 0.
template <int DIM, int INF, bool NOFAIL>
__device__ __forceinline__ void bond_fp64_fast(
    const RnpDev& m, const double* xi, const double* ui, const double* xj,
    const double* uj, double volj, const SlotMasks& k, unsigned long long& nb,
    unsigned long long& band, unsigned long long& ex, double* F) {
  double dx0 = xj[0] - xi[0], dx1 = xj[1] - xi[1];
  double dx2 = DIM == 3 ? xj[2] - xi[2] : 0.0;
  double du0 = uj[0] - ui[0], du1 = uj[1] - ui[1];
  double du2 = DIM == 3 ? uj[2] - ui[2] : 0.0;
  double dd = __builtin_fma(dx1, dx1, dx0 * dx0);
  double num = __builtin_fma(dx1, du1, dx0 * du0);
  if (DIM == 3) {
    dd = __builtin_fma(dx2, dx2, dd);
    num = __builtin_fma(dx2, du2, num);
  }
  double rinv = rsqrt_f64(dd);
  double r = dd * rinv;
  double S = num * (rinv * rinv);
  // (check_up - r)/h >= 1 exactly when r <= check_low, so the clamp gives
  // the reference's factor away from the relative-epsilon edge (bonds near
  // the edge are in k.vex)
  double fac = fmin((m.check_up - r) * m.rcp_h, 1.0);
  // rho^2 against the guard band [break_lo, break_hi] (+inf when bonds
  // never break)
  const double rho2 = (S * S) * r * m.rcp_krbar_sq;
  const unsigned long long v_hi = __builtin_amdgcn_ballot_w64(rho2 > m.break_hi);
  const unsigned long long v_lo =
      __builtin_amdgcn_ballot_w64(rho2 >= m.break_lo);
  const unsigned long long cand = k.act & k.brk & ~k.w;
  nb = cand & v_hi;
  band = cand & v_lo & ~v_hi;
  const unsigned long long fs = k.w | nb;
  ex = (k.vex | band | (fs & k.brk & m.contact_mask)) & k.act;
  const unsigned long long keep = k.act & ~(fs & k.brk) & ~ex;
  double infl = influence_f64<INF>(m, r);
  double ex_ = exp(-m.beta * r * S * S);
  // bonds touching a no-fail node use the linearised force whatever their
  // bit (rnpBond.cpp:296-300)
  if (NOFAIL) ex_ = __builtin_amdgcn_inverse_ballot_w64(k.brk) ? ex_ : 1.0;
  double sr = (infl * m.c4) * S * ex_ * (volj * fac) * rinv;
  sr = __builtin_amdgcn_inverse_ballot_w64(keep) ? sr : 0.0;
  F[0] = sr * dx0;
  F[1] = sr * dx1;
  F[2] = DIM == 3 ? sr * dx2 : 0.0;
}

// FP32-force common path on the bond's position and displacement
// differences (FP32).  cvj = vol_j c4 (J) in FP32 (c4f_infl): the bond's
// constant factor.  The force comes out in FP32.
template <int DIM, int INF, bool NOFAIL>
__device__ __forceinline__ void bond_fp32_core(
    const RnpDev& m, float dx0, float dx1, float dx2, float du0, float du1,
    float du2, float cvj, const SlotMasks& k, unsigned long long& nb,
    unsigned long long& band, unsigned long long& ex, float* F) {
  float dd = __builtin_fmaf(dx1, dx1, dx0 * dx0);
  float num = __builtin_fmaf(dx1, du1, dx0 * du0);
  if (DIM == 3) {
    dd = __builtin_fmaf(dx2, dx2, dd);
    num = __builtin_fmaf(dx2, du2, num);
  }
  float rinv = __builtin_amdgcn_rsqf(dd);
  float r = dd * rinv;
  float q = num * rinv;           // S r
  float S = q * rinv;
  float rS2 = q * S;              // r S^2: exponent and break ratio
  // volume-correction factor (check_up - r)/h clamped to [0, 1] (the
  // clamp is an output modifier of the multiply-add; r < check_up for every
  // bond); bonds near the correction's edge are in k.vex
  const float fac =
      __builtin_amdgcn_fmed3f(__builtin_fmaf(r, -m.rcp_h_f, m.up_rcp_h_f),
                              0.f, 1.f);
  // r S^2 against the guard band [break_lo_f, break_hi_f] (+inf when bonds
  // never break)
  const unsigned long long v_hi =
      __builtin_amdgcn_ballot_w64(rS2 > m.break_hi_f);
  const unsigned long long v_lo =
      __builtin_amdgcn_ballot_w64(rS2 >= m.break_lo_f);
  const unsigned long long cand = k.act & k.brk & ~k.w;
  nb = cand & v_hi;
  band = cand & v_lo & ~v_hi;
  const unsigned long long fs = k.w | nb;
  const unsigned long long v_vex =
      __builtin_amdgcn_ballot_w64(fabsf(r - m.low_band_f) <= m.vex_tol_f);
  ex = (v_vex | band | (fs & k.brk & m.contact_mask)) & k.act;
  const unsigned long long keep = k.act & ~(fs & k.brk) & ~ex;
  float ex_ = __builtin_amdgcn_exp2f(rS2 * m.nbeta_log2e_f);
  if (NOFAIL) ex_ = __builtin_amdgcn_inverse_ballot_w64(k.brk) ? ex_ : 1.0f;
  float sr = (S * ex_) * ((cvj * fac) * rinv);
  if constexpr (INF != 0) sr *= influence_f32<INF>(m, r);
  sr = __builtin_amdgcn_inverse_ballot_w64(keep) ? sr : 0.f;
  F[0] = sr * dx0;
  F[1] = sr * dx1;
  F[2] = DIM == 3 ? sr * dx2 : 0.f;
}

// FP32-force common path of the LDS-staged kernel, on scaled differences:
// dx = sigma (x_j - x_i) and dv = tau (u_j - u_i) with tau a power of two and
// tau^2 / sigma = beta log2(e) (the halo records and the node hold sigma
// (x - origin) in FP32 and tau u in FP64).  Then q = dx.dv / |dx| = tau S r,
// Ss = q / |dx| = tau S / sigma and b = q Ss = beta log2(e) r S^2 is at once
// the exponent of the bond force, exp(-beta r S^2) = exp2(-b), and the break
// test's quantity (its guard band's lower bound break_lo_s scaled on the
// host).
// The force S exp(-beta r S^2) V_j c4 J (fac / r) (x_j - x_i) equals
// Ss exp2(-b) cvs (fac / |dx|) dx with cvs = V_j c4 J sigma / tau.  The
// volume-correction factor enters as fac / |dx| =
// clamp((check_up/h) / |dx| - 1/(sigma h), 0, 1/|dx|), whose upper bound holds
// exactly when r <= check_low; the edge band is tested on |dx|^2.  ACC: the
// force is added to F with fused multiply-adds (F = sr dx + F), else written.
template <int DIM, int INF, bool NOFAIL, bool ACC>
__device__ __forceinline__ void bond_fp32_scaled(
    const RnpDev& m, float dx0, float dx1, float dx2, float dv0, float dv1,
    float dv2, float cvs, const SlotMasks& k, unsigned long long& nb,
    unsigned long long& band, unsigned long long& ex, float* F) {
  float dd = __builtin_fmaf(dx1, dx1, dx0 * dx0);
  float num = __builtin_fmaf(dx1, dv1, dx0 * dv0);
  if (DIM == 3) {
    dd = __builtin_fmaf(dx2, dx2, dd);
    num = __builtin_fmaf(dx2, dv2, num);
  }
  float rinv = __builtin_amdgcn_rsqf(dd);
  float q = num * rinv;           // tau S r
  float Ss = q * rinv;            // tau S / sigma
  float b = q * Ss;               // beta log2(e) r S^2
  const float facr = __builtin_amdgcn_fmed3f(
      __builtin_fmaf(m.up_rcp_h_f, rinv, -m.rcp_h_s), 0.f, rinv);
  // One vote on the guard band's lower bound: a bond that may break (inside
  // the band, or beyond it: breaking in this step) is decided, and its force
  // taken, by the deferred evaluation, which also reports the new broken
  // bits; band carries those bonds, nb stays empty here.
  const unsigned long long v_lo =
      __builtin_amdgcn_ballot_w64(b >= m.break_lo_s);
  band = k.act & k.brk & ~k.w & v_lo;
  nb = 0ull;
  const unsigned long long fsb = k.w & k.brk;
  const unsigned long long v_vex =
      __builtin_amdgcn_ballot_w64(fabsf(dd - m.low_dd_f) <= m.vex_dd_tol_f);
  ex = (v_vex | band | (fsb & m.contact_mask)) & k.act;
  const unsigned long long keep = k.act & ~fsb & ~ex;
  float ex_ = __builtin_amdgcn_exp2f(-b);
  if (NOFAIL) ex_ = __builtin_amdgcn_inverse_ballot_w64(k.brk) ? ex_ : 1.0f;
  float sr = (Ss * ex_) * (cvs * facr);
  if constexpr (INF != 0)
    sr *= influence_f32<INF>(m, (dd * rinv) * m.rcp_sigma_f);
  sr = __builtin_amdgcn_inverse_ballot_w64(keep) ? sr : 0.f;
  if constexpr (ACC) {
    F[0] = __builtin_fmaf(sr, dx0, F[0]);
    F[1] = __builtin_fmaf(sr, dx1, F[1]);
    if (DIM == 3) F[2] = __builtin_fmaf(sr, dx2, F[2]);
  } else {
    F[0] = sr * dx0;
    F[1] = sr * dx1;
    F[2] = DIM == 3 ? sr * dx2 : 0.f;
  }
}

// The same from FP64 positions and displacements: the differences are
// formed in FP64 (no cancellation) and rounded to FP32.
template <int DIM, int INF, bool NOFAIL>
__device__ __forceinline__ void bond_fp32_fast(
    const RnpDev& m, const double* xi, const double* ui, const double* xj,
    const double* uj, float cvj, const SlotMasks& k, unsigned long long& nb,
    unsigned long long& band, unsigned long long& ex, float* F) {
  bond_fp32_core<DIM, INF, NOFAIL>(
      m, float(xj[0] - xi[0]), float(xj[1] - xi[1]),
      DIM == 3 ? float(xj[2] - xi[2]) : 0.f, float(uj[0] - ui[0]),
      float(uj[1] - ui[1]), DIM == 3 ? float(uj[2] - ui[2]) : 0.f, cvj, k, nb,
      band, ex, F);
}

}  // namespace nlm
