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
//  FP32F  displacement difference and stretch in FP64, bond force magnitude
//         and direction in FP32, accumulated in FP64 by the caller.
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

// FP64 fast variant.
template <int DIM>
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
  // volume correction with guard band around the relative-epsilon boundary
  double lo = m.check_low * (1.0 + 1e-5);
  double band = r - lo;
  if (fabs(band) <= 1e-9 * lo)
    volj = corrected_volj_exact<DIM>(m, xi, xj, volj);
  else if (band > 0.)
    volj *= (m.check_up - r) * m.rcp_h;
  double infl = m.inf_a;
  if (m.inf_type == 1)
    infl = __builtin_fma(m.inf_b, r * m.rcp_horizon, m.inf_a);
  else if (m.inf_type == 2) {
    double q = r * m.rcp_horizon;
    infl = m.inf_a * exp(-q * q / m.inf_b);
  }
  F[0] = F[1] = F[2] = 0.;
  if (brk) {
    if (m.irreversible && !fs) {
      double rho2 = (S * S) * r / m.krbar_sq;
      double d = rho2 - kBreakRho2;
      if (fabs(d) <= kGuard64 * kBreakRho2)
        fs = break_exact<DIM>(m, xi, ui, xj, uj);
      else
        fs = d > 0.;
    }
    if (!fs) {
      double ex = exp(-m.beta * r * S * S);
      double s = (infl * m.c4) * S * ex * volj;
      double sr = s * rinv;
      F[0] = sr * dx0;
      F[1] = sr * dx1;
      if (DIM == 3) F[2] = sr * dx2;
      return;
    }
    if (m.contact) {
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
    return;
  }
  double s = (infl * m.c4) * S * volj;
  double sr = s * rinv;
  F[0] = sr * dx0;
  F[1] = sr * dx1;
  if (DIM == 3) F[2] = sr * dx2;
}

// FP32-force variant: FP64 differences, FP32 geometry and force.
template <int DIM>
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
  float lo = float(m.check_low * (1.0 + 1e-5));
  float band = r - lo;
  float vj = float(volj);
  if (fabsf(band) <= 1e-5f * lo)
    vj = float(corrected_volj_exact<DIM>(m, xi, xj, volj));
  else if (band > 0.f)
    vj *= (float(m.check_up) - r) * float(m.rcp_h);
  float infl = float(m.inf_a);
  if (m.inf_type == 1)
    infl = float(m.inf_a) + float(m.inf_b) * (r * float(m.rcp_horizon));
  else if (m.inf_type == 2) {
    float q = r * float(m.rcp_horizon);
    infl = float(m.inf_a) * __expf(-q * q / float(m.inf_b));
  }
  F[0] = F[1] = F[2] = 0.;
  if (brk) {
    if (m.irreversible && !fs) {
      float rho2 = (S * S) * r / float(m.krbar_sq);
      float d = rho2 - float(kBreakRho2);
      if (fabsf(d) <= kGuard32 * float(kBreakRho2))
        fs = break_exact<DIM>(m, xi, ui, xj, uj);
      else
        fs = d > 0.f;
    }
    if (!fs) {
      float ex = __expf(-m.beta_f * r * S * S);
      float s = infl * m.c4f * S * ex * vj;
      float sr = s * rinv;
      F[0] = double(sr * dx0);
      F[1] = double(sr * dx1);
      if (DIM == 3) F[2] = double(sr * dx2);
      return;
    }
    if (m.contact) {
      double y0 = (xj[0] + uj[0]) - (xi[0] + ui[0]);
      double y1 = (xj[1] + uj[1]) - (xi[1] + ui[1]);
      double y2 = DIM == 3 ? (xj[2] + uj[2]) - (xi[2] + ui[2]) : 0.0;
      float R2 = float(y0 * y0 + y1 * y1 + y2 * y2);
      float Rinv = __builtin_amdgcn_rsqf(R2);
      float R = R2 * Rinv;
      float sf = float(m.kn * (voli * volj / (voli + volj))) *
                 (float(m.rc) - R) * Rinv;
      if (sf < 0.f) sf = 0.f;
      F[0] = -double(sf) * y0;
      F[1] = -double(sf) * y1;
      if (DIM == 3) F[2] = -double(sf) * y2;
    }
    return;
  }
  float s = infl * m.c4f * S * vj;
  float sr = s * rinv;
  F[0] = double(sr * dx0);
  F[1] = double(sr * dx1);
  if (DIM == 3) F[2] = double(sr * dx2);
}

}  // namespace nlm
