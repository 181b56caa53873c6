// bond_es.h -- per-bond arithmetic of the linear state-based material
// (material::pd::ElasticState) on the device.
//
// A force evaluation has two passes over the bonds:
//   pass 1  extension e_ij = |Y| - |X| (StateBasedHelperFunctions::dilatation,
//           stateBasedHelperFunctions.cpp:26-83), the critical-stretch break
//           test, and the dilatation
//             theta_i = sum_j (3/m_i) |X| e_ij c_ij V_j          (3D)
//             theta_i = sum_j (2/m_i) factor2D |X| e_ij c_ij V_j (2D)
//           with the volume correction c_ij (VolumeCorrection::correctVolume,
//           volumeCorrection.cpp:25-49) and the weighted volume m_i;
//   pass 2  the force: ElasticState::getBondEF (ElasticState.cpp:160-271)
//           gives T_ij = t_ij M_ij c_ij with M = Y/|Y| and the scalar state
//             3D  t_ij = (9 K / m_i) theta_i |X| / 3        (e_d stays 0)
//             2D  t_ij = C_i (theta_i |X| / 3) / 3 + alpha_d,i (e_ij - theta_i|X|/3)
//               C_i = 2 factor2D alpha_s,i - (3 - 2 factor2D) alpha_d,i,
//               alpha_s,i = (9/m_i) Kc, alpha_d,i = (8/m_i) mu;
//           QuasiStaticModel::computeForces (QuasiStaticModel.cpp:169-239)
//           adds T_ij V_j to f_i and subtracts it, times V_i, from f_j, so
//           node i collects f_i = sum_j V_j (T_ij - T_ji) with T_ji =
//           -t_ji M_ij c_ij.
//
// Arithmetic variants (as for the bond-based material):
//   EXACT  the reference's expressions and order, IEEE sqrt/division;
//   FP64   e from the cancellation-free form e = du.(2 dx + du) / (|Y| + |X|),
//          rsqrt with Newton steps, fused multiply-adds;
//   FP32F  differences dx, du formed in FP64, then the same cancellation-free
//          extension and the force in FP32; sums in FP64.
// The break decision definitelyGreaterThan(e/|X|, sc) and the volume
// correction test definitelyGreaterThan(|X|, horizon - h/2) are taken
// directly outside a guard band around their switching points and by the
// EXACT expression inside it, so they equal the reference's on the same
// displacement field.
#pragma once

#include <hip/hip_runtime.h>

#include "bond_rnp.h"
#include "device_common.h"
#include "nlm_internal.h"

namespace nlm {

constexpr double kEsGuard64 = 1e-9;
constexpr float kEsGuard32 = 2e-5f;

// Y = (xj + uj) - (xi + ui), X = xj - xi (Point3 operators, point.h), their
// lengths sqrt(x*x + y*y + z*z) (point.h:119) and the extension.
__device__ __forceinline__ void es_geom_exact(const double* xi,
                                              const double* ui,
                                              const double* xj,
                                              const double* uj, double* Y,
                                              double& lx, double& ly,
                                              double& ext) {
  double X[3];
  for (int d = 0; d < 3; d++) {
    Y[d] = (xj[d] + uj[d]) - (xi[d] + ui[d]);
    X[d] = xj[d] - xi[d];
  }
  ly = sqrt(Y[0] * Y[0] + Y[1] * Y[1] + Y[2] * Y[2]);
  lx = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  ext = ly - lx;
}

// VolumeCorrection::correctVolume, volumeCorrection.cpp:40-44
__device__ __forceinline__ double es_corr_exact(const EsDev& m, double lx) {
  double c = 1.;
  if (def_greater(lx, m.check_low)) c = (m.check_up - lx) / m.h;
  return c;
}

// Pass 1, EXACT: break test (fs in/out) and the dilatation term
// ((((c3 * |X|) * e) * c) * V_j), c3 = 3/m_i (3D) or (2/m_i) factor2D (2D).
__device__ __forceinline__ double es_dil_exact(const EsDev& m,
                                               const double* xi,
                                               const double* ui,
                                               const double* xj,
                                               const double* uj, double volj,
                                               double c3, bool& fs) {
  double Y[3], lx, ly, ext;
  es_geom_exact(xi, ui, xj, uj, Y, lx, ly, ext);
  if (m.breaking && !fs && def_greater(ext / lx, m.sc)) fs = true;
  double c = es_corr_exact(m, lx);
  return c3 * lx * ext * c * volj;
}

// Pass 1, FP64: same decision, cancellation-free extension.  exn: the
// decision lies in a guard band and must be redone by es_dil_exact.
template <int DIM>
__device__ __forceinline__ double es_dil_fp64(const EsDev& m, const double* xi,
                                              const double* ui,
                                              const double* xj,
                                              const double* uj, double volj,
                                              double c3, bool fs0, bool& fs,
                                              bool& exn) {
  double dx0 = xj[0] - xi[0], dx1 = xj[1] - xi[1];
  double dx2 = DIM == 3 ? xj[2] - xi[2] : 0.;
  double du0 = uj[0] - ui[0], du1 = uj[1] - ui[1];
  double du2 = DIM == 3 ? uj[2] - ui[2] : 0.;
  double dd = __builtin_fma(dx1, dx1, dx0 * dx0);
  double g = __builtin_fma(du1, __builtin_fma(2., dx1, du1),
                           du0 * __builtin_fma(2., dx0, du0));
  if (DIM == 3) {
    dd = __builtin_fma(dx2, dx2, dd);
    g = __builtin_fma(du2, __builtin_fma(2., dx2, du2), g);
  }
  double rx = rsqrt_f64(dd);
  double lx = dd * rx;
  double yy = dd + g;
  double ly = yy * rsqrt_f64(yy);
  double ext = g / (ly + lx);
  double s = ext * rx;
  bool bex = m.breaking && !fs0 &&
             fabs(s - m.sc_break) <= kEsGuard64 * m.sc_break;
  fs = fs0 || (m.breaking && s > m.sc_break);
  bool vex = fabs(lx - m.low_band) <= kEsGuard64 * m.low_band;
  exn = bex || vex;
  double c = lx > m.low_band ? (m.check_up - lx) * m.rcp_h : 1.;
  return c3 * lx * ext * c * volj;
}

// Pass 1, FP32F.
template <int DIM>
__device__ __forceinline__ double es_dil_fp32(const EsDev& m, const double* xi,
                                              const double* ui,
                                              const double* xj,
                                              const double* uj, double volj,
                                              float c3, bool fs0, bool& fs,
                                              bool& exn) {
  float dx0 = float(xj[0] - xi[0]), dx1 = float(xj[1] - xi[1]);
  float dx2 = DIM == 3 ? float(xj[2] - xi[2]) : 0.f;
  float du0 = float(uj[0] - ui[0]), du1 = float(uj[1] - ui[1]);
  float du2 = DIM == 3 ? float(uj[2] - ui[2]) : 0.f;
  float dd = __builtin_fmaf(dx1, dx1, dx0 * dx0);
  float g = __builtin_fmaf(du1, __builtin_fmaf(2.f, dx1, du1),
                           du0 * __builtin_fmaf(2.f, dx0, du0));
  if (DIM == 3) {
    dd = __builtin_fmaf(dx2, dx2, dd);
    g = __builtin_fmaf(du2, __builtin_fmaf(2.f, dx2, du2), g);
  }
  float rx = __builtin_amdgcn_rsqf(dd);
  float lx = dd * rx;
  float yy = dd + g;
  float ly = yy * __builtin_amdgcn_rsqf(yy);
  float ext = g * __builtin_amdgcn_rcpf(ly + lx);
  float s = ext * rx;
  bool bex = m.breaking && !fs0 &&
             fabsf(s - m.sc_break_f) <= kEsGuard32 * m.sc_break_f;
  fs = fs0 || (m.breaking && s > m.sc_break_f);
  bool vex = fabsf(lx - m.low_band_f) <= 3e-5f * m.low_band_f;
  exn = bex || vex;
  float c = fminf((m.check_up_f - lx) * m.rcp_h_f, 1.0f);
  return double(c3 * lx * ext * c * float(volj));
}

// Pass 2, EXACT: V_j T_ij and V_j T_ji of getBondEF for the two directions
// of one bond, given the per-node coefficients (3D: P = (9/m) K; 2D: P = C,
// Q = alpha_d) and dilatations of both ends.
template <int DIM>
__device__ __forceinline__ void es_force_exact(
    const EsDev& m, const double* xi, const double* ui, const double* xj,
    const double* uj, double volj, double Pi, double Qi, double thi, double Pj,
    double Qj, double thj, double* own, double* oth) {
  double Y[3], lx, ly, ext;
  es_geom_exact(xi, ui, xj, uj, Y, lx, ly, ext);
  double M[3] = {Y[0] / ly, Y[1] / ly, Y[2] / ly};
  double c = es_corr_exact(m, lx);
  const double w = 1.;
  double tij, tji;
  if constexpr (DIM == 3) {
    double es_i = thi * lx / 3.;
    double es_j = thj * lx / 3.;
    tij = Pi * w * es_i;
    tji = Pj * w * es_j;
  } else {
    double es_i = thi * lx / 3.;
    double ed_i = ext - es_i;
    tij = Pi * w * es_i / 3. + Qi * w * ed_i;
    double es_j = thj * lx / 3.;
    double ed_j = ext - es_j;
    tji = Pj * w * es_j / 3. + Qj * w * ed_j;
  }
  for (int d = 0; d < 3; d++) {
    own[d] = volj * ((M[d] * tij) * c);
    oth[d] = volj * ((-M[d] * tji) * c);
  }
}

}  // namespace nlm
