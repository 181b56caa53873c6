// device_common.h -- device helpers shared by the nlmech_amd kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nlm_internal.h"

namespace nlm {

// Relative-epsilon comparisons of util::compare (src/util/compare.cpp:25-33,
// COMPARE_EPS = 1e-5 at src/util/compare.h:13).  The library is compiled with
// -ffp-contract=off, so these round every operation like the reference.
__host__ __device__ __forceinline__ bool def_greater(double a, double b) {
  double aa = fabs(a), ab = fabs(b);
  return (a - b) > ((aa < ab ? ab : aa) * 1e-5);
}
__host__ __device__ __forceinline__ bool def_less(double a, double b) {
  double aa = fabs(a), ab = fabs(b);
  return (b - a) > ((aa < ab ? ab : aa) * 1e-5);
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup b runs on XCD b % 8).  This maps the b-th dispatched workgroup to
// a logical block so that each XCD walks one contiguous eighth of the nodes
// and its private L2 keeps the neighbor band of that eighth.  The grid is
// padded to a multiple of 8 (grid_xcd); logical blocks >= nblocks exit.
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks) {
  int64_t g = gridDim.x;
  int64_t per = g >> 3;
  int64_t b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

template <int DIM>
struct VT;
template <>
struct VT<2> {
  using T = double2;
};
template <>
struct VT<3> {
  using T = double4;
};

template <class V>
__device__ __forceinline__ double comp(const V& a, int d) {
  return d == 0 ? a.x : (d == 1 ? a.y : a.z);
}
__device__ __forceinline__ double comp(const double2& a, int d) {
  return d == 0 ? a.x : a.y;
}

// Displacement BC value at time t for a node at position (px,py,pz):
// ULoading::apply (src/loading/uLoading.cpp:166-253).  A "constant" spatial
// function without parameters multiplies by 1.
__device__ __forceinline__ void ubc_eval(const BcSetDev& s, double px,
                                         double py, double pz, double t,
                                         double& du, double& dv) {
  double umax = s.tp[0];
  du = 0.;
  dv = 0.;
  switch (s.space_fn) {
    case NLM_SFN_SIN_X: {
      double a = M_PI * s.sp[0];
      umax = umax * sin(a * px);
    } break;
    case NLM_SFN_SIN_Y: {
      double a = M_PI * s.sp[0];
      umax = umax * sin(a * py);
    } break;
    case NLM_SFN_SIN_Z: {
      double a = M_PI * s.sp[0];
      umax = umax * sin(a * pz);
    } break;
    case NLM_SFN_LINEAR_X: {
      double a = s.sp[0];
      umax = umax * a * px;
    } break;
    case NLM_SFN_LINEAR_Y: {
      double a = s.sp[0];
      umax = umax * a * py;
    } break;
    case NLM_SFN_LINEAR_Z: {
      double a = s.sp[0];
      umax = umax * a * pz;
    } break;
    case NLM_SFN_CONSTANT: {
      double a = s.n_space_params > 0 ? s.sp[0] : 1.0;
      umax = umax * a;
    } break;
    default:
      break;
  }
  const double* p = s.tp;
  switch (s.time_fn) {
    case NLM_TFN_CONSTANT:
      du = umax;
      break;
    case NLM_TFN_LINEAR:
      du = umax * t;
      dv = umax;
      break;
    case NLM_TFN_QUADRATIC:
      du = umax * t + p[1] * t * t;
      dv = umax + p[1] * t;
      break;
    case NLM_TFN_SIN: {
      double a = M_PI * p[1];
      du = umax * sin(a * t);
      dv = umax * a * cos(a * t);
    } break;
    case NLM_TFN_LINEAR_STEP: {
      // util::function::linearStepFunc / derLinearStepFunc,
      // src/util/utilFunction.cpp:46-64
      double period = floor(t / (p[1] + p[2]));
      if (def_less(t, period * (p[1] + p[2]) + p[1])) {
        du = umax * (t - period * p[2]);
        dv = umax * 1.;
      } else {
        du = umax * ((period + 1.) * p[1]);
        dv = umax * 0.;
      }
    } break;
    case NLM_TFN_LINEAR_SLOW_FAST:
      if (def_greater(t, p[1])) {
        du = umax * p[3] * t;
        dv = umax * p[3];
      } else {
        du = umax * p[2] * t;
        dv = umax * p[2];
      }
      break;
    default:
      break;
  }
}

// Force BC value (force density) at time t: FLoading::apply
// (src/loading/fLoading.cpp:117-279).
__device__ __forceinline__ double fbc_eval(const BcSetDev& s, double px,
                                           double py, double pz, double t) {
  const double* sp = s.sp;
  const double* p = s.tp;
  double fmax = 1.0;
  auto hat = [](double x, double x_min, double x_max) -> double {
    // util::function::hatFunction, src/util/utilFunction.cpp:14-30
    if (def_greater(x, x_min - 1.0E-12) && def_less(x, x_max + 1.0E-12)) {
      double x_mid = 0.5 * (x_min + x_max);
      double l = x_mid - x_min;
      if (l < 1.0E-12) return 1.0;
      if (def_less(x, x_mid)) return (x - x_min) / l;
      return (x_max - x) / l;
    }
    return 0.0;
  };
  switch (s.space_fn) {
    case NLM_SFN_HAT_X:
      fmax = sp[0] * hat(px, s.x1, s.x2);
      break;
    case NLM_SFN_HAT_Y:
      fmax = sp[0] * hat(py, s.y1, s.y2);
      break;
    case NLM_SFN_HAT_Z:
      fmax = sp[0] * hat(pz, s.z1, s.z2);
      break;
    case NLM_SFN_SIN_X: {
      double a = M_PI * sp[0];
      fmax = sp[0] * sin(a * px);
    } break;
    case NLM_SFN_SIN_Y: {
      double a = M_PI * sp[0];
      fmax = sp[0] * sin(a * py);
    } break;
    case NLM_SFN_SIN_Z: {
      double a = M_PI * sp[0];
      fmax = sp[0] * sin(a * pz);
    } break;
    case NLM_SFN_LINEAR_X: {
      double a = sp[0];
      fmax = sp[0] * a * px;
    } break;
    case NLM_SFN_LINEAR_Y: {
      double a = sp[0];
      fmax = sp[0] * a * py;
    } break;
    case NLM_SFN_LINEAR_Z: {
      double a = sp[0];
      fmax = sp[0] * a * pz;
    } break;
    case NLM_SFN_LINE_LOAD:
      // fLoading.cpp:172-228: per loaded direction, the node's position in
      // whole mesh steps over the extent's whole number of steps (size_t
      // truncations as written; the reference's swap for params[0] == -1
      // only shadows min and max, so the extent keeps its order)
      for (int d = 0; d < 3; d++) {
        if (!((s.dirmask >> d) & 1)) continue;
        const double mn = d == 0 ? s.x1 : (d == 1 ? s.y1 : s.z1);
        const double mx = d == 0 ? s.x2 : (d == 1 ? s.y2 : s.z2);
        const double xc = d == 0 ? px : (d == 1 ? py : pz);
        const uint64_t nodes = uint64_t(fabs(mx - mn) / s.h);
        const double scale = 1. / double(nodes);
        uint64_t pos = uint64_t((xc - mn) / s.h);
        if (s.n_space_params > 0 && sp[0] == -1) pos = uint64_t((mx - xc) / s.h);
        fmax *= double(pos) * scale;
      }
      break;
    case NLM_SFN_CONSTANT:
      fmax = s.n_space_params > 0 ? sp[0] : 1.0;
      break;
    default:
      break;
  }
  if (s.time_fn == NLM_TFN_LINEAR) {
    fmax *= t;
  } else if (s.time_fn == NLM_TFN_LINEAR_STEP) {
    double period = floor(t / (p[1] + p[2]));
    double ls = def_less(t, period * (p[1] + p[2]) + p[1])
                    ? t - period * p[2]
                    : (period + 1.) * p[1];
    fmax *= ls;
  } else if (s.time_fn == NLM_TFN_LINEAR_SLOW_FAST) {
    if (def_greater(t, p[1]))
      fmax *= p[3] * t;
    else
      fmax *= p[2] * t;
  } else if (s.time_fn == NLM_TFN_SIN) {
    double a = M_PI * p[1];
    fmax *= sin(a * t);
  }
  fmax *= p[0];
  return fmax;
}

}  // namespace nlm
