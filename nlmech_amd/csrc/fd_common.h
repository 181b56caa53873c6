// fd_common.h -- kernel arguments and node-level helpers shared by the
// explicit-dynamics kernels of both materials (fd.hip, es.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "nlm_internal.h"

namespace nlm {

template <int DIM>
struct FdArgs {
  using V = typename VT<DIM>::T;
  int64_t n;
  int64_t nslices;
  const V* __restrict__ x;
  const double* __restrict__ vol;
  const uint32_t* __restrict__ flags;
  const int64_t* __restrict__ sl_ptr;
  const int32_t* __restrict__ col;
  const uint16_t* __restrict__ rowlen;
  unsigned long long* __restrict__ bits;
  const V* __restrict__ u;
  const V* __restrict__ v;
  V* __restrict__ u_out;
  V* __restrict__ v_out;
  const V* __restrict__ f_in;
  V* __restrict__ f_out;
  const BcTable* __restrict__ bc;
  const int64_t* __restrict__ halo_ptr;
  const int32_t* __restrict__ halo;
  const uint16_t* __restrict__ colL;
  // per 256-node block: origin of the FP32 block-relative coordinates and
  // the node whose displacement is the displacement origin
  const unsigned long long* __restrict__ vexw;
  const double2* __restrict__ blk_org;
  int64_t halo_cap;  // LDS entries per halo array
  RnpDev m;
  double dt;
  double fact;
  int has_fbc;
  int has_nofail;
  const double* __restrict__ node_rate;  // node-list BC rates (or null)
  const double* __restrict__ tdev;       // step times from device memory
};

// Displacement and velocity of a prescribed dof at time t: ULoading::apply
// for BC sets, u = rate t for node-list BCs.
template <int DIM>
__device__ __forceinline__ void ubc_dof(const FdArgs<DIM>& a, uint32_t id,
                                        int64_t i, int d, const double* xi,
                                        double t, double& du, double& dv) {
  if (id == kNodeListBc) {
    const double r = a.node_rate[3 * i + d];
    du = r * t;
    dv = r;
  } else {
    ubc_eval(a.bc->u[id - 1], xi[0], xi[1], xi[2], t, du, dv);
  }
}

template <int DIM>
__device__ __forceinline__ void load3(const typename VT<DIM>::T& a,
                                      double* o) {
  o[0] = a.x;
  o[1] = a.y;
  o[2] = 0.;
  if constexpr (DIM == 3) o[2] = a.z;
}

template <int DIM>
__device__ __forceinline__ typename VT<DIM>::T make_v(const double* o) {
  typename VT<DIM>::T r;
  r.x = o[0];
  r.y = o[1];
  if constexpr (DIM == 3) {
    r.z = o[2];
    r.w = 0.;
  }
  return r;
}

// External force density of node i at time t (FLoading::apply order: sets in
// order, each adding to every dof it loads).
template <int DIM>
__device__ __forceinline__ void fext_node(const BcTable* bc, uint32_t fl,
                                          const double* xi, double t,
                                          double* b) {
  uint32_t mem = (fl >> kFlagFbcShift) & 0xffu;
  for (int s = 0; s < bc->nf; s++) {
    if (!((mem >> s) & 1u)) continue;
    const BcSetDev& set = bc->f[s];
    double val = fbc_eval(set, xi[0], xi[1], xi[2], t);
    for (int d = 0; d < DIM; d++)
      if ((set.dirmask >> d) & 1) b[d] += val;
  }
}

// KIND_FORCE: f_out = base + internal force (base = f_in or f_ext(t_force)).
// KIND_CD: fused force at t_force + central-difference update + BCs at t_next.
constexpr int KIND_FORCE = 0;
constexpr int KIND_CD = 1;
constexpr int KIND_VERLET = 2;

// Fourth dword of a raw buffer resource on gfx9 targets (32-bit untyped
// accesses, range-checked against the record count).
constexpr int kBufferRsrcWord3 = 0x00020000;

// Element idx of a device array addressed as base + 32-bit byte offset, so
// the load uses the scalar-base + vector-offset form (no 64-bit address
// arithmetic per lane).
template <class T>
__device__ __forceinline__ T ld32(const T* __restrict__ base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     idx * uint32_t(sizeof(T)));
}

// Central-difference tail of a fused step for node i: f is the total force
// density (external + internal).  KIND_FORCE stores f; KIND_CD performs
// FDModel::integrateCD's update on free dofs (fDModel.cpp:386-394) into the
// other u/v buffer and ULoading::apply at t_next on fixed dofs
// (uLoading.cpp:166-253).
template <int DIM, int KIND>
__device__ __forceinline__ void node_tail(const FdArgs<DIM>& a, int64_t i,
                                          uint32_t fl, const double* xi,
                                          const double* ui, const double* f,
                                          double t_next) {
  if (KIND == KIND_FORCE) {
    a.f_out[i] = make_v<DIM>(f);
    return;
  }
  double vi[3];
  load3<DIM>(a.v[i], vi);
  double un[3] = {0., 0., 0.}, vn[3] = {0., 0., 0.};
  for (int d = 0; d < DIM; d++) {
    uint32_t id = (fl >> (kFlagUbcShift + 4 * d)) & 0xfu;
    if (id == 0) {
      double uo = ui[d];
      un[d] = uo + (a.fact * f[d] + a.dt * vi[d]);
      vn[d] = (un[d] - uo) / a.dt;
    } else {
      double du, dv;
      ubc_dof(a, id, i, d, xi, t_next, du, dv);
      un[d] = du;
      vn[d] = dv;
    }
  }
  a.u_out[i] = make_v<DIM>(un);
  a.v_out[i] = make_v<DIM>(vn);
}

// Kernel arguments of the current model state (host side).
template <int DIM>
inline FdArgs<DIM> make_args(nlm_ctx* c) {
  using V = typename VT<DIM>::T;
  FdState& s = c->fd;
  FdArgs<DIM> a;
  a.n = s.n;
  a.nslices = s.sell.nslices;
  a.x = s.x.as<V>();
  a.vol = s.vol.as<double>();
  a.flags = s.flags.as<uint32_t>();
  a.sl_ptr = s.sell.sl_ptr.as<int64_t>();
  a.col = s.sell.col.as<int32_t>();
  a.rowlen = s.sell.rowlen.as<uint16_t>();
  a.bits = s.sell.bits.as<unsigned long long>();
  a.u = s.u[s.cur].as<V>();
  a.v = s.v[s.cur].as<V>();
  a.u_out = s.u[1 - s.cur].as<V>();
  a.v_out = s.v[1 - s.cur].as<V>();
  a.f_in = s.f.as<V>();
  a.f_out = s.f.as<V>();
  a.bc = s.bctab.as<BcTable>();
  a.halo_ptr = s.halo_ptr.as<int64_t>();
  a.halo = s.halo.as<int32_t>();
  a.colL = s.colL.as<uint16_t>();
  a.vexw = s.vexw.as<unsigned long long>();
  a.blk_org = s.blk_org.as<double2>();
  a.halo_cap = (s.halo_max + 1) & ~int64_t(1);
  a.m = s.rnp;
  a.dt = s.dt;
  a.fact = s.scheme == NLM_TIME_CENTRAL_DIFFERENCE
               ? s.dt * s.dt / s.density
               : 0.5 * s.dt / s.density;
  a.has_fbc = s.has_fbc ? 1 : 0;
  a.has_nofail = s.has_nofail ? 1 : 0;
  a.node_rate = s.node_rate.as<double>();
  a.tdev = nullptr;
  return a;
}

inline int64_t force_blocks(const FdState& s) {
  return (s.sell.nslices * 64 + 255) / 256;
}

// Force of the configured material at the current u (fd.hip).
void fd_launch_force(nlm_ctx* c, bool base_f, double t);

// State-based material (es.hip).
int es_prepare(nlm_ctx* c);
// LDS bytes of a staged state-based launch for a halo of halo_max nodes
size_t es_halo_lds_bytes(int precision, int64_t halo_max);
void es_launch_force(nlm_ctx* c, bool base_f, double t);
void es_launch_fused(nlm_ctx* c, double tf, double tn);
// fused pass pair reading the step times from tdev (graph capture)
void es_launch_fused_dev(nlm_ctx* c, const double* tdev);
// us3 (optional): microseconds per step of the pass pair with force-pass
// chunk width 2, 4, 8
int es_autotune(nlm_ctx* c, int* chosen, double* best_us,
                double* us3 = nullptr);

}  // namespace nlm
