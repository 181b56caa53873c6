// fd.hip -- explicit peridynamics on the GPU: FDModel<RNPBond>'s force
// computation and time integration (src/model/fd/fDModel.cpp).
//
// Per time step one fused kernel does, for every node i (one lane per node,
// one wavefront per 64-node slice of the neighbor lists):
//   - RNPBond::getBondEF over the row of i in list order (rnpBond.cpp:216-301)
//     including the break test and the bond-bit update, summed sequentially
//     like FDModel::computeForce (fDModel.cpp:544-582);
//   - the force BCs of FLoading::apply at the force's time
//     (fLoading.cpp:117-279), f_i = f_ext + sum (fDModel.cpp:536-537);
//   - the central-difference update of FDModel::integrateCD
//     (fDModel.cpp:377-425) into the other u/v buffer, and the displacement BCs
//     of ULoading::apply at t+dt (uLoading.cpp:166-253) on fixed dofs.
// The neighbor displacements are read from the current buffer only, so the
// update of node i never races with the force of another node.
//
// Bond bits: slot k of the 64 nodes of a slice is one 64-bit word (bit =
// lane); a wavefront reads it once, and when bonds break it ORs the ballot of
// newly broken lanes into the word with one vector store from lane 0.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "bond_es.h"
#include "bond_rnp.h"
#include "device_common.h"
#include "fd_common.h"
#include "layout.h"
#include "nlm_internal.h"

// Staging of nlm_fd_postproc_noalloc, the entry point that allocates
// nothing: Z (FP32), phi (FP32) and strain energy (FP64) of up to
// kNoAllocMaxNodes nodes, 16 bytes a node (16 MiB).  Each GPU the library
// runs on holds its own copy until the process ends.  In host code the
// array's name is not a device address: the per-device address comes from
// hipGetSymbolAddress (nlm::scratch_setup) and only that address is used.
static __device__ __align__(256) unsigned char scratch[nlm::kNoAllocBytes];

namespace nlm {

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int scratch_setup(int device, void** out, std::string& err) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) {
    err = "scratch_setup: device index beyond the library's table";
    return NLM_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    // the device is current (nlm_ctx_create set it)
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") +
            hipGetErrorString(e);
      return NLM_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return NLM_OK;
}

// Quantum of the state-based FP32-force passes' fixed-point positions and
// the lengths in its units: a power of two with every offset of a block's
// node or halo node from the block's origin (at most org_range) below 2^29 q,
// so the difference of two (at most 2^30 q) is an exact int32 and the
// quantization (q/2) is ~1e-9 of the range.
void es_set_quantum(EsDev& e, double org_range) {
  const double rng = std::max(org_range, e.h);
  e.q = std::ldexp(1.0, std::ilogb(rng) + 1 - 29);
  e.inv_q = 1.0 / e.q;
  e.q_f = float(e.q);
  e.q2_f = float(e.q * e.q);
  e.check_up_qf = float(e.check_up * e.inv_q);
  e.low_band_qf = float(e.low_band * e.inv_q);
  e.rcp_h_qf = float(e.q / e.h);
  e.up_rcp_h_qf = float(e.check_up / e.h);
  e.vex_tol_qf = 3e-5f * e.low_band_qf;
}

namespace {

template <int DIM, int MODE, int INF>
__device__ __forceinline__ void bond_mode(const RnpDev& m, const double* xi,
                                          const double* ui, const double* xj,
                                          const double* uj, double voli,
                                          double volj, bool brk, bool& fs,
                                          double* F) {
  if constexpr (MODE == kModeExact)
    bond_exact<DIM, false>(m, xi, ui, xj, uj, voli, volj, brk, fs, F, nullptr);
  else if constexpr (MODE == kModeFp64)
    bond_fp64<DIM, INF>(m, xi, ui, xj, uj, voli, volj, brk, fs, F);
  else
    bond_fp32<DIM, INF>(m, xi, ui, xj, uj, voli, volj, brk, fs, F);
}

// One halo node staged in LDS by the tiled step kernel (2D): reference
// position, displacement, volume and flags in one 48-byte record, so a
// neighbor costs one address computation and three LDS reads.
struct alignas(16) HaloRec {
  double2 x;
  double2 u;
  double vol;
  uint32_t flags;
  float cv;  // FP32-force mode: vol * c4 (* J when constant), in FP32
};

// The FP32-force mode's record (40 bytes: a 10-dword stride puts record
// starts on 32 of the 64 LDS banks where an 8-dword stride reaches 8): the
// displacement in FP64 and the position relative to the block's origin in
// FP32.  A bond's displacement
// difference is formed in FP64 and rounded once to FP32, so its error is
// relative to the difference itself however large the displacements are
// (cancellation-free).  Positions within a block and its halo lie within
// about (16 + 2 horizon/h) h of the origin, so the relative coordinates
// carry FP32 rounding of that extent: a relative error of the bond's
// position difference of order 1e-6, far inside the guard bands of the
// break and volume-correction decisions (which the exact path re-takes).
// 8-byte aligned pair of doubles (the record's FP64 displacement)
struct alignas(8) Dbl2 {
  double x, y;
  __device__ Dbl2& operator=(const double2& v) {
    x = v.x;
    y = v.y;
    return *this;
  }
};
struct alignas(8) HaloRec32 {
  Dbl2 u;
  float2 xr;
  float cv;  // vol c4 (J)
  uint32_t flags;
  uint32_t pad[2];
};
static_assert(sizeof(HaloRec32) == 40, "FP32-force record");

template <int MODE>
using HaloRecOf =
    typename std::conditional<MODE == kModeFp32, HaloRec32, HaloRec>::type;

template <int DIM, int MODE, int KIND, bool BASE_F, int INF, int U,
          bool NOFAIL, bool TILED>
__device__ __forceinline__ void fd_step_body(
    FdArgs<DIM> a, const unsigned long long* __restrict__ bits_ro,
    const unsigned long long* __restrict__ vex_ro, double t_force,
    double t_next) {
  if (a.tdev) {  // graph-captured steps read the step's times
    t_force = a.tdev[0];
    t_next = a.tdev[1];
  }
  // TILED: the workgroup's nodes (its size: 128, 256 or 512, the tuned
  // tile) are one block of the layout's halos; the halo (every node its rows
  // reference) is staged in LDS before the bond loop and neighbors are read
  // through halo-local entries (LDS byte offsets).
  const int bn = TILED ? int(blockDim.x) : 256;
  int64_t nblk = (a.nslices * 64 + bn - 1) / bn;
  int64_t lb = xcd_block(nblk);
  if (lb >= nblk) return;
  static_assert(!TILED || DIM == 2, "the LDS-staged kernel is 2D");
  constexpr bool REL = TILED && MODE == kModeFp32;  // block-relative FP32
  typedef HaloRecOf<MODE> Rec;
  extern __shared__ HaloRec smem_halo[];
  Rec* const halo_lds = reinterpret_cast<Rec*>(smem_halo);
  double2 org = make_double2(0., 0.);
  int64_t hb = 0;
  if constexpr (TILED) {
    if constexpr (REL) org = a.blk_org[lb];
    hb = a.halo_ptr[lb];
    const int hn = int(a.halo_ptr[lb + 1] - hb);
    if constexpr (REL) {
      // record 0: zero force constant, far from every node (its bonds come
      // out as exactly zero forces, never near a decision)
      if (threadIdx.x == 0) {
        HaloRec32 r0;
        r0.u = make_double2(0., 0.);
        r0.xr = make_float2(1e15f, 1e15f);
        r0.cv = 0.f;
        r0.flags = 0u;
        r0.pad[0] = r0.pad[1] = 0u;
        halo_lds[0] = r0;
      }
    }
    for (int t = threadIdx.x; t < hn; t += bn) {
      const uint32_t g = uint32_t(a.halo[hb + t]);
      Rec r;
      const double2 xg = ld32(a.x, g), ug = ld32(a.u, g);
      const double vg = ld32(a.vol, g);
      r.flags = NOFAIL ? ld32(a.flags, g) : 0u;
      if constexpr (REL) {
        // sigma (x - origin), tau u, vol c4 (J) sigma / tau
        // (bond_fp32_scaled)
        r.cv = float(vg) * a.m.c4f_infl_s;
        r.xr = make_float2(float((xg.x - org.x) * a.m.sigma),
                           float((xg.y - org.y) * a.m.sigma));
        r.u = make_double2(ug.x * a.m.tau, ug.y * a.m.tau);
      } else {
        r.cv = float(vg) * a.m.c4f_infl;
        r.x = xg;
        r.u = ug;
        r.vol = vg;
      }
      halo_lds[REL ? t + 1 : t] = r;
    }
    __syncthreads();
  }
  // The slice index is the same for the 64 lanes of a wavefront; reading it
  // from the first lane makes the slice bounds, the slot loop and the
  // bond-bit addresses scalar.
  const int lane = threadIdx.x & 63;
  const int64_t s =
      lb * (bn >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (s >= a.nslices) return;  // whole wavefront
  const int64_t i = s * 64 + lane;
  bool valid = i < a.n;
  const int64_t cb = a.sl_ptr[s];
  const int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  const int64_t bb = cb >> 6;
  int len = valid ? int(a.rowlen[i]) : 0;
  double xi[3] = {0., 0., 0.}, ui[3] = {0., 0., 0.};
  double voli = 1.;
  uint32_t fl = 0;
  if (valid) {
    load3<DIM>(a.x[i], xi);
    load3<DIM>(a.u[i], ui);
    voli = a.vol[i];
    fl = a.flags[i];
  }
  bool int_i = !(fl & kFlagNonInterior);
  // the node's scaled block-relative FP32 position and scaled displacement
  float xir[2] = {0.f, 0.f};
  double uis[2] = {0., 0.};
  if constexpr (REL) {
    xir[0] = float((xi[0] - org.x) * a.m.sigma);
    xir[1] = float((xi[1] - org.y) * a.m.sigma);
    uis[0] = ui[0] * a.m.tau;
    uis[1] = ui[1] * a.m.tau;
  }
  // this slice's neighbor slots: slot k of this lane at entry 64 k + lane
  const int32_t* __restrict__ colS = a.col + cb;
  const uint16_t* __restrict__ colL = a.colL + cb;
  // bond-bit words of this slice, read through the scalar path (each word is
  // read once, before this wavefront may rewrite it).  A chunk may read up
  // to U - 1 words past the slice (the next slice's, or the array's
  // padding); those slots are inactive.
  const unsigned long long* __restrict__ bits_in = bits_ro + bb;
  // static per-slot masks of the bonds near the volume-correction edge
  // (padded like the bond bits)
  const unsigned long long* __restrict__ vex_in = vex_ro + bb;
  double acc[3] = {0., 0., 0.};
  // Neighbor indices are fetched one chunk of U slots ahead of the chunk
  // whose positions and displacements are gathered, so the index latency
  // overlaps the previous chunk's arithmetic.  Every load is unconditional:
  // index fetches past the slice return 0 (a valid halo entry or node), and
  // slots at or past the row length are masked out of the arithmetic (in the
  // staged FP32-force kernel entry 0 and the padded entries address the
  // zero-weight record 0 instead, which needs no mask).  A
  // slot's 64 entries start at a wavefront-uniform address; the lane offset
  // is the only per-lane part.
  // The indices are fetched through a buffer resource over the slice: the
  // slot offset is a scalar and the lane offset a per-lane constant, so a
  // fetch costs no 64-bit vector address arithmetic.
  const __amdgpu_buffer_rsrc_t col_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      TILED ? (void*)colL : (void*)colS, 0,
      width * 64 * int(TILED ? sizeof(uint16_t) : sizeof(int32_t)),
      kBufferRsrcWord3);
  auto slot_index = [&](int k) -> uint32_t {
    if constexpr (TILED)
      return uint32_t(__builtin_amdgcn_raw_buffer_load_b16(
          col_rsrc, lane * int(sizeof(uint16_t)),
          k * 64 * int(sizeof(uint16_t)), 0));
    else
      return uint32_t(__builtin_amdgcn_raw_buffer_load_b32(
          col_rsrc, lane * int(sizeof(int32_t)),
          k * 64 * int(sizeof(int32_t)), 0));
  };
  // wavefront masks that do not change along the row
  const unsigned long long int_m =
      NOFAIL ? __builtin_amdgcn_ballot_w64(int_i) : ~0ull;
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
    float xrj[U][2], duj[U][2];
    float cvj[U];
    unsigned long long w[U], vx[U];
    uint32_t flj[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t j = jc[q];
      if constexpr (REL) {
        // The slot entry is the record's LDS address: the kernel has no
        // static LDS, so the dynamic area starts at address 0 (the host
        // checks the kernel's static LDS size before it selects this path).
        typedef const __attribute__((address_space(3))) HaloRec32 LdsRec;
        LdsRec* r = (LdsRec*)j;
        xrj[q][0] = r->xr.x;
        xrj[q][1] = r->xr.y;
        // scaled displacement difference in FP64 (tau a power of two: exact),
        // rounded once
        duj[q][0] = float(r->u.x - uis[0]);
        duj[q][1] = float(r->u.y - uis[1]);
        cvj[q] = r->cv;
        if (NOFAIL) flj[q] = r->flags;
      } else if constexpr (TILED) {
        typedef const __attribute__((address_space(3))) HaloRec LdsRec;
        LdsRec* r = (LdsRec*)j;
        xj[q][0] = r->x.x;
        xj[q][1] = r->x.y;
        xj[q][2] = 0.;
        uj[q][0] = r->u.x;
        uj[q][1] = r->u.y;
        uj[q][2] = 0.;
        vj[q] = r->vol;
        cvj[q] = r->cv;
        if (NOFAIL) flj[q] = r->flags;
      } else {
        load3<DIM>(ld32(a.x, j), xj[q]);
        load3<DIM>(ld32(a.u, j), uj[q]);
        vj[q] = ld32(a.vol, j);
        cvj[q] = float(vj[q]) * a.m.c4f_infl;
        if (NOFAIL) flj[q] = ld32(a.flags, j);
      }
      w[q] = bits_in[k0 + q];
      vx[q] = MODE == kModeFp32 ? 0ull : vex_in[k0 + q];
    }
    // newly broken bonds of each slot (wavefront masks)
    unsigned long long nb[U];
    if constexpr (MODE == kModeExact) {
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const bool act = k0 + q < len;
        bool brk = NOFAIL ? (int_i && !(flj[q] & kFlagNonInterior)) : true;
        // w[q] is wavefront-uniform: its bits are the lanes' broken flags
        bool fs0 = __builtin_amdgcn_inverse_ballot_w64(w[q]);
        bool fs = fs0;
        double F[3];
        bond_mode<DIM, MODE, INF>(a.m, xi, ui, xj[q], uj[q], voli, vj[q], brk,
                                  fs, F);
        acc[0] += act ? F[0] : 0.0;
        acc[1] += act ? F[1] : 0.0;
        if (DIM == 3) acc[2] += act ? F[2] : 0.0;
        nb[q] = __builtin_amdgcn_ballot_w64(act && fs && !fs0);
      }
    } else {
      // The arithmetic runs on every lane with no branch; its decisions are
      // wavefront masks (votes on single comparisons).  Bonds with a
      // decision inside a guard band, or a contact force, are left out of
      // the common path and evaluated after the chunk.  Forces are summed in
      // pairs of slots (2m, 2m+1) -- in FP32 in the FP32-force mode -- and
      // the pairs are added to the FP64 accumulator in order (the staged
      // FP32-force kernel fuses the odd slot's force into its pair,
      // fma(s, dx, F_2m); a deferred evaluation is added to its pair after
      // the chunk); a pair never spans two chunks, so the chunk width does
      // not change the result.
      typedef typename std::conditional<MODE == kModeFp32, float, double>::type
          Real;
      constexpr int P = (U + 1) / 2;
      unsigned long long band[U], ex[U];
      unsigned long long ex_any = 0ull;
      Real pf[P][3];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        SlotMasks km;
        // (staged FP32-force: slots past a row's length address the zero-
        // weight record 0, whose force is exactly zero)
        km.act = REL ? ~0ull : __builtin_amdgcn_ballot_w64(k0 + q < len);
        km.brk = NOFAIL ? (int_m & __builtin_amdgcn_ballot_w64(
                                       (flj[q] & kFlagNonInterior) == 0u))
                        : ~0ull;
        km.w = w[q];
        km.vex = vx[q];
        Real F[3];
        if constexpr (REL) {
          // written into (even slot) or fused into (odd slot) the pair sum
          if (q % 2 == 1)
            bond_fp32_scaled<DIM, INF, NOFAIL, true>(
                a.m, xrj[q][0] - xir[0], xrj[q][1] - xir[1], 0.f, duj[q][0],
                duj[q][1], 0.f, cvj[q], km, nb[q], band[q], ex[q], pf[q / 2]);
          else
            bond_fp32_scaled<DIM, INF, NOFAIL, false>(
                a.m, xrj[q][0] - xir[0], xrj[q][1] - xir[1], 0.f, duj[q][0],
                duj[q][1], 0.f, cvj[q], km, nb[q], band[q], ex[q], pf[q / 2]);
        } else if constexpr (MODE == kModeFp64)
          bond_fp64_fast<DIM, INF, NOFAIL>(a.m, xi, ui, xj[q], uj[q], vj[q],
                                           km, nb[q], band[q], ex[q], F);
        else
          bond_fp32_fast<DIM, INF, NOFAIL>(a.m, xi, ui, xj[q], uj[q], cvj[q],
                                           km, nb[q], band[q], ex[q], F);
        if constexpr (REL) {
        } else if (q % 2 == 0) {
          pf[q / 2][0] = F[0];
          pf[q / 2][1] = F[1];
          pf[q / 2][2] = F[2];
        } else {
          pf[q / 2][0] += F[0];
          pf[q / 2][1] += F[1];
          if (DIM == 3) pf[q / 2][2] += F[2];
        }
        ex_any |= ex[q];
      }
      if (ex_any != 0ull) {  // rare: whole wavefronts skip it
#pragma unroll
        for (int q = 0; q < U; ++q) {
          if (ex[q] == 0ull) continue;
          bool fsn = false;
          // (lanes whose slot is past their row, in the staged FP32-force
          // kernel, evaluate nothing)
          if (__builtin_amdgcn_inverse_ballot_w64(ex[q]) &&
              (!REL || k0 + q < len)) {
            bool brk = NOFAIL ? (int_i && !(flj[q] & kFlagNonInterior)) : true;
            bool fs0 = __builtin_amdgcn_inverse_ballot_w64(w[q]);
            bool fs = fs0;
            double Fd[3];
            double xs[3], us[3], vs;
            if constexpr (REL) {
              // FP64 values of the neighbor from its halo entry
              const uint32_t g = uint32_t(
                  a.halo[hb + jc[q] / uint32_t(sizeof(HaloRec32)) - 1]);
              load3<DIM>(ld32(a.x, g), xs);
              load3<DIM>(ld32(a.u, g), us);
              vs = ld32(a.vol, g);
            } else if constexpr (TILED) {
              // the neighbor's record is re-read from LDS (volatile, so the
              // common path need not keep its FP64 values alive)
              typedef volatile const __attribute__((address_space(3))) HaloRec
                  LdsRecV;
              LdsRecV* r = (LdsRecV*)jc[q];
              xs[0] = r->x.x;
              xs[1] = r->x.y;
              xs[2] = 0.;
              us[0] = r->u.x;
              us[1] = r->u.y;
              us[2] = 0.;
              vs = r->vol;
            } else {
              // re-read from global memory (volatile, so the common path
              // need not keep the FP64 values alive)
              const uint32_t g = jc[q];
              const volatile double* px = (const volatile double*)(a.x + g);
              const volatile double* pu = (const volatile double*)(a.u + g);
              for (int d = 0; d < 3; d++) {
                xs[d] = d < DIM ? px[d] : 0.;
                us[d] = d < DIM ? pu[d] : 0.;
              }
              vs = *(const volatile double*)(a.vol + g);
            }
            bond_mode<DIM, MODE, INF>(a.m, xi, ui, xs, us, voli, vs, brk, fs,
                                      Fd);
            fsn = fs && !fs0;
            // (in the FP32-force mode a bond force comes out in FP32, so its
            // conversion is exact; a contact force of a broken bond is
            // computed in FP64 and rounded to FP32 here, like every bond
            // force of that mode)
            pf[q / 2][0] += Real(Fd[0]);
            pf[q / 2][1] += Real(Fd[1]);
            if (DIM == 3) pf[q / 2][2] += Real(Fd[2]);
          }
          // the guard band's bonds were decided exactly here
          nb[q] |= band[q] & __builtin_amdgcn_ballot_w64(fsn);
        }
      }
#pragma unroll
      for (int m = 0; m < P; ++m) {
        acc[0] += double(pf[m][0]);
        acc[1] += double(pf[m][1]);
        if (DIM == 3) acc[2] += double(pf[m][2]);
      }
    }
    // new broken bits (rare): one scalar test per chunk; every lane stores
    // the same word
    unsigned long long nb_any = 0ull;
#pragma unroll
    for (int q = 0; q < U; ++q) nb_any |= nb[q];
    if (nb_any != 0ull) {
#pragma unroll
      for (int q = 0; q < U; ++q)
        if (nb[q] != 0ull && k0 + q < width)
          a.bits[bb + k0 + q] = w[q] | nb[q];
    }
  }
  if (!valid) return;
  double f[3] = {0., 0., 0.};
  if (BASE_F) {
    load3<DIM>(a.f_in[i], f);
  } else if (a.has_fbc) {
    fext_node<DIM>(a.bc, fl, xi, t_force, f);
  }
  for (int d = 0; d < DIM; d++) f[d] += acc[d];
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
      if (KIND == KIND_CD) {
        // u += fact*f + dt*v; v = (u - u_old)/dt  (fDModel.cpp:386-394)
        double uo = ui[d];
        un[d] = uo + (a.fact * f[d] + a.dt * vi[d]);
        vn[d] = (un[d] - uo) / a.dt;
      }
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

// The fused step kernel (one integrateCD iteration or a force evaluation),
// 256-node workgroups.
template <int DIM, int MODE, int KIND, bool BASE_F, int INF, int U,
          bool NOFAIL, bool TILED>
__global__ __launch_bounds__(256) void k_fd(
    FdArgs<DIM> a, const unsigned long long* __restrict__ bits_ro,
    const unsigned long long* __restrict__ vex_ro, double t_force,
    double t_next) {
  fd_step_body<DIM, MODE, KIND, BASE_F, INF, U, NOFAIL, TILED>(
      a, bits_ro, vex_ro, t_force, t_next);
}

// The staged kernel with tiles of up to 512 nodes (the workgroup size).
template <int DIM, int MODE, int KIND, bool BASE_F, int INF, int U,
          bool NOFAIL>
__global__ __launch_bounds__(512) void k_fd_t(
    FdArgs<DIM> a, const unsigned long long* __restrict__ bits_ro,
    const unsigned long long* __restrict__ vex_ro, double t_force,
    double t_next) {
  fd_step_body<DIM, MODE, KIND, BASE_F, INF, U, NOFAIL, true>(
      a, bits_ro, vex_ro, t_force, t_next);
}

// The staged kernel held to at least 6 wavefronts per SIMD (at most 80
// vector registers): the FP32-force staged kernel with two slots per chunk
// fits without spilling, and the compiler's own register choice can miss
// that occupancy step by a register.  (HIP's second launch-bounds argument
// is a minimum of wavefronts per SIMD, so the bound is stated directly.)
template <int DIM, int MODE, int KIND, bool BASE_F, int INF, int U,
          bool NOFAIL>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6)))
void k_fd_w6(
    FdArgs<DIM> a, const unsigned long long* __restrict__ bits_ro,
    const unsigned long long* __restrict__ vex_ro, double t_force,
    double t_next) {
  fd_step_body<DIM, MODE, KIND, BASE_F, INF, U, NOFAIL, true>(
      a, bits_ro, vex_ro, t_force, t_next);
}

// Update from a stored force (the first iteration of a stepping call):
// integrateCD's update loop and the BCs at t_next.
template <int DIM>
__global__ __launch_bounds__(256) void k_update_cd(FdArgs<DIM> a,
                                                   double t_next) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  double xi[3], ui[3], vi[3], f[3];
  load3<DIM>(a.x[i], xi);
  load3<DIM>(a.u[i], ui);
  load3<DIM>(a.v[i], vi);
  load3<DIM>(a.f_in[i], f);
  uint32_t fl = a.flags[i];
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

// Velocity-Verlet pieces (FDModel::integrateVerlet, fDModel.cpp:446-525).
// Step 1-2: v += fact f; u += dt v on free dofs; BCs at t_next.
template <int DIM>
__global__ __launch_bounds__(256) void k_verlet_a(FdArgs<DIM> a,
                                                  double t_next) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  double xi[3], ui[3], vi[3], f[3];
  load3<DIM>(a.x[i], xi);
  load3<DIM>(a.u[i], ui);
  load3<DIM>(a.v[i], vi);
  load3<DIM>(a.f_in[i], f);
  uint32_t fl = a.flags[i];
  for (int d = 0; d < DIM; d++) {
    uint32_t id = (fl >> (kFlagUbcShift + 4 * d)) & 0xfu;
    if (id == 0) {
      vi[d] += a.fact * f[d];
      ui[d] += a.dt * vi[d];
    } else {
      double du, dv;
      ubc_dof(a, id, i, d, xi, t_next, du, dv);
      ui[d] = du;
      vi[d] = dv;
    }
  }
  a.u_out[i] = make_v<DIM>(ui);
  a.v_out[i] = make_v<DIM>(vi);
}
// Step 3: v += fact f on free dofs (in place).
template <int DIM>
__global__ __launch_bounds__(256) void k_verlet_b(FdArgs<DIM> a) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  double vi[3], f[3];
  load3<DIM>(a.v[i], vi);
  load3<DIM>(a.f_in[i], f);
  uint32_t fl = a.flags[i];
  for (int d = 0; d < DIM; d++) {
    uint32_t id = (fl >> (kFlagUbcShift + 4 * d)) & 0xfu;
    if (id == 0) vi[d] += a.fact * f[d];
  }
  a.v_out[i] = make_v<DIM>(vi);
}

// Displacement BCs at time t on fixed dofs (ULoading::apply).
template <int DIM>
__global__ __launch_bounds__(256) void k_apply_ubc(FdArgs<DIM> a, double t) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  uint32_t fl = a.flags[i];
  if ((fl & 0xfffu) == 0) return;
  double xi[3], ui[3], vi[3];
  load3<DIM>(a.x[i], xi);
  load3<DIM>(a.u[i], ui);
  load3<DIM>(a.v[i], vi);
  for (int d = 0; d < DIM; d++) {
    uint32_t id = (fl >> (kFlagUbcShift + 4 * d)) & 0xfu;
    if (id) ubc_dof(a, id, i, d, xi, t, ui[d], vi[d]);
  }
  a.u_out[i] = make_v<DIM>(ui);
  a.v_out[i] = make_v<DIM>(vi);
}

// f_out = f_ext(t) (FLoading::apply after the force reset).
template <int DIM>
__global__ __launch_bounds__(256) void k_fext(FdArgs<DIM> a, double t,
                                              int add) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  double xi[3], f[3] = {0., 0., 0.};
  load3<DIM>(a.x[i], xi);
  if (add) load3<DIM>(a.f_in[i], f);
  fext_node<DIM>(a.bc, a.flags[i], xi, t, f);
  a.f_out[i] = make_v<DIM>(f);
}

// ---- edge cracks: Fracture::computeFracturedBondFd (fracture.cpp:74-160) --

__device__ __forceinline__ bool pt_outside(double px, double py, int o,
                                           const nlm_crack& c) {
  // EdgeCrack::ptOutside with its default theta = 0 (fractureDeck.h:115-147;
  // computeFracturedBondFd does not pass the crack angle)
  if (o == -1) return def_less(py, c.pb[1]) || def_greater(py, c.pt[1]);
  if (o == 1) return def_less(px, c.pb[0]) || def_greater(px, c.pt[0]);
  if (o == 0) {
    double th = 0.0;
    double qx = px - c.pb[0], qy = py - c.pb[1];
    double pmx = qx * cos(th) + qy * sin(th);
    double tx = c.pt[0] - c.pb[0], ty = c.pt[1] - c.pb[1];
    double ptx = tx * cos(th) + ty * sin(th);
    return def_less(pmx, 0.0) || def_greater(pmx, ptx);
  }
  return true;
}
__device__ __forceinline__ bool pt_left(double px, double py,
                                        const nlm_crack& c) {
  // EdgeCrack::ptLeftside, fractureDeck.h:156-174
  double ar = 0.5 * ((c.pt[0] - c.pb[0]) * (py - c.pb[1]) -
                     (px - c.pb[0]) * (c.pt[1] - c.pb[1]));
  return !def_less(ar, 0.0);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_precrack(FdArgs<DIM> a,
                                                  nlm_crack c) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  int lane = threadIdx.x & 63;
  int64_t s = i >> 6;
  if (s >= a.nslices) return;
  bool valid = i < a.n;
  int64_t cb = a.sl_ptr[s];
  int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  int64_t bb = cb >> 6;
  int len = valid ? int(a.rowlen[i]) : 0;
  double xi[3] = {0., 0., 0.};
  if (valid) load3<DIM>(a.x[i], xi);
  bool active = valid && !pt_outside(xi[0], xi[1], c.o, c);
  bool left = active && pt_left(xi[0], xi[1], c);
  for (int k = 0; k < width; ++k) {
    bool set = false;
    if (active && k < len) {
      int32_t j = a.col[cb + int64_t(k) * 64 + lane];
      double xj[3];
      load3<DIM>(a.x[j], xj);
      bool modify = !pt_outside(xj[0], xj[1], c.o, c);
      bool jl = pt_left(xj[0], xj[1], c);
      set = modify && (left ? !jl : jl);
    }
    unsigned long long m = __ballot(set);
    if (m != 0ull && lane == 0) a.bits[bb + k] |= m;
  }
}

// ---- bond-bit conversions ------------------------------------------------

template <int DIM>
__global__ __launch_bounds__(256) void k_bits_to_bytes(
    FdArgs<DIM> a, const int64_t* __restrict__ row_ptr,
    uint8_t* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  int lane = int(i & 63);
  int64_t s = i >> 6;
  int64_t bb = a.sl_ptr[s] >> 6;
  int len = a.rowlen[i];
  int64_t r0 = row_ptr[i];
  for (int k = 0; k < len; k++)
    out[r0 + k] = uint8_t((a.bits[bb + k] >> lane) & 1ull);
}

template <int DIM>
__global__ __launch_bounds__(64) void k_bytes_to_bits(
    FdArgs<DIM> a, const int64_t* __restrict__ row_ptr,
    const uint8_t* __restrict__ in) {
  int64_t s = blockIdx.x;
  if (s >= a.nslices) return;
  int lane = threadIdx.x;
  int64_t i = s * 64 + lane;
  bool valid = i < a.n;
  int64_t cb = a.sl_ptr[s];
  int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  int len = valid ? int(a.rowlen[i]) : 0;
  int64_t r0 = valid ? row_ptr[i] : 0;
  for (int k = 0; k < width; k++) {
    bool b = k < len && in[r0 + k] != 0;
    unsigned long long m = __ballot(b);
    if (lane == 0) a.bits[(cb >> 6) + k] = m;
  }
}

// ---- post-processing: computePostProcFields (fDModel.cpp:652-778) -------

template <int DIM>
__global__ __launch_bounds__(256) void k_postproc(
    FdArgs<DIM> a, float* __restrict__ Z, float* __restrict__ phi,
    double* __restrict__ energy, const int32_t* __restrict__ perm) {
  int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  int lane = threadIdx.x & 63;
  int64_t s = i >> 6;
  if (s >= a.nslices) return;
  bool valid = i < a.n;
  int64_t cb = a.sl_ptr[s];
  int width = int((a.sl_ptr[s + 1] - cb) >> 6);
  int64_t bb = cb >> 6;
  int len = valid ? int(a.rowlen[i]) : 0;
  double xi[3] = {0., 0., 0.}, ui[3] = {0., 0., 0.};
  double voli = 1.;
  uint32_t fl = 0;
  if (valid) {
    load3<DIM>(a.x[i], xi);
    load3<DIM>(a.u[i], ui);
    voli = a.vol[i];
    fl = a.flags[i];
  }
  bool int_i = !(fl & kFlagNonInterior);
  const RnpDev& m = a.m;
  double energy_i = 0., sa = 0., sb = 0., z = 0.;
  for (int k = 0; k < width; ++k) {
    unsigned long long w = a.bits[bb + k];
    bool newly = false;
    if (k < len) {
      int32_t j = a.col[cb + int64_t(k) * 64 + lane];
      double xj[3], uj[3];
      load3<DIM>(a.x[j], xj);
      load3<DIM>(a.u[j], uj);
      double volj0 = a.vol[j];
      bool brk = int_i;
      if (a.has_nofail) brk = brk && !(a.flags[j] & kFlagNonInterior);
      bool fs0 = (w >> lane) & 1ull;
      bool fs = fs0;
      double F[3], E;
      bond_exact<DIM, true>(m, xi, ui, xj, uj, voli, volj0, brk, fs, F, &E);
      newly = fs && !fs0;
      energy_i += E;
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
      double volj = volj0;
      if (def_greater(rji, m.check_low)) volj *= (m.check_up - rji) / m.h;
      if (!fs) sa += volj;
      sb += volj;
      double sr = 0.;
      if (def_greater(rji, 1.0E-12)) sr = fabs(S) / (m.rbar / sqrt(rji));
      if (def_less(z, sr)) z = sr;
    }
    unsigned long long nb = __ballot(newly);
    if (nb != 0ull && lane == 0) a.bits[bb + k] = w | nb;
  }
  if (!valid) return;
  // perm (optional): outputs at the node's reference id instead of its
  // internal index
  const int64_t o = perm ? int64_t(perm[i]) : i;
  if (Z) Z[o] = float(z);
  if (phi) phi[o] = float(1. - sa / sb);
  if (energy) energy[o] = energy_i * voli;
}

// External force density of every node at time t: FLoading::apply on a zero
// force vector (fLoading.cpp:117-279), as FDModel::computePostProcFields
// takes it for the work done (fDModel.cpp:655-662).  Internal order, 3
// components per node.
template <int DIM>
__global__ __launch_bounds__(256) void k_fext(FdArgs<DIM> a, double t,
                                              double* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  double xi[3] = {0., 0., 0.}, f[3] = {0., 0., 0.};
  load3<DIM>(a.x[i], xi);
  if (a.has_fbc) fext_node<DIM>(a.bc, a.flags[i], xi, t, f);
  out[3 * i] = f[0];
  out[3 * i + 1] = f[1];
  out[3 * i + 2] = f[2];
}

// util::geometry::direction (utilGeom.cpp:317-326) as the reference writes
// it: the orientation determinant is stored in an int (truncation toward
// zero), so only determinants of magnitude >= 1 count as turns.
__device__ __forceinline__ int ref_direction(double ax, double ay, double bx,
                                             double by, double cx, double cy) {
  const int val = int((by - ay) * (cx - bx) - (bx - ax) * (cy - by));
  if (val == 0) return 0;
  return val < 0 ? 2 : 1;
}

// util::geometry::onLine (utilGeom.cpp:308-315), as written (C below or at
// both the maximum and the minimum of each coordinate).
__device__ __forceinline__ bool ref_on_line(double ax, double ay, double bx,
                                            double by, double cx, double cy) {
  return cx <= fmax(ax, bx) && cx <= fmin(ax, bx) && cy <= fmax(ay, by) &&
         cy <= fmin(ay, by);
}

// util::geometry::doLinesIntersect (utilGeom.cpp:281-306).
__device__ __forceinline__ bool ref_lines_intersect(double ax, double ay,
                                                    double bx, double by,
                                                    double cx, double cy,
                                                    double dx, double dy) {
  const int d1 = ref_direction(ax, ay, bx, by, cx, cy);
  const int d2 = ref_direction(ax, ay, bx, by, dx, dy);
  const int d3 = ref_direction(cx, cy, dx, dy, ax, ay);
  const int d4 = ref_direction(cx, cy, dx, dy, bx, by);
  if (d1 != d2 && d3 != d4) return true;
  if (d1 == 0 && ref_on_line(ax, ay, bx, by, cx, cy)) return true;
  if (d2 == 0 && ref_on_line(ax, ay, bx, by, dx, dy)) return true;
  if (d3 == 0 && ref_on_line(cx, cy, dx, dy, ax, ay)) return true;
  if (d4 == 0 && ref_on_line(cx, cy, dx, dy, bx, by)) return true;
  return false;
}

// Reaction force of FDModel::computeForce (fDModel.cpp:548-580): for a node
// in the layer max_y - 1.5 delta < y < max_y - delta, the volume-weighted
// forces of its bonds that cross the line y = max_y - delta + 1e-6 between
// min_x and max_x (FDModel::is_reaction_force, fDModel.cpp:585-605), at the
// current state (the bond bits are read, not changed).  R: internal order,
// 3 components per node; R_len: its length (Total_Reaction_Force per node).
template <int DIM>
__global__ __launch_bounds__(256) void k_reaction(
    FdArgs<DIM> a, double min_x, double max_x, double max_y, double delta,
    double* __restrict__ R, double* __restrict__ R_len) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int64_t s = i >> 6;
  if (s >= a.nslices) return;
  if (i >= a.n) return;
  const int64_t cb = a.sl_ptr[s];
  const int64_t bb = cb >> 6;
  const int len = int(a.rowlen[i]);
  double xi[3] = {0., 0., 0.}, ui[3] = {0., 0., 0.};
  load3<DIM>(a.x[i], xi);
  load3<DIM>(a.u[i], ui);
  const double voli = a.vol[i];
  const uint32_t fl = a.flags[i];
  const bool int_i = !(fl & kFlagNonInterior);
  double r[3] = {0., 0., 0.};
  const double eps = 1e-6;
  if (xi[1] > max_y - 1.5 * delta && xi[1] < max_y - delta) {
    const double ly = max_y - delta + eps;
    for (int k = 0; k < len; ++k) {
      const int32_t j = a.col[cb + int64_t(k) * 64 + lane];
      double xj[3] = {0., 0., 0.}, uj[3] = {0., 0., 0.};
      load3<DIM>(a.x[j], xj);
      if (!ref_lines_intersect(min_x, ly, max_x, ly, xi[0], xi[1], xj[0],
                               xj[1]))
        continue;
      load3<DIM>(a.u[j], uj);
      bool brk = int_i;
      if (a.has_nofail) brk = brk && !(a.flags[j] & kFlagNonInterior);
      bool fs = (a.bits[bb + k] >> lane) & 1ull;
      double F[3];
      bond_exact<DIM, false>(a.m, xi, ui, xj, uj, voli, a.vol[j], brk, fs, F,
                             nullptr);
      for (int d = 0; d < 3; d++) r[d] += voli * F[d];
    }
  }
  R[3 * i] = r[0];
  R[3 * i + 1] = r[1];
  R[3 * i + 2] = r[2];
  R_len[i] = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

// ---- host side ------------------------------------------------------------

// Launch helpers.  The neighbor-slot chunk U of the fused step kernel is a
// tuning knob (nlm_fd_autotune); force-only launches use U = 4.
template <int DIM, int MODE, int KIND, bool BASE_F, int INF, int U,
          bool NOFAIL>
void launch_t(nlm_ctx* c, const FdArgs<DIM>& a, double tf, double tn) {
  unsigned g = grid_xcd(force_blocks(c->fd) * 256, 256);
  const unsigned long long* bits_ro = a.bits;
  if constexpr (DIM == 2) {
    if (c->fd.tiled && c->fd.use_tiled) {
      // one workgroup per tile of the layout's halos
      const int bn = c->fd.tile;
      const unsigned gt = grid_xcd(c->fd.sell.nslices * 64, bn);
      // (the FP32-force records follow a zero-weight record 0)
      size_t lds = size_t(a.halo_cap + (MODE == kModeFp32 ? 1 : 0)) *
                   sizeof(HaloRecOf<MODE>);
      if constexpr (MODE == kModeFp32 && U == 2) {
        hipLaunchKernelGGL((k_fd_w6<DIM, MODE, KIND, BASE_F, INF, U, NOFAIL>),
                           dim3(gt), dim3(bn), lds, c->stream, a, bits_ro,
                           a.vexw, tf, tn);
        return;
      }
      hipLaunchKernelGGL((k_fd_t<DIM, MODE, KIND, BASE_F, INF, U, NOFAIL>),
                         dim3(gt), dim3(bn), lds, c->stream, a, bits_ro,
                         a.vexw, tf, tn);
      return;
    }
  }
  {
    hipLaunchKernelGGL((k_fd<DIM, MODE, KIND, BASE_F, INF, U, NOFAIL, false>),
                       dim3(g), dim3(256), 0, c->stream, a, bits_ro, a.vexw,
                       tf, tn);
  }
}

template <int DIM, int MODE, int KIND, bool BASE_F, int INF, int U>
void launch_k(nlm_ctx* c, const FdArgs<DIM>& a, double tf, double tn) {
  if (a.has_nofail)
    launch_t<DIM, MODE, KIND, BASE_F, INF, U, true>(c, a, tf, tn);
  else
    launch_t<DIM, MODE, KIND, BASE_F, INF, U, false>(c, a, tf, tn);
}

template <int DIM, int MODE, int KIND, bool BASE_F>
void launch_inf(nlm_ctx* c, const FdArgs<DIM>& a, double tf, double tn,
                int U) {
  const int inf = c->fd.rnp.inf_type;
  if constexpr (KIND == KIND_CD && MODE != kModeExact) {
    if (U == 2) {
      if (inf == 0) launch_k<DIM, MODE, KIND, BASE_F, 0, 2>(c, a, tf, tn);
      else if (inf == 1) launch_k<DIM, MODE, KIND, BASE_F, 1, 2>(c, a, tf, tn);
      else launch_k<DIM, MODE, KIND, BASE_F, 2, 2>(c, a, tf, tn);
      return;
    }
    if (U == 8) {
      if (inf == 0) launch_k<DIM, MODE, KIND, BASE_F, 0, 8>(c, a, tf, tn);
      else if (inf == 1) launch_k<DIM, MODE, KIND, BASE_F, 1, 8>(c, a, tf, tn);
      else launch_k<DIM, MODE, KIND, BASE_F, 2, 8>(c, a, tf, tn);
      return;
    }
  }
  if (inf == 0) launch_k<DIM, MODE, KIND, BASE_F, 0, 4>(c, a, tf, tn);
  else if (inf == 1) launch_k<DIM, MODE, KIND, BASE_F, 1, 4>(c, a, tf, tn);
  else launch_k<DIM, MODE, KIND, BASE_F, 2, 4>(c, a, tf, tn);
}

template <int DIM, int KIND, bool BASE_F>
void launch_mode(nlm_ctx* c, const FdArgs<DIM>& a, double tf, double tn,
                 int U) {
  switch (c->fd.precision) {
    case kModeExact:
      launch_inf<DIM, kModeExact, KIND, BASE_F>(c, a, tf, tn, U);
      break;
    case kModeFp64:
      launch_inf<DIM, kModeFp64, KIND, BASE_F>(c, a, tf, tn, U);
      break;
    default:
      launch_inf<DIM, kModeFp32, KIND, BASE_F>(c, a, tf, tn, U);
      break;
  }
}

template <int DIM>
void launch_force_dispatch(nlm_ctx* c, bool base_f, double t) {
  if (c->fd.material == NLM_MATERIAL_ELASTIC_STATE)
    return es_launch_force(c, base_f, t);
  FdArgs<DIM> a = make_args<DIM>(c);
  if (base_f)
    launch_mode<DIM, KIND_FORCE, true>(c, a, t, t, 4);
  else
    launch_mode<DIM, KIND_FORCE, false>(c, a, t, t, 4);
}

}  // namespace

// f = f_ext(t) + f_int(u) (base_f: f += f_int(u)), bond bits updated.
void fd_launch_force(nlm_ctx* c, bool base_f, double t) {
  if (c->fd.dim == 2)
    launch_force_dispatch<2>(c, base_f, t);
  else
    launch_force_dispatch<3>(c, base_f, t);
}

namespace {

template <int DIM>
void launch_fused_dispatch(nlm_ctx* c, double tf, double tn) {
  if (c->fd.material == NLM_MATERIAL_ELASTIC_STATE)
    return es_launch_fused(c, tf, tn);
  FdArgs<DIM> a = make_args<DIM>(c);
  launch_mode<DIM, KIND_CD, false>(c, a, tf, tn, c->fd.unroll);
}

unsigned node_blocks(const FdState& s) { return unsigned((s.n + 255) / 256); }

// Fracture::addCrack (fracture.cpp:45-72): at most one pending crack per call
// becomes active once definitelyLessThan(activation_time, t).
template <int DIM>
void check_cracks(nlm_ctx* c, double t) {
  FdState& s = c->fd;
  for (size_t k = 0; k < s.timed_cracks.size(); k++) {
    if (s.crack_done[k]) continue;
    if (def_less(s.timed_cracks[k].activation_time, t)) {
      FdArgs<DIM> a = make_args<DIM>(c);
      hipLaunchKernelGGL(k_precrack<DIM>,
                         dim3(unsigned((s.sell.nslices * 64 + 255) / 256)),
                         dim3(256), 0, c->stream, a, s.timed_cracks[k]);
      s.crack_done[k] = true;
      return;
    }
  }
}

template <int DIM>
int do_steps(nlm_ctx* c, int nsteps, double* ms_fused) {
  FdState& s = c->fd;
  const unsigned nb = node_blocks(s);
  if (s.scheme == NLM_TIME_VELOCITY_VERLET) {
    for (int k = 0; k < nsteps; k++) {
      double tn = s.time + s.dt;
      FdArgs<DIM> a = make_args<DIM>(c);
      hipLaunchKernelGGL(k_verlet_a<DIM>, dim3(nb), dim3(256), 0, c->stream, a,
                         tn);
      s.cur ^= 1;
      s.time = tn;
      s.step++;
      launch_force_dispatch<DIM>(c, false, s.time);
      FdArgs<DIM> b = make_args<DIM>(c);
      b.v_out = const_cast<typename VT<DIM>::T*>(b.v);
      hipLaunchKernelGGL(k_verlet_b<DIM>, dim3(nb), dim3(256), 0, c->stream, b);
      check_cracks<DIM>(c, s.time);
    }
    if (ms_fused) *ms_fused = 0.;
    return NLM_OK;
  }
  // first iteration: update from the stored force
  {
    double tn = s.time + s.dt;
    FdArgs<DIM> a = make_args<DIM>(c);
    hipLaunchKernelGGL(k_update_cd<DIM>, dim3(nb), dim3(256), 0, c->stream, a,
                       tn);
    s.cur ^= 1;
    s.time = tn;
    s.step++;
  }
  if (ms_fused) NLM_HIP(c, hipEventRecord(s.ev0, c->stream));
  for (int k = 1; k < nsteps; k++) {
    double tn = s.time + s.dt;
    launch_fused_dispatch<DIM>(c, s.time, tn);
    check_cracks<DIM>(c, s.time);
    s.cur ^= 1;
    s.time = tn;
    s.step++;
  }
  if (ms_fused) NLM_HIP(c, hipEventRecord(s.ev1, c->stream));
  launch_force_dispatch<DIM>(c, false, s.time);
  check_cracks<DIM>(c, s.time);
  NLM_HIP(c, hipGetLastError());
  if (ms_fused) {
    NLM_HIP(c, hipEventSynchronize(s.ev1));
    float ms = 0.f;
    NLM_HIP(c, hipEventElapsedTime(&ms, s.ev0, s.ev1));
    *ms_fused = nsteps > 1 ? double(ms) / double(nsteps - 1) : 0.;
  }
  return NLM_OK;
}

// 1D models: the y dof of every node is held at zero through the node-list
// displacement BC (per-node rate 0), kept across BC updates; the fixity the
// library reports (the reference's 1D fixity) does not include it.
void pin_y_dim1(const FdState& s, std::vector<uint32_t>& fl) {
  if (!s.dim1) return;
  for (auto& f : fl)
    f = (f & ~(0xfu << (kFlagUbcShift + 4))) |
        (kNodeListBc << (kFlagUbcShift + 4));
}

int check_ready(nlm_ctx* c, const char* fn) {
  if (!c) return NLM_ERR_INVALID;
  if (!c->fd.ready)
    return fail(c, NLM_ERR_STATE, std::string(fn) + ": nlm_fd_setup not done");
  return NLM_OK;
}

// Host-side region tests for BC node selection: ULoading::ULoading
// (uLoading.cpp:86-163) and FLoading::FLoading (fLoading.cpp:86-113).
void rotate_cw(double x, double y, double th, double& ox, double& oy) {
  // util::transformation::rotateCW2D, transformation.cpp:20-24
  ox = x * std::cos(th) + y * std::sin(th);
  oy = -x * std::sin(th) + y * std::cos(th);
}
bool h_in_rect(const double* p, double x1, double x2, double y1, double y2) {
  // util::geometry::isPointInsideRectangle, utilGeom.cpp:43-50
  return !(def_less(p[0], x1 - 1.0E-12) || def_less(p[1], y1 - 1.0E-12) ||
           def_greater(p[0], x2 + 1.0E-12) || def_greater(p[1], y2 + 1.0E-12));
}
bool h_in_angled(const double* p, const nlm_bc& b) {
  // util::geometry::isPointInsideAngledRectangle, utilGeom.cpp:60-95
  double lx, ly, mx, my;
  rotate_cw(b.x2 - b.x1, b.y2 - b.y1, b.theta, lx, ly);
  rotate_cw(p[0] - b.x1, p[1] - b.y1, b.theta, mx, my);
  return !(def_less(mx, -1.0E-12) || def_less(my, -1.0E-12) ||
           def_greater(mx, lx + 1.0E-12) || def_greater(my, ly + 1.0E-12));
}
bool h_in_circle(const double* p, double cx, double cy, double r) {
  // util::geometry::isPointinCircle, utilGeom.cpp:328-333 ((A-c).length())
  double a = p[0] - cx, b = p[1] - cy, z = p[2] - 0.;
  return std::sqrt(a * a + b * b + z * z) <= r;
}
bool h_in_cuboid(const double* p, const nlm_bc& b) {
  // util::geometry::isPointInsideCuboid (dim 3), utilGeom.cpp:97-115
  return !(def_less(p[0], b.x1 - 1.0E-12) || def_less(p[1], b.y1 - 1.0E-12) ||
           def_less(p[2], b.z1 - 1.0E-12) ||
           def_greater(p[0], b.x2 + 1.0E-12) ||
           def_greater(p[1], b.y2 + 1.0E-12) ||
           def_greater(p[2], b.z2 + 1.0E-12));
}

BcSetDev to_dev(const nlm_bc& b) {
  BcSetDev d{};
  d.time_fn = b.time_fn;
  d.space_fn = b.space_fn;
  d.dirmask = 0;
  for (int k = 0; k < b.ndir && k < 3; k++)
    if (b.dir[k] >= 1 && b.dir[k] <= 3) d.dirmask |= 1 << (b.dir[k] - 1);
  d.n_space_params = b.n_space_params;
  for (int k = 0; k < 4; k++) {
    d.tp[k] = b.time_params[k];
    d.sp[k] = b.space_params[k];
  }
  d.x1 = b.x1;
  d.x2 = b.x2;
  d.y1 = b.y1;
  d.y2 = b.y2;
  d.z1 = b.z1;
  d.z2 = b.z2;
  return d;
}

int n_time_params_u(int fn) {
  // ULoading::ULoading parameter counts, uLoading.cpp:58-71
  if (fn == NLM_TFN_QUADRATIC || fn == NLM_TFN_SIN) return 2;
  if (fn == NLM_TFN_LINEAR_STEP) return 3;
  if (fn == NLM_TFN_LINEAR_SLOW_FAST) return 4;
  return 1;
}

}  // namespace
}  // namespace nlm

using namespace nlm;

namespace {

// LDS byte offsets of the staged step kernel's slot entries (halo-local
// record index times the record size).  In the FP32-force mode the records
// start at index 1: record 0 is a zero-weight record far from every node,
// addressed by the padded slots (past a row's length) and by the fetches
// past a slice, so the common path needs no per-slot row-length test.
std::vector<uint16_t> staged_entries(const std::vector<uint16_t>& colL,
                                     int precision, const int64_t* sl_ptr,
                                     int64_t nslices, const uint16_t* rowlen,
                                     int64_t n) {
  std::vector<uint16_t> out(colL.size(), 0);
  if (precision != NLM_PREC_FP32_FORCE) {
    for (size_t e = 0; e < colL.size(); e++)
      out[e] = uint16_t(uint32_t(colL[e]) * uint32_t(sizeof(HaloRec)));
    return out;
  }
  for (int64_t sl = 0; sl < nslices; sl++) {
    const int64_t e0 = sl_ptr[sl];
    const int64_t width = (sl_ptr[sl + 1] - e0) / 64;
    for (int64_t k = 0; k < width; k++)
      for (int lane = 0; lane < 64; lane++) {
        const int64_t e = e0 + 64 * k + lane, a = sl * 64 + lane;
        if (a < n && k < int64_t(rowlen[a]))
          out[size_t(e)] = uint16_t((uint32_t(colL[size_t(e)]) + 1u) *
                                    uint32_t(sizeof(HaloRec32)));
      }
  }
  return out;
}

}  // namespace

extern "C" int nlm_fd_setup(nlm_ctx* c, const nlm_fd_problem* p) {
  if (!c) return NLM_ERR_INVALID;
  if (!p || p->n <= 0 || !p->x || !p->vol)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: missing mesh data");
  if (p->dim < 1 || p->dim > 3)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: dim must be 1, 2 or 3");
  if (p->precision < NLM_PREC_FP64_EXACT || p->precision > NLM_PREC_FP32_FORCE)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: unknown precision mode");
  if (p->time_scheme != NLM_TIME_CENTRAL_DIFFERENCE &&
      p->time_scheme != NLM_TIME_VELOCITY_VERLET)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: unknown time scheme");
  if (p->material != NLM_MATERIAL_RNP_BOND &&
      p->material != NLM_MATERIAL_ELASTIC_STATE)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: unknown material");
  const bool es_mat = p->material == NLM_MATERIAL_ELASTIC_STATE;
  const double density = es_mat ? p->es.density : p->mat.density;
  const double horizon = es_mat ? p->es.horizon : p->mat.horizon;
  const double mesh_h = es_mat ? p->es.h : p->mat.h;
  if (!(p->dt > 0.) || !(density > 0.) || !(horizon > 0.) || !(mesh_h > 0.))
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_setup: dt, density, horizon and h must be > 0");
  if (es_mat && p->time_scheme != NLM_TIME_CENTRAL_DIFFERENCE)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_setup: the state-based material uses central "
                "difference");
  if (es_mat && (p->es.dim != p->dim ||
                 (p->dim == 1 ? !(p->es.E > 0.)
                              : (!(p->es.K > 0.) || !(p->es.mu > 0.) ||
                                 !(p->es.factor2D > 0.)))))
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_setup: state-based material needs dim and E (1D) or "
                "K, G (2D, 3D) (nlm_es_params_from_deck)");
  if (p->n >= (int64_t(1) << 31))
    return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: n >= 2^31");
  FdState& s = c->fd;
  s.ready = false;
  // the quasi-static solver's captured batch refers to the buffers a new
  // setup replaces
  if (c->qs.bicg_graph) {
    (void)hipGraphExecDestroy(c->qs.bicg_graph);
    c->qs.bicg_graph = nullptr;
    c->qs.bicg_key.clear();
  }
  const int64_t n = p->n;
  // a 1D model runs on the 2D kernels (nodes on y = 0, y dofs held at 0)
  const int dim = p->dim == 1 ? 2 : p->dim;
  s.udim = p->dim;
  s.dim1 = p->dim == 1;
  hipStream_t st = c->stream;
  int rc;
  // neighbor lists: caller's CSR or the context's last build
  const int64_t* h_row = p->row_ptr;
  std::vector<int64_t> row_tmp;
  if (!h_row) {
    if (!c->nbr.ready || c->nbr.n != n)
      return fail(c, NLM_ERR_STATE,
                  "nlm_fd_setup: no CSR given and no matching neighbor build");
    row_tmp.resize(n + 1);
    NLM_HIP(c, hipMemcpy(row_tmp.data(), c->nbr.row_ptr.p,
                         sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    h_row = row_tmp.data();
  }
  int64_t nnz = h_row[n];
  for (int64_t i = 0; i < n; i++)
    if (h_row[i + 1] - h_row[i] > 65535 || h_row[i + 1] < h_row[i])
      return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: bad row_ptr");
  s.row_ptr_host.assign(h_row, h_row + n + 1);
  std::vector<int32_t> col_tmp;
  const int32_t* h_col = p->col;
  if (h_col) {
    for (int64_t k = 0; k < nnz; k++)
      if (h_col[k] < 0 || h_col[k] >= n)
        return fail(c, NLM_ERR_INVALID, "nlm_fd_setup: column out of range");
  } else {
    if (!c->nbr.ready || c->nbr.nnz != nnz)
      return fail(c, NLM_ERR_STATE, "nlm_fd_setup: neighbor build mismatch");
    col_tmp.resize(nnz + 1);
    NLM_HIP(c, hipMemcpy(col_tmp.data(), c->nbr.col.p, sizeof(int32_t) * nnz,
                         hipMemcpyDeviceToHost));
    h_col = col_tmp.data();
  }
  // internal order, slices and halos (layout.cpp)
  HostLayout L;
  build_layout(dim, n, p->x, mesh_h, h_row, h_col, L);
  s.perm_host = L.perm;
  s.inv_host = L.inv;
  s.sell.nslices = L.nslices;
  s.sell.nentries = L.sl_ptr[L.nslices];
  const int64_t nent = s.sell.nentries;
  auto up = [&](DevBuf& d, const void* src, size_t bytes) -> int {
    int r = alloc(c, d, bytes + 256);
    if (r) return r;
    NLM_HIP(c, hipMemsetAsync(d.p, 0, d.bytes, st));
    if (src && bytes)
      NLM_HIP(c, hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, st));
    return NLM_OK;
  };
  if ((rc = up(s.row_ptr, L.row_ptr.data(), sizeof(int64_t) * (n + 1)))) return rc;
  if ((rc = up(s.col, L.col.data(), sizeof(int32_t) * nnz))) return rc;
  if ((rc = up(s.ref_off, L.ref_off.data(), sizeof(int64_t) * n))) return rc;
  if ((rc = up(s.perm_dev, L.perm.data(), sizeof(int32_t) * n))) return rc;
  if ((rc = up(s.sell.sl_ptr, L.sl_ptr.data(), sizeof(int64_t) * (L.nslices + 1))))
    return rc;
  if ((rc = up(s.sell.col, L.colS.data(), sizeof(int32_t) * nent))) return rc;
  if ((rc = up(s.sell.rowlen, L.rowlen.data(), sizeof(uint16_t) * n))) return rc;
  // bond-bit words: one per (slice, slot), plus 8 words of padding that the
  // step kernel's chunked reads may touch past the last slice
  if ((rc = up(s.sell.bits, nullptr, sizeof(uint64_t) * (nent / 64 + 8))))
    return rc;
  if ((rc = up(s.halo_ptr, L.halo_ptr.data(), sizeof(int64_t) * (L.nblocks + 1))))
    return rc;
  if ((rc = up(s.halo, L.halo.data(), sizeof(int32_t) * L.halo.size()))) return rc;
  s.halo_max = L.halo_max;
  s.tile = 256;
  // LDS budget of the staged kernel: one 48-byte record per halo node
  s.tiled = !es_mat && dim == 2 && L.halo_max > 0 &&
            (L.halo_max + 1) * int64_t(sizeof(HaloRec)) <= 64 * 1024;
  if (s.tiled) {
    // the staged kernel addresses its dynamic LDS from 0, which holds while
    // it has no static LDS (one instance checked: all share the source)
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(
            &fa, reinterpret_cast<const void*>(
                     &k_fd_w6<2, kModeFp32, KIND_CD, false, 0, 2, false>)) !=
            hipSuccess ||
        fa.sharedSizeBytes != 0)
      s.tiled = false;
  }
  // the bond-based staged kernel reads halo-local entries as LDS byte
  // offsets of its records (below 64 KiB by the budget above; the FP32-force
  // mode's records are the 40-byte block-relative ones after a zero-weight
  // record 0)
  if (s.tiled) {
    const std::vector<uint16_t> colL_off =
        staged_entries(L.colL, p->precision, L.sl_ptr.data(), L.nslices,
                       L.rowlen.data(), n);
    if ((rc = up(s.colL, colL_off.data(), sizeof(uint16_t) * nent))) return rc;
  } else if ((rc = up(s.colL, L.colL.data(), sizeof(uint16_t) * nent))) {
    return rc;
  }
  // node arrays in internal order
  const int stride = dim == 2 ? 2 : 4;
  std::vector<double> hx(size_t(n) * stride, 0.);
  std::vector<double> hv(n);
  for (int64_t a = 0; a < n; a++) {
    int64_t i = L.perm[a];
    for (int d = 0; d < dim; d++) hx[a * stride + d] = p->x[3 * i + d];
    hv[a] = p->vol[i];
  }
  // Per 256-node block: the centre of its nodes' bounding box, the origin of
  // the FP32 block-relative positions (double2 per block in 2D, double4 in
  // 3D).
  std::vector<double> borg(size_t(L.nblocks) * stride, 0.);
  for (int64_t b = 0; b < L.nblocks; b++) {
    const int64_t a0 = 256 * b, a1 = std::min<int64_t>(n, a0 + 256);
    if (a0 >= n) break;
    double lo[3], hi[3];
    for (int d = 0; d < dim; d++) lo[d] = hi[d] = hx[a0 * stride + d];
    for (int64_t a = a0; a < a1; a++)
      for (int d = 0; d < dim; d++) {
        lo[d] = std::min(lo[d], hx[a * stride + d]);
        hi[d] = std::max(hi[d], hx[a * stride + d]);
      }
    for (int d = 0; d < dim; d++) borg[stride * b + d] = 0.5 * (lo[d] + hi[d]);
  }
  // largest offset of a block's node or halo node from the block's origin
  // (per component): the range of the fixed-point positions of the
  // state-based FP32-force passes
  double org_range = 0.;
  for (int64_t b = 0; b < L.nblocks; b++) {
    auto span = [&](int64_t a) {
      for (int d = 0; d < dim; d++)
        org_range = std::max(
            org_range, std::fabs(hx[a * stride + d] - borg[stride * b + d]));
    };
    for (int64_t a = 256 * b; a < std::min<int64_t>(n, 256 * b + 256); a++)
      span(a);
    for (int64_t t = L.halo_ptr[b]; t < L.halo_ptr[b + 1]; t++)
      span(L.halo[size_t(t)]);
  }
  if ((rc = up(s.blk_org, borg.data(), sizeof(double) * borg.size())))
    return rc;
  size_t vbytes = sizeof(double) * stride * size_t(n);
  if ((rc = alloc(c, s.x, vbytes))) return rc;
  if ((rc = alloc(c, s.vol, sizeof(double) * n))) return rc;
  if ((rc = alloc(c, s.flags, sizeof(uint32_t) * n))) return rc;
  for (int b = 0; b < 2; b++) {
    if ((rc = alloc(c, s.u[b], vbytes))) return rc;
    if ((rc = alloc(c, s.v[b], vbytes))) return rc;
    NLM_HIP(c, hipMemsetAsync(s.u[b].p, 0, vbytes, st));
    NLM_HIP(c, hipMemsetAsync(s.v[b].p, 0, vbytes, st));
  }
  if ((rc = alloc(c, s.f, vbytes))) return rc;
  NLM_HIP(c, hipMemsetAsync(s.f.p, 0, vbytes, st));
  NLM_HIP(c, hipMemcpyAsync(s.x.p, hx.data(), vbytes, hipMemcpyHostToDevice,
                            st));
  NLM_HIP(c, hipMemcpyAsync(s.vol.p, hv.data(), sizeof(double) * n,
                            hipMemcpyHostToDevice, st));
  std::vector<uint32_t> fl(n, 0u);
  s.has_nofail = false;
  if (p->interior)
    for (int64_t a = 0; a < n; a++)
      if (!p->interior[L.perm[a]]) {
        fl[a] |= kFlagNonInterior;
        s.has_nofail = true;
      }
  pin_y_dim1(s, fl);
  NLM_HIP(c, hipMemcpyAsync(s.flags.p, fl.data(), sizeof(uint32_t) * n,
                            hipMemcpyHostToDevice, st));
  if ((rc = alloc(c, s.node_rate, sizeof(double) * 3 * size_t(n) + 256)))
    return rc;
  NLM_HIP(c, hipMemsetAsync(s.node_rate.p, 0, s.node_rate.bytes, st));
  if ((rc = alloc(c, s.tdev, 64))) return rc;
  if (s.graph) {
    (void)hipGraphExecDestroy(s.graph);
    s.graph = nullptr;
  }
  s.graph_chunk = 0;
  if ((rc = alloc(c, s.bctab, sizeof(BcTable)))) return rc;
  std::memset(&s.bc_host, 0, sizeof(BcTable));
  NLM_HIP(c, hipMemcpyAsync(s.bctab.p, &s.bc_host, sizeof(BcTable),
                            hipMemcpyHostToDevice, st));
  NLM_HIP(c, hipStreamSynchronize(st));
  s.fix_host.assign(n, 0);
  s.n = n;
  s.dim = dim;
  s.nnz = nnz;
  s.precision = p->precision;
  s.scheme = p->time_scheme;
  s.dt = p->dt;
  s.density = density;
  s.material = p->material;
  s.time = 0.;
  s.step = 0;
  s.cur = 0;
  s.has_fbc = false;
  s.mat = p->mat;
  s.timed_cracks.clear();
  s.crack_done.clear();
  RnpDev& m = s.rnp;
  std::memset(&m, 0, sizeof(m));
  m.C = p->mat.C;
  m.beta = p->mat.beta;
  m.rbar = p->mat.rbar;
  m.inv_factor = p->mat.inv_factor;
  m.factor_sc = p->mat.factor_sc;
  m.horizon = p->mat.horizon;
  m.h = p->mat.h;
  // rnpBond.cpp:248-250
  m.check_up = p->mat.horizon + 0.5 * p->mat.h;
  m.check_low = p->mat.horizon - 0.5 * p->mat.h;
  m.inf_type = p->mat.inf_type;
  m.inf_a = p->mat.inf_a;
  m.inf_b = p->mat.inf_b;
  m.irreversible = p->mat.irreversible;
  m.contact = p->mat.contact;
  m.kn = p->mat.contact_kn;
  m.rc = p->mat.contact_rc;
  m.rcp_inv_factor = 1.0 / m.inv_factor;
  m.rcp_h = 1.0 / m.h;
  m.rcp_horizon = 1.0 / m.horizon;
  m.krbar_sq = (m.factor_sc * m.rbar) * (m.factor_sc * m.rbar);
  m.rcp_krbar_sq = 1.0 / m.krbar_sq;
  m.c4 = 4.0 * m.C * m.beta * m.rcp_inv_factor;
  m.low_band = m.check_low * (1.0 + 1e-5);
  m.rcp_inf_b = m.inf_type == 2 ? 1.0 / m.inf_b : 0.0;
  m.c4f = float(m.c4);
  m.beta_f = float(m.beta);
  m.check_up_f = float(m.check_up);
  m.low_band_f = float(m.low_band);
  m.rcp_h_f = float(m.rcp_h);
  m.rcp_horizon_f = float(m.rcp_horizon);
  m.inf_a_f = float(m.inf_a);
  m.inf_b_f = float(m.inf_b);
  m.rcp_inf_b_f = float(m.rcp_inf_b);
  m.rcp_krbar_sq_f = float(m.rcp_krbar_sq);
  m.nbeta_log2e_f = float(-m.beta * 1.4426950408889634);
  m.c4f_infl = m.inf_type == 0 ? float(m.c4 * m.inf_a) : float(m.c4);
  m.break_rs2_f = m.irreversible ? float(kBreakRho2 * m.krbar_sq)
                                 : std::numeric_limits<float>::infinity();
  m.break_rho2 = m.irreversible ? kBreakRho2
                                : std::numeric_limits<double>::infinity();
  m.contact_mask = m.contact ? ~0ull : 0ull;
  m.break_guard_f = float(double(kGuard32) * kBreakRho2 * m.krbar_sq);
  m.break_lo = m.break_rho2 - kGuard64 * kBreakRho2;
  m.break_hi = m.break_rho2 + kGuard64 * kBreakRho2;
  m.break_lo_f = m.irreversible ? float(double(m.break_rs2_f) -
                                        double(m.break_guard_f))
                                : std::numeric_limits<float>::infinity();
  m.break_hi_f = m.irreversible ? float(double(m.break_rs2_f) +
                                        double(m.break_guard_f))
                                : std::numeric_limits<float>::infinity();
  m.up_rcp_h_f = float(m.check_up * m.rcp_h);
  m.vex_tol_f = float(3e-5 * m.low_band);
  {
    const double b2 = m.beta * 1.4426950408889634;  // beta log2 e
    if (b2 > 0.) {  // (beta > 0 for RNP bonds)
      m.tau = std::ldexp(1.0, int(std::lround(0.5 * std::log2(b2))));
      m.sigma = m.tau * m.tau / b2;
    } else {
      m.tau = m.sigma = 1.0;
    }
    const double bl = m.tau * m.tau / m.sigma;  // beta log2 e
    m.c4f_infl_s = float(double(m.c4f_infl) * m.sigma / m.tau);
    m.rcp_h_s = float(1.0 / (m.sigma * m.h));
    m.rcp_sigma_f = float(1.0 / m.sigma);
    m.break_lo_s = m.irreversible ? float(double(m.break_lo_f) * bl)
                                  : std::numeric_limits<float>::infinity();
    // r within 3e-5 (relative) of low_band <=> r^2 within about 6e-5 of
    // low_band^2; the band on r^2 is taken a little wider
    const double lows = m.sigma * m.low_band;
    m.low_dd_f = float(lows * lows);
    m.vex_dd_tol_f = float(6.1e-5 * lows * lows);
  }
  {
    // Bonds whose reference length lies within 3e-5 (relative) of the
    // volume-correction edge: a static mask per (slice, slot), padded like
    // the bond bits; the step kernels send them to the exact evaluation.
    // The band is far wider than the fast paths' rounding of the length.
    std::vector<unsigned long long> vw(size_t(nent / 64 + 8), 0ull);
    if (!es_mat) {
      const double tol = 3e-5 * m.low_band;
      for (int64_t sl = 0; sl < L.nslices; sl++) {
        const int64_t e0 = L.sl_ptr[sl];
        const int64_t width = (L.sl_ptr[sl + 1] - e0) / 64;
        for (int64_t k = 0; k < width; k++) {
          unsigned long long word = 0ull;
          for (int lane = 0; lane < 64; lane++) {
            const int64_t a = sl * 64 + lane;
            if (a >= n || k >= int64_t(L.rowlen[a])) continue;
            const int64_t g = L.colS[e0 + 64 * k + lane];
            double d2 = 0.;
            for (int d = 0; d < dim; d++) {
              const double e = hx[g * stride + d] - hx[a * stride + d];
              d2 += e * e;
            }
            if (std::fabs(std::sqrt(d2) - m.low_band) <= tol)
              word |= 1ull << lane;
          }
          vw[size_t(e0 / 64 + k)] = word;
        }
      }
    }
    if ((rc = up(s.vexw, vw.data(), sizeof(unsigned long long) * vw.size())))
      return rc;
    NLM_HIP(c, hipStreamSynchronize(st));
  }
  s.unroll = 4;
  if (es_mat) {
    // ElasticState constants (ElasticState.cpp:47-158, volumeCorrection.cpp)
    const nlm_es_params& q = p->es;
    s.es_par = q;
    EsDev& e = s.es;
    std::memset(&e, 0, sizeof(e));
    e.K = q.K;
    e.mu = q.mu;
    e.f2d = q.factor2D;
    e.E = q.E;
    e.one_d = s.dim1 ? 1 : 0;
    // 1D on the 2D kernels: (2/m_i) f2d = 1/m_i exactly (a power-of-two
    // factor), the reference's 1D dilatation factor
    if (s.dim1) e.f2d = 0.5;
    // alpha_s = (9/m) (K + ((nu+1)/(2 nu-1))^2 mu / 9) for plane stress,
    // (9/m) (K + mu/9) for plane strain (ElasticState.cpp:201-212)
    e.Kc = q.plane_strain
               ? q.K + q.mu / 9.
               : q.K + std::pow((q.nu + 1.) / (2. * q.nu - 1.), 2) * q.mu / 9.;
    e.horizon = q.horizon;
    e.h = q.h;
    const double r = q.h * 0.5;
    e.check_low = q.horizon - r;
    e.check_up = q.horizon + r;
    e.rcp_h = 1.0 / q.h;
    e.low_band = e.check_low * (1.0 + 1e-5);
    e.breaking = q.critical_stretch > 0. ? 1 : 0;
    e.sc = q.critical_stretch;
    e.sc_break = e.breaking ? q.critical_stretch / (1.0 - 1e-5) : 0.;
    e.check_up_f = float(e.check_up);
    e.low_band_f = float(e.low_band);
    e.rcp_h_f = float(e.rcp_h);
    e.sc_break_f = float(e.sc_break);
    // The fast first pass takes the factor as min((check_up - |X|)/h, 1),
    // which is below the reference's 1 for |X| in (check_low, low_band]
    // (definitelyGreaterThan's relative window, 1e-5 check_low wide): the
    // exact-evaluation band of both fast modes covers that whole window.
    e.vex_tol_f = 3e-5f * e.low_band_f;
    e.vex_tol = 3e-5 * e.low_band;
    es_set_quantum(e, org_range);
    e.brk_lo = e.breaking ? e.sc_break - kEsGuard64 * e.sc_break
                          : std::numeric_limits<double>::infinity();
    e.brk_lo_f = e.breaking ? e.sc_break_f - kEsGuard32 * e.sc_break_f
                            : std::numeric_limits<float>::infinity();
    if ((rc = es_prepare(c))) return rc;
  }
  if (!s.ev0) NLM_HIP(c, hipEventCreate(&s.ev0));
  if (!s.ev1) NLM_HIP(c, hipEventCreate(&s.ev1));
  s.ready = true;
  return NLM_OK;
}

extern "C" int nlm_fd_set_ubc(nlm_ctx* c, int nsets, const nlm_bc* sets) {
  int rc = check_ready(c, "nlm_fd_set_ubc");
  if (rc) return rc;
  FdState& s = c->fd;
  if (nsets < 0 || nsets > kMaxUbcSets || (nsets > 0 && !sets))
    return fail(c, NLM_ERR_INVALID, "nlm_fd_set_ubc: 0..14 sets");
  std::vector<double> hx(size_t(s.n) * 3);
  {
    const int stride = s.dim == 2 ? 2 : 4;
    std::vector<double> tmp(size_t(s.n) * stride);
    NLM_HIP(c, hipMemcpy(tmp.data(), s.x.p, s.x.bytes, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < s.n; i++)
      for (int d = 0; d < 3; d++)
        hx[3 * i + d] = d < s.dim ? tmp[i * stride + d] : 0.;
  }
  std::vector<uint32_t> fl(s.n);
  NLM_HIP(c, hipMemcpy(fl.data(), s.flags.p, sizeof(uint32_t) * s.n,
                       hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < s.n; i++) fl[i] &= ~0xfffu;
  s.fix_host.assign(s.n, 0);
  for (int k = 0; k < nsets; k++) {
    const nlm_bc& b = sets[k];
    if (b.region < NLM_REGION_RECTANGLE || b.region > NLM_REGION_CUBOID)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_set_ubc: unknown region");
    if (b.time_fn < NLM_TFN_CONSTANT || b.time_fn > NLM_TFN_LINEAR_SLOW_FAST)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_set_ubc: unknown time function");
    if (b.space_fn < NLM_SFN_CONSTANT || b.space_fn > NLM_SFN_LINEAR_Z)
      return fail(c, NLM_ERR_INVALID,
                  "nlm_fd_set_ubc: unknown spatial function");
    if (b.n_time_params != n_time_params_u(b.time_fn))
      return fail(c, NLM_ERR_INVALID,
                  "nlm_fd_set_ubc: wrong number of time parameters");
    if (b.space_fn != NLM_SFN_CONSTANT && b.n_space_params < 1)
      return fail(c, NLM_ERR_INVALID,
                  "nlm_fd_set_ubc: spatial function needs a parameter");
    if (b.ndir < 1 || b.ndir > 3)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_set_ubc: 1..3 directions");
    for (int d = 0; d < b.ndir; d++)
      if (b.dir[d] < 1 || b.dir[d] > s.udim)
        return fail(c, NLM_ERR_INVALID, "nlm_fd_set_ubc: bad direction");
    s.bc_host.u[k] = to_dev(b);
    for (int64_t i = 0; i < s.n; i++) {
      const double* p = &hx[3 * i];
      bool sel = false;
      switch (b.region) {
        case NLM_REGION_RECTANGLE:
          sel = h_in_rect(p, b.x1, b.x2, b.y1, b.y2);
          break;
        case NLM_REGION_ANGLED_RECTANGLE:
          sel = h_in_angled(p, b);
          break;
        case NLM_REGION_CIRCLE:
          sel = h_in_circle(p, b.x1, b.y1, b.r1);
          break;
        case NLM_REGION_TORUS:
          sel = h_in_circle(p, b.x1, b.y1, b.r1) &&
                !h_in_circle(p, b.x1, b.y1, b.r2);
          break;
        case NLM_REGION_LINE:
          sel = def_greater(p[0], b.x1) && def_less(p[0], b.x2);
          break;
        case NLM_REGION_CUBOID:
          // ULoading selects no node for a cuboid in the reference; this
          // library selects the nodes inside it (3D pulls).
          sel = h_in_cuboid(p, b);
          break;
      }
      if (!sel) continue;
      for (int d = 0; d < b.ndir; d++) {
        int dof = b.dir[d] - 1;
        fl[i] = (fl[i] & ~(0xfu << (4 * dof))) | (uint32_t(k + 1) << (4 * dof));
        s.fix_host[s.perm_host[i]] |= uint8_t(1u << dof);
      }
    }
  }
  s.bc_host.nu = nsets;
  pin_y_dim1(s, fl);
  NLM_HIP(c, hipMemcpy(s.flags.p, fl.data(), sizeof(uint32_t) * s.n,
                       hipMemcpyHostToDevice));
  NLM_HIP(c, hipMemcpy(s.bctab.p, &s.bc_host, sizeof(BcTable),
                       hipMemcpyHostToDevice));
  return NLM_OK;
}

extern "C" int nlm_fd_set_fbc(nlm_ctx* c, int nsets, const nlm_bc* sets) {
  int rc = check_ready(c, "nlm_fd_set_fbc");
  if (rc) return rc;
  FdState& s = c->fd;
  if (nsets < 0 || nsets > kMaxFbcSets || (nsets > 0 && !sets))
    return fail(c, NLM_ERR_INVALID, "nlm_fd_set_fbc: 0..8 sets");
  std::vector<double> hx(size_t(s.n) * 3);
  {
    const int stride = s.dim == 2 ? 2 : 4;
    std::vector<double> tmp(size_t(s.n) * stride);
    NLM_HIP(c, hipMemcpy(tmp.data(), s.x.p, s.x.bytes, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < s.n; i++)
      for (int d = 0; d < 3; d++)
        hx[3 * i + d] = d < s.dim ? tmp[i * stride + d] : 0.;
  }
  std::vector<uint32_t> fl(s.n);
  NLM_HIP(c, hipMemcpy(fl.data(), s.flags.p, sizeof(uint32_t) * s.n,
                       hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < s.n; i++) fl[i] &= ~(0xffu << kFlagFbcShift);
  for (int k = 0; k < nsets; k++) {
    const nlm_bc& b = sets[k];
    if (b.region != NLM_REGION_LINE && b.region != NLM_REGION_RECTANGLE &&
        b.region != NLM_REGION_ANGLED_RECTANGLE &&
        b.region != NLM_REGION_CUBOID)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_set_fbc: unsupported region");
    if (b.time_fn == NLM_TFN_QUADRATIC || b.time_fn < 0 ||
        b.time_fn > NLM_TFN_LINEAR_SLOW_FAST)
      return fail(c, NLM_ERR_INVALID,
                  "nlm_fd_set_fbc: unsupported time function");
    int ntp = b.time_fn == NLM_TFN_LINEAR_STEP
                  ? 3
                  : (b.time_fn == NLM_TFN_LINEAR_SLOW_FAST
                         ? 4
                         : (b.time_fn == NLM_TFN_SIN ? 2 : 1));
    if (b.n_time_params != ntp)
      return fail(c, NLM_ERR_INVALID,
                  "nlm_fd_set_fbc: wrong number of time parameters");
    if (b.space_fn < 0 || b.space_fn > NLM_SFN_LINE_LOAD)
      return fail(c, NLM_ERR_INVALID,
                  "nlm_fd_set_fbc: unknown spatial function");
    if (b.space_fn != NLM_SFN_CONSTANT && b.space_fn != NLM_SFN_LINE_LOAD &&
        b.n_space_params < 1)
      return fail(c, NLM_ERR_INVALID,
                  "nlm_fd_set_fbc: spatial function needs a parameter");
    if (b.ndir < 1 || b.ndir > 3)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_set_fbc: 1..3 directions");
    for (int d = 0; d < b.ndir; d++)
      if (b.dir[d] < 1 || b.dir[d] > s.udim)
        return fail(c, NLM_ERR_INVALID, "nlm_fd_set_fbc: bad direction");
    s.bc_host.f[k] = to_dev(b);
    // mesh size of FLoading's line_load
    s.bc_host.f[k].h =
        s.material == NLM_MATERIAL_ELASTIC_STATE ? s.es.h : s.rnp.h;
    for (int64_t i = 0; i < s.n; i++) {
      const double* p = &hx[3 * i];
      bool sel = false;
      if (b.region == NLM_REGION_LINE)
        sel = def_greater(p[0], b.x1) && def_less(p[0], b.x2);
      else if (b.region == NLM_REGION_RECTANGLE)
        sel = h_in_rect(p, b.x1, b.x2, b.y1, b.y2);
      else if (b.region == NLM_REGION_ANGLED_RECTANGLE)
        sel = h_in_angled(p, b);
      else
        sel = h_in_cuboid(p, b);
      if (sel) fl[i] |= 1u << (kFlagFbcShift + k);
    }
  }
  s.bc_host.nf = nsets;
  s.has_fbc = nsets > 0;
  NLM_HIP(c, hipMemcpy(s.flags.p, fl.data(), sizeof(uint32_t) * s.n,
                       hipMemcpyHostToDevice));
  NLM_HIP(c, hipMemcpy(s.bctab.p, &s.bc_host, sizeof(BcTable),
                       hipMemcpyHostToDevice));
  return NLM_OK;
}

extern "C" int nlm_fd_add_cracks(nlm_ctx* c, int ncracks,
                                 const nlm_crack* cracks) {
  int rc = check_ready(c, "nlm_fd_add_cracks");
  if (rc) return rc;
  FdState& s = c->fd;
  if (ncracks < 0 || (ncracks > 0 && !cracks))
    return fail(c, NLM_ERR_INVALID, "nlm_fd_add_cracks: bad arguments");
  if (ncracks > 0 && s.udim != 2)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_add_cracks: pre-cracks need dimension 2");
  for (int k = 0; k < ncracks; k++) {
    const nlm_crack& cr = cracks[k];
    if (cr.o != -1 && cr.o != 0 && cr.o != 1)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_add_cracks: bad orientation");
    if (cr.activation_time < 0.) {
      if (s.dim == 2) {
        FdArgs<2> a = make_args<2>(c);
        hipLaunchKernelGGL(k_precrack<2>,
                           dim3(unsigned((s.sell.nslices * 64 + 255) / 256)),
                           dim3(256), 0, c->stream, a, cr);
      }
    } else {
      s.timed_cracks.push_back(cr);
      s.crack_done.push_back(false);
    }
  }
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

namespace {
// AoS reference-order host vectors <-> internal-order device vectors
int to_dev_vec(nlm_ctx* c, const double* src, DevBuf& dst) {
  FdState& s = c->fd;
  const int stride = s.dim == 2 ? 2 : 4;
  std::vector<double> tmp(size_t(s.n) * stride, 0.);
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    for (int d = 0; d < s.dim; d++) tmp[a * stride + d] = src[3 * i + d];
  }
  NLM_HIP(c, hipMemcpy(dst.p, tmp.data(), sizeof(double) * tmp.size(),
                       hipMemcpyHostToDevice));
  return NLM_OK;
}
int from_dev_vec(nlm_ctx* c, const DevBuf& src, double* dst) {
  FdState& s = c->fd;
  const int stride = s.dim == 2 ? 2 : 4;
  std::vector<double> tmp(size_t(s.n) * stride, 0.);
  NLM_HIP(c, hipMemcpy(tmp.data(), src.p, sizeof(double) * tmp.size(),
                       hipMemcpyDeviceToHost));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    for (int d = 0; d < 3; d++)
      dst[3 * i + d] = d < s.dim ? tmp[a * stride + d] : 0.;
  }
  return NLM_OK;
}
}  // namespace

extern "C" int nlm_fd_set_state(nlm_ctx* c, const double* u, const double* v,
                                const double* f) {
  int rc = check_ready(c, "nlm_fd_set_state");
  if (rc) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  FdState& s = c->fd;
  if (u && (rc = to_dev_vec(c, u, s.u[s.cur]))) return rc;
  if (v && (rc = to_dev_vec(c, v, s.v[s.cur]))) return rc;
  if (f && (rc = to_dev_vec(c, f, s.f))) return rc;
  return NLM_OK;
}

extern "C" int nlm_fd_get_state(nlm_ctx* c, double* u, double* v, double* f) {
  int rc = check_ready(c, "nlm_fd_get_state");
  if (rc) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  FdState& s = c->fd;
  if (u && (rc = from_dev_vec(c, s.u[s.cur], u))) return rc;
  if (v && (rc = from_dev_vec(c, s.v[s.cur], v))) return rc;
  if (f && (rc = from_dev_vec(c, s.f, f))) return rc;
  return NLM_OK;
}

extern "C" int nlm_fd_get_bonds(nlm_ctx* c, uint8_t* out) {
  int rc = check_ready(c, "nlm_fd_get_bonds");
  if (rc) return rc;
  if (!out) return fail(c, NLM_ERR_INVALID, "nlm_fd_get_bonds: null output");
  FdState& s = c->fd;
  DevBuf tmp;
  if ((rc = alloc(c, tmp, size_t(s.nnz) + 1))) return rc;
  if (s.dim == 2) {
    FdArgs<2> a = make_args<2>(c);
    hipLaunchKernelGGL(k_bits_to_bytes<2>, dim3(node_blocks(s)), dim3(256), 0,
                       c->stream, a, s.ref_off.as<int64_t>(),
                       tmp.as<uint8_t>());
  } else {
    FdArgs<3> a = make_args<3>(c);
    hipLaunchKernelGGL(k_bits_to_bytes<3>, dim3(node_blocks(s)), dim3(256), 0,
                       c->stream, a, s.ref_off.as<int64_t>(),
                       tmp.as<uint8_t>());
  }
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipMemcpyAsync(out, tmp.p, size_t(s.nnz), hipMemcpyDeviceToHost,
                            c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_fd_set_bonds(nlm_ctx* c, const uint8_t* in) {
  int rc = check_ready(c, "nlm_fd_set_bonds");
  if (rc) return rc;
  if (!in) return fail(c, NLM_ERR_INVALID, "nlm_fd_set_bonds: null input");
  FdState& s = c->fd;
  DevBuf tmp;
  if ((rc = alloc(c, tmp, size_t(s.nnz) + 1))) return rc;
  NLM_HIP(c, hipMemcpyAsync(tmp.p, in, size_t(s.nnz), hipMemcpyHostToDevice,
                            c->stream));
  if (s.dim == 2) {
    FdArgs<2> a = make_args<2>(c);
    hipLaunchKernelGGL(k_bytes_to_bits<2>, dim3(unsigned(s.sell.nslices)),
                       dim3(64), 0, c->stream, a, s.ref_off.as<int64_t>(),
                       tmp.as<uint8_t>());
  } else {
    FdArgs<3> a = make_args<3>(c);
    hipLaunchKernelGGL(k_bytes_to_bits<3>, dim3(unsigned(s.sell.nslices)),
                       dim3(64), 0, c->stream, a, s.ref_off.as<int64_t>(),
                       tmp.as<uint8_t>());
  }
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_fd_bond_bytes_size(nlm_ctx* c, int64_t* size) {
  int rc = check_ready(c, "nlm_fd_bond_bytes_size");
  if (rc) return rc;
  FdState& s = c->fd;
  int64_t t = 0;
  for (int64_t i = 0; i < s.n; i++) {
    int64_t ns = s.row_ptr_host[i + 1] - s.row_ptr_host[i];
    t += (ns + 7) / 8;
  }
  *size = t;
  return NLM_OK;
}

extern "C" int nlm_fd_get_bond_bytes(nlm_ctx* c, uint8_t* bytes) {
  int rc = check_ready(c, "nlm_fd_get_bond_bytes");
  if (rc) return rc;
  FdState& s = c->fd;
  std::vector<uint8_t> st(size_t(s.nnz) + 1);
  if ((rc = nlm_fd_get_bonds(c, st.data()))) return rc;
  int64_t o = 0;
  for (int64_t i = 0; i < s.n; i++) {
    int64_t r0 = s.row_ptr_host[i], ns = s.row_ptr_host[i + 1] - r0;
    int64_t nb = (ns + 7) / 8;
    for (int64_t b = 0; b < nb; b++) bytes[o + b] = 0;
    for (int64_t k = 0; k < ns; k++)
      if (st[r0 + k]) bytes[o + k / 8] |= uint8_t(1u << (k % 8));
    o += nb;
  }
  return NLM_OK;
}

extern "C" int nlm_fd_get_fixity(nlm_ctx* c, uint8_t* fix) {
  int rc = check_ready(c, "nlm_fd_get_fixity");
  if (rc) return rc;
  std::memcpy(fix, c->fd.fix_host.data(), size_t(c->fd.n));
  return NLM_OK;
}

extern "C" int nlm_fd_get_time(nlm_ctx* c, double* t, int64_t* step) {
  int rc = check_ready(c, "nlm_fd_get_time");
  if (rc) return rc;
  if (t) *t = c->fd.time;
  if (step) *step = c->fd.step;
  return NLM_OK;
}

extern "C" int nlm_fd_set_time(nlm_ctx* c, double t, int64_t step) {
  int rc = check_ready(c, "nlm_fd_set_time");
  if (rc) return rc;
  c->fd.time = t;
  c->fd.step = step;
  return NLM_OK;
}

extern "C" int nlm_fd_init(nlm_ctx* c) {
  int rc = check_ready(c, "nlm_fd_init");
  if (rc) return rc;
  FdState& s = c->fd;
  unsigned nb = node_blocks(s);
  // ULoading::apply in place (u_out = u)
  if (s.dim == 2) {
    FdArgs<2> a = make_args<2>(c);
    a.u_out = const_cast<double2*>(a.u);
    a.v_out = const_cast<double2*>(a.v);
    hipLaunchKernelGGL(k_apply_ubc<2>, dim3(nb), dim3(256), 0, c->stream, a,
                       s.time);
    launch_force_dispatch<2>(c, false, s.time);
  } else {
    FdArgs<3> a = make_args<3>(c);
    a.u_out = const_cast<double4*>(a.u);
    a.v_out = const_cast<double4*>(a.v);
    hipLaunchKernelGGL(k_apply_ubc<3>, dim3(nb), dim3(256), 0, c->stream, a,
                       s.time);
    launch_force_dispatch<3>(c, false, s.time);
  }
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_fd_step(nlm_ctx* c, int nsteps) {
  int rc = check_ready(c, "nlm_fd_step");
  if (rc) return rc;
  if (nsteps < 0) return fail(c, NLM_ERR_INVALID, "nlm_fd_step: nsteps < 0");
  if (nsteps == 0) return NLM_OK;
  rc = c->fd.dim == 2 ? do_steps<2>(c, nsteps, nullptr)
                      : do_steps<3>(c, nsteps, nullptr);
  if (rc) return rc;
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_fd_timed_steps(nlm_ctx* c, int nsteps, double* ms_total,
                                  double* ms_per_step) {
  int rc = check_ready(c, "nlm_fd_timed_steps");
  if (rc) return rc;
  if (nsteps < 2)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_timed_steps: nsteps >= 2");
  hipEvent_t a0, a1;
  NLM_HIP(c, hipEventCreate(&a0));
  NLM_HIP(c, hipEventCreate(&a1));
  NLM_HIP(c, hipEventRecord(a0, c->stream));
  double per = 0.;
  rc = c->fd.dim == 2 ? do_steps<2>(c, nsteps, &per)
                      : do_steps<3>(c, nsteps, &per);
  if (rc) return rc;
  NLM_HIP(c, hipEventRecord(a1, c->stream));
  NLM_HIP(c, hipEventSynchronize(a1));
  float ms = 0.f;
  NLM_HIP(c, hipEventElapsedTime(&ms, a0, a1));
  (void)hipEventDestroy(a0);
  (void)hipEventDestroy(a1);
  if (ms_total) *ms_total = ms;
  if (ms_per_step) *ms_per_step = per;
  return NLM_OK;
}

extern "C" int nlm_fd_compute_forces(nlm_ctx* c) {
  int rc = check_ready(c, "nlm_fd_compute_forces");
  if (rc) return rc;
  if (c->fd.dim == 2)
    launch_force_dispatch<2>(c, true, c->fd.time);
  else
    launch_force_dispatch<3>(c, true, c->fd.time);
  NLM_HIP(c, hipGetLastError());
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_fd_postproc(nlm_ctx* c, float* Z, float* phi,
                               double* energy) {
  int rc = check_ready(c, "nlm_fd_postproc");
  if (rc) return rc;
  FdState& s = c->fd;
  if (s.material != NLM_MATERIAL_RNP_BOND)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_postproc: bond-based material only");
  DevBuf dz, dp, de;
  if ((rc = alloc(c, dz, sizeof(float) * s.n))) return rc;
  if ((rc = alloc(c, dp, sizeof(float) * s.n))) return rc;
  if ((rc = alloc(c, de, sizeof(double) * s.n))) return rc;
  unsigned g = unsigned((s.sell.nslices * 64 + 255) / 256);
  if (s.dim == 2) {
    FdArgs<2> a = make_args<2>(c);
    hipLaunchKernelGGL(k_postproc<2>, dim3(g), dim3(256), 0, c->stream, a,
                       dz.as<float>(), dp.as<float>(), de.as<double>(),
                       nullptr);
  } else {
    FdArgs<3> a = make_args<3>(c);
    hipLaunchKernelGGL(k_postproc<3>, dim3(g), dim3(256), 0, c->stream, a,
                       dz.as<float>(), dp.as<float>(), de.as<double>(),
                       nullptr);
  }
  NLM_HIP(c, hipGetLastError());
  std::vector<float> hz(s.n), hp(s.n);
  std::vector<double> he(s.n);
  NLM_HIP(c, hipMemcpyAsync(hz.data(), dz.p, sizeof(float) * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipMemcpyAsync(hp.data(), dp.p, sizeof(float) * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipMemcpyAsync(he.data(), de.p, sizeof(double) * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    if (Z) Z[i] = hz[a];
    if (phi) phi[i] = hp[a];
    if (energy) energy[i] = he[a];
  }
  return NLM_OK;
}

extern "C" int nlm_fd_postproc_noalloc(nlm_ctx* c, float* Z, float* phi,
                                       double* energy) {
  int rc = check_ready(c, "nlm_fd_postproc_noalloc");
  if (rc) return rc;
  FdState& s = c->fd;
  if (s.material != NLM_MATERIAL_RNP_BOND)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_postproc_noalloc: bond-based material only");
  if (s.n > kNoAllocMaxNodes)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_postproc_noalloc: " + std::to_string(s.n) +
                    " nodes exceed the static staging limit of " +
                    std::to_string(kNoAllocMaxNodes) +
                    " nodes; use nlm_fd_postproc");
  if (!c->scratch)
    return fail(c, NLM_ERR_STATE,
                "nlm_fd_postproc_noalloc: no staging address for the device");
  // the context's device copy of the staging array (its address from
  // hipGetSymbolAddress at context creation): Z, phi, energy regions at
  // fixed offsets, each written at every reference id 0..n-1 by the kernel
  // (perm is a bijection), so nothing needs clearing
  char* const base = static_cast<char*>(c->scratch);
  float* const dz = reinterpret_cast<float*>(base);
  float* const dp = reinterpret_cast<float*>(base + kNoAllocZBytes);
  double* const de =
      reinterpret_cast<double*>(base + kNoAllocZBytes + kNoAllocPhiBytes);
  const unsigned g = unsigned((s.sell.nslices * 64 + 255) / 256);
  const int32_t* perm = s.perm_dev.as<int32_t>();
  if (s.dim == 2)
    hipLaunchKernelGGL(k_postproc<2>, dim3(g), dim3(256), 0, c->stream,
                       make_args<2>(c), dz, dp, de, perm);
  else
    hipLaunchKernelGGL(k_postproc<3>, dim3(g), dim3(256), 0, c->stream,
                       make_args<3>(c), dz, dp, de, perm);
  NLM_HIP(c, hipGetLastError());
  if (Z)
    NLM_HIP(c, hipMemcpyAsync(Z, dz, sizeof(float) * s.n,
                              hipMemcpyDeviceToHost, c->stream));
  if (phi)
    NLM_HIP(c, hipMemcpyAsync(phi, dp, sizeof(float) * s.n,
                              hipMemcpyDeviceToHost, c->stream));
  if (energy)
    NLM_HIP(c, hipMemcpyAsync(energy, de, sizeof(double) * s.n,
                              hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_fd_external_force(nlm_ctx* c, double t, double* f_ext) {
  int rc = check_ready(c, "nlm_fd_external_force");
  if (rc) return rc;
  if (!f_ext)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_external_force: null output");
  FdState& s = c->fd;
  DevBuf d;
  if ((rc = alloc(c, d, sizeof(double) * 3 * s.n))) return rc;
  const unsigned g = unsigned((s.n + 255) / 256);
  if (s.dim == 2)
    hipLaunchKernelGGL(k_fext<2>, dim3(g), dim3(256), 0, c->stream,
                       make_args<2>(c), t, d.as<double>());
  else
    hipLaunchKernelGGL(k_fext<3>, dim3(g), dim3(256), 0, c->stream,
                       make_args<3>(c), t, d.as<double>());
  NLM_HIP(c, hipGetLastError());
  std::vector<double> h(size_t(3 * s.n));
  NLM_HIP(c, hipMemcpyAsync(h.data(), d.p, sizeof(double) * 3 * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    for (int k = 0; k < 3; k++) f_ext[3 * i + k] = h[3 * a + k];
  }
  return NLM_OK;
}

extern "C" int nlm_fd_reaction_force(nlm_ctx* c, const double* box,
                                     double horizon, double* R,
                                     double* R_len) {
  int rc = check_ready(c, "nlm_fd_reaction_force");
  if (rc) return rc;
  FdState& s = c->fd;
  if (s.material != NLM_MATERIAL_RNP_BOND)
    return fail(c, NLM_ERR_INVALID,
                "nlm_fd_reaction_force: bond-based material only");
  if (!box)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_reaction_force: null box");
  DevBuf dr, dl;
  if ((rc = alloc(c, dr, sizeof(double) * 3 * s.n))) return rc;
  if ((rc = alloc(c, dl, sizeof(double) * s.n))) return rc;
  const unsigned g = unsigned((s.sell.nslices * 64 + 255) / 256);
  if (s.dim == 2)
    hipLaunchKernelGGL(k_reaction<2>, dim3(g), dim3(256), 0, c->stream,
                       make_args<2>(c), box[0], box[1], box[2], horizon,
                       dr.as<double>(), dl.as<double>());
  else
    hipLaunchKernelGGL(k_reaction<3>, dim3(g), dim3(256), 0, c->stream,
                       make_args<3>(c), box[0], box[1], box[2], horizon,
                       dr.as<double>(), dl.as<double>());
  NLM_HIP(c, hipGetLastError());
  std::vector<double> hr(size_t(3 * s.n)), hl(size_t(s.n));
  NLM_HIP(c, hipMemcpyAsync(hr.data(), dr.p, sizeof(double) * 3 * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipMemcpyAsync(hl.data(), dl.p, sizeof(double) * s.n,
                            hipMemcpyDeviceToHost, c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t a = 0; a < s.n; a++) {
    const int64_t i = s.perm_host[a];
    if (R)
      for (int k = 0; k < 3; k++) R[3 * i + k] = hr[3 * a + k];
    if (R_len) R_len[i] = hl[a];
  }
  return NLM_OK;
}

namespace {
// Times the fused step kernel for each neighbor-slot chunk width on the
// current state.  The trials write the spare u/v buffers (overwritten by the
// next step anyway) and a scratch copy of the bond bits, so the model state
// is unchanged.
// us[k]: microseconds per step with chunk width {2, 4, 8}[k] (0 = not
// timed: the FP64-exact mode runs U = 4 only).
template <int DIM>
int time_chunks(nlm_ctx* c, double us[3]) {
  FdState& s = c->fd;
  DevBuf bits_copy;
  int rc;
  if ((rc = alloc(c, bits_copy, s.sell.bits.bytes))) return rc;
  NLM_HIP(c, hipMemcpyAsync(bits_copy.p, s.sell.bits.p, s.sell.bits.bytes,
                            hipMemcpyDeviceToDevice, c->stream));
  FdArgs<DIM> a = make_args<DIM>(c);
  a.bits = bits_copy.as<unsigned long long>();
  const int cands[3] = {2, 4, 8};
  hipEvent_t e0, e1;
  NLM_HIP(c, hipEventCreate(&e0));
  NLM_HIP(c, hipEventCreate(&e1));
  const int reps = 6;
  for (int k = 0; k < 3; k++) {
    const int U = cands[k];
    us[k] = 0.;
    if (s.precision == NLM_PREC_FP64_EXACT && U != 4) continue;
    for (int r = 0; r < 2; r++)
      launch_mode<DIM, KIND_CD, false>(c, a, s.time, s.time + s.dt, U);
    NLM_HIP(c, hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++)
      launch_mode<DIM, KIND_CD, false>(c, a, s.time, s.time + s.dt, U);
    NLM_HIP(c, hipEventRecord(e1, c->stream));
    NLM_HIP(c, hipEventSynchronize(e1));
    float ms = 0.f;
    NLM_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    us[k] = 1e3 * double(ms) / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  NLM_HIP(c, hipGetLastError());
  return NLM_OK;
}

int best_chunk(const double us[3], double* best) {
  const int cands[3] = {2, 4, 8};
  int pick = 4;
  *best = 1e300;
  for (int k = 0; k < 3; k++)
    if (us[k] > 0. && us[k] < *best) {
      *best = us[k];
      pick = cands[k];
    }
  return pick;
}

template <int DIM>
int autotune(nlm_ctx* c, int* chosen, double* best_us) {
  double us[3];
  int rc = time_chunks<DIM>(c, us);
  if (rc) return rc;
  double best;
  const int pick = best_chunk(us, &best);
  c->fd.unroll = pick;
  if (chosen) *chosen = pick;
  if (best_us) *best_us = best;
  return NLM_OK;
}

// Rebuilds the staged kernels' halos for tiles of bn nodes (128, 256 or
// 512): halo lists and halo-local entries (the 2D bond-based kernel's: LDS
// byte offsets), the tile origins, the LDS budget and, for the state-based
// FP32-force passes, the fixed-point quantum (the origins' range changes
// with the tile).  NLM_ERR_INVALID (state unchanged) when the largest halo
// does not fit the LDS budget.
int retile(nlm_ctx* c, int bn) {
  FdState& s = c->fd;
  if (bn == s.tile) return NLM_OK;
  const bool es = s.material == NLM_MATERIAL_ELASTIC_STATE;
  const int64_t n = s.n, ns = s.sell.nslices, nent = s.sell.nentries;
  std::vector<int64_t> slp(ns + 1);
  std::vector<int32_t> colS(nent);
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  NLM_HIP(c, hipMemcpy(slp.data(), s.sell.sl_ptr.p, sizeof(int64_t) * (ns + 1),
                       hipMemcpyDeviceToHost));
  NLM_HIP(c, hipMemcpy(colS.data(), s.sell.col.p, sizeof(int32_t) * nent,
                       hipMemcpyDeviceToHost));
  std::vector<int64_t> hp;
  std::vector<int32_t> halo;
  std::vector<uint16_t> colL;
  int64_t hmax = 0;
  build_halos(n, ns, slp.data(), colS.data(), bn / 64, hp, halo, colL, hmax);
  if (!es) {
    if ((hmax + 1) * int64_t(sizeof(HaloRec)) > 64 * 1024)
      return fail(c, NLM_ERR_INVALID,
                  "tile of " + std::to_string(bn) +
                      " nodes: its largest halo exceeds the LDS budget");
    std::vector<uint16_t> rl(static_cast<size_t>(n));
    NLM_HIP(c, hipMemcpy(rl.data(), s.sell.rowlen.p, sizeof(uint16_t) * n,
                         hipMemcpyDeviceToHost));
    colL = staged_entries(colL, s.precision, slp.data(), ns, rl.data(), n);
  } else {
    int lds_max = 0;
    NLM_HIP(c, hipDeviceGetAttribute(&lds_max,
                                     hipDeviceAttributeMaxSharedMemoryPerBlock,
                                     c->device));
    if (hmax > 65535 ||
        es_halo_lds_bytes(s.precision, hmax) > size_t(lds_max))
      return fail(c, NLM_ERR_INVALID,
                  "tile of " + std::to_string(bn) +
                      " nodes: its largest halo exceeds the LDS budget");
  }
  // origins: the centre of each block's bounding box
  const int dim = s.dim == 3 ? 3 : 2;
  const int stride = dim == 2 ? 2 : 4;
  const int64_t nb = int64_t(hp.size()) - 1;
  std::vector<double> hx(size_t(n) * stride);
  NLM_HIP(c, hipMemcpy(hx.data(), s.x.p, sizeof(double) * stride * size_t(n),
                       hipMemcpyDeviceToHost));
  std::vector<double> borg(size_t(nb) * stride, 0.);
  double org_range = 0.;
  for (int64_t b = 0; b < nb; b++) {
    const int64_t a0 = int64_t(bn) * b, a1 = std::min<int64_t>(n, a0 + bn);
    if (a0 >= n) break;
    double lo[3], hi[3];
    for (int d = 0; d < dim; d++) lo[d] = hi[d] = hx[stride * a0 + d];
    for (int64_t a = a0; a < a1; a++)
      for (int d = 0; d < dim; d++) {
        lo[d] = std::min(lo[d], hx[stride * a + d]);
        hi[d] = std::max(hi[d], hx[stride * a + d]);
      }
    for (int d = 0; d < dim; d++) borg[stride * b + d] = 0.5 * (lo[d] + hi[d]);
    auto span = [&](int64_t a) {
      for (int d = 0; d < dim; d++)
        org_range = std::max(
            org_range, std::fabs(hx[stride * a + d] - borg[stride * b + d]));
    };
    for (int64_t a = a0; a < a1; a++) span(a);
    for (int64_t t = hp[b]; t < hp[b + 1]; t++) span(halo[size_t(t)]);
  }
  // a captured step graph references the buffers replaced below
  if (s.graph) {
    (void)hipGraphExecDestroy(s.graph);
    s.graph = nullptr;
  }
  s.graph_chunk = 0;
  int rc;
  auto up = [&](DevBuf& d, const void* src, size_t bytes) -> int {
    if (!d.p || d.bytes < bytes + 256)
      if (int r = alloc(c, d, bytes + 256)) return r;
    NLM_HIP(c, hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
    return NLM_OK;
  };
  if ((rc = up(s.halo_ptr, hp.data(), sizeof(int64_t) * hp.size())) ||
      (rc = up(s.halo, halo.data(), sizeof(int32_t) * halo.size())) ||
      (rc = up(s.colL, colL.data(), sizeof(uint16_t) * nent)) ||
      (rc = up(s.blk_org, borg.data(), sizeof(double) * borg.size())))
    return rc;
  s.halo_max = hmax;
  s.tile = bn;
  if (es) es_set_quantum(s.es, org_range);
  return NLM_OK;
}

}  // namespace

extern "C" int nlm_fd_autotune(nlm_ctx* c, int* chosen, double* best_us) {
  int rc = check_ready(c, "nlm_fd_autotune");
  if (rc) return rc;
  c->fd.graph_chunk = 0;  // a captured step graph is stale
  if (c->fd.precision == NLM_PREC_FP64_EXACT) {
    c->fd.unroll = 4;
    if (chosen) *chosen = 4;
    if (best_us) *best_us = 0.;
    return NLM_OK;
  }
  if (c->fd.material == NLM_MATERIAL_ELASTIC_STATE)
    return es_autotune(c, chosen, best_us);
  return c->fd.dim == 2 ? autotune<2>(c, chosen, best_us)
                        : autotune<3>(c, chosen, best_us);
}

extern "C" int nlm_fd_autotune_tiles(nlm_ctx* c, int* tile, int* chunk,
                                     double* us) {
  int rc = check_ready(c, "nlm_fd_autotune_tiles");
  if (rc) return rc;
  FdState& s = c->fd;
  s.graph_chunk = 0;  // a captured step graph is stale
  const int tiles[3] = {128, 256, 512};
  double tab[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  const bool es = s.material == NLM_MATERIAL_ELASTIC_STATE;
  const bool staged = es ? (s.es_tiled && s.use_tiled && s.dim == 3)
                         : (s.tiled && s.use_tiled && s.dim == 2);
  if (!staged) {
    // no staged kernel: the chunk width alone
    int pick = 4;
    double best = 0.;
    if ((rc = nlm_fd_autotune(c, &pick, &best))) return rc;
    tab[3 + (pick == 2 ? 0 : pick == 4 ? 1 : 2)] = best;
    if (tile) *tile = s.tile;
    if (chunk) *chunk = pick;
    if (us) std::memcpy(us, tab, sizeof(tab));
    return NLM_OK;
  }
  double best = 1e300;
  int bt = s.tile, bu = s.unroll;
  for (int k = 0; k < 3; k++) {
    if ((rc = retile(c, tiles[k]))) {
      if (rc != NLM_ERR_INVALID) return rc;
      c->err.clear();  // over the LDS budget: not a candidate
      continue;
    }
    if (es) {
      int pick = 4;
      double bu = 0.;
      if ((rc = es_autotune(c, &pick, &bu, tab + 3 * k))) return rc;
    } else if ((rc = time_chunks<2>(c, tab + 3 * k))) {
      return rc;
    }
    double b;
    const int u = best_chunk(tab + 3 * k, &b);
    if (b < best) {
      best = b;
      bt = tiles[k];
      bu = u;
    }
  }
  if ((rc = retile(c, bt))) return rc;
  s.unroll = bu;
  if (tile) *tile = bt;
  if (chunk) *chunk = bu;
  if (us) std::memcpy(us, tab, sizeof(tab));
  return NLM_OK;
}

extern "C" int nlm_fd_set_tile(nlm_ctx* c, int tile) {
  int rc = check_ready(c, "nlm_fd_set_tile");
  if (rc) return rc;
  if (tile != 128 && tile != 256 && tile != 512)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_set_tile: 128, 256 or 512");
  const bool es = c->fd.material == NLM_MATERIAL_ELASTIC_STATE;
  if (es ? !(c->fd.es_tiled && c->fd.dim == 3)
         : !(c->fd.tiled && c->fd.dim == 2))
    return fail(c, NLM_ERR_STATE,
                "nlm_fd_set_tile: the model has no LDS-staged step kernel");
  c->fd.graph_chunk = 0;
  return retile(c, tile);
}

extern "C" int nlm_fd_set_unroll(nlm_ctx* c, int unroll) {
  int rc = check_ready(c, "nlm_fd_set_unroll");
  if (rc) return rc;
  if (unroll != 2 && unroll != 4 && unroll != 8)
    return fail(c, NLM_ERR_INVALID, "nlm_fd_set_unroll: 2, 4 or 8");
  c->fd.unroll = unroll;
  c->fd.graph_chunk = 0;  // a captured step graph is stale
  return NLM_OK;
}

extern "C" int nlm_fd_num_bonds(nlm_ctx* c, int64_t* nnz) {
  int rc = check_ready(c, "nlm_fd_num_bonds");
  if (rc) return rc;
  *nnz = c->fd.nnz;
  return NLM_OK;
}

extern "C" const char* nlm_fd_kernel_name(nlm_ctx* c) {
  if (!c || !c->fd.ready) return "";
  if (c->fd.material == NLM_MATERIAL_ELASTIC_STATE) return "k_es_force";
  return "k_fd";
}

// Node-list displacement BCs: flags field kNodeListBc and per-node rates.
extern "C" int nlm_fd_set_node_ubc(nlm_ctx* c, int64_t count,
                                   const int64_t* nodes, const uint8_t* dofs,
                                   const double* rates) {
  int rc = check_ready(c, "nlm_fd_set_node_ubc");
  if (rc) return rc;
  FdState& s = c->fd;
  if (count < 0 || (count > 0 && (!nodes || !dofs || !rates)))
    return fail(c, NLM_ERR_INVALID, "nlm_fd_set_node_ubc: bad arguments");
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<uint32_t> fl(s.n);
  NLM_HIP(c, hipMemcpy(fl.data(), s.flags.p, sizeof(uint32_t) * s.n,
                       hipMemcpyDeviceToHost));
  std::vector<double> rate(3 * size_t(s.n), 0.);
  // drop earlier node-list BCs
  for (int64_t a = 0; a < s.n; a++)
    for (int d = 0; d < 3; d++)
      if (((fl[a] >> (kFlagUbcShift + 4 * d)) & 0xfu) == kNodeListBc) {
        fl[a] &= ~(0xfu << (kFlagUbcShift + 4 * d));
        s.fix_host[s.perm_host[a]] &= uint8_t(~(1u << d));
      }
  for (int64_t k = 0; k < count; k++) {
    const int64_t i = nodes[k];
    if (i < 0 || i >= s.n)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_set_node_ubc: node out of range");
    if (dofs[k] >> s.udim)
      return fail(c, NLM_ERR_INVALID, "nlm_fd_set_node_ubc: bad dof mask");
    const int64_t a = s.inv_host[i];
    for (int d = 0; d < s.dim; d++) {
      if (!((dofs[k] >> d) & 1)) continue;
      fl[a] = (fl[a] & ~(0xfu << (kFlagUbcShift + 4 * d))) |
              (kNodeListBc << (kFlagUbcShift + 4 * d));
      rate[3 * a + d] = rates[3 * k + d];
      s.fix_host[i] |= uint8_t(1u << d);
    }
  }
  pin_y_dim1(s, fl);
  NLM_HIP(c, hipMemcpy(s.flags.p, fl.data(), sizeof(uint32_t) * s.n,
                       hipMemcpyHostToDevice));
  NLM_HIP(c, hipMemcpy(s.node_rate.p, rate.data(), sizeof(double) * rate.size(),
                       hipMemcpyHostToDevice));
  return NLM_OK;
}

namespace {
// t_force <- t_next; t_next <- t_next + dt (the host's accumulation, on the
// device, between graph-captured steps).
__global__ void k_time_advance(double* __restrict__ t, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    t[0] = t[1];
    t[1] = t[1] + dt;
  }
}

template <int DIM>
void launch_fused_graph(nlm_ctx* c) {
  FdState& s = c->fd;
  if (s.material == NLM_MATERIAL_ELASTIC_STATE) {
    es_launch_fused_dev(c, s.tdev.as<double>());
    return;
  }
  FdArgs<DIM> a = make_args<DIM>(c);
  a.tdev = s.tdev.as<double>();
  launch_mode<DIM, KIND_CD, false>(c, a, 0., 0., s.unroll);
}

// Captures `chunk` fused steps (even, ping-pong buffers return to the
// starting parity) into an executable graph.
template <int DIM>
int capture_steps(nlm_ctx* c, int chunk) {
  FdState& s = c->fd;
  if (s.graph) {
    (void)hipGraphExecDestroy(s.graph);
    s.graph = nullptr;
  }
  hipGraph_t gr = nullptr;
  NLM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int cur0 = s.cur;
  for (int k = 0; k < chunk; k++) {
    launch_fused_graph<DIM>(c);
    hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(64), 0, c->stream,
                       s.tdev.as<double>(), s.dt);
    s.cur ^= 1;
  }
  s.cur = cur0;
  NLM_HIP(c, hipStreamEndCapture(c->stream, &gr));
  hipError_t e = hipGraphInstantiate(&s.graph, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) {
    s.graph = nullptr;
    return fail(c, NLM_ERR_HIP,
                std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  s.graph_chunk = chunk;
  s.graph_cur = cur0;
  return NLM_OK;
}

// nsteps of integrate() with the fused steps replayed from a graph: the first
// update from the stored force, (nsteps - 1) fused steps (graph chunks, the
// remainder launched directly), the trailing force.
template <int DIM>
int do_steps_graph(nlm_ctx* c, int nsteps, int chunk, double* ms_fused) {
  FdState& s = c->fd;
  const unsigned nb = node_blocks(s);
  {
    double tn = s.time + s.dt;
    FdArgs<DIM> a = make_args<DIM>(c);
    hipLaunchKernelGGL(k_update_cd<DIM>, dim3(nb), dim3(256), 0, c->stream, a,
                       tn);
    s.cur ^= 1;
    s.time = tn;
    s.step++;
  }
  int rc;
  const int fused = nsteps - 1;
  const int reps = fused / chunk;
  if (reps > 0 && (!s.graph || s.graph_chunk != chunk || s.graph_cur != s.cur))
    if ((rc = capture_steps<DIM>(c, chunk))) return rc;
  // device times for the first graph step
  double t2[2] = {s.time, s.time + s.dt};
  NLM_HIP(c, hipMemcpyAsync(s.tdev.p, t2, sizeof(t2), hipMemcpyHostToDevice,
                            c->stream));
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  if (ms_fused) NLM_HIP(c, hipEventRecord(s.ev0, c->stream));
  for (int r = 0; r < reps; r++) {
    NLM_HIP(c, hipGraphLaunch(s.graph, c->stream));
    for (int k = 0; k < chunk; k++) {  // the host keeps the same clock
      s.time = s.time + s.dt;
      s.step++;
    }
  }
  if (ms_fused) NLM_HIP(c, hipEventRecord(s.ev1, c->stream));
  for (int k = reps * chunk; k < fused; k++) {
    double tn = s.time + s.dt;
    launch_fused_dispatch<DIM>(c, s.time, tn);
    s.cur ^= 1;
    s.time = tn;
    s.step++;
  }
  launch_force_dispatch<DIM>(c, false, s.time);
  NLM_HIP(c, hipGetLastError());
  if (ms_fused) {
    NLM_HIP(c, hipEventSynchronize(s.ev1));
    float ms = 0.f;
    NLM_HIP(c, hipEventElapsedTime(&ms, s.ev0, s.ev1));
    *ms_fused = reps > 0 ? double(ms) / double(reps * chunk) : 0.;
  }
  return NLM_OK;
}

int graph_check(nlm_ctx* c, int nsteps, int chunk, const char* fn) {
  int rc = check_ready(c, fn);
  if (rc) return rc;
  FdState& s = c->fd;
  if (nsteps < 1 || chunk < 2 || (chunk & 1))
    return fail(c, NLM_ERR_INVALID,
                std::string(fn) + ": nsteps >= 1, even chunk >= 2");
  if (s.scheme != NLM_TIME_CENTRAL_DIFFERENCE)
    return fail(c, NLM_ERR_INVALID, std::string(fn) + ": central difference");
  for (size_t k = 0; k < s.crack_done.size(); k++)
    if (!s.crack_done[k])
      return fail(c, NLM_ERR_STATE,
                  std::string(fn) + ": pending timed cracks need nlm_fd_step");
  return NLM_OK;
}
}  // namespace

extern "C" int nlm_fd_step_graph(nlm_ctx* c, int nsteps, int chunk) {
  int rc = graph_check(c, nsteps, chunk, "nlm_fd_step_graph");
  if (rc) return rc;
  rc = c->fd.dim == 2 ? do_steps_graph<2>(c, nsteps, chunk, nullptr)
                      : do_steps_graph<3>(c, nsteps, chunk, nullptr);
  if (rc) return rc;
  NLM_HIP(c, hipStreamSynchronize(c->stream));
  return NLM_OK;
}

extern "C" int nlm_fd_timed_steps_graph(nlm_ctx* c, int nsteps, int chunk,
                                        double* ms_total, double* ms_per_step) {
  int rc = graph_check(c, nsteps, chunk, "nlm_fd_timed_steps_graph");
  if (rc) return rc;
  hipEvent_t a0, a1;
  NLM_HIP(c, hipEventCreate(&a0));
  NLM_HIP(c, hipEventCreate(&a1));
  NLM_HIP(c, hipEventRecord(a0, c->stream));
  double per = 0.;
  rc = c->fd.dim == 2 ? do_steps_graph<2>(c, nsteps, chunk, &per)
                      : do_steps_graph<3>(c, nsteps, chunk, &per);
  if (rc) return rc;
  NLM_HIP(c, hipEventRecord(a1, c->stream));
  NLM_HIP(c, hipEventSynchronize(a1));
  float ms = 0.f;
  NLM_HIP(c, hipEventElapsedTime(&ms, a0, a1));
  (void)hipEventDestroy(a0);
  (void)hipEventDestroy(a1);
  if (ms_total) *ms_total = ms;
  if (ms_per_step) *ms_per_step = per;
  return NLM_OK;
}
