// nlm_internal.h -- context and device-side data of the nlmech_amd library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nlmech_amd.h"

namespace nlm {

// Raw device buffer owned by a context (freed in the destructor).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Bits of the per-node flag word.
constexpr uint32_t kFlagUbcShift = 0;         // 4 bits per dof: u-BC set id + 1
constexpr uint32_t kFlagNonInterior = 1u << 12;  // node never breaks bonds
constexpr uint32_t kFlagFbcShift = 16;        // 8 bits: force-BC set membership
constexpr int kMaxUbcSets = 14;
constexpr uint32_t kNodeListBc = 15;  // u-BC id of node-list BCs (per-node rate)
constexpr int kMaxFbcSets = 8;

// One BC set as evaluated on the device (inp::BCData with codes).
struct BcSetDev {
  int time_fn;
  int space_fn;
  int dirmask;  // bit d set: dof d (0-based) is loaded
  int n_space_params;
  double tp[4];
  double sp[4];
  double x1, x2, y1, y2, z1, z2;
  double h;  // mesh size (line_load)
};

struct BcTable {
  int nu;
  int nf;
  int pad[2];
  BcSetDev u[kMaxUbcSets];
  BcSetDev f[kMaxFbcSets];
};

// RNPBond constants as used by the kernels.
struct RnpDev {
  double C, beta, rbar, inv_factor, factor_sc;
  double horizon, h, check_up, check_low;
  double inf_a, inf_b;
  int inf_type;
  int irreversible;
  int contact;
  int pad;
  double kn, rc;
  // derived for the FP64/FP32 fast paths
  double rcp_inv_factor;   // 1 / inv_factor
  double rcp_h;            // 1 / h
  double rcp_horizon;      // 1 / horizon
  double krbar_sq;         // (factor_sc * rbar)^2
  double rcp_krbar_sq;     // 1 / krbar_sq
  double c4;               // 4 * C * beta / inv_factor
  double low_band;         // check_low * (1 + 1e-5): volume-correction edge
  double rcp_inf_b;        // 1 / inf_b (gaussian influence)
  float c4f, beta_f, check_up_f, low_band_f, rcp_h_f, rcp_horizon_f;
  float inf_a_f, inf_b_f, rcp_inf_b_f, rcp_krbar_sq_f;
  float nbeta_log2e_f;     // -beta * log2(e): exp(-beta t) = exp2(t * this)
  float c4f_infl;          // c4 * J for the constant influence, else c4
  float break_rs2_f;       // r S^2 at which |S| / (k rbar / sqrt(r)) reaches
                           // 1 / (1 - 1e-5): krbar_sq / (1 - 1e-5)^2
  float break_guard_f;     // half-width of the exact-decision band around it
  float up_rcp_h_f;        // check_up / h: fac = min(up/h - r/h, 1)
  float vex_tol_f;         // half-width of the volume-correction edge band
  // fast paths: the break threshold of rho^2 (+inf when bonds never break,
  // which removes the break test) and an all-ones mask when contact forces
  // are on (else 0)
  double break_rho2;
  unsigned long long contact_mask;
  // bounds of the break test's guard band: on rho^2 (FP64 path) and on
  // r S^2 (FP32 path); +inf when bonds never break
  double break_lo, break_hi;
  float break_lo_f, break_hi_f;
  // FP32-force staged kernels: displacements are held as tau u (tau a power
  // of two, so tau u is exact) and block-relative positions as sigma x with
  // tau^2 / sigma = beta log2(e), so the product (tau S / sigma)(tau S r)
  // they form is beta log2(e) r S^2: the exponent of exp2 and the break
  // ratio at once
  double tau;              // power of two nearest sqrt(beta log2 e)
  double sigma;            // tau^2 / (beta log2 e), in [1/2, 2]
  float c4f_infl_s;        // c4f_infl sigma / tau
  float rcp_h_s;           // 1 / (sigma h)
  float rcp_sigma_f;       // 1 / sigma
  float break_lo_s;        // break_lo_f times beta log2 e
  float low_dd_f;          // (sigma low_band)^2
  float vex_dd_tol_f;      // half-width of the edge band on (sigma r)^2
};

// ElasticState constants as used by the kernels.
struct EsDev {
  double K, mu, f2d;      // bulk and shear modulus, factor2D
  double Kc;              // 2D: alpha_s = (9/m) Kc (ElasticState.cpp:201-212)
  double horizon, h;
  double check_low;       // horizon - h/2 (volumeCorrection.cpp:40-42)
  double check_up;        // horizon + h/2
  double rcp_h;
  double low_band;        // check_low (1 + 1e-5): definitelyGreaterThan edge
  double sc;              // critical stretch
  double sc_break;        // sc / (1 - 1e-5): definitelyGreaterThan edge
  int breaking;           // sc > 0
  int pad;
  float check_up_f, low_band_f, rcp_h_f, sc_break_f;
  // FP32 pass 1: half-width of the volume-correction edge band; lower bound
  // of the break test's guard band on e/|X| (+inf when bonds never break;
  // stretches at or above it are decided by the exact evaluation)
  float vex_tol_f, brk_lo_f;
  // the same in FP64 for the FP64 mode
  double vex_tol, brk_lo;
  // FP32-force staged passes (3D): positions are fixed-point offsets from
  // the block's origin in units of the quantum q (a power of two), and the
  // bond arithmetic runs in units of q: 1/q, q^2 (dilatation factor) and q
  // (force coefficient) scale it back; lengths of the edge tests in q
  double q, inv_q;
  float q2_f, q_f;
  float check_up_qf, low_band_qf, rcp_h_qf, vex_tol_qf;
  // 1D model on the 2D kernels: dilatation (1/m_i) sum |X| e c V (f2d is
  // then 1/2, so (2/m_i) f2d = 1/m_i exactly) and t = (E/m_i) e
  // (ElasticState.cpp:187-193), i.e. no dilatation term in the force
  int one_d;
  // check_up / h (FP32): the clamped volume-correction factor as one
  // multiply-add, fma(|X|, -1/h, check_up/h), in either unit
  float up_rcp_h_qf;
  double E;
};

// Slice layout of the neighbor lists: 64 consecutive nodes (one wavefront)
// share a slice; entry (k, lane) of slice s sits at col[sl_ptr[s] + 64k +
// lane]; the broken bits of slot k of slice s form one 64-bit word
// bits[sl_ptr[s]/64 + k], bit `lane`.
struct SellDev {
  int64_t nslices = 0;
  int64_t nentries = 0;  // padded entries (multiple of 64)
  DevBuf sl_ptr;         // int64[nslices+1]
  DevBuf col;            // int32[nentries]
  DevBuf rowlen;         // uint16[n]
  DevBuf bits;           // uint64[nentries/64]
};

struct FdState {
  bool ready = false;
  int material = NLM_MATERIAL_RNP_BOND;
  // internal dimension of the kernels (2 or 3); a 1D model (udim 1) runs as
  // a 2D one with every node on y = 0 and every y dof held at zero
  int dim = 2;
  int udim = 2;  // the model's own dimension (1, 2 or 3)
  bool dim1 = false;
  int64_t n = 0;
  int64_t nnz = 0;
  int precision = NLM_PREC_FP64_EXACT;
  int scheme = NLM_TIME_CENTRAL_DIFFERENCE;
  double dt = 0.;
  double density = 1.;
  double time = 0.;
  int64_t step = 0;
  bool has_fbc = false;
  bool has_nofail = false;
  nlm_rnp_params mat{};
  RnpDev rnp{};
  nlm_es_params es_par{};
  EsDev es{};
  // state-based material: weighted volume m_i, dilatation theta_i and the
  // per-node force coefficients of the last pass 1 (internal order)
  DevBuf es_mw, es_theta, es_P, es_Q;
  // node data; 2D: double2, 3D: double4 (w unused)
  DevBuf x, vol, flags;
  DevBuf u[2], v[2], f;
  int cur = 0;  // which of u[]/v[] holds the current state
  int unroll = 4;  // neighbor slots per chunk in the fused step kernel
  int tile = 256;  // nodes per workgroup of the LDS-staged kernel (its halo
                   // block): 128, 256 or 512
  DevBuf row_ptr, col;  // CSR in internal ids (int64, int32)
  SellDev sell;
  // internal node order (see layout.cpp): perm[a] = reference id of a
  std::vector<int32_t> perm_host, inv_host;
  DevBuf ref_off;   // int64 per internal node: row start in reference CSR
  DevBuf perm_dev;  // int32 per internal node: its reference id
  DevBuf halo_ptr;  // int64 per 256-node block
  DevBuf halo;      // int32 internal ids
  DevBuf blk_org;   // double2 per block: origin of block-relative positions
  DevBuf vexw;      // uint64 per (slice, slot): bonds whose length lies in
                    // the volume-correction edge band (exact evaluation)
  DevBuf colL;      // uint16 halo-local slice entries (bond-based staged
                    // kernel: LDS byte offsets of the halo records)
  int64_t halo_max = 0;
  bool tiled = false;   // LDS-staged step kernel usable
  bool es_tiled = false;  // LDS-staged state-based passes usable
  bool use_tiled = true;
  DevBuf node_rate;  // double[3n] internal order: node-list BC rates
  DevBuf tdev;       // (t_force, t_next) read by graph-captured steps
  hipGraphExec_t graph = nullptr;
  int graph_chunk = 0;
  int graph_cur = -1;  // value of cur when the graph was captured
  DevBuf bctab;  // BcTable
  BcTable bc_host{};
  std::vector<uint8_t> fix_host;
  std::vector<nlm_crack> timed_cracks;
  std::vector<bool> crack_done;
  std::vector<int64_t> row_ptr_host;  // for host-side conversions
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// Quasi-static solver data (2D): tangent and preconditioner blocks in the
// slice layout (double4 = row-major 2x2 per slice entry), per-node diagonal
// blocks, Krylov vectors (double2 per node, internal order) and the
// per-block partial sums and per-iteration scalars of BiCGSTAB.
struct QsState {
  bool assembled = false;
  bool has_m = false;
  int tangent = 0;
  DevBuf blk, diag, dinv;    // A = -J blocks, diagonal blocks, Jacobi 1/a_dd
  // analytic bond tangent: one scalar per slot, -B_ij = g dx dx^T (bs_mode
  // 1) or its diagonal (bs_mode 2, component tangent); 0 = blocks only
  DevBuf bg;
  int bs_mode = 0;
  bool bond_scalar = true;   // products of the tangent from bg
  DevBuf mblk, mdiag;        // preconditioner blocks (FP32)
  DevBuf mhalf;              // BF16 copy of mblk read by the solver
  bool mhalf_ok = false;     // mhalf matches mblk
  DevBuf r, rh, p, v, s, t, x, ph, sh, b, y, dup, u0;
  DevBuf part;               // 4 * nb partial sums
  DevBuf scal;               // per-iteration scalars (rho, alpha, omega, rr)
  DevBuf ctl;                // int control words (done flag, iteration)
  int64_t nb = 0;            // 256-node blocks
  int max_iter_alloc = 0;
  // two-level: aggregate adjacency, dense coarse matrix and inverse, coarse
  // restriction / correction vectors
  DevBuf adj_ptr, adj, cA, cB, rc, zcp, zcs;
  double* cinv = nullptr;
  int m = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;  // BiCGSTAB batches in flight
  int* hctl = nullptr;       // pinned: control words of two batches
  // finite-difference tangent: node colors (no two nodes of one color
  // within 4 horizons), perturbation factor (eps = factor * h)
  DevBuf color;
  int ncolors = 0;
  double fd_perturbation = 1e-7;
  // BiCGSTAB iterations run in batches replayed from a captured HIP graph;
  // the key holds the kernel arguments of the capture
  bool use_graph = true;
  hipGraphExec_t bicg_graph = nullptr;
  std::vector<unsigned char> bicg_key;
};

// Graph-network preconditioner: FP32 master weights, gradient, Adam
// moments, BF16 copies of the layer matrices (forward: transposed [64][K],
// backward: [K][64]) and the activations of the last forward pass.
struct GnnState {
  bool ready = false;
  int64_t np = 0;            // parameters
  DevBuf w, grad, am, av;    // FP32
  DevBuf wfwd, wbwd;         // BF16 copies
  int64_t adam_t = 0;
  int64_t n = 0, ne = 0;     // nodes, slice entries of the sized buffers
  DevBuf xn, xe;             // features (BF16)
  DevBuf h, g, agg;          // per-layer node / edge states (BF16)
  DevBuf P, Q;               // FP32 node projections
  DevBuf scale, isq, deg;    // node diagonal scale, its 1/sqrt, degree
  DevBuf erow, rev, valid;   // per slice entry: node, reverse entry, mask
  DevBuf dh, dh16, dg, dg2, dpre, dP, dQ, dPQ16, dagg;  // backward
  DevBuf part;               // partial sums of weight gradients
  DevBuf vr, ve, vz;         // training vectors
  DevBuf dme, dmn;           // dL/d(decoder outputs)
  DevBuf lpart;              // loss partials
};

struct NbrState {
  bool ready = false;
  int64_t n = 0;
  int64_t nnz = 0;
  DevBuf row_ptr;  // int64[n+1]
  DevBuf col;      // int32[nnz]
};

}  // namespace nlm

struct nlm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // the device's copy of the static staging array of the allocation-free
  // entry point (hipGetSymbolAddress at creation; fd.hip)
  void* scratch = nullptr;
  std::string err;
  std::string arch;
  nlm::NbrState nbr;
  nlm::FdState fd;
  nlm::QsState qs;
  nlm::GnnState gnn;
};

namespace nlm {

// Static staging of nlm_fd_postproc_noalloc (fd.hip): Z and phi (FP32) and
// the strain energy (FP64) of at most kNoAllocMaxNodes nodes.
constexpr int64_t kNoAllocMaxNodes = int64_t(1) << 20;
constexpr size_t kNoAllocZBytes = sizeof(float) * size_t(kNoAllocMaxNodes);
constexpr size_t kNoAllocPhiBytes = sizeof(float) * size_t(kNoAllocMaxNodes);
constexpr size_t kNoAllocEBytes = sizeof(double) * size_t(kNoAllocMaxNodes);
constexpr size_t kNoAllocBytes =
    kNoAllocZBytes + kNoAllocPhiBytes + kNoAllocEBytes;
constexpr int kMaxDevices = 64;
// Address of this device's staging array (once per device, with the device
// current); *out stays valid until the process ends.
int scratch_setup(int device, void** out, std::string& err);

inline int fail(nlm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define NLM_HIP(ctx, expr)                                                   \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::nlm::fail((ctx), NLM_ERR_HIP,                                 \
                         std::string(#expr ": ") + hipGetErrorString(e_));   \
  } while (0)

inline int alloc(nlm_ctx* c, DevBuf& b, size_t bytes) {
  b.release();
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(c, NLM_ERR_MEMORY,
                std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  b.bytes = bytes;
  return NLM_OK;
}

// Grid helper: ceil-divide and pad to a multiple of 8 for the XCD remap.
inline unsigned grid_xcd(int64_t items, int block) {
  int64_t g = (items + block - 1) / block;
  if (g < 1) g = 1;
  return unsigned((g + 7) / 8 * 8);
}


}  // namespace nlm
