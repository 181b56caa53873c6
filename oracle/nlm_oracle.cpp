// nlm_oracle.cpp -- CPU restatement of NLMech's explicit-dynamics hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the nlmech_amd HIP library.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Each function follows the reference function cited next to it, in the same
// floating-point operation order (Point3 arithmetic is kept 3-component, the z
// component being 0 in 2D).  Build with -ffp-contract=off (see Makefile): no
// fused multiply-add, IEEE double everywhere, libm exp/sqrt/sin/cos/pow.
//
// Parity status of each part (how it is pinned) is listed in tests/ and in the
// module docstrings of tests/test_oracle_*.py.

#include "nlm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// src/util/compare.h:13
constexpr double kCompareEps = 1e-5;

// util::compare::definitelyGreaterThan, src/util/compare.cpp:25-28
inline bool dgt(double a, double b) {
  return (a - b) >
         ((std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a)) * kCompareEps);
}
// util::compare::definitelyLessThan, src/util/compare.cpp:30-33
inline bool dlt(double a, double b) {
  return (b - a) >
         ((std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a)) * kCompareEps);
}

// util::Point3, src/util/point.h (3-component arithmetic, left-to-right)
struct P3 {
  double x = 0., y = 0., z = 0.;
};
inline P3 p3(const double* a) { return P3{a[0], a[1], a[2]}; }
inline P3 operator+(P3 a, const P3& b) {  // point.h:287 (+= :430)
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  return a;
}
inline P3 operator-(P3 a, const P3& b) {  // point.h:298 (-= :443)
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
  return a;
}
inline P3 operator*(double s, P3 a) {  // point.h:370 (*= :417)
  a.x *= s;
  a.y *= s;
  a.z *= s;
  return a;
}
inline P3 operator/(P3 a, double s) {  // point.h:381 (/= :469)
  a.x /= s;
  a.y /= s;
  a.z /= s;
  return a;
}
inline double dot(const P3& a, const P3& b) {  // point.h:157
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
inline double length(const P3& a) {  // point.h:119
  return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
inline double dist(const P3& a, const P3& b) {  // point.h:165-169
  return std::sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) +
                   (a.z - b.z) * (a.z - b.z));
}

// util::transformation::rotateCW2D, src/util/transformation.cpp:20-24
inline P3 rotate_cw_2d(const P3& x, double theta) {
  return P3{x.x * std::cos(theta) + x.y * std::sin(theta),
            -x.x * std::sin(theta) + x.y * std::cos(theta), 0.0};
}

// ---------------------------------------------------------------------------
// Influence functions, src/material/pd/influenceFn.h:85-408
// ---------------------------------------------------------------------------
struct Influence {
  int type = 0;
  double a = 0., b = 0.;
  double value(double r) const {
    if (type == 0) return a;                         // :104-106
    if (type == 1) return a + b * r;                 // LinearInfluenceFn
    return a * std::exp(-r * r / b);                 // GaussianInfluenceFn
  }
  double moment(size_t i) const {
    if (type == 0) return a / double(i + 1);         // :117-119
    if (type == 1) return (a / double(i + 1)) + (b / double(i + 2));
    double sq1 = std::sqrt(b);
    double sq2 = std::sqrt(M_PI);
    if (i == 0) return 0.5 * a * sq1 * sq2 * std::erf(1. / sq1);
    if (i == 1) return 0.5 * a * b * (1. - std::exp(-1. / b));
    if (i == 2)
      return 0.5 * a * b * sq1 *
             (0.5 * sq2 * std::erf(1. / sq1) - (1. / sq1) * std::exp(-1. / b));
    return 0.5 * a * b * b * (1. - (1. + 1. / b) * std::exp(-1. / b));
  }
};

Influence make_influence(int type, int np, const double* p, int dim) {
  Influence f;
  f.type = type;
  if (type == 0) {
    f.a = np == 0 ? double(dim + 1) : p[0];  // influenceFn.h:95
  } else if (type == 1) {
    if (np == 0) {
      f.a = dim == 1 ? 6. : (dim == 2 ? 12. : 20.);
      f.b = -f.a;
    } else {
      f.a = p[0];
      f.b = np < 2 ? -f.a : p[1];
    }
  } else {
    if (np == 0) {
      f.b = 0.2;
      double beta = f.b;
      if (dim == 1)
        f.a = 2. / (beta * (1. - std::exp(-1. / beta)));
      else if (dim == 2)
        f.a = (4.0 / beta) * 1.0 /
              (std::sqrt(M_PI * beta) * std::erf(1.0 / std::sqrt(beta)) -
               2.0 * std::exp(-1.0 / beta));
      else
        f.a = (2.0 / beta) * 1.0 / (beta - (beta + 1.) * std::exp(-1.0 / beta));
    } else {
      f.a = p[0];
      f.b = p[1];
    }
  }
  return f;
}

Influence influence_of(const ora_rnp& m) {
  Influence f;
  f.type = m.inf_type;
  f.a = m.inf_a;
  f.b = m.inf_b;
  return f;
}

// ---------------------------------------------------------------------------
// RNPBond::getBondEF, src/material/pd/rnpBond.cpp:216-301
// ---------------------------------------------------------------------------
inline void bond_ef(const ora_rnp& m, const Influence& J, const P3& xi,
                    const P3& ui, const P3& xj, const P3& uj, double voli,
                    double volj, bool interior_i, bool interior_j, bool& fs,
                    P3& force, double& energy) {
  force = P3{};
  energy = 0.;
  bool break_bonds = true;
  if (!interior_i || !interior_j) break_bonds = false;  // :239-240

  double rji = dist(xj, xi);                        // :243
  P3 dx = xj - xi;
  P3 du = uj - ui;
  double Sji = dot(dx, du) / dot(dx, dx);          // getS :303-306
  P3 eij = dx / length(dx);                         // :330-333

  double h = m.h;                                   // :248-250
  double check_up = m.horizon + 0.5 * h;
  double check_low = m.horizon - 0.5 * h;
  if (dgt(rji, check_low)) volj *= (check_up - rji) / h;  // :255-256

  double influence = J.value(rji / m.horizon);      // :259

  if (break_bonds) {
    if (m.irreversible && !fs &&
        dgt(std::abs(Sji), m.factor_sc * (m.rbar / std::sqrt(rji))))  // :263-266
      fs = true;
    if (!fs) {                                      // :273-280
      energy = (influence * m.C *
                (1. - std::exp(-m.beta * rji * Sji * Sji)) / m.inv_factor) *
               volj;
      force = (influence * 4. * Sji * m.C * m.beta *
               std::exp(-m.beta * rji * Sji * Sji) / m.inv_factor) *
              volj * eij;
      return;
    }
    energy = influence * m.C / m.inv_factor * volj;  // :283
    P3 yji = (xj + uj) - (xi + ui);                  // :286
    double Rji = length(yji);
    double scalar_f = m.contact_kn * (voli * volj / (voli + volj)) *
                      (m.contact_rc - Rji) / Rji;
    if (scalar_f < 0.) scalar_f = 0.;
    P3 c = (-scalar_f) * yji;
    force = force + c;                               // :291 force += ...
    return;
  }
  energy = (influence * m.C * m.beta * rji * Sji * Sji / m.inv_factor) * volj;
  force = (influence * 4. * Sji * m.C * m.beta / m.inv_factor) * volj * eij;
}

// Derivative of getBondEF's force with respect to u_j at fixed bond state:
// dF_ij/du_j = k e e^T, dF_ij/du_i = -k e e^T, with S = dx.du/dx.dx and the
// reference direction e = dx/|dx| (rnpBond.cpp:243-300).  The quasi-static
// tangent of the bond-based material is assembled from these scalars; the
// contact force of broken bonds is not differentiated (k = 0).
inline double bond_tangent(const ora_rnp& m, const Influence& J, const P3& xi,
                           const P3& ui, const P3& xj, const P3& uj,
                           double volj, bool interior_i, bool interior_j,
                           bool fs) {
  bool break_bonds = interior_i && interior_j;
  double rji = dist(xj, xi);
  P3 dx = xj - xi;
  P3 du = uj - ui;
  double Sji = dot(dx, du) / dot(dx, dx);
  double h = m.h;
  double check_up = m.horizon + 0.5 * h;
  double check_low = m.horizon - 0.5 * h;
  if (dgt(rji, check_low)) volj *= (check_up - rji) / h;
  double influence = J.value(rji / m.horizon);
  double a = influence * 4. * m.C * m.beta / m.inv_factor * volj;
  if (break_bonds) {
    if (fs) return 0.;
    double q = m.beta * rji * Sji * Sji;
    return a * std::exp(-q) * (1. - 2. * q) / rji;
  }
  return a / rji;
}

// ---------------------------------------------------------------------------
// Geometry predicates, src/util/utilGeom.cpp
// ---------------------------------------------------------------------------
bool inside_rectangle(const P3& x, double x_min, double x_max, double y_min,
                      double y_max) {  // :43-50
  return !(dlt(x.x, x_min - 1.0E-12) or dlt(x.y, y_min - 1.0E-12) or
           dgt(x.x, x_max + 1.0E-12) or dgt(x.y, y_max + 1.0E-12));
}
bool inside_angled_rectangle(const P3& x, double x1, double x2, double y1,
                             double y2, double theta) {  // :60-95
  P3 lam = rotate_cw_2d(P3{x2 - x1, y2 - y1, 0.0}, theta);
  P3 xmap = rotate_cw_2d(P3{x.x - x1, x.y - y1, 0.0}, theta);
  return !(dlt(xmap.x, -1.0E-12) or dlt(xmap.y, -1.0E-12) or
           dgt(xmap.x, lam.x + 1.0E-12) or dgt(xmap.y, lam.y + 1.0E-12));
}
bool inside_cuboid3(const P3& x, const P3& lo, const P3& hi) {  // :97-115
  return !(dlt(x.x, lo.x - 1.0E-12) or dlt(x.y, lo.y - 1.0E-12) or
           dlt(x.z, lo.z - 1.0E-12) or dgt(x.x, hi.x + 1.0E-12) or
           dgt(x.y, hi.y + 1.0E-12) or dgt(x.z, hi.z + 1.0E-12));
}
bool in_circle(const P3& a, const P3& c, double r) {  // :328-333
  return length(a - c) <= r;
}

// util::function, src/util/utilFunction.cpp:14-64
double hat_function(double x, double x_min, double x_max) {
  if (dgt(x, x_min - 1.0E-12) and dlt(x, x_max + 1.0E-12)) {
    double x_mid = 0.5 * (x_min + x_max);
    double l = x_mid - x_min;
    if (l < 1.0E-12) return 1.0;
    if (dlt(x, x_mid))
      return (x - x_min) / l;
    else
      return (x_max - x) / l;
  }
  return 0.0;
}
double linear_step(double x, double x1, double x2) {
  double period = std::floor(x / (x1 + x2));
  if (dlt(x, period * (x1 + x2) + x1))
    return x - period * x2;
  else
    return (period + 1.) * x1;
}
double der_linear_step(double x, double x1, double x2) {
  double period = std::floor(x / (x1 + x2));
  if (dlt(x, period * (x1 + x2) + x1))
    return 1.;
  else
    return 0.;
}

// ---------------------------------------------------------------------------
// Edge crack predicates, src/inp/decks/fractureDeck.h:115-186
// ---------------------------------------------------------------------------
bool pt_outside(const P3& p, int o, const P3& pb, const P3& pt,
                double theta = 0.0) {
  if (o == -1) return dlt(p.y, pb.y) or dgt(p.y, pt.y);
  if (o == 1) return dlt(p.x, pb.x) or dgt(p.x, pt.x);
  if (o == 0) {
    P3 pmap = rotate_cw_2d(p - pb, theta);
    P3 ptmap = rotate_cw_2d(pt - pb, theta);
    return dlt(pmap.x, 0.0) or dgt(pmap.x, ptmap.x);
  }
  return true;
}
bool pt_leftside(const P3& p, const P3& pb, const P3& pt) {
  double a = 0.5 * ((pt.x - pb.x) * (p.y - pb.y) - (p.x - pb.x) * (pt.y - pb.y));
  return !dlt(a, 0.0);
}
bool pt_rightside(const P3& p, const P3& pb, const P3& pt) {
  return !pt_leftside(p, pb, pt);
}

// ---------------------------------------------------------------------------
// Neighbor lists
// ---------------------------------------------------------------------------
struct Nbr {
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
};

// geometry::Neighbor::Neighbor brute force, src/geometry/neighbor.cpp:69-89
Nbr* neighbor_bruteforce(int64_t n, const double* x, double horizon) {
  std::vector<std::vector<int32_t>> rows(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; i++) {
    P3 xi = p3(x + 3 * i);
    std::vector<int32_t> neighs;
    for (int64_t j = 0; j < n; j++) {
      if (j == i) continue;
      if (dlt(dist(xi, p3(x + 3 * j)), horizon + 1.0E-10))
        neighs.push_back(int32_t(j));
    }
    rows[i] = std::move(neighs);
  }
  Nbr* out = new Nbr;
  out->row_ptr.resize(n + 1, 0);
  for (int64_t i = 0; i < n; i++)
    out->row_ptr[i + 1] = out->row_ptr[i] + int64_t(rows[i].size());
  out->col.resize(out->row_ptr[n]);
  for (int64_t i = 0; i < n; i++)
    std::copy(rows[i].begin(), rows[i].end(),
              out->col.begin() + out->row_ptr[i]);
  return out;
}

// Same predicate with candidates from a uniform grid of cells of edge
// 1.01*(horizon+1e-10) >= any accepted distance; rows sorted ascending so the
// lists equal the brute-force ones.
Nbr* neighbor_cells(int64_t n, const double* x, double horizon) {
  double cs = 1.01 * (horizon + 1.0E-10);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], x[3 * i + d]);
      hi[d] = std::max(hi[d], x[3 * i + d]);
    }
  int64_t nc[3];
  for (int d = 0; d < 3; d++) nc[d] = int64_t((hi[d] - lo[d]) / cs) + 1;
  auto cell_of = [&](const double* p, int64_t* c) {
    for (int d = 0; d < 3; d++) {
      int64_t k = int64_t((p[d] - lo[d]) / cs);
      c[d] = std::min(std::max<int64_t>(k, 0), nc[d] - 1);
    }
  };
  int64_t ncell = nc[0] * nc[1] * nc[2];
  std::vector<int64_t> head(ncell + 1, 0);
  std::vector<int64_t> cid(n);
  for (int64_t i = 0; i < n; i++) {
    int64_t c[3];
    cell_of(x + 3 * i, c);
    cid[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
    head[cid[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) head[c + 1] += head[c];
  std::vector<int64_t> members(n);
  {
    std::vector<int64_t> fill(head.begin(), head.end() - 1);
    for (int64_t i = 0; i < n; i++) members[fill[cid[i]]++] = i;
  }
  std::vector<std::vector<int32_t>> rows(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; i++) {
    P3 xi = p3(x + 3 * i);
    int64_t c[3];
    cell_of(x + 3 * i, c);
    std::vector<int32_t> neighs;
    for (int64_t dz = -1; dz <= 1; dz++)
      for (int64_t dy = -1; dy <= 1; dy++)
        for (int64_t dxc = -1; dxc <= 1; dxc++) {
          int64_t cx = c[0] + dxc, cy = c[1] + dy, cz = c[2] + dz;
          if (cx < 0 || cy < 0 || cz < 0 || cx >= nc[0] || cy >= nc[1] ||
              cz >= nc[2])
            continue;
          int64_t cc = (cz * nc[1] + cy) * nc[0] + cx;
          for (int64_t k = head[cc]; k < head[cc + 1]; k++) {
            int64_t j = members[k];
            if (j == i) continue;
            if (dlt(dist(xi, p3(x + 3 * j)), horizon + 1.0E-10))
              neighs.push_back(int32_t(j));
          }
        }
    std::sort(neighs.begin(), neighs.end());
    rows[i] = std::move(neighs);
  }
  Nbr* out = new Nbr;
  out->row_ptr.resize(n + 1, 0);
  for (int64_t i = 0; i < n; i++)
    out->row_ptr[i + 1] = out->row_ptr[i] + int64_t(rows[i].size());
  out->col.resize(out->row_ptr[n]);
  for (int64_t i = 0; i < n; i++)
    std::copy(rows[i].begin(), rows[i].end(),
              out->col.begin() + out->row_ptr[i]);
  return out;
}

// ---------------------------------------------------------------------------
// Explicit FD model
// ---------------------------------------------------------------------------
struct Crack {
  ora_crack c;
  bool activated = false;
};

// Interior flags behind a virtual call, as RNPBond::getBondEF reads them
// (geometry::InteriorFlags, interiorFlags.cpp:27-30, 228-231): used by the
// reference-faithful timing variant only.
struct InteriorIface {
  virtual ~InteriorIface() = default;
  virtual bool get(int64_t i) const = 0;
};
struct InteriorFromData : InteriorIface {
  const uint8_t* flags = nullptr;
  bool get(int64_t i) const override { return flags[i] != 0; }
};

struct FD {
  int dim = 2;
  int64_t n = 0;
  std::vector<P3> x;
  std::vector<double> vol;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
  ora_rnp m{};
  Influence J;
  double dt = 0.;
  std::vector<ora_bc> ubc, fbc;
  std::vector<std::vector<int64_t>> ubc_nodes, fbc_nodes;
  std::vector<Crack> cracks;
  std::vector<uint8_t> interior;
  std::vector<uint8_t> fix;
  // geometry::Fracture storage, src/geometry/fracture.cpp:24-33
  std::vector<std::vector<uint8_t>> frac;
  std::vector<P3> u, v, f;
  double time = 0.;
  int64_t nstep = 0;

  // Fracture::setBondState / getBondState, fracture.cpp:162-187
  void set_bond(int64_t i, int64_t jl, bool state) {
    state ? (frac[i][jl / 8] |= uint8_t(1UL << (jl % 8)))
          : (frac[i][jl / 8] &= uint8_t(~(1UL << (jl % 8))));
  }
  bool get_bond(int64_t i, int64_t jl) const {
    return (frac[i][jl / 8] >> (jl % 8)) & 1UL;
  }

  bool node_free(int64_t i, int dof) const {  // mesh.h:194-200
    return !(fix[i] >> dof & 1UL);
  }

  // Fracture::computeFracturedBondFd, fracture.cpp:74-160 (ptOutside is
  // called with its default theta = 0, as the reference does)
  void fractured_bonds(int64_t i, const ora_crack& c) {
    P3 inode = x[i];
    P3 pb{c.pb[0], c.pb[1], 0.};
    P3 pt{c.pt[0], c.pt[1], 0.};
    if (pt_outside(inode, c.o, pb, pt)) return;
    bool left_side = pt_leftside(inode, pb, pt);
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      int64_t jl = k - row_ptr[i];
      P3 jnode = x[col[k]];
      bool modify = true;
      if (pt_outside(jnode, c.o, pb, pt)) modify = false;
      if (left_side) {
        if (pt_rightside(jnode, pb, pt) and modify) set_bond(i, jl, modify);
      } else {
        if (pt_leftside(jnode, pb, pt) and modify) set_bond(i, jl, modify);
      }
    }
  }

  // Fracture::Fracture, fracture.cpp:20-43
  void init_fracture() {
    frac.resize(n);
    for (int64_t i = 0; i < n; i++) {
      size_t ns = size_t(row_ptr[i + 1] - row_ptr[i]);
      size_t s = ns / 8;
      if (s * 8 < ns) s++;
      frac[i] = std::vector<uint8_t>(s, uint8_t(0));
      for (auto& cr : cracks)
        if (cr.c.activation_time < 0.) {
          fractured_bonds(i, cr.c);
          cr.activated = true;
        }
    }
  }

  // Fracture::addCrack, fracture.cpp:45-72
  bool add_crack(double t) {
    for (auto& cr : cracks) {
      if (!cr.activated) {
        if (dlt(cr.c.activation_time, t)) {
          for (int64_t i = 0; i < n; i++) fractured_bonds(i, cr.c);
          cr.activated = true;
          return true;
        }
      }
    }
    return false;
  }

  // ULoading::ULoading node selection, src/loading/uLoading.cpp:86-163
  void init_ubc() {
    for (const auto& bc : ubc) {
      std::vector<int64_t> nodes;
      for (int64_t i = 0; i < n; i++) {
        bool fixed = false;
        const P3& xi = x[i];
        if (bc.region == ORA_REGION_RECTANGLE &&
            inside_rectangle(xi, bc.x1, bc.x2, bc.y1, bc.y2))
          fixed = true;
        else if (bc.region == ORA_REGION_ANGLED_RECTANGLE &&
                 inside_angled_rectangle(xi, bc.x1, bc.x2, bc.y1, bc.y2,
                                         bc.theta))
          fixed = true;
        else if (bc.region == ORA_REGION_CIRCLE &&
                 in_circle(xi, P3{bc.x1, bc.y1, 0.}, bc.r1))
          fixed = true;
        else if (bc.region == ORA_REGION_TORUS &&
                 in_circle(xi, P3{bc.x1, bc.y1, 0.}, bc.r1) &&
                 !in_circle(xi, P3{bc.x1, bc.y1, 0.}, bc.r2))
          fixed = true;
        else if (bc.region == ORA_REGION_LINE && dgt(xi.x, bc.x1) &&
                 dlt(xi.x, bc.x2))
          fixed = true;
        // ULoading has no cuboid region (uLoading.cpp:27-31); the 3D
        // specimens select their face layers with the FLoading cuboid test
        // (utilGeom.cpp:97-115) as an extension.
        else if (bc.region == ORA_REGION_CUBOID &&
                 inside_cuboid3(xi, P3{bc.x1, bc.y1, bc.z1},
                                P3{bc.x2, bc.y2, bc.z2}))
          fixed = true;
        if (fixed) {
          for (int k = 0; k < bc.ndir; k++)
            fix[i] |= uint8_t(1U << (bc.dir[k] - 1));
          nodes.push_back(i);
        }
      }
      ubc_nodes.push_back(nodes);
    }
  }

  // FLoading::FLoading node selection, src/loading/fLoading.cpp:18-116
  void init_fbc() {
    for (const auto& bc : fbc) {
      std::vector<int64_t> nodes;
      for (int64_t i = 0; i < n; i++) {
        const P3& xi = x[i];
        bool sel = false;
        if (bc.region == ORA_REGION_LINE)
          sel = dgt(xi.x, bc.x1) && dlt(xi.x, bc.x2);
        else if (bc.region == ORA_REGION_RECTANGLE)
          sel = inside_rectangle(xi, bc.x1, bc.x2, bc.y1, bc.y2);
        else if (bc.region == ORA_REGION_ANGLED_RECTANGLE)
          sel = inside_angled_rectangle(xi, bc.x1, bc.x2, bc.y1, bc.y2,
                                        bc.theta);
        else if (bc.region == ORA_REGION_CUBOID)
          sel = inside_cuboid3(xi, P3{bc.x1, bc.y1, bc.z1},
                               P3{bc.x2, bc.y2, bc.z2});
        if (sel) nodes.push_back(i);
      }
      fbc_nodes.push_back(nodes);
    }
  }

  // ULoading::apply, src/loading/uLoading.cpp:166-253.  A "constant" spatial
  // function without parameters reads params[0] in the reference (undefined
  // behaviour); the multiplier 1 is used here.
  void apply_ubc(double t) {
    for (size_t s = 0; s < ubc.size(); s++) {
      const ora_bc& bc = ubc[s];
      for (int64_t i : ubc_nodes[s]) {
        const P3& xx = x[i];
        double umax = bc.time_params[0];
        double du = 0., dv = 0.;
        switch (bc.space_fn) {
          case ORA_SFN_SIN_X: {
            double a = M_PI * bc.space_params[0];
            umax = umax * std::sin(a * xx.x);
          } break;
          case ORA_SFN_SIN_Y: {
            double a = M_PI * bc.space_params[0];
            umax = umax * std::sin(a * xx.y);
          } break;
          case ORA_SFN_SIN_Z: {
            double a = M_PI * bc.space_params[0];
            umax = umax * std::sin(a * xx.z);
          } break;
          case ORA_SFN_LINEAR_X: {
            double a = bc.space_params[0];
            umax = umax * a * xx.x;
          } break;
          case ORA_SFN_LINEAR_Y: {
            double a = bc.space_params[0];
            umax = umax * a * xx.y;
          } break;
          case ORA_SFN_LINEAR_Z: {
            double a = bc.space_params[0];
            umax = umax * a * xx.z;
          } break;
          case ORA_SFN_CONSTANT: {
            double a = bc.n_space_params > 0 ? bc.space_params[0] : 1.0;
            umax = umax * a;
          } break;
          default:
            break;
        }
        const double* p = bc.time_params;
        if (bc.time_fn == ORA_TFN_CONSTANT)
          du = umax;
        else if (bc.time_fn == ORA_TFN_LINEAR) {
          du = umax * t;
          dv = umax;
        } else if (bc.time_fn == ORA_TFN_QUADRATIC) {
          du = umax * t + p[1] * t * t;
          dv = umax + p[1] * t;
        } else if (bc.time_fn == ORA_TFN_SIN) {
          double a = M_PI * p[1];
          du = umax * std::sin(a * t);
          dv = umax * a * std::cos(a * t);
        } else if (bc.time_fn == ORA_TFN_LINEAR_STEP) {
          du = umax * linear_step(t, p[1], p[2]);
          dv = umax * der_linear_step(t, p[1], p[2]);
        } else if (bc.time_fn == ORA_TFN_LINEAR_SLOW_FAST) {
          if (dgt(t, p[1])) {
            du = umax * p[3] * t;
            dv = umax * p[3];
          } else {
            du = umax * p[2] * t;
            dv = umax * p[2];
          }
        }
        for (int k = 0; k < bc.ndir; k++) {
          int d = bc.dir[k];
          if (d == 1) {
            u[i].x = du;
            v[i].x = dv;
          } else if (d == 2) {
            u[i].y = du;
            v[i].y = dv;
          } else if (d == 3) {
            u[i].z = du;
            v[i].z = dv;
          }
        }
      }
    }
  }

  // FLoading::apply, src/loading/fLoading.cpp:117-279 (a "constant" spatial
  // function without parameters uses 1)
  void apply_fbc(double t, std::vector<P3>& ff) {
    for (size_t s = 0; s < fbc.size(); s++) {
      const ora_bc& bc = fbc[s];
      const double* sp = bc.space_params;
      const double* p = bc.time_params;
      for (int64_t i : fbc_nodes[s]) {
        const P3& xx = x[i];
        double fmax = 1.0;
        switch (bc.space_fn) {
          case ORA_SFN_HAT_X:
            fmax = sp[0] * hat_function(xx.x, bc.x1, bc.x2);
            break;
          case ORA_SFN_HAT_Y:
            fmax = sp[0] * hat_function(xx.y, bc.y1, bc.y2);
            break;
          case ORA_SFN_HAT_Z:
            fmax = sp[0] * hat_function(xx.z, bc.z1, bc.z2);
            break;
          case ORA_SFN_SIN_X: {
            double a = M_PI * sp[0];
            fmax = sp[0] * std::sin(a * xx.x);
          } break;
          case ORA_SFN_SIN_Y: {
            double a = M_PI * sp[0];
            fmax = sp[0] * std::sin(a * xx.y);
          } break;
          case ORA_SFN_SIN_Z: {
            double a = M_PI * sp[0];
            fmax = sp[0] * std::sin(a * xx.z);
          } break;
          case ORA_SFN_LINEAR_X: {
            double a = sp[0];
            fmax = sp[0] * a * xx.x;
          } break;
          case ORA_SFN_LINEAR_Y: {
            double a = sp[0];
            fmax = sp[0] * a * xx.y;
          } break;
          case ORA_SFN_LINEAR_Z: {
            double a = sp[0];
            fmax = sp[0] * a * xx.z;
          } break;
          case ORA_SFN_CONSTANT:
            fmax = bc.n_space_params > 0 ? sp[0] : 1.0;
            break;
          case ORA_SFN_LINE_LOAD:
            // fLoading.cpp:172-228; the block that would swap min and max
            // for params[0] == -1 only declares shadowing locals
            for (int k = 0; k < bc.ndir; k++) {
              const int d = bc.dir[k];
              double mn = d == 1 ? bc.x1 : (d == 2 ? bc.y1 : bc.z1);
              double mx = d == 1 ? bc.x2 : (d == 2 ? bc.y2 : bc.z2);
              double xd = d == 1 ? xx.x : (d == 2 ? xx.y : xx.z);
              double length = std::abs(mx - mn);
              size_t nodes = length / m.h;
              double scale = 1. / nodes;
              size_t pos = (xd - mn) / m.h;
              if (bc.n_space_params > 0 && sp[0] == -1) pos = (mx - xd) / m.h;
              fmax *= pos * scale;
            }
            break;
          default:
            break;
        }
        if (bc.time_fn == ORA_TFN_LINEAR)
          fmax *= t;
        else if (bc.time_fn == ORA_TFN_LINEAR_STEP)
          fmax *= linear_step(t, p[1], p[2]);
        else if (bc.time_fn == ORA_TFN_LINEAR_SLOW_FAST) {
          if (dgt(t, p[1]))
            fmax *= p[3] * t;
          else
            fmax *= p[2] * t;
        } else if (bc.time_fn == ORA_TFN_SIN) {
          double a = M_PI * p[1];
          fmax *= std::sin(a * t);
        }
        fmax *= p[0];
        for (int k = 0; k < bc.ndir; k++) {
          int d = bc.dir[k];
          if (d == 1)
            ff[i].x += fmax;
          else if (d == 2)
            ff[i].y += fmax;
          else if (d == 3)
            ff[i].z += fmax;
        }
      }
    }
  }

  // FDModel::computeForce, fDModel.cpp:544-582 (reaction-force bookkeeping
  // is output-only and not restated)
  P3 compute_force(int64_t i, double* energy_out) {
    P3 force_i{};
    double energy_i = 0.;
    const P3& xi = x[i];
    const P3& ui = u[i];
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      int64_t jl = k - row_ptr[i];
      int64_t j = col[k];
      bool fs = get_bond(i, jl);                  // rnpBond.cpp:225
      P3 F;
      double E;
      bond_ef(m, J, xi, ui, x[j], u[j], vol[i], vol[j], interior[i] != 0,
              interior[j] != 0, fs, F, E);
      if (interior[i] != 0 && interior[j] != 0) set_bond(i, jl, fs);  // :269
      force_i = force_i + F;
      energy_i += E;
    }
    if (energy_out) *energy_out = energy_i;
    return force_i;
  }

  // Reference-faithful timing variant (the CPU baseline that keeps the
  // reference's per-node and per-bond bookkeeping; forces identical to
  // compute_force's): FDModel::computeForce (fDModel.cpp:544-582) looks up
  // two output tags per node (OutputDeck::isTagInOutput, a string compare
  // over the tag list, outputDeck.h:95-103) and tests is_reaction_force per
  // bond (fDModel.cpp:585-605); RNPBond::getBondEF reads both interior flags
  // through a virtual call per bond (rnpBond.cpp:236-240).
  bool faithful = false;
  std::vector<std::string> out_tags;
  std::vector<P3> react;
  std::vector<double> react_tot;
  double bb_min_x = 0., bb_max_x = 0., bb_max_y = 0.;
  InteriorFromData interior_if;

  bool tag_in_output(const std::string& tag) const {
    for (const auto& t : out_tags)
      if (tag == t) return true;
    return false;
  }

  bool is_reaction_force(int64_t i, int64_t j) const {
    const double delta = m.horizon, eps = 1e-6;
    const P3 xi = x[i], xj = x[j];
    if (xi.y > bb_max_y - 1.5 * delta && xi.y < bb_max_y - delta) {
      P3 A{bb_min_x, bb_max_y - delta + eps, 0.};
      P3 B{bb_max_x, bb_max_y - delta + eps, 0.};
      if (ref_lines_intersect(A, B, xi, xj)) return true;
    }
    return false;
  }

  P3 compute_force_faithful(int64_t i) {
    P3 force_i{0., 0., 0.};
    double energy_i = 0.;
    const InteriorIface& inter = interior_if;
    if (tag_in_output("Reaction_Force") or
        tag_in_output("Total_Reaction_Force")) {
      react[i] = P3{};
      react_tot[i] = 0.;
    }
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      int64_t jl = k - row_ptr[i];
      int64_t j = col[k];
      bool fs = get_bond(i, jl);
      const bool int_i = inter.get(i), int_j = inter.get(j);
      P3 F;
      double E;
      bond_ef(m, J, x[i], u[i], x[j], u[j], vol[i], vol[j], int_i, int_j, fs,
              F, E);
      if (int_i && int_j) set_bond(i, jl, fs);
      force_i = force_i + F;
      energy_i += E;
      if (is_reaction_force(i, j) and
          (tag_in_output("Reaction_Force") or
           tag_in_output("Total_Reaction_Force")))
        react[i] = react[i] + vol[i] * F;
    }
    if (tag_in_output("Total_Reaction_Force"))
      react_tot[i] = length(react[i]);
    return force_i;
  }

  void set_faithful(bool on) {
    faithful = on;
    if (!on) return;
    // the output tags of the reference's example explicit deck
    // (examples/fdModel/crack_propagation/input.yaml:66-70)
    out_tags = {"Displacement", "Velocity", "Force", "Damage_Z"};
    react.assign(size_t(n), P3{});
    react_tot.assign(size_t(n), 0.);
    bb_min_x = bb_max_x = x[0].x;
    bb_max_y = x[0].y;
    for (int64_t i = 0; i < n; i++) {
      bb_min_x = std::min(bb_min_x, x[i].x);
      bb_max_x = std::max(bb_max_x, x[i].x);
      bb_max_y = std::max(bb_max_y, x[i].y);
    }
    interior_if.flags = interior.data();
  }

  // per-bond tangent scalars k (CSR order) at the current u and bond state
  void tangent(double* kout) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        int64_t jl = k - row_ptr[i];
        int64_t j = col[k];
        kout[k] = bond_tangent(m, J, x[i], u[i], x[j], u[j], vol[j],
                               interior[i] != 0, interior[j] != 0,
                               get_bond(i, jl));
      }
  }

  // FDModel::computeForces, fDModel.cpp:528-541
  void compute_forces() {
    if (faithful) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) f[i] = f[i] + compute_force_faithful(i);
      return;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) f[i] = f[i] + compute_force(i, nullptr);
  }

  // FDModel::integrate prologue, fDModel.cpp:318-333 (no initial condition
  // deck is supported here)
  void init() {
    apply_ubc(time);
    apply_fbc(time, f);
    compute_forces();
  }

  // FDModel::integrateCD, fDModel.cpp:377-443, followed by the crack
  // activation check of integrate(), :366-372
  void step_cd() {
    double delta_t = dt;
    double fact = delta_t * delta_t / m.density;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      if (node_free(i, 0)) {
        double u_old = u[i].x;
        u[i].x += fact * f[i].x + delta_t * v[i].x;
        v[i].x = (u[i].x - u_old) / delta_t;
      }
      if (dim > 1)
        if (node_free(i, 1)) {
          double u_old = u[i].y;
          u[i].y += fact * f[i].y + delta_t * v[i].y;
          v[i].y = (u[i].y - u_old) / delta_t;
        }
      if (dim > 2)
        if (node_free(i, 2)) {
          double u_old = u[i].z;
          u[i].z += fact * f[i].z + delta_t * v[i].z;
          v[i].z = (u[i].z - u_old) / delta_t;
        }
      f[i] = P3{};
    }
    nstep++;
    time += dt;
    apply_ubc(time);
    apply_fbc(time, f);
    compute_forces();
    add_crack(time);
  }

  // FDModel::integrateVerlet, fDModel.cpp:446-525
  void step_verlet() {
    double delta_t = dt;
    double fact = 0.5 * delta_t / m.density;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      if (node_free(i, 0)) {
        v[i].x += fact * f[i].x;
        u[i].x += delta_t * v[i].x;
      }
      if (dim > 1)
        if (node_free(i, 1)) {
          v[i].y += fact * f[i].y;
          u[i].y += delta_t * v[i].y;
        }
      if (dim > 2)
        if (node_free(i, 2)) {
          v[i].z += fact * f[i].z;
          u[i].z += delta_t * v[i].z;
        }
      f[i] = P3{};
    }
    nstep++;
    time += dt;
    apply_ubc(time);
    apply_fbc(time, f);
    compute_forces();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double fact2 = 0.5 * dt / m.density;
      if (node_free(i, 0)) v[i].x += fact2 * f[i].x;
      if (dim > 1)
        if (node_free(i, 1)) v[i].y += fact2 * f[i].y;
      if (dim > 2)
        if (node_free(i, 2)) v[i].z += fact2 * f[i].z;
    }
    add_crack(time);
  }

  // Per node sum_j |F_ij| (Euclidean norm of each bond force) at the current
  // state without touching the bond bits: the scale of the force-parity
  // metric |f_gpu - f_ref| <= tol * sum_j |F_ij|.
  void force_abs_sum(double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double s = 0.;
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        int64_t jl = k - row_ptr[i];
        int64_t j = col[k];
        bool fs = get_bond(i, jl);
        P3 F;
        double E;
        bond_ef(m, J, x[i], u[i], x[j], u[j], vol[i], vol[j],
                interior[i] != 0, interior[j] != 0, fs, F, E);
        s += length(F);
      }
      out[i] = s;
    }
  }

  // External force density at time t: FLoading::apply on zeros
  // (fLoading.cpp:117-279), fDModel.cpp:655-662.
  void external_force(double t, double* out) {
    std::vector<P3> ff(size_t(n), P3{0., 0., 0.});
    apply_fbc(t, ff);
    for (int64_t i = 0; i < n; i++) {
      out[3 * i] = ff[i].x;
      out[3 * i + 1] = ff[i].y;
      out[3 * i + 2] = ff[i].z;
    }
  }

  // util::geometry::direction, utilGeom.cpp:317-326 (int determinant).
  static int ref_direction(const P3& A, const P3& B, const P3& C) {
    int val = int((B.y - A.y) * (C.x - B.x) - (B.x - A.x) * (C.y - B.y));
    if (val == 0) return 0;
    if (val < 0) return 2;
    return 1;
  }
  // util::geometry::onLine, utilGeom.cpp:308-315.
  static bool ref_on_line(const P3& A, const P3& B, const P3& C) {
    return C.x <= std::max(A.x, B.x) && C.x <= std::min(A.x, B.x) &&
           (C.y <= std::max(A.y, B.y) && C.y <= std::min(A.y, B.y));
  }
  // util::geometry::doLinesIntersect, utilGeom.cpp:281-306.
  static bool ref_lines_intersect(const P3& A, const P3& B, const P3& C,
                                  const P3& D) {
    int dir1 = ref_direction(A, B, C);
    int dir2 = ref_direction(A, B, D);
    int dir3 = ref_direction(C, D, A);
    int dir4 = ref_direction(C, D, B);
    if (dir1 != dir2 && dir3 != dir4) return true;
    if (dir1 == 0 && ref_on_line(A, B, C)) return true;
    if (dir2 == 0 && ref_on_line(A, B, D)) return true;
    if (dir3 == 0 && ref_on_line(C, D, A)) return true;
    if (dir4 == 0 && ref_on_line(C, D, B)) return true;
    return false;
  }

  // Reaction force of FDModel::computeForce (fDModel.cpp:548-580) with
  // FDModel::is_reaction_force (fDModel.cpp:585-605), at the current state;
  // box = {min_x, max_x, max_y}.  The bond bits are read, not changed.
  void reaction_force(const double* box, double delta, double* R,
                      double* R_len) {
    const double min_x = box[0], max_x = box[1], max_y = box[2];
    const double eps = 1e-6;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      P3 r{0., 0., 0.};
      const P3& xi = x[i];
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        int64_t jl = k - row_ptr[i];
        int64_t j = col[k];
        bool crosses = false;
        if (xi.y > max_y - 1.5 * delta && xi.y < max_y - delta) {
          P3 A{min_x, max_y - delta + eps, 0.};
          P3 B{max_x, max_y - delta + eps, 0.};
          crosses = ref_lines_intersect(A, B, xi, x[j]);
        }
        if (!crosses) continue;
        bool fs = get_bond(i, jl);
        P3 F;
        double E;
        bond_ef(m, J, xi, u[i], x[j], u[j], vol[i], vol[j], interior[i] != 0,
                interior[j] != 0, fs, F, E);
        r = r + vol[i] * F;  // R_i += V_i F (fDModel.cpp:572-574)
      }
      R[3 * i] = r.x;
      R[3 * i + 1] = r.y;
      R[3 * i + 2] = r.z;
      R_len[i] = length(r);
    }
  }

  // FDModel::computePostProcFields damage/energy part, fDModel.cpp:652-751
  void postproc(float* Z, float* phi, double* energy) {
    double h = m.h;
    double check_up = m.horizon + 0.5 * h;
    double check_low = m.horizon - 0.5 * h;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double energy_i = 0.0, a = 0., b = 0., z = 0.;
      const P3& xi = x[i];
      const P3& ui = u[i];
      double voli = vol[i];
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        int64_t jl = k - row_ptr[i];
        int64_t j = col[k];
        bool fs = get_bond(i, jl);
        P3 F;
        double E;
        bond_ef(m, J, xi, ui, x[j], u[j], voli, vol[j], interior[i] != 0,
                interior[j] != 0, fs, F, E);
        if (interior[i] != 0 && interior[j] != 0) set_bond(i, jl, fs);
        fs = get_bond(i, jl);
        energy_i += E;
        const P3& xj = x[j];
        const P3& uj = u[j];
        double rji = dist(xj, xi);
        P3 dx = xj - xi;
        P3 du = uj - ui;
        double Sji = dot(dx, du) / dot(dx, dx);
        double volj = vol[j];
        if (dgt(rji, check_low)) volj *= (check_up - rji) / h;
        if (!fs) a += volj;
        b += volj;
        double sr = 0.;
        if (dgt(rji, 1.0E-12)) sr = std::abs(Sji) / (m.rbar / std::sqrt(rji));
        if (dlt(z, sr)) z = sr;
      }
      if (Z) Z[i] = float(z);
      if (phi) phi[i] = float(1. - a / b);
      if (energy) energy[i] = energy_i * voli;
    }
  }
};

// ---------------------------------------------------------------------------
// Explicit linear state-based model (ElasticState) -- the reference runs this
// material only in QuasiStaticModel; the explicit model here combines its
// force (scatter form of QuasiStaticModel::computeForces) with integrateCD.
// Broken bonds (stretch definitelyGreaterThan sc, irreversible) are left out
// of the dilatation and the force; the reference ElasticState has no damage.
// ---------------------------------------------------------------------------
struct ES {
  int dim = 3;
  int64_t n = 0;
  std::vector<P3> x;
  std::vector<double> vol;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
  ora_es m{};
  double dt = 0.;
  std::vector<ora_bc> ubc, fbc;
  std::vector<std::vector<int64_t>> ubc_nodes, fbc_nodes;
  std::vector<uint8_t> fix;
  std::vector<uint8_t> broken;          // per bond in CSR order
  std::vector<double> corr;             // volume correction per bond
  std::vector<double> mw;               // weighted volume per node
  std::vector<double> theta;            // dilatation per node
  std::vector<double> ext;              // extension per bond
  std::vector<P3> u, v, f;
  double time = 0.;
  FD* loader = nullptr;                 // BC selection/application helper

  // VolumeCorrection::correctVolume / weightedVolume, volumeCorrection.cpp
  void init_volume() {
    corr.assign(col.size(), 1.);
    mw.assign(n, 1.);
    for (int64_t i = 0; i < n; i++) {
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        P3 X = x[col[k]] - x[i];
        double r = m.h * 0.5;
        if (dgt(length(X), m.horizon - r))
          corr[k] = (m.horizon + r - length(X)) / m.h;
      }
      double tmp = 0;
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        P3 X = x[col[k]] - x[i];
        tmp += (length(X) * length(X) * corr[k] * vol[col[k]]);
      }
      mw[i] = tmp;
    }
  }

  // StateBasedHelperFunctions::dilatation, stateBasedHelperFunctions.cpp:
  // 26-83, with the break test of the explicit model first
  void dilatation() {
    theta.assign(n, 0.);
    ext.assign(col.size(), 0.);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double w = 1;
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        int64_t j = col[k];
        P3 Y = (x[j] + u[j]) - (x[i] + u[i]);
        P3 X = x[j] - x[i];
        ext[k] = length(Y) - length(X);
        if (m.sc > 0. && !broken[k] && dgt(ext[k] / length(X), m.sc))
          broken[k] = 1;
        if (broken[k]) continue;
        switch (dim) {
          case 1:
            theta[i] += (1. / mw[i]) * w * length(X) * ext[k] * corr[k] *
                        vol[j];
            break;
          case 2:
            theta[i] += (2. / mw[i]) * m.factor2D * w * length(X) * ext[k] *
                        corr[k] * vol[j];
            break;
          default:
            theta[i] += (3. / mw[i]) * w * length(X) * ext[k] * corr[k] *
                        vol[j];
            break;
        }
      }
    }
  }

  // ElasticState::getBondEF, ElasticState.cpp:160-271 (force part)
  P3 bond_force(int64_t i, int64_t k) {
    int64_t j = col[k];
    double w = 1;
    double t = 0., alpha_s = 0., alpha_d = 0., e_s = 0., e_d = 0., alpha = 0.;
    double t_s = 0., t_d = 0.;
    P3 Y = (x[j] + u[j]) - (x[i] + u[i]);
    P3 X = x[j] - x[i];
    P3 M = Y / length(Y);
    switch (dim) {
      case 1:
        alpha = m.E / mw[i];
        t = alpha * w * ext[k];
        break;
      case 2:
        if (!m.plane_strain)
          alpha_s = (9. / mw[i]) *
                    (m.K + std::pow((m.nu + 1.) / (2. * m.nu - 1.), 2) * m.mu /
                               9.);
        else
          alpha_s = (9. / mw[i]) * (m.K + m.mu / 9.);
        alpha_d = (8. / mw[i]) * m.mu;
        e_s = theta[i] * length(X) / 3.;
        e_d = ext[k] - e_s;
        t_s = (2. * m.factor2D * alpha_s - (3. - 2. * m.factor2D) * alpha_d) *
              w * e_s / 3.;
        t_d = alpha_d * w * e_d;
        t = t_s + t_d;
        break;
      default:
        alpha_s = (9. / mw[i]) * m.K;
        alpha_d = (15. / mw[i]) * m.mu;
        e_s = theta[i] * length(X) / 3.;
        t_s = alpha_s * w * e_s;
        t_d = alpha_d * w * e_d;  // e_d stays 0 in 3D (reference)
        t = t_s + t_d;
        break;
    }
    P3 r = t * M;
    return corr[k] * r;
  }

  // Strain energy of bond k of node i: the second value of
  // ElasticState::getBondEF (ElasticState.cpp:250-262).
  double bond_energy(int64_t i, int64_t k) {
    int64_t j = col[k];
    double w = 1;
    double alpha_s = 0., alpha_d = 0., e_s = 0., e_d = 0., alpha = 0.;
    P3 X = x[j] - x[i];
    switch (dim) {
      case 1:
        alpha = m.E / mw[i];
        break;
      case 2:
        if (!m.plane_strain)
          alpha_s = (9. / mw[i]) *
                    (m.K + std::pow((m.nu + 1.) / (2. * m.nu - 1.), 2) * m.mu /
                               9.);
        else
          alpha_s = (9. / mw[i]) * (m.K + m.mu / 9.);
        alpha_d = (8. / mw[i]) * m.mu;
        e_s = theta[i] * length(X) / 3.;
        e_d = ext[k] - e_s;
        break;
      default:
        alpha_s = (9. / mw[i]) * m.K;
        alpha_d = (15. / mw[i]) * m.mu;
        e_s = theta[i] * length(X) / 3.;
        break;
    }
    if (dim == 1)
      return 0.5 * alpha * w * ext[k] * ext[k] * corr[k] * vol[j];
    return 0.5 * w * (alpha_s * e_s * e_s + alpha_d * e_d * e_d) * corr[k] *
           vol[j];
  }

  // 3x3 matrices of ElasticState's strain and stress (blaze::StaticMatrix
  // in the reference)
  struct M33 {
    double a[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
  };
  static M33 outer(const P3& p, const P3& q) {  // Point3::toMatrix, point.h
    const double pv[3] = {p.x, p.y, p.z}, qv[3] = {q.x, q.y, q.z};
    M33 r;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) r.a[a][b] = pv[a] * qv[b];
    return r;
  }

  // ElasticState::getStrain (ElasticState.cpp:336-345) with
  // deformation_gradient (:304-334): F = (sum_j Y X^T w c_ij V_j) (K^-1)^T
  // with the shape tensor K = sum_j X X^T w c V_j of K_shape_tensor
  // (:287-302), whose volume-correction index is never
  // advanced (every neighbor takes the first one's correction: kept);
  // strain = (F + F^T)/2 - I on the dim x dim block.
  M33 strain(int64_t i) {
    const double w = 1;
    const int64_t k0 = row_ptr[i];
    M33 K, T;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      const int64_t j = col[k];
      const P3 X = x[j] - x[i];
      const P3 Y = (x[j] + u[j]) - (x[i] + u[i]);
      const M33 xx = outer(X, X), yx = outer(Y, X);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          K.a[a][b] += xx.a[a][b] * w * corr[k0] * vol[j];
          T.a[a][b] += yx.a[a][b] * w * corr[k] * vol[j];
        }
    }
    // blaze::inv of the dim x dim block, transposed into a zero 3x3
    M33 Si;
    if (dim == 1) {
      Si.a[0][0] = 1. / K.a[0][0];
    } else if (dim == 2) {
      const double det = K.a[0][0] * K.a[1][1] - K.a[0][1] * K.a[1][0];
      const double inv[2][2] = {{K.a[1][1] / det, -K.a[0][1] / det},
                                {-K.a[1][0] / det, K.a[0][0] / det}};
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) Si.a[a][b] = inv[b][a];
    } else {
      const auto& A = K.a;
      const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
      const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
      const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
      const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
      double inv[3][3];
      inv[0][0] = c00 / det;
      inv[1][0] = c01 / det;
      inv[2][0] = c02 / det;
      inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
      inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
      inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
      inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
      inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
      inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Si.a[a][b] = inv[b][a];
    }
    M33 F, E;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        for (int c = 0; c < 3; c++) F.a[a][b] += T.a[a][c] * Si.a[c][b];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        E.a[a][b] = (F.a[a][b] + F.a[b][a]) * 0.5 -
                    ((a == b && a < dim) ? 1. : 0.);
    return E;
  }

  // ElasticState::getStress (ElasticState.cpp:432-471) with
  // K_modulus_tensor (:359-430) and dirac_delta (:347-357):
  // stress = sum_j sum_k (K(j, k) (strain X_k)) X_j^T c_k V_k c_j V_j.
  M33 stress(int64_t i, const M33& eps_i) {
    const double w = 1;
    M33 S;
    for (int64_t kn = row_ptr[i]; kn < row_ptr[i + 1]; kn++) {
      const int64_t j = col[kn];
      const P3 Xj = x[j] - x[i];
      const P3 M = Xj / length(Xj);
      for (int64_t km = row_ptr[i]; km < row_ptr[i + 1]; km++) {
        const int64_t kk = col[km];
        const P3 Xk = x[kk] - x[i];
        const double volume = corr[km] * vol[kk] * corr[kn] * vol[j];
        // dirac_delta: essentiallyEqual(|Xk - Xj|, 0) holds only at 0
        const double delta =
            length(Xk - Xj) == 0. ? 1. / vol[kk] * corr[km] : 0.;
        M33 Km;
        if (dim == 1) {
          const double alpha = m.E / mw[i];
          const M33 mm = outer(M, M);
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
              Km.a[a][b] = alpha * w * mm.a[a][b] * delta;
        } else {
          double alpha_s, alpha_d, lead;
          if (dim == 2) {
            const double Nu = (3. * m.K - 2. * m.mu) / (2. * (3. * m.K + m.mu));
            alpha_s = (9. / mw[i]) *
                      (m.K + std::pow((Nu + 1.) / (2. * Nu - 1.), 2) * m.mu / 9.);
            alpha_d = (8. / mw[i]) * m.mu;
            const double alpha_sb = (2. * m.factor2D * alpha_s -
                                     (3. - 2. * m.factor2D) * alpha_d) /
                                    3.;
            lead = (alpha_sb - alpha_d) / mw[i];
          } else {
            alpha_s = (9. / mw[i]) * m.K;
            alpha_d = (15. / mw[i]) * m.mu;
            lead = (alpha_s - alpha_d) / mw[i];
          }
          const M33 jk = outer(Xj, Xk), mm = outer(M, M);
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
              Km.a[a][b] = lead * w * w * jk.a[a][b] +
                           alpha_d * w * mm.a[a][b] * delta;
        }
        // res = K (strain X_k); tmp(o, :) = res[o] X_j
        const double xk[3] = {Xk.x, Xk.y, Xk.z};
        double ex[3], res[3];
        for (int a = 0; a < 3; a++) {
          ex[a] = 0.;
          for (int b = 0; b < 3; b++) ex[a] += eps_i.a[a][b] * xk[b];
        }
        for (int a = 0; a < 3; a++) {
          res[a] = 0.;
          for (int b = 0; b < 3; b++) res[a] += Km.a[a][b] * ex[b];
        }
        const double xj[3] = {Xj.x, Xj.y, Xj.z};
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) S.a[a][b] += res[a] * xj[b] * volume;
      }
    }
    return S;
  }

  // The output fields of QuasiStaticModel::computeForces(true)
  // (QuasiStaticModel.cpp:206-229) at the current state: the strain energy
  // of every bond is added in float to energy_acc[i] (the reference keeps a
  // std::vector<float> and adds (float) per bond, never clearing it), the
  // strain and stress tensors are written as (00, 11, 22, 12, 02, 01)
  // (vtkWriter.cpp:213-233).
  void fields(float* energy_acc, double* strain6, double* stress6) {
    dilatation();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      if (energy_acc) {
        float e = energy_acc[i];
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
          if (broken[k]) continue;
          e += float(bond_energy(i, k));
        }
        energy_acc[i] = e;
      }
      if (!strain6 && !stress6) continue;
      const M33 E = strain(i);
      auto put = [&](double* o, const M33& A) {
        o[6 * i + 0] = A.a[0][0];
        o[6 * i + 1] = A.a[1][1];
        o[6 * i + 2] = A.a[2][2];
        o[6 * i + 3] = A.a[1][2];
        o[6 * i + 4] = A.a[0][2];
        o[6 * i + 5] = A.a[0][1];
      };
      if (strain6) put(strain6, E);
      if (stress6) put(stress6, stress(i, E));
    }
  }

  // QuasiStaticModel::computeForces (QuasiStaticModel.cpp:169-239), run
  // serially: f_i += T_ij V_j, f_j -= T_ij V_i.  Adds into ff.
  void forces(std::vector<P3>& ff) {
    dilatation();
    for (int64_t i = 0; i < n; i++) {
      P3 force_i{};
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        if (broken[k]) continue;
        int64_t j = col[k];
        P3 T = bond_force(i, k);
        force_i = force_i + vol[j] * T;
        ff[j] = ff[j] - vol[i] * T;
      }
      ff[i] = ff[i] + force_i;
    }
  }

  // per node sum_j |V_j T_ij| + |V_j T_ji| over unbroken bonds at the
  // current u (dilatation recomputed, bond states untouched)
  void force_abs_sum(double* out) {
    std::vector<uint8_t> keep = broken;
    double sc = m.sc;
    m.sc = -1.;
    dilatation();
    m.sc = sc;
    broken = keep;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double acc = 0.;
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        if (broken[k]) continue;
        int64_t j = col[k];
        P3 T = bond_force(i, k);
        acc += length(vol[j] * T);
        for (int64_t kk = row_ptr[j]; kk < row_ptr[j + 1]; kk++)
          if (col[kk] == i) {
            P3 T2 = bond_force(j, kk);
            acc += length(vol[j] * T2);
            break;
          }
      }
      out[i] = acc;
    }
  }

  // Test tolerance (3D): a bound on the per-node force difference that the
  // reference's own rounding of the extension e = |(x_j + u_j) - (x_i +
  // u_i)| - |x_j - x_i| (stateBasedHelperFunctions.cpp:39-47) can cause
  // against a cancellation-free evaluation of the same state.  The sums x +
  // u round at the coordinates' magnitude, so per bond
  //   de <= 8 eps (max(|x_i + u_i|_inf, |x_j + u_j|_inf) + |X|);
  // the dilatation then carries dth_i <= (3/m_i) sum_j |X| c V_j de, and
  // the force f_i = sum_j V_j c (P_i + P_j) |X| Y/|Y| with P = 3 K th/m
  // (e_d = 0 in 3D, ElasticState.cpp:227-245) carries
  //   df_i <= sum_j V_j c |X| (3K/m_i dth_i + 3K/m_j dth_j).
  void extension_rounding_bound(double* out) {
    const double eps = 0x1p-53;
    std::vector<double> dth(size_t(n), 0.);
    auto inf_norm = [&](int64_t a) {
      const P3 y = x[a] + u[a];
      return std::max(std::fabs(y.x), std::max(std::fabs(y.y), std::fabs(y.z)));
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double acc = 0.;
      const double yi = inf_norm(i);
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        const int64_t j = col[k];
        const double lx = length(x[j] - x[i]);
        const double de = 8. * eps * (std::max(yi, inf_norm(j)) + lx);
        acc += lx * corr[k] * vol[j] * de;
      }
      dth[size_t(i)] = 3. / mw[i] * acc;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double acc = 0.;
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        const int64_t j = col[k];
        const double lx = length(x[j] - x[i]);
        acc += vol[j] * corr[k] * lx *
               (3. * m.K / mw[i] * dth[size_t(i)] +
                3. * m.K / mw[j] * dth[size_t(j)]);
      }
      out[i] = acc;
    }
  }

  void apply_ubc(double t) {
    loader->u = u;
    loader->v = v;
    loader->apply_ubc(t);
    u = loader->u;
    v = loader->v;
  }
  void apply_fbc(double t) { loader->apply_fbc(t, f); }

  void init() {
    apply_ubc(time);
    apply_fbc(time);
    forces(f);
  }

  void step_cd() {
    double delta_t = dt;
    double fact = delta_t * delta_t / m.density;
    for (int64_t i = 0; i < n; i++) {
      double* ui = &u[i].x;
      double* vi = &v[i].x;
      double* fi = &f[i].x;
      for (int d = 0; d < dim; d++) {
        if (!(fix[i] >> d & 1UL)) {
          double u_old = ui[d];
          ui[d] += fact * fi[d] + delta_t * vi[d];
          vi[d] = (ui[d] - u_old) / delta_t;
        }
      }
      f[i] = P3{};
    }
    time += dt;
    apply_ubc(time);
    apply_fbc(time);
    forces(f);
  }
};

}  // namespace

extern "C" {

int ora_es_params(const ora_es_deck* d, ora_es* o) {
  std::memset(o, 0, sizeof(*o));
  o->dim = d->dim;
  o->horizon = d->horizon;
  o->h = d->h;
  o->density = d->density;
  o->E = d->E;
  o->K = d->K;
  o->G = d->G;
  o->plane_strain = d->plane_strain;
  o->sc = d->sc;
  if (d->dim == 1) {
    if (dlt(d->E, 0.)) return -1;
  } else {
    if (d->K <= 0. || d->G <= 0.) return -2;
  }
  // inp::MatData::toNuClassical, materialDeck.h:120
  o->nu = (3 * d->K - 2 * d->G) / (2 * (3 * d->K + d->G));
  o->mu = d->G;
  if (d->plane_strain == 0 && d->dim == 2)
    o->factor2D = (2. * o->nu - 1.) / (o->nu - 1.);
  else
    o->factor2D = 1;
  return 0;
}

void* ora_es_create(int dim, int64_t n, const double* x, const double* vol,
                    const int64_t* row_ptr, const int32_t* col,
                    const ora_es* mat, double dt, int n_ubc, const ora_bc* ubc,
                    int n_fbc, const ora_bc* fbc) {
  ES* s = new ES;
  s->dim = dim;
  s->n = n;
  s->x.resize(n);
  for (int64_t i = 0; i < n; i++) s->x[i] = p3(x + 3 * i);
  s->vol.assign(vol, vol + n);
  s->row_ptr.assign(row_ptr, row_ptr + n + 1);
  s->col.assign(col, col + row_ptr[n]);
  s->m = *mat;
  s->dt = dt;
  s->broken.assign(s->col.size(), 0);
  s->u.assign(n, P3{});
  s->v.assign(n, P3{});
  s->f.assign(n, P3{});
  // BC node selection through the FD loader (same ULoading/FLoading code)
  FD* L = new FD;
  L->dim = dim;
  L->n = n;
  L->x = s->x;
  L->ubc.assign(ubc, ubc + n_ubc);
  L->fbc.assign(fbc, fbc + n_fbc);
  L->fix.assign(n, 0);
  L->u.assign(n, P3{});
  L->v.assign(n, P3{});
  L->init_ubc();
  L->init_fbc();
  s->fix = L->fix;
  s->loader = L;
  s->init_volume();
  return s;
}
void ora_es_free(void* h) {
  ES* s = static_cast<ES*>(h);
  delete s->loader;
  delete s;
}
void ora_es_init(void* h) { static_cast<ES*>(h)->init(); }
void ora_es_step_cd(void* h, int nsteps) {
  ES* s = static_cast<ES*>(h);
  for (int k = 0; k < nsteps; k++) s->step_cd();
}
void ora_es_compute_forces(void* h) {
  ES* s = static_cast<ES*>(h);
  s->f.assign(s->n, P3{});
  s->forces(s->f);
}
void ora_es_get_state(void* h, double* u, double* v, double* f) {
  ES* s = static_cast<ES*>(h);
  for (int64_t i = 0; i < s->n; i++)
    for (int d = 0; d < 3; d++) {
      if (u) u[3 * i + d] = (&s->u[i].x)[d];
      if (v) v[3 * i + d] = (&s->v[i].x)[d];
      if (f) f[3 * i + d] = (&s->f[i].x)[d];
    }
}
void ora_es_set_state(void* h, const double* u, const double* v,
                      const double* f) {
  ES* s = static_cast<ES*>(h);
  for (int64_t i = 0; i < s->n; i++) {
    if (u) s->u[i] = p3(u + 3 * i);
    if (v) s->v[i] = p3(v + 3 * i);
    if (f) s->f[i] = p3(f + 3 * i);
  }
}
void ora_es_get_bonds(void* h, uint8_t* b) {
  ES* s = static_cast<ES*>(h);
  std::copy(s->broken.begin(), s->broken.end(), b);
}
void ora_es_set_bonds(void* h, const uint8_t* b) {
  ES* s = static_cast<ES*>(h);
  for (size_t k = 0; k < s->broken.size(); k++) s->broken[k] = b[k] ? 1 : 0;
}
void ora_es_get_node_fields(void* h, double* theta, double* mvol) {
  ES* s = static_cast<ES*>(h);
  for (int64_t i = 0; i < s->n; i++) {
    if (theta) theta[i] = s->theta.empty() ? 0. : s->theta[i];
    if (mvol) mvol[i] = s->mw[i];
  }
}
double ora_es_time(void* h) { return static_cast<ES*>(h)->time; }
void ora_es_set_time(void* h, double t) { static_cast<ES*>(h)->time = t; }
void ora_es_fields(void* h, float* energy_acc, double* strain6,
                   double* stress6) {
  static_cast<ES*>(h)->fields(energy_acc, strain6, stress6);
}
void ora_es_extension_rounding_bound(void* h, double* out) {
  static_cast<ES*>(h)->extension_rounding_bound(out);
}
void ora_fd_tangent(void* h, double* k) { static_cast<FD*>(h)->tangent(k); }
void ora_es_force_abs_sum(void* h, double* out) {
  static_cast<ES*>(h)->force_abs_sum(out);
}

int ora_definitely_greater(double a, double b) { return dgt(a, b) ? 1 : 0; }
int ora_definitely_less(double a, double b) { return dlt(a, b) ? 1 : 0; }

// RNPBond::RNPBond (rnpBond.cpp:23-97), computeParameters (:99-189),
// computeMaterialProperties (:191-212); conversions materialDeck.h:112-215.
int ora_rnp_params(const ora_mat_deck* d, ora_rnp* o) {
  std::memset(o, 0, sizeof(*o));
  o->dim = d->dim;
  o->horizon = d->horizon;
  o->h = d->h;
  o->density = d->density;
  if (d->inf_type < 0 || d->inf_type > 2) return -1;
  Influence J = make_influence(d->inf_type, d->n_inf_params, d->inf_params,
                               d->dim);
  o->inf_type = J.type;
  o->inf_a = J.a;
  o->inf_b = J.b;
  double M = J.moment(size_t(d->dim));
  o->inf_moment = M;
  o->irreversible = d->irreversible;
  o->factor_sc = d->check_sc_factor;
  if (d->dim == 1)
    o->inv_factor = std::pow(d->horizon, 2) * 2.;
  else if (d->dim == 2)
    o->inv_factor = std::pow(d->horizon, 3) * M_PI;
  else if (d->dim == 3)
    o->inv_factor = std::pow(d->horizon, 4) * 4. * M_PI / 3.;

  double E = d->E, K = d->K, Gc = d->Gc, KIc = d->KIc, nu, lambda, G;
  if (d->compute_from_classical) {
    if (dlt(E, 0.) && dlt(K, 0.)) return -2;
    if (dlt(Gc, 0.) && dlt(KIc, 0.)) return -3;
    nu = d->plane_strain ? 0.25 : 1. / 3.;
    if (E > 0.) K = E / (3. * (1. - 2. * nu));
    if (K > 0. && E < 0.) E = K * (3. * (1. - 2. * nu));
    if (Gc > 0.) KIc = std::sqrt(Gc * E);
    if (KIc > 0. && Gc < 0.) Gc = KIc * KIc / E;
    lambda = E * nu / ((1. + nu) * (1. - 2. * nu));
    G = E / (2. * (1. + nu));
    if (d->dim == 2) {
      o->C = M_PI * Gc / (4. * M);
      o->beta = 4. * lambda / (o->C * M);
    } else if (d->dim == 3) {
      o->C = 2. * Gc / (3. * M);
      o->beta = 5. * lambda / (o->C * M);
    }
    o->rbar = std::sqrt(0.5 / o->beta);
  } else {
    if (d->n_bond_params < 2) return -4;
    o->C = d->bond_params[0];
    o->beta = d->bond_params[1];
    o->rbar = std::sqrt(0.5 / o->beta);
    nu = 0.33;
    if (d->dim == 2) {
      Gc = 4. * M * o->C / M_PI;
      lambda = o->C * M * o->beta / 4.;
    } else {
      Gc = 3. * M * o->C / 2.;
      lambda = o->C * M * o->beta / 5.;
    }
    G = lambda;
    E = lambda * (1. + nu) * (1. - 2. * nu) / nu;
    K = E / (3. * (1. - 2. * nu));
    KIc = std::sqrt(Gc * E);
  }
  o->E = E;
  o->K = K;
  o->nu = nu;
  o->lambda = lambda;
  o->G = G;
  o->Gc = Gc;
  o->KIc = KIc;
  o->contact = d->no_penetration;
  if (d->no_penetration) {  // rnpBond.cpp:85-94
    o->contact_rc = 0.9 * d->h;
    o->contact_kn = K * 18. / (M_PI * std::pow(d->horizon, 5));
  }
  return 0;
}

void ora_bond_ef(const ora_rnp* m, const double* xi, const double* ui,
                 const double* xj, const double* uj, double voli, double volj,
                 int interior_i, int interior_j, int* fs, double* force3,
                 double* energy) {
  Influence J = influence_of(*m);
  bool b = *fs != 0;
  P3 F;
  double E;
  bond_ef(*m, J, p3(xi), p3(ui), p3(xj), p3(uj), voli, volj, interior_i != 0,
          interior_j != 0, b, F, E);
  if (interior_i && interior_j) *fs = b ? 1 : 0;
  force3[0] = F.x;
  force3[1] = F.y;
  force3[2] = F.z;
  *energy = E;
}

void* ora_neighbor_bruteforce(int64_t n, const double* x, double horizon) {
  return neighbor_bruteforce(n, x, horizon);
}
void* ora_neighbor_cells(int64_t n, const double* x, double horizon) {
  return neighbor_cells(n, x, horizon);
}
int64_t ora_nbr_nnz(void* h) { return int64_t(static_cast<Nbr*>(h)->col.size()); }
void ora_nbr_copy(void* h, int64_t* row_ptr, int32_t* col) {
  Nbr* b = static_cast<Nbr*>(h);
  std::copy(b->row_ptr.begin(), b->row_ptr.end(), row_ptr);
  std::copy(b->col.begin(), b->col.end(), col);
}
void ora_nbr_free(void* h) { delete static_cast<Nbr*>(h); }

void* ora_fd_create(int dim, int64_t n, const double* x, const double* vol,
                    const int64_t* row_ptr, const int32_t* col,
                    const ora_rnp* mat, double dt, int n_ubc, const ora_bc* ubc,
                    int n_fbc, const ora_bc* fbc, int n_cracks,
                    const ora_crack* cracks, const uint8_t* interior) {
  FD* s = new FD;
  s->dim = dim;
  s->n = n;
  s->x.resize(n);
  for (int64_t i = 0; i < n; i++) s->x[i] = p3(x + 3 * i);
  s->vol.assign(vol, vol + n);
  s->row_ptr.assign(row_ptr, row_ptr + n + 1);
  s->col.assign(col, col + row_ptr[n]);
  s->m = *mat;
  s->J = influence_of(*mat);
  s->dt = dt;
  s->ubc.assign(ubc, ubc + n_ubc);
  s->fbc.assign(fbc, fbc + n_fbc);
  for (int k = 0; k < n_cracks; k++) {
    Crack c;
    c.c = cracks[k];
    s->cracks.push_back(c);
  }
  s->interior.assign(n, 1);
  if (interior) s->interior.assign(interior, interior + n);
  s->fix.assign(n, 0);
  s->u.assign(n, P3{});
  s->v.assign(n, P3{});
  s->f.assign(n, P3{});
  s->init_fracture();
  s->init_ubc();
  s->init_fbc();
  return s;
}
void ora_fd_free(void* h) { delete static_cast<FD*>(h); }
void ora_fd_init(void* h) { static_cast<FD*>(h)->init(); }
void ora_fd_step_cd(void* h, int nsteps) {
  FD* s = static_cast<FD*>(h);
  for (int k = 0; k < nsteps; k++) s->step_cd();
}
void ora_fd_step_verlet(void* h, int nsteps) {
  FD* s = static_cast<FD*>(h);
  for (int k = 0; k < nsteps; k++) s->step_verlet();
}
void ora_fd_compute_forces(void* h) { static_cast<FD*>(h)->compute_forces(); }
void ora_fd_get_state(void* h, double* u, double* v, double* f) {
  FD* s = static_cast<FD*>(h);
  for (int64_t i = 0; i < s->n; i++) {
    if (u) {
      u[3 * i] = s->u[i].x;
      u[3 * i + 1] = s->u[i].y;
      u[3 * i + 2] = s->u[i].z;
    }
    if (v) {
      v[3 * i] = s->v[i].x;
      v[3 * i + 1] = s->v[i].y;
      v[3 * i + 2] = s->v[i].z;
    }
    if (f) {
      f[3 * i] = s->f[i].x;
      f[3 * i + 1] = s->f[i].y;
      f[3 * i + 2] = s->f[i].z;
    }
  }
}
void ora_fd_set_state(void* h, const double* u, const double* v,
                      const double* f) {
  FD* s = static_cast<FD*>(h);
  for (int64_t i = 0; i < s->n; i++) {
    if (u) s->u[i] = p3(u + 3 * i);
    if (v) s->v[i] = p3(v + 3 * i);
    if (f) s->f[i] = p3(f + 3 * i);
  }
}
void ora_fd_get_bonds(void* h, uint8_t* out) {
  FD* s = static_cast<FD*>(h);
  for (int64_t i = 0; i < s->n; i++)
    for (int64_t k = s->row_ptr[i]; k < s->row_ptr[i + 1]; k++)
      out[k] = s->get_bond(i, k - s->row_ptr[i]) ? 1 : 0;
}
void ora_fd_set_bonds(void* h, const uint8_t* in) {
  FD* s = static_cast<FD*>(h);
  for (int64_t i = 0; i < s->n; i++)
    for (int64_t k = s->row_ptr[i]; k < s->row_ptr[i + 1]; k++)
      s->set_bond(i, k - s->row_ptr[i], in[k] != 0);
}
int64_t ora_fd_bond_bytes_size(void* h) {
  FD* s = static_cast<FD*>(h);
  int64_t t = 0;
  for (auto& r : s->frac) t += int64_t(r.size());
  return t;
}
void ora_fd_get_bond_bytes(void* h, uint8_t* bytes) {
  FD* s = static_cast<FD*>(h);
  int64_t o = 0;
  for (auto& r : s->frac)
    for (uint8_t b : r) bytes[o++] = b;
}
void ora_fd_get_fixity(void* h, uint8_t* fix) {
  FD* s = static_cast<FD*>(h);
  std::copy(s->fix.begin(), s->fix.end(), fix);
}
double ora_fd_time(void* h) { return static_cast<FD*>(h)->time; }
int64_t ora_fd_step_count(void* h) { return static_cast<FD*>(h)->nstep; }
void ora_fd_external_force(void* h, double t, double* out) {
  static_cast<FD*>(h)->external_force(t, out);
}
void ora_fd_reaction_force(void* h, const double* box, double delta,
                           double* R, double* R_len) {
  static_cast<FD*>(h)->reaction_force(box, delta, R, R_len);
}
void ora_fd_postproc(void* h, float* Z, float* phi, double* energy) {
  static_cast<FD*>(h)->postproc(Z, phi, energy);
}
void ora_fd_force_abs_sum(void* h, double* out) {
  static_cast<FD*>(h)->force_abs_sum(out);
}
void ora_fd_set_faithful(void* h, int on) {
  static_cast<FD*>(h)->set_faithful(on != 0);
}
void ora_fd_set_time(void* h, double t, int64_t step) {
  static_cast<FD*>(h)->time = t;
  static_cast<FD*>(h)->nstep = step;
}

void ora_interior_flags(int64_t n, const double* x, const double* lo,
                        const double* hi, double tol, int nreg,
                        const double* regions, uint8_t* out) {
  for (int64_t i = 0; i < n; i++) {
    P3 xx = p3(x + 3 * i);
    bool in = true;
    if (dlt(xx.x, lo[0] + tol) || dlt(xx.y, lo[1] + tol) ||
        dgt(xx.x, hi[0] - tol) || dgt(xx.y, hi[1] - tol))
      in = false;
    for (int r = 0; in && r < nreg; r++) {
      const double* p = regions + 5 * r;
      if (p[0] == 1.0) {  // Circle {cx, cy, cz, radius}
        P3 c{p[1], p[2], p[3]};
        if (dlt(dist(c, xx), p[4])) in = false;
      } else {  // Rectangle {x1, y1, x2, y2}
        if (inside_rectangle(xx, p[1], p[3], p[2], p[4])) in = false;
      }
    }
    out[i] = in ? 1 : 0;
  }
}

void ora_bbox(int64_t n, const double* x, double* lo, double* hi) {
  double p1[3] = {0., 0., 0.}, p2[3] = {0., 0., 0.};
  for (int64_t i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      if (dlt(x[3 * i + d], p1[d])) p1[d] = x[3 * i + d];
      if (dlt(p2[d], x[3 * i + d])) p2[d] = x[3 * i + d];
    }
  }
  for (int d = 0; d < 3; d++) {
    lo[d] = p1[d];
    hi[d] = p2[d];
  }
}

int ora_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
