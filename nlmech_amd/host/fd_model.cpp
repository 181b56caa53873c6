// model::FDModel<T> on the GPU (see model.h).  Each method cites the
// reference step it replaces in src/model/fd/fDModel.cpp.
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>

#include "model.h"
#include "vtu.h"

namespace nlm {

namespace {

[[noreturn]] void deck_error(const std::string& m) {
  throw inp::DeckError(m);
}

// definitelyLessThan / definitelyGreaterThan (src/util/compare.cpp:25-33)
bool less_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (b - a) > m * 1e-5;
}
bool greater_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (a - b) > m * 1e-5;
}

// util::geometry::isPointInsideRectangle (src/util/utilGeom.cpp:43-50)
bool inside_rect(const double* x, double x1, double x2, double y1, double y2) {
  return !(less_than(x[0], x1 - 1.0E-12) || less_than(x[1], y1 - 1.0E-12) ||
           greater_than(x[0], x2 + 1.0E-12) ||
           greater_than(x[1], y2 + 1.0E-12));
}

int region_code(const std::string& r) {
  if (r == "rectangle") return NLM_REGION_RECTANGLE;
  if (r == "angled_rectangle") return NLM_REGION_ANGLED_RECTANGLE;
  if (r == "circle") return NLM_REGION_CIRCLE;
  if (r == "torus") return NLM_REGION_TORUS;
  if (r == "line") return NLM_REGION_LINE;
  if (r == "cuboid") return NLM_REGION_CUBOID;
  deck_error("Error: BC region type " + r + " is not supported.");
}

int time_code(const std::string& t) {
  if (t == "constant") return NLM_TFN_CONSTANT;
  if (t == "linear") return NLM_TFN_LINEAR;
  if (t == "quadratic") return NLM_TFN_QUADRATIC;
  if (t == "sin") return NLM_TFN_SIN;
  if (t == "linear_step") return NLM_TFN_LINEAR_STEP;
  if (t == "linear_slow_fast") return NLM_TFN_LINEAR_SLOW_FAST;
  deck_error("Error: Time function type " + t + " not recognized.");
}

int space_code(const std::string& s) {
  static const char* names[] = {"constant", "sin_x",    "sin_y", "sin_z",
                                "linear_x", "linear_y", "linear_z",
                                "hat_x",    "hat_y",    "hat_z", "line_load"};
  for (int k = 0; k < 11; k++)
    if (s == names[k]) return k;
  deck_error("Error: Spatial function type " + s + " not recognized.");
}

nlm_bc to_bc(const inp::BCData& b) {
  nlm_bc s;
  std::memset(&s, 0, sizeof(s));
  s.region = region_code(b.d_regionType);
  s.x1 = b.d_x1;
  s.x2 = b.d_x2;
  s.y1 = b.d_y1;
  s.y2 = b.d_y2;
  s.z1 = b.d_z1;
  s.z2 = b.d_z2;
  s.r1 = b.d_r1;
  s.r2 = b.d_r2;
  s.theta = b.d_theta;
  if (b.d_direction.size() > 3) deck_error("Error: at most 3 BC directions");
  s.ndir = int(b.d_direction.size());
  for (int k = 0; k < s.ndir; k++) s.dir[k] = int(b.d_direction[k]);
  s.time_fn = b.d_timeFnType.empty() ? NLM_TFN_CONSTANT
                                     : time_code(b.d_timeFnType);
  if (b.d_timeFnParams.size() > 4) deck_error("Error: too many parameters");
  s.n_time_params = int(b.d_timeFnParams.size());
  for (int k = 0; k < s.n_time_params; k++)
    s.time_params[k] = b.d_timeFnParams[k];
  s.space_fn = b.d_spatialFnType.empty() ? NLM_SFN_CONSTANT
                                         : space_code(b.d_spatialFnType);
  if (b.d_spatialFnParams.size() > 4) deck_error("Error: too many parameters");
  s.n_space_params = int(b.d_spatialFnParams.size());
  for (int k = 0; k < s.n_space_params; k++)
    s.space_params[k] = b.d_spatialFnParams[k];
  return s;
}

int precision_code(const std::string& p) {
  if (p == "fp64") return NLM_PREC_FP64;
  if (p == "fp32_force") return NLM_PREC_FP32_FORCE;
  return NLM_PREC_FP64_EXACT;
}

// util::function::gaussian (src/util/utilFunction.cpp:66-69)
double gaussian(double r, double a, double beta) {
  return a * std::exp(-std::pow(r, 2) / beta);
}

// InitialCondition::getICFormula (src/loading/initialCondition.cpp:79-94)
double ic_formula(const inp::ICData& d, const double* x, int dof, size_t dim) {
  const auto& p = d.d_params;
  auto r_to = [&](double cx, double cy) {
    const double dx = x[0] - cx, dy = x[1] - cy, dz = x[2] - 0.;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  if (d.d_type == "gaussian" && dim == 2) {
    if (p.size() < 6)
      deck_error("Error: Not enough parameters to compute guassian 2-d "
                 "function.");
    return gaussian(r_to(p[0], p[1]), p[5], p[4]) * p[2 + dof];
  }
  if (d.d_type == "double_gaussian" && dim == 2) {
    if (p.size() < 10)
      deck_error("Error: Not enough parameters to compute guassian 2-d "
                 "function.");
    return gaussian(r_to(p[0], p[1]), p[9], p[8]) * p[4 + dof] +
           gaussian(r_to(p[2], p[3]), p[9], p[8]) * p[6 + dof];
  }
  deck_error("Error: Check initial condition flag = " + d.d_type +
             ". Currently only guassian and double_guassian type functions "
             "are supported in 2-d.");
}

}  // namespace

namespace material {
namespace pd {

RNPBond::RNPBond(const inp::MaterialDeck& deck, size_t dim, double horizon,
                 double h) {
  nlm_material_deck d;
  std::memset(&d, 0, sizeof(d));
  d.dim = int(dim);
  d.horizon = horizon;
  d.h = h;
  d.density = deck.d_density;
  d.E = deck.d_matData.d_E;
  d.K = deck.d_matData.d_K;
  d.Gc = deck.d_matData.d_Gc;
  d.KIc = deck.d_matData.d_KIc;
  d.plane_strain = deck.d_isPlaneStrain ? 1 : 0;
  d.compute_from_classical = deck.d_computeParamsFromElastic ? 1 : 0;
  if (deck.d_bondPotentialParams.size() > 2 ||
      deck.d_influenceFnParams.size() > 2)
    deck_error("Error: too many bond potential or influence parameters");
  d.n_bond_params = int(deck.d_bondPotentialParams.size());
  for (int k = 0; k < d.n_bond_params; k++)
    d.bond_params[k] = deck.d_bondPotentialParams[k];
  d.inf_type = int(deck.d_influenceFnType);
  d.n_inf_params = int(deck.d_influenceFnParams.size());
  for (int k = 0; k < d.n_inf_params; k++)
    d.inf_params[k] = deck.d_influenceFnParams[k];
  d.check_sc_factor = deck.d_checkScFactor;
  d.irreversible = deck.d_irreversibleBondBreak ? 1 : 0;
  d.no_penetration = deck.d_applyContact ? 1 : 0;
  if (nlm_rnp_params_from_deck(&d, &d_params) != NLM_OK)
    deck_error("Error: RNPBond parameters: check E/K, Gc/KIc, the influence "
               "function and the horizon in the material deck.");
}

void RNPBond::fill(nlm_fd_problem* p) const {
  p->material = NLM_MATERIAL_RNP_BOND;
  p->mat = d_params;
}

ElasticState::ElasticState(const inp::MaterialDeck& deck, size_t dim,
                           double horizon, double h) {
  if (!deck.d_computeParamsFromElastic)
    deck_error("Parameters from classical continuum mechanics are necessary "
               "to compute the elastic state-based peridynamic properties.");
  nlm_es_deck d;
  std::memset(&d, 0, sizeof(d));
  d.dim = int(dim);
  d.horizon = horizon;
  d.h = h;
  d.density = deck.d_density;
  d.E = deck.d_matData.d_E;
  d.K = deck.d_matData.d_K;
  d.G = deck.d_matData.d_G;
  d.plane_strain = deck.d_isPlaneStrain ? 1 : 0;
  d.critical_stretch = deck.d_criticalStretch;
  if (nlm_es_params_from_deck(&d, &d_params) != NLM_OK)
    deck_error("Error: ElasticState parameters: 1D needs E, 2D/3D need K and "
               "G.");
}

void ElasticState::fill(nlm_fd_problem* p) const {
  p->material = NLM_MATERIAL_ELASTIC_STATE;
  p->es = d_params;
  p->mat.density = d_params.density;
}

}  // namespace pd
}  // namespace material

namespace model {

template <class T>
FDModel<T>::FDModel(inp::Input* deck) : FDModel(deck, true) {}

template <class T>
FDModel<T>::FDModel(inp::Input* deck, bool run_now)
    : d_input_p(deck),
      d_modelDeck_p(deck->getModelDeck()),
      d_outputDeck_p(deck->getOutputDeck()) {
  if (!run_now) return;
  if (d_modelDeck_p->d_isRestartActive)
    restart();
  else
    run();
}

template <class T>
FDModel<T>::~FDModel() {
  if (d_ctx) nlm_ctx_destroy(d_ctx);
}

template <class T>
void FDModel<T>::check(int rc, const char* what) {
  if (rc != NLM_OK)
    throw Error(std::string(what) + " failed: " + nlm_last_error(d_ctx));
}

template <class T>
void FDModel<T>::run() {
  const std::string& td = d_modelDeck_p->d_timeDiscretization;
  if (td != "central_difference" && td != "velocity_verlet")
    deck_error("Error: FDModel needs Time: central_difference or "
               "velocity_verlet.");
  initHObjects();
  init();
  integrate();
}

// InteriorFlags (src/geometry/interiorFlags.cpp:57-100, 122-183): with a
// No_Fail_Region deck a node within noFailTol of the bounding box, or inside
// a listed circle / rectangle, is not breakable.
template <class T>
std::vector<uint8_t> FDModel<T>::interiorFlags() const {
  const inp::InteriorFlagsDeck* d = d_input_p->getInteriorFlagsDeck();
  const size_t n = d_mesh->getNumNodes();
  std::vector<uint8_t> flag(n, 1);
  const auto& bb = d_mesh->getBoundingBox();
  const double tol = d->d_noFailTol;
  const double* X = d_mesh->getNodes().data();
  for (size_t i = 0; i < n; i++) {
    const double* x = X + 3 * i;
    bool f = true;
    if (less_than(x[0], bb.first[0] + tol)) f = false;
    if (less_than(x[1], bb.first[1] + tol)) f = false;
    if (greater_than(x[0], bb.second[0] - tol)) f = false;
    if (greater_than(x[1], bb.second[1] - tol)) f = false;
    for (const auto& r : d->d_noFailRegions) {
      if (r.first == "Circle") {
        if (r.second.size() < 4) deck_error("Error: Circle needs 4 parameters");
        const double dx = r.second[0] - x[0], dy = r.second[1] - x[1],
                     dz = r.second[2] - x[2];
        if (less_than(std::sqrt(dx * dx + dy * dy + dz * dz), r.second[3]))
          f = false;
      } else if (r.first == "Rectangle") {
        if (r.second.size() < 4)
          deck_error("Error: Rectangle needs 4 parameters");
        if (inside_rect(x, r.second[0], r.second[2], r.second[1],
                        r.second[3]))
          f = false;
      }
    }
    flag[i] = f ? 1 : 0;
  }
  return flag;
}

// initHObjects (fDModel.cpp:134-190): mesh, neighbor list, pre-cracks,
// interior flags, loading, material, all handed to the GPU library.
template <class T>
void FDModel<T>::initHObjects() {
  std::cout << "FDModel: Initializing high level objects.\n";
  d_mesh.reset(new fe::Mesh(*d_input_p->getMeshDeck()));
  const size_t dim = d_mesh->getDimension();
  std::cout << "number of nodes = " << d_mesh->getNumNodes()
            << " number of elements = " << d_mesh->getNumElements() << "\n";
  if (dim < 1 || dim > 3)
    deck_error("Error: the GPU engine runs 1D, 2D and 3D models.");
  const std::string& td = d_modelDeck_p->d_timeDiscretization;
  if (td != "central_difference" && td != "velocity_verlet" &&
      td != "quasi_static")
    deck_error("Error: Time discretization " + td + " is not known.");
  d_material.reset(new T(*d_input_p->getMaterialDeck(), dim,
                         d_modelDeck_p->d_horizon, d_mesh->getMeshSize()));

  const inp::GpuDeck* gpu = d_input_p->getGpuDeck();
  check(nlm_ctx_create(&d_ctx, gpu->d_device), "nlm_ctx_create");
  const int64_t n = int64_t(d_mesh->getNumNodes());
  std::cout << "FDModel: Creating neighbor list.\n";
  check(nlm_neighbor_build(d_ctx, int(dim), n, d_mesh->getNodes().data(),
                           d_modelDeck_p->d_horizon, &d_nnz),
        "nlm_neighbor_build");

  nlm_fd_problem p;
  std::memset(&p, 0, sizeof(p));
  p.dim = int(dim);
  p.n = n;
  p.x = d_mesh->getNodes().data();
  p.vol = d_mesh->getNodalVolumes().data();
  std::vector<uint8_t> interior;
  if (d_input_p->getInteriorFlagsDeck()->d_noFailActive) {
    interior = interiorFlags();
    p.interior = interior.data();
  }
  p.dt = d_modelDeck_p->d_dt;
  p.precision = precision_code(gpu->d_precision);
  p.time_scheme = td == "velocity_verlet" ? NLM_TIME_VELOCITY_VERLET
                                          : NLM_TIME_CENTRAL_DIFFERENCE;
  d_material->fill(&p);
  check(nlm_fd_setup(d_ctx, &p), "nlm_fd_setup");

  const inp::LoadingDeck* ld = d_input_p->getLoadingDeck();
  std::vector<nlm_bc> ubc, fbc;
  for (const auto& b : ld->d_uBCData) ubc.push_back(to_bc(b));
  for (const auto& b : ld->d_fBCData) fbc.push_back(to_bc(b));
  if (!ubc.empty())
    check(nlm_fd_set_ubc(d_ctx, int(ubc.size()), ubc.data()),
          "nlm_fd_set_ubc");
  if (!fbc.empty())
    check(nlm_fd_set_fbc(d_ctx, int(fbc.size()), fbc.data()),
          "nlm_fd_set_fbc");

  std::cout << "FDModel: Creating edge crack if any and modifying the "
               "fracture state of bonds.\n";
  std::vector<nlm_crack> cr;
  for (const auto& c : d_input_p->getFractureDeck()->d_cracks) {
    nlm_crack k;
    std::memset(&k, 0, sizeof(k));
    k.o = c.d_o;
    k.pb[0] = c.d_pb[0];
    k.pb[1] = c.d_pb[1];
    k.pt[0] = c.d_pt[0];
    k.pt[1] = c.d_pt[1];
    k.theta = c.d_theta;
    k.activation_time = c.d_activationTime;
    if (c.d_activationTime >= 0.) d_pendingCracks = true;
    cr.push_back(k);
  }
  if (!cr.empty())
    check(nlm_fd_add_cracks(d_ctx, int(cr.size()), cr.data()),
          "nlm_fd_add_cracks");
  d_fix.assign(size_t(n), 0);
  check(nlm_fd_get_fixity(d_ctx, d_fix.data()), "nlm_fd_get_fixity");
  if (gpu->d_autotune) {
    // staged kernels: tile and chunk width together; others: the chunk
    int tile = 0, chosen = 0;
    double us[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
    check(nlm_fd_autotune_tiles(d_ctx, &tile, &chosen, us),
          "nlm_fd_autotune_tiles");
    const int t = tile == 128 ? 0 : tile == 512 ? 2 : 1;
    const int k = chosen == 2 ? 0 : chosen == 8 ? 2 : 1;
    std::cout << "FDModel: autotuner chose tiles of " << tile << " nodes, "
              << chosen << " slots per pass (" << us[3 * t + k]
              << " us per step)\n";
  }
}

// init (fDModel.cpp:193-315): the state lives on the device (u = v = f = 0
// after nlm_fd_setup); here only the output-criteria checks remain.
template <class T>
void FDModel<T>::init() {
  std::cout << "FDModel: Initializing basic datas.\n";
  d_n = 0;
  d_time = 0.;
  inp::OutputDeck* od = d_outputDeck_p;
  if (od->d_dtOut == 0)
    deck_error("Error: Output: Output_Interval must be given and positive.");
  if (od->d_outCriteria == "max_Z" || od->d_outCriteria == "max_Z_stop") {
    const size_t need = od->d_outCriteria == "max_Z_stop" ? 6 : 1;
    if (od->d_outCriteriaParams.size() < need)
      deck_error("Error: Output criteria " + od->d_outCriteria + " requires " +
                 std::to_string(need) + " parameters. ");
    if (!d_input_p->getPolicyDeck()->d_enablePostProcessing) {
      std::cout << "Warning: Output criteria " << od->d_outCriteria
                << " requires Damage data Z but either postprocessing is set "
                   "to off. Therefore setting output criteria to null.\n";
      od->d_outCriteria.clear();
    }
  }
}

// InitialCondition::apply (src/loading/initialCondition.cpp:20-64) on the
// free dofs.
template <class T>
void FDModel<T>::applyInitialCondition() {
  const inp::InitialConditionDeck* ic = d_input_p->getInitialConditionDeck();
  const bool do_u = !ic->d_uICData.d_type.empty() &&
                    ic->d_uICData.d_type != "file";
  const bool do_v = !ic->d_vICData.d_type.empty() &&
                    ic->d_vICData.d_type != "file";
  if (!do_u && !do_v) return;
  const size_t n = d_mesh->getNumNodes(), dim = d_mesh->getDimension();
  std::vector<double> u(3 * n, 0.), v(3 * n, 0.);
  const double* X = d_mesh->getNodes().data();
  for (int pass = 0; pass < 2; pass++) {
    const bool on = pass == 0 ? do_u : do_v;
    if (!on) continue;
    const inp::ICData& d = pass == 0 ? ic->d_uICData : ic->d_vICData;
    std::vector<double>& w = pass == 0 ? u : v;
    for (size_t i = 0; i < n; i++)
      for (size_t k = 0; k < dim; k++)
        if (!(d_fix[i] >> k & 1))
          w[3 * i + k] = ic_formula(d, X + 3 * i, int(k), dim);
  }
  check(nlm_fd_set_state(d_ctx, do_u ? u.data() : nullptr,
                         do_v ? v.data() : nullptr, nullptr),
        "nlm_fd_set_state");
}

// integrate (fDModel.cpp:318-374): initial condition and loading, forces,
// output at step 0, then the time loop.  The steps between two outputs run
// as one nlm_fd_step call (optionally through a captured HIP graph).
template <class T>
void FDModel<T>::integrate() {
  if (d_n == 0) applyInitialCondition();
  check(nlm_fd_set_time(d_ctx, d_time, int64_t(d_n)), "nlm_fd_set_time");
  check(nlm_fd_init(d_ctx), "nlm_fd_init");
  const bool pp = d_input_p->getPolicyDeck()->d_enablePostProcessing;
  if (d_n == 0) {
    if (pp) computePostProcFields();
    output();
  }
  const size_t Nt = size_t(d_modelDeck_p->d_Nt);
  const int chunk = d_input_p->getGpuDeck()->d_graphChunk;
  const bool graph = chunk > 0 && !d_pendingCracks &&
                     d_modelDeck_p->d_timeDiscretization ==
                         "central_difference";
  while (d_n < Nt) {
    const size_t dt_out = d_outputDeck_p->d_dtOut;
    size_t next = (d_n / dt_out + 1) * dt_out;
    if (next > Nt) next = Nt;
    const int steps = int(next - d_n);
    if (graph && steps >= 2)
      check(nlm_fd_step_graph(d_ctx, steps, chunk & ~1 ? chunk & ~1 : 2),
            "nlm_fd_step_graph");
    else
      check(nlm_fd_step(d_ctx, steps), "nlm_fd_step");
    int64_t step = 0;
    check(nlm_fd_get_time(d_ctx, &d_time, &step), "nlm_fd_get_time");
    d_n = size_t(step);
    if (d_n % d_outputDeck_p->d_dtOut == 0 &&
        d_n >= d_outputDeck_p->d_dtOut) {
      if (pp) computePostProcFields();
      output();
      if (d_stop) return;
      checkOutputCriteria();
    }
  }
}

template <class T>
void FDModel<T>::computeForces() {
  check(nlm_fd_compute_forces(d_ctx), "nlm_fd_compute_forces");
}

template <class T>
bool FDModel<T>::populate(const std::string& tag) const {
  // tags left out at memory-control levels 1..3 (policy.cpp:38-56)
  const int lvl = d_input_p->getPolicyDeck()->d_memControlFlag;
  static const char* const l1[] = {"Model_d_e", "Model_d_w", "Model_d_phi",
                                   "Model_d_strain", "Model_d_stress"};
  static const char* const l2[] = {"Model_d_eF", "Model_d_eFB"};
  if (lvl >= 1)
    for (const char* t : l1)
      if (tag == t) return false;
  if (lvl >= 2)
    for (const char* t : l2)
      if (tag == t) return false;
  if (lvl >= 3 && tag == "Model_d_Z") return false;
  return true;
}

template <class T>
void FDModel<T>::computePostProcFields() {
  const size_t n = d_mesh->getNumNodes();
  d_Z.assign(n, 0.f);
  d_phi.assign(n, 0.f);
  d_energy.assign(n, 0.);
  if (T::kMaterial != NLM_MATERIAL_RNP_BOND) return;
  check(nlm_fd_postproc(d_ctx, d_Z.data(), d_phi.data(), d_energy.data()),
        "nlm_fd_postproc");
  const inp::OutputDeck* od = d_outputDeck_p;
  const auto& V = d_mesh->getNodalVolumes();
  std::vector<double> u, v;
  getState(&u, &v, nullptr);
  // work done u . f_ext (fDModel.cpp:655-662, 734-735)
  d_work.assign(n, 0.);
  if (populate("Model_d_w")) {
    std::vector<double> fe(3 * n, 0.);
    check(nlm_fd_external_force(d_ctx, d_time, fe.data()),
          "nlm_fd_external_force");
    for (size_t i = 0; i < n; i++)
      d_work[i] = u[3 * i] * fe[3 * i] + u[3 * i + 1] * fe[3 * i + 1] +
                  u[3 * i + 2] * fe[3 * i + 2];
  }
  // kinetic energy (fDModel.cpp:755-761)
  d_kinetic.assign(n, 0.);
  if (populate("Model_d_e")) {
    const double rho = d_input_p->getMaterialDeck()->d_density;
    for (size_t i = 0; i < n; i++)
      d_kinetic[i] = 0.5 * rho *
                     (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                      v[3 * i + 2] * v[3 * i + 2]) *
                     V[i];
  }
  // fracture energies of fully damaged nodes; kept from earlier outputs
  // elsewhere (fDModel.cpp:737-745)
  if (d_fracE.size() != n) d_fracE.assign(n, 0.);
  if (d_fracEB.size() != n) d_fracEB.assign(n, 0.);
  for (size_t i = 0; i < n; i++)
    if (greater_than(double(d_Z[i]), 1.0 - 1.0E-10)) {
      if (populate("Model_d_eFB")) d_fracEB[i] = d_energy[i];
      if (populate("Model_d_eF")) d_fracE[i] = d_energy[i];
    }
  // reaction force of computeForce (fDModel.cpp:548-580)
  if (od->isTagInOutput("Reaction_Force") ||
      od->isTagInOutput("Total_Reaction_Force")) {
    const auto& bb = d_mesh->getBoundingBox();
    const double box[3] = {bb.first[0], bb.second[0], bb.second[1]};
    d_reaction.assign(3 * n, 0.);
    d_reactionLen.assign(n, 0.);
    check(nlm_fd_reaction_force(d_ctx, box, d_modelDeck_p->d_horizon,
                                d_reaction.data(), d_reactionLen.data()),
          "nlm_fd_reaction_force");
  }
}

template <class T>
void FDModel<T>::getState(std::vector<double>* u, std::vector<double>* v,
                          std::vector<double>* f) {
  const size_t n3 = 3 * d_mesh->getNumNodes();
  if (u) u->assign(n3, 0.);
  if (v) v->assign(n3, 0.);
  if (f) f->assign(n3, 0.);
  check(nlm_fd_get_state(d_ctx, u ? u->data() : nullptr,
                         v ? v->data() : nullptr, f ? f->data() : nullptr),
        "nlm_fd_get_state");
}

template <class T>
void FDModel<T>::setState(const std::vector<double>* u,
                          const std::vector<double>* v) {
  check(nlm_fd_set_state(d_ctx, u ? u->data() : nullptr,
                         v ? v->data() : nullptr, nullptr),
        "nlm_fd_set_state");
}

template <class T>
std::vector<uint8_t> FDModel<T>::getBondBytes() {
  int64_t size = 0;
  check(nlm_fd_bond_bytes_size(d_ctx, &size), "nlm_fd_bond_bytes_size");
  std::vector<uint8_t> b(size_t(size), 0);
  check(nlm_fd_get_bond_bytes(d_ctx, b.data()), "nlm_fd_get_bond_bytes");
  return b;
}

// model::Output (src/model/util.cpp:4-190).  Tags the GPU engine does not
// compute (Work_Done, reaction forces, fracture energies, strain / stress
// tensors) are skipped with a warning.  With Output: Bond_State (default on)
// the file also carries the broken-bond bits in the reference's packing and
// so is a complete restart point.
template <class T>
void FDModel<T>::output() {
  std::cout << "Output: time step = " << d_n << "\n";
  const inp::OutputDeck* od = d_outputDeck_p;
  const size_t n = d_mesh->getNumNodes();
  std::vector<double> u, v, f;
  getState(&u, &v, &f);
  const std::string file =
      od->d_path + "output_" + std::to_string(d_n / od->d_dtOutCriteria);
  rw::VtuWriter w(file, od->d_compressType, od->d_outFormat == "legacy_vtk");
  w.appendNodes(d_mesh->getNodes(), u.data());
  if (d_mesh->getNumElements() != 0 && od->d_performFEOut)
    w.appendCells(d_mesh->getElementType(),
                  d_mesh->getElementConnectivities());
  if (od->isTagInOutput("Displacement"))
    w.appendPointData("Displacement", u.data(), 3);
  if (od->isTagInOutput("Velocity"))
    w.appendPointData("Velocity", v.data(), 3);
  if (od->isTagInOutput("Force")) {
    std::vector<double> F(3 * n);
    const auto& V = d_mesh->getNodalVolumes();
    for (size_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) F[3 * i + k] = f[3 * i + k] * V[i];
    w.appendPointData("Force", F.data(), 3);
  }
  w.addTimeStep(d_time);
  if (od->d_bondState) w.appendFieldData("Bond_State", getBondBytes());
  if (d_input_p->getPolicyDeck()->d_enablePostProcessing) {
    if (od->isTagInOutput("Force_Density"))
      w.appendPointData("Force_Density", f.data(), 3);
    if (od->isTagInOutput("Fixity"))
      w.appendPointData("Fixity", d_fix.data());
    if (od->isTagInOutput("Node_Volume"))
      w.appendPointData("Node_Volume", d_mesh->getNodalVolumes().data(), 1);
    if (d_Z.size() == n) {
      if (od->isTagInOutput("Damage_Phi") && populate("Model_d_phi"))
        w.appendPointData("Damage_Phi", d_phi.data(), 1);
      if (od->isTagInOutput("Damage_Z") && populate("Model_d_Z"))
        w.appendPointData("Damage_Z", d_Z.data(), 1);
      if (od->isTagInOutput("Strain_Energy"))
        w.appendPointData("Strain_Energy", d_energy.data(), 1);
    }
    if (d_reaction.size() == 3 * n) {
      if (od->isTagInOutput("Reaction_Force"))
        w.appendPointData("Reaction_Force", d_reaction.data(), 3);
      if (od->isTagInOutput("Total_Reaction_Force")) {
        // model/util.cpp:79-96: std::accumulate with an int initial value
        // accumulates in int (each partial sum truncated toward zero); times
        // the area (|max_x - min_x| - horizon) (|max_y - min_y| - horizon)
        int acc = 0;
        for (size_t i = 0; i < n; i++) acc = int(acc + d_reactionLen[i]);
        const auto& bb = d_mesh->getBoundingBox();
        const double delta = d_modelDeck_p->d_horizon;
        const double area = (std::abs(bb.second[0] - bb.first[0]) - delta) *
                            (std::abs(bb.second[1] - bb.first[1]) - delta);
        w.appendFieldData("Total_Reaction_Force", double(acc) * area);
      }
    }
    if (d_work.size() == n && od->isTagInOutput("Work_Done") &&
        populate("Model_d_w"))
      w.appendPointData("Work_Done", d_work.data(), 1);
    if (d_fracEB.size() == n && od->isTagInOutput("Fracture_Perienergy_Bond") &&
        populate("Model_d_eFB"))
      w.appendPointData("Fracture_Perienergy_Bond", d_fracEB.data(), 1);
    if (d_fracE.size() == n && od->isTagInOutput("Fracture_Perienergy_Total") &&
        populate("Model_d_eF"))
      w.appendPointData("Fracture_Perienergy_Total", d_fracE.data(), 1);
    auto total = [](const std::vector<double>& a) {
      double s = 0.;
      for (double x : a) s += x;
      return s;
    };
    // model/util.cpp:140-164
    if (d_energy.size() == n && od->isTagInOutput("Total_Energy") &&
        populate("Model_d_e"))
      w.appendFieldData("Total_Energy", total(d_energy) - total(d_work) +
                                            total(d_kinetic));
    if (d_fracEB.size() == n &&
        od->isTagInOutput("Total_Fracture_Perienergy_Bond") &&
        populate("Model_d_eFB"))
      w.appendFieldData("Total_Fracture_Perienergy_Bond", total(d_fracEB));
    if (d_fracE.size() == n &&
        od->isTagInOutput("Total_Fracture_Perienergy_Total") &&
        populate("Model_d_eF"))
      w.appendFieldData("Total_Fracture_Perienergy_Total", total(d_fracE));
    if (od->isTagInOutput("Neighbors")) {
      std::vector<int64_t> rp(n + 1);
      check(nlm_neighbor_get(d_ctx, rp.data(), nullptr), "nlm_neighbor_get");
      std::vector<double> cnt(n);
      for (size_t i = 0; i < n; i++) cnt[i] = double(rp[i + 1] - rp[i]);
      w.appendPointData("Neighbors", cnt.data(), 1);
    }
    for (const char* t : {"Strain_Tensor", "Stress_Tensor"})
      if (od->isTagInOutput(t) && d_n == 0)
        std::cout << "Warning: output tag " << t
                  << " is not computed by the GPU engine.\n";
  }
  w.close();
}

// checkOutputCriteria (fDModel.cpp:781-915): switch to the small output
// interval once max Z exceeds params[0] (max_Z, max_Z_stop); with max_Z_stop
// switch back and stop once a node in the rectangle params[2..] has Z above
// params[1].
template <class T>
void FDModel<T>::checkOutputCriteria() {
  inp::OutputDeck* od = d_outputDeck_p;
  if (od->d_outCriteria.empty() || od->d_dtOutOld == od->d_dtOutCriteria)
    return;
  if (d_n % od->d_dtOutOld != 0) return;
  if (od->d_outCriteria != "max_Z" && od->d_outCriteria != "max_Z_stop")
    return;
  if (od->d_outCriteria == "max_Z" && od->d_dtOut == od->d_dtOutCriteria)
    return;
  bool changed_now = false;
  if (od->d_dtOut > od->d_dtOutCriteria && !d_changedToSmall) {
    float mx = 0.f;
    for (size_t i = 0; i < d_Z.size(); i++)
      if (i == 0 || d_Z[i] > mx) mx = d_Z[i];
    if (greater_than(mx, od->d_outCriteriaParams[0])) {
      od->d_dtOut = od->d_dtOutCriteria;
      std::cout << "Message: Changing output interval to smaller value.\n";
      std::ofstream fd(od->d_path + "dt_out_change.info");
      fd << "Large_Dt_To_Small_Dt:\n  N: " << d_n
         << "\n  dN: " << d_n / od->d_dtOutCriteria << "\n";
      d_changedToSmall = true;
      changed_now = true;
    }
  }
  if (!changed_now && !d_changedBackToLarge &&
      od->d_outCriteria == "max_Z_stop" && od->d_dtOut < od->d_dtOutOld) {
    const auto& ps = od->d_outCriteriaParams;
    double lo[3] = {0., 0., 0.}, hi[3] = {0., 0., 0.};
    if (ps.size() == 6) {
      lo[0] = ps[2];
      lo[1] = ps[3];
      hi[0] = ps[4];
      hi[1] = ps[5];
    } else if (ps.size() == 8) {
      lo[0] = ps[2];
      lo[1] = ps[3];
      lo[2] = ps[4];
      hi[0] = ps[5];
      hi[1] = ps[6];
      hi[2] = ps[7];
    }
    const double refZ = ps[1];
    const double* X = d_mesh->getNodes().data();
    bool found = false;
    for (size_t i = 0; i < d_Z.size() && !found; i++)
      if (inside_rect(X + 3 * i, lo[0], hi[0], lo[1], hi[1]) &&
          greater_than(d_Z[i], refZ))
        found = true;
    if (found) {
      od->d_dtOut = od->d_dtOutOld;
      std::cout << "Message: Changing output interval to larger value.\n";
      std::ofstream fd(od->d_path + "dt_out_change.info", std::ios::app);
      fd << "Small_Dt_To_Large_Dt:\n  N: " << d_n
         << "\n  dN: " << d_n / od->d_dtOutCriteria << "\n";
      d_changedBackToLarge = true;
      d_stop = true;
    }
  }
}

// restart (fDModel.cpp:102-131): the step from the restart deck, u and v
// from the file (u from the points when the file has no Displacement).  A
// file written with Bond_State also restores the broken bonds and the exact
// time of that step, so the run continues as if it had not stopped; other
// files keep the reference's behaviour (time = step * dt, bonds as after
// the pre-cracks).
template <class T>
void FDModel<T>::restart() {
  const inp::RestartDeck* rd = d_input_p->getRestartDeck();
  const std::string& td = d_modelDeck_p->d_timeDiscretization;
  if (td != "central_difference" && td != "velocity_verlet")
    deck_error("Error: FDModel needs Time: central_difference or "
               "velocity_verlet.");
  initHObjects();
  init();
  d_n = rd->d_step;
  d_time = double(d_n) * d_modelDeck_p->d_dt;
  const size_t n = d_mesh->getNumNodes();
  std::vector<double> u, v;
  std::vector<uint8_t> bonds;
  std::vector<double> tfile;
  try {
    rw::VtuReader r(rd->d_file);
    if (r.numPoints() != int64_t(n))
      deck_error("Error: Number of nodes in input file = " + rd->d_file +
                 " and number nodes in data X are not same.");
    if (!r.readPointData("Displacement", &u)) {
      std::vector<double> y;
      r.readNodes(&y);
      u.resize(3 * n);
      const auto& X = d_mesh->getNodes();
      for (size_t k = 0; k < 3 * n; k++) u[k] = y[k] - X[k];
    }
    if (!r.readPointData("Velocity", &v)) v.assign(3 * n, 0.);
    r.readFieldBytes("Bond_State", &bonds);
    r.readFieldData("TIME", &tfile);
  } catch (const std::runtime_error& e) {
    if (dynamic_cast<const inp::DeckError*>(&e)) throw;
    deck_error(std::string("Error: ") + e.what());
  }
  if (u.size() != 3 * n || v.size() != 3 * n)
    deck_error("Error: restart file " + rd->d_file + " has bad point data");
  setState(&u, &v);
  if (!bonds.empty()) {
    int64_t size = 0;
    check(nlm_fd_bond_bytes_size(d_ctx, &size), "nlm_fd_bond_bytes_size");
    if (int64_t(bonds.size()) != size)
      deck_error("Error: Bond_State of " + rd->d_file +
                 " does not match the neighbor lists.");
    // unpack the reference packing into one byte per bond (CSR order)
    std::vector<int64_t> rp(n + 1);
    check(nlm_neighbor_get(d_ctx, rp.data(), nullptr), "nlm_neighbor_get");
    std::vector<uint8_t> per_bond(size_t(rp[n]), 0);
    size_t off = 0;
    for (size_t i = 0; i < n; i++) {
      const int64_t ni = rp[i + 1] - rp[i];
      for (int64_t j = 0; j < ni; j++)
        per_bond[size_t(rp[i] + j)] = (bonds[off + size_t(j / 8)] >> (j % 8)) & 1;
      off += size_t((ni + 7) / 8);
    }
    check(nlm_fd_set_bonds(d_ctx, per_bond.data()), "nlm_fd_set_bonds");
    if (tfile.size() == 1) d_time = tfile[0];
  }
  integrate();
}

template class FDModel<material::pd::RNPBond>;
template class FDModel<material::pd::ElasticState>;

}  // namespace model
}  // namespace nlm
