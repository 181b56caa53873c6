// model::QuasiStaticModel<T> on the GPU (see qs_model.h).
#include "qs_model.h"

#include <cstring>
#include <iostream>

#include "vtu.h"

namespace nlm {
namespace model {

namespace {

[[noreturn]] void deck_error(const std::string& m) {
  throw inp::DeckError(m);
}

int precond_code(const std::string& p) {
  if (p == "jacobi") return NLM_QS_PRECOND_JACOBI;
  if (p == "two_level") return NLM_QS_PRECOND_JACOBI | NLM_QS_PRECOND_COARSE;
  if (p == "none") return NLM_QS_PRECOND_NONE;
  deck_error("Error: Solver Preconditioner must be none, jacobi or "
             "two_level.");
}

}  // namespace

template <class T>
QuasiStaticModel<T>::QuasiStaticModel(inp::Input* deck)
    : QuasiStaticModel(deck, true) {}

template <class T>
QuasiStaticModel<T>::QuasiStaticModel(inp::Input* deck, bool run_now)
    : d_input_p(deck), d_fd(deck, false) {
  if (run_now) run();
}

template <class T>
void QuasiStaticModel<T>::run() {
  initHObjects();
  solver();
}

template <class T>
void QuasiStaticModel<T>::initHObjects() {
  if (d_input_p->getModelDeck()->d_timeDiscretization != "quasi_static")
    deck_error("Error: QuasiStaticModel needs Time: quasi_static.");
  if (d_input_p->getOutputDeck()->d_dtOut == 0)
    deck_error("Error: Output: Output_Interval must be given and positive.");
  d_fd.initHObjects();
  const size_t dim = d_fd.getMesh().getDimension();
  if (dim != 1 && dim != 2)
    deck_error("Error: the GPU quasi-static solver runs 1D and 2D models.");
  const size_t n = d_fd.getMesh().getNumNodes();
  d_strainEnergy.assign(n, 0.f);
  d_strain.assign(6 * n, 0.);
  d_stress.assign(6 * n, 0.);
}

template <class T>
void QuasiStaticModel<T>::solver() {
  const inp::SolverDeck* sd = d_input_p->getSolverDeck();
  const inp::ModelDeck* md = d_input_p->getModelDeck();
  nlm_qs_options o;
  std::memset(&o, 0, sizeof(o));
  o.nsteps = 1;
  o.dt = md->d_dt;
  o.tol_abs = sd->d_tol;
  o.tol_rel = sd->d_tolRel;
  o.max_newton = sd->d_maxIters;
  o.tangent = T::kMaterial == NLM_MATERIAL_ELASTIC_STATE ? NLM_QS_TANGENT_FD
                                                         : NLM_QS_TANGENT_FULL;
  o.precond = precond_code(sd->d_preconditioner);
  o.lin_tol = sd->d_linTol;
  o.lin_max_iter = sd->d_linMaxIter;
  o.predictor = 0;
  o.gnn_refresh = 0;
  o.perturbation = sd->d_perturbation > 0. ? sd->d_perturbation : 1e-7;
  nlm_ctx* c = d_fd.context();
  d_n = 0;
  d_time = 0.;
  output();
  for (d_n = 1; d_n <= size_t(md->d_Nt); d_n++) {
    d_time = double(d_n) * md->d_dt;
    if (nlm_qs_solve(c, &o, &d_stats) != NLM_OK)
      throw Error(std::string("nlm_qs_solve failed: ") + nlm_last_error(c));
    std::cout << "Load step " << d_n << ": " << d_stats.newton_iters
              << " Newton iterations, " << d_stats.linear_iters
              << " BiCGSTAB iterations, Res: " << d_stats.residual << "\n";
    if (d_stats.residual > o.tol_abs && d_stats.residual > o.tol_rel *
                                                               d_stats.residual0)
      std::cout << "Warning: Newton stopped at Max_Iteration before the "
                   "tolerance.\n";
    // computeForces(true) after the load step's Newton iterations
    // (QuasiStaticModel.cpp:589): strain energy, strain and stress
    if (T::kMaterial == NLM_MATERIAL_ELASTIC_STATE) {
      const inp::OutputDeck* od = d_input_p->getOutputDeck();
      const bool e = od->isTagInOutput("Strain_Energy");
      const bool st = od->isTagInOutput("Strain_Tensor");
      const bool ss = od->isTagInOutput("Stress_Tensor");
      if ((e || st || ss) &&
          nlm_fd_es_fields(c, e ? d_strainEnergy.data() : nullptr,
                           st ? d_strain.data() : nullptr,
                           ss ? d_stress.data() : nullptr) != NLM_OK)
        throw Error(std::string("nlm_fd_es_fields: ") + nlm_last_error(c));
    }
    output();
  }
}

template <class T>
void QuasiStaticModel<T>::output() {
  const inp::OutputDeck* od = d_input_p->getOutputDeck();
  const fe::Mesh& mesh = d_fd.getMesh();
  const size_t n = mesh.getNumNodes();
  nlm_ctx* c = d_fd.context();
  std::vector<double> u, v, f_keep;
  d_fd.getState(&u, &v, &f_keep);
  // internal force alone (computeForces(true), QuasiStaticModel.cpp:169)
  std::vector<double> zero(3 * n, 0.), f(3 * n, 0.);
  if (nlm_fd_set_state(c, nullptr, nullptr, zero.data()) != NLM_OK ||
      nlm_fd_compute_forces(c) != NLM_OK ||
      nlm_fd_get_state(c, nullptr, nullptr, f.data()) != NLM_OK ||
      nlm_fd_set_state(c, nullptr, nullptr, f_keep.data()) != NLM_OK)
    throw Error(std::string("output forces: ") + nlm_last_error(c));
  std::cout << "Output: time step = " << d_n << "\n";
  const std::string file =
      od->d_path + "output_" + std::to_string(d_n / od->d_dtOutCriteria);
  rw::VtuWriter w(file, od->d_compressType, od->d_outFormat == "legacy_vtk");
  w.appendNodes(mesh.getNodes(), u.data());
  if (mesh.getNumElements() != 0 && od->d_performFEOut)
    w.appendCells(mesh.getElementType(), mesh.getElementConnectivities());
  if (od->isTagInOutput("Displacement"))
    w.appendPointData("Displacement", u.data(), 3);
  if (od->isTagInOutput("Velocity")) w.appendPointData("Velocity", v.data(), 3);
  if (od->isTagInOutput("Force")) {
    std::vector<double> F(3 * n);
    const auto& V = mesh.getNodalVolumes();
    for (size_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) F[3 * i + k] = f[3 * i + k] * V[i];
    w.appendPointData("Force", F.data(), 3);
  }
  w.addTimeStep(d_time);
  if (d_input_p->getPolicyDeck()->d_enablePostProcessing) {
    if (od->isTagInOutput("Force_Density"))
      w.appendPointData("Force_Density", f.data(), 3);
    if (od->isTagInOutput("Fixity")) {
      std::vector<uint8_t> fix(n);
      if (nlm_fd_get_fixity(c, fix.data()) != NLM_OK)
        throw Error(nlm_last_error(c));
      w.appendPointData("Fixity", fix.data());
    }
    if (od->isTagInOutput("Node_Volume"))
      w.appendPointData("Node_Volume", mesh.getNodalVolumes().data(), 1);
    if (od->isTagInOutput("Neighbors")) {
      std::vector<int64_t> rp(n + 1);
      if (nlm_neighbor_get(c, rp.data(), nullptr) != NLM_OK)
        throw Error(nlm_last_error(c));
      std::vector<double> cnt(n);
      for (size_t i = 0; i < n; i++) cnt[i] = double(rp[i + 1] - rp[i]);
      w.appendPointData("Neighbors", cnt.data(), 1);
    }
    if (T::kMaterial == NLM_MATERIAL_ELASTIC_STATE) {
      // model/util.cpp:176-186 (zeros before the first load step, as the
      // reference's freshly allocated vectors)
      if (od->isTagInOutput("Strain_Energy")) {
        std::vector<double> e(d_strainEnergy.begin(), d_strainEnergy.end());
        w.appendPointData("Strain_Energy", e.data(), 1);
      }
      if (od->isTagInOutput("Strain_Tensor"))
        w.appendPointData("Strain_Tensor", d_strain.data(), 6);
      if (od->isTagInOutput("Stress_Tensor"))
        w.appendPointData("Stress_Tensor", d_stress.data(), 6);
    } else if (d_n == 0) {
      for (const char* t : {"Strain_Energy", "Strain_Tensor", "Stress_Tensor"})
        if (od->isTagInOutput(t))
          std::cout << "Warning: output tag " << t
                    << " is not computed by the GPU engine for the "
                       "bond-based material.\n";
    }
  }
  w.close();
}

template class QuasiStaticModel<material::pd::RNPBond>;
template class QuasiStaticModel<material::pd::ElasticState>;

}  // namespace model
}  // namespace nlm
