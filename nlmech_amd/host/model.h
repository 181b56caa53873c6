// Host side of the explicit finite-difference model: NLMech's
// model::FDModel<T> (src/model/fd/fDModel.h:81-283, fDModel.cpp) with T a
// material class, on top of the C-ABI of libnlmech_amd (include/
// nlmech_amd.h).  The class reads the same input deck, builds the mesh,
// neighbor lists, pre-cracks, no-fail flags, loading and material, runs the
// integrate() loop on the GPU and writes the reference's .vtu output files.
//
// Where the reference exits on a deck error, the constructor throws
// inp::DeckError; a failing C-ABI call throws nlm::Error with the library's
// message.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "input.h"
#include "mesh.h"
#include "nlmech_amd.h"

namespace nlm {

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

namespace material {
namespace pd {

// material::pd::RNPBond (src/material/pd/rnpBond.cpp:23-189): the
// constructor turns the material deck into the per-bond constants the GPU
// bond-force kernels take.
class RNPBond {
 public:
  static constexpr int kMaterial = NLM_MATERIAL_RNP_BOND;
  RNPBond(const inp::MaterialDeck& deck, size_t dim, double horizon,
          double h);
  const nlm_rnp_params& params() const { return d_params; }
  double getDensity() const { return d_params.density; }
  double getHorizon() const { return d_params.horizon; }
  bool isStateActive() const { return false; }
  void fill(nlm_fd_problem* p) const;

 private:
  nlm_rnp_params d_params;
};

// material::pd::ElasticState (src/material/pd/ElasticState.cpp:17-158);
// in explicit dynamics with the optional critical-stretch bond breaking
// (MaterialDeck Critical_Stretch).
class ElasticState {
 public:
  static constexpr int kMaterial = NLM_MATERIAL_ELASTIC_STATE;
  ElasticState(const inp::MaterialDeck& deck, size_t dim, double horizon,
               double h);
  const nlm_es_params& params() const { return d_params; }
  double getDensity() const { return d_params.density; }
  double getHorizon() const { return d_params.horizon; }
  bool isStateActive() const { return true; }
  void fill(nlm_fd_problem* p) const;

 private:
  nlm_es_params d_params;
};

}  // namespace pd
}  // namespace material

namespace model {

// model::Model (src/model/model.h): step counter and time.
class Model {
 public:
  size_t getN() const { return d_n; }
  double getTime() const { return d_time; }

 protected:
  size_t d_n = 0;
  double d_time = 0.;
};

template <class T>
class FDModel : public Model {
 public:
  // FDModel::FDModel (fDModel.cpp:47-67): runs the whole simulation (or the
  // restart) as the reference's constructor does.
  explicit FDModel(inp::Input* deck);
  // Builds the model without running it (run() / restart() / the steps
  // below drive it).
  FDModel(inp::Input* deck, bool run_now);
  ~FDModel();
  FDModel(const FDModel&) = delete;
  FDModel& operator=(const FDModel&) = delete;

  void run();      // initHObjects, init, integrate (fDModel.cpp:90-99)
  void restart();  // fDModel.cpp:102-131 plus the bond state (see .cpp)

  void initHObjects();
  void init();
  void integrate();
  // computeForces (fDModel.cpp:528-541) on the current state.
  void computeForces();
  // computePostProcFields (fDModel.cpp:652-778): damage Z, phi, strain
  // energy, work done, kinetic and fracture energies, and the reaction
  // force of computeForce (fDModel.cpp:548-580) when output asks for it.
  void computePostProcFields();
  // model::Output (src/model/util.cpp:4-190) at the current step.
  void output();
  void checkOutputCriteria();

  // Host copies of the device state (AoS [3n]).
  void getState(std::vector<double>* u, std::vector<double>* v,
                std::vector<double>* f);
  void setState(const std::vector<double>* u, const std::vector<double>* v);
  std::vector<uint8_t> getBondBytes();
  const fe::Mesh& getMesh() const { return *d_mesh; }
  nlm_ctx* context() { return d_ctx; }
  const T& getMaterial() const { return *d_material; }
  bool stopped() const { return d_stop; }

 private:
  void check(int rc, const char* what);
  void applyInitialCondition();
  std::vector<uint8_t> interiorFlags() const;

  inp::Input* d_input_p;
  inp::ModelDeck* d_modelDeck_p;
  inp::OutputDeck* d_outputDeck_p;
  std::unique_ptr<fe::Mesh> d_mesh;
  std::unique_ptr<T> d_material;
  nlm_ctx* d_ctx = nullptr;
  int64_t d_nnz = 0;
  bool d_stop = false;
  bool d_changedToSmall = false;
  bool d_changedBackToLarge = false;
  bool d_pendingCracks = false;
  // inp::Policy::populateData (policy.cpp:90-97) for the deck's
  // Memory_Control_Flag.
  bool populate(const std::string& tag) const;

  std::vector<float> d_Z, d_phi;
  std::vector<double> d_energy;
  std::vector<double> d_work, d_kinetic;     // u . f_ext, (1/2) rho v.v V
  std::vector<double> d_fracE, d_fracEB;     // kept across outputs
  std::vector<double> d_reaction, d_reactionLen;
  std::vector<uint8_t> d_fix;
};

extern template class FDModel<material::pd::RNPBond>;
extern template class FDModel<material::pd::ElasticState>;

}  // namespace model
}  // namespace nlm
