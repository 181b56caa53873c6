// Input deck: the YAML sections NLMech's inp::Input reads
// (src/inp/input.cpp:42-899, decks in src/inp/decks/), with the same keys,
// defaults and checks.  Where the reference prints an error and exits, the
// reader throws nlm::inp::DeckError with the reference's message; the
// nlm_fd program turns it into the same message and exit status 1.
//
// One section has no counterpart in the reference: "GPU" selects the
// precision mode, the device and the stepping options of the GPU engine.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "yaml.h"

namespace nlm {
namespace inp {

class DeckError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// inp::ModelDeck (src/inp/decks/modelDeck.h), read by setModelDeck
// (input.cpp:124-188).
struct ModelDeck {
  size_t d_dim = 0;
  std::string d_timeDiscretization;
  std::string d_spatialDiscretization;
  std::string d_simType;
  double d_horizon = 0.;
  int d_rh = 0;
  double d_h = 0.;
  double d_tFinal = 0.;
  int64_t d_Nt = 0;
  double d_dt = 0.;
  bool d_isRestartActive = false;
};

// inp::RestartDeck (input.cpp:190-212).
struct RestartDeck {
  std::string d_file;
  size_t d_step = 0;
};

// inp::MeshDeck (input.cpp:214-254).
struct MeshDeck {
  size_t d_dim = 0;
  std::string d_spatialDiscretization;
  std::string d_filename;
  bool d_computeMeshSize = false;
  double d_h = 0.;
  bool d_isCentroidBasedDiscretization = false;
  bool d_keepElementConn = false;
};

// inp::EdgeCrack (src/inp/decks/fractureDeck.h:22-104), setFractureDeck
// (input.cpp:290-355).
struct EdgeCrack {
  int d_o = 0;
  double d_pb[2] = {0., 0.};
  double d_pt[2] = {0., 0.};
  double d_l = 0.;
  double d_theta = 0.;
  double d_activationTime = -1.;
};
struct FractureDeck {
  std::vector<EdgeCrack> d_cracks;
};

// inp::InteriorFlagsDeck (input.cpp:357-401).
struct InteriorFlagsDeck {
  bool d_noFailActive = false;
  double d_noFailTol = 0.;
  bool d_computeAndNotStoreFlag = false;
  std::vector<std::pair<std::string, std::vector<double>>> d_noFailRegions;
};

// inp::ICData / InitialConditionDeck (input.cpp:403-436).
struct ICData {
  std::string d_type;
  std::string d_file;
  std::vector<double> d_params;
};
struct InitialConditionDeck {
  ICData d_uICData, d_vICData;
};

// inp::BCData (src/inp/decks/loadingDeck.h), setLoadingDeck
// (input.cpp:438-612).
struct BCData {
  std::string d_regionType;
  double d_x1 = 0., d_x2 = 0., d_y1 = 0., d_y2 = 0., d_z1 = 0., d_z2 = 0.;
  double d_r1 = 0., d_r2 = 0., d_theta = 0.;
  std::vector<size_t> d_direction;
  std::string d_timeFnType;
  std::vector<double> d_timeFnParams;
  std::string d_spatialFnType;
  std::vector<double> d_spatialFnParams;
};
struct LoadingDeck {
  std::vector<BCData> d_uBCData, d_fBCData;
};

// inp::MaterialDeck / MatData (src/inp/decks/materialDeck.h:20-288),
// setMaterialDeck (input.cpp:614-699).  MatData entries are -1 when absent.
struct MatData {
  double d_E = -1., d_G = -1., d_K = -1., d_lambda = -1., d_mu = -1.,
         d_nu = -1., d_Gc = -1., d_KIc = -1.;
};
struct MaterialDeck {
  std::string d_materialType;
  bool d_isPlaneStrain = false;
  bool d_computeParamsFromElastic = true;
  MatData d_matData;
  size_t d_bondPotentialType = 0;
  double d_checkScFactor = 1.;
  bool d_irreversibleBondBreak = true;
  std::vector<double> d_bondPotentialParams;
  size_t d_statePotentialType = 0;
  bool d_stateContributionFromBrokenBond = true;
  std::vector<double> d_statePotentialParams;
  size_t d_influenceFnType = 0;
  std::vector<double> d_influenceFnParams;
  double d_density = 1.;
  bool d_applyContact = false;
  // ElasticState in explicit dynamics: critical stretch of the bond-breaking
  // rule (Material: Critical_Stretch); the reference's ElasticState has no
  // damage, so <= 0 (the default) keeps every bond.
  double d_criticalStretch = -1.;
};

// inp::OutputDeck (src/inp/decks/outputDeck.h), setOutputDeck
// (input.cpp:701-735).
struct OutputDeck {
  std::string d_outFormat = "vtu";
  std::string d_path = "./";
  std::vector<std::string> d_outTags;
  size_t d_dtOut = 0;
  size_t d_dtOutOld = 0;
  size_t d_dtOutCriteria = 0;
  size_t d_debug = 0;
  bool d_performFEOut = true;
  std::string d_compressType;
  std::string d_outCriteria;
  std::vector<double> d_outCriteriaParams;
  // Output: Bond_State (no counterpart in the reference): write the broken
  // bond bits into each output file so a restart from it is exact.
  bool d_bondState = true;

  bool isTagInOutput(const std::string& tag) const {
    for (const auto& t : d_outTags)
      if (t == tag) return true;
    return false;
  }
};

// inp::PolicyDeck (input.cpp:737-746).
struct PolicyDeck {
  int d_memControlFlag = 0;
  bool d_enablePostProcessing = true;
};

// inp::SolverDeck (input.cpp:748-762).
struct SolverDeck {
  std::string d_solverType;
  int d_maxIters = 0;
  double d_tol = 0.;
  double d_perturbation = 0.;
  // GPU linear solve of the quasi-static Newton systems (no counterpart in
  // the reference): preconditioner "jacobi", "two_level", "learned" or
  // "learned_two_level", relative tolerance and iteration cap of BiCGSTAB.
  std::string d_preconditioner = "jacobi";
  double d_linTol = 1e-8;
  int d_linMaxIter = 20000;
  double d_tolRel = 0.;
};

// GPU options (no counterpart in the reference).
struct GpuDeck {
  std::string d_precision = "fp64_exact";  // fp64_exact, fp64, fp32_force
  int d_device = 0;
  int d_graphChunk = 0;  // > 0: steps between outputs go through a captured
                         // HIP graph of this many steps
  bool d_autotune = false;
};

class Input {
 public:
  // Reads every section of the deck; throws DeckError.
  explicit Input(const std::string& filename);
  // Deck from YAML text (tests and embedding).
  static Input fromText(const std::string& text);

  ModelDeck* getModelDeck() { return &d_model; }
  RestartDeck* getRestartDeck() { return &d_restart; }
  MeshDeck* getMeshDeck() { return &d_mesh; }
  FractureDeck* getFractureDeck() { return &d_fracture; }
  InteriorFlagsDeck* getInteriorFlagsDeck() { return &d_interior; }
  InitialConditionDeck* getInitialConditionDeck() { return &d_ic; }
  LoadingDeck* getLoadingDeck() { return &d_loading; }
  MaterialDeck* getMaterialDeck() { return &d_material; }
  OutputDeck* getOutputDeck() { return &d_output; }
  PolicyDeck* getPolicyDeck() { return &d_policy; }
  SolverDeck* getSolverDeck() { return &d_solver; }
  GpuDeck* getGpuDeck() { return &d_gpu; }
  const std::string& getSpatialDiscretization() const {
    return d_model.d_spatialDiscretization;
  }

 private:
  Input() = default;
  void read(const yaml::Node& config);
  void setModelDeck(const yaml::Node& c);
  void setRestartDeck(const yaml::Node& c);
  void setMeshDeck(const yaml::Node& c);
  void setFractureDeck(const yaml::Node& c);
  void setInteriorFlagsDeck(const yaml::Node& c);
  void setInitialConditionDeck(const yaml::Node& c);
  void setLoadingDeck(const yaml::Node& c);
  void setMaterialDeck(const yaml::Node& c);
  void setOutputDeck(const yaml::Node& c);
  void setPolicyDeck(const yaml::Node& c);
  void setSolverDeck(const yaml::Node& c);
  void setGpuDeck(const yaml::Node& c);

  ModelDeck d_model;
  RestartDeck d_restart;
  MeshDeck d_mesh;
  FractureDeck d_fracture;
  InteriorFlagsDeck d_interior;
  InitialConditionDeck d_ic;
  LoadingDeck d_loading;
  MaterialDeck d_material;
  OutputDeck d_output;
  PolicyDeck d_policy;
  SolverDeck d_solver;
  GpuDeck d_gpu;
};

}  // namespace inp
}  // namespace nlm
