// Host side of the quasi-static model: NLMech's model::QuasiStaticModel<T>
// (src/model/quasistatic/QuasiStaticModel.cpp) on the GPU solver of
// libnlmech_amd (nlm_qs_solve).  Load step n is at t = n dt; each load step
// runs Newton iterations until ||R||_2 <= Solver: Tolerance (or
// Max_Iteration), then writes the output file as solver() does
// (QuasiStaticModel.cpp:527-592).
//
// The bond-based material uses the analytic tangent; the state-based
// material the reference's central-difference Jacobian (same-component
// entries on the neighbor sparsity, eps = Perturbation * h).  The Newton
// systems are solved by BiCGSTAB on the GPU with the deck's
// Preconditioner (jacobi or two_level) to Linear_Tolerance.
#pragma once

#include "model.h"

namespace nlm {
namespace model {

template <class T>
class QuasiStaticModel : public Model {
 public:
  // Runs the whole solve, as the reference's constructor does
  // (QuasiStaticModel.cpp:24-62).
  explicit QuasiStaticModel(inp::Input* deck);
  QuasiStaticModel(inp::Input* deck, bool run_now);

  void run();
  void initHObjects();
  // solver (QuasiStaticModel.cpp:527-592).
  void solver();
  // model::Output (src/model/util.cpp:4-190) of the current load step; the
  // Force array is f_int V like the reference's computeForces(true).
  void output();

  const nlm_qs_stats& lastStats() const { return d_stats; }
  FDModel<T>& fd() { return d_fd; }

 private:
  inp::Input* d_input_p;
  FDModel<T> d_fd;
  nlm_qs_stats d_stats{};
  // state-based material: the strain energy (float, added to at every load
  // step like the reference's vector) and the strain / stress tensors of
  // the last load step (6 per node)
  std::vector<float> d_strainEnergy;
  std::vector<double> d_strain, d_stress;
};

extern template class QuasiStaticModel<material::pd::RNPBond>;
extern template class QuasiStaticModel<material::pd::ElasticState>;

}  // namespace model
}  // namespace nlm
