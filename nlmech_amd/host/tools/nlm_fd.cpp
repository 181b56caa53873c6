// nlm_fd: runs an NLMech input deck on the GPU, as the reference's NLMech
// executable does (src/main.cpp:40-91): "nlm_fd -i input.yaml".
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "model.h"
#include "qs_model.h"

namespace {

std::string q(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o + "\"";
}

std::string num(double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", v);
  return b;
}

template <class V>
std::string list(const V& v) {
  std::string o = "[";
  for (size_t k = 0; k < v.size(); k++)
    o += (k ? ", " : "") + num(double(v[k]));
  return o + "]";
}

std::string bc(const nlm::inp::BCData& b) {
  return "{\"region\": " + q(b.d_regionType) + ", \"x1\": " + num(b.d_x1) +
         ", \"x2\": " + num(b.d_x2) + ", \"y1\": " + num(b.d_y1) +
         ", \"y2\": " + num(b.d_y2) + ", \"z1\": " + num(b.d_z1) +
         ", \"z2\": " + num(b.d_z2) + ", \"r1\": " + num(b.d_r1) +
         ", \"r2\": " + num(b.d_r2) + ", \"theta\": " + num(b.d_theta) +
         ", \"direction\": " + list(b.d_direction) +
         ", \"time_fn\": " + q(b.d_timeFnType) +
         ", \"time_params\": " + list(b.d_timeFnParams) +
         ", \"space_fn\": " + q(b.d_spatialFnType) +
         ", \"space_params\": " + list(b.d_spatialFnParams) + "}";
}

// The parsed deck as JSON (nlm_fd --dump-deck): checks a deck without a GPU.
void dump(nlm::inp::Input& d) {
  const auto* m = d.getModelDeck();
  const auto* mat = d.getMaterialDeck();
  const auto* o = d.getOutputDeck();
  const auto* l = d.getLoadingDeck();
  std::string s = "{";
  s += "\"model\": {\"dim\": " + num(double(m->d_dim)) +
       ", \"time\": " + q(m->d_timeDiscretization) +
       ", \"spatial\": " + q(m->d_spatialDiscretization) +
       ", \"horizon\": " + num(m->d_horizon) + ", \"rh\": " +
       num(m->d_rh) + ", \"h\": " + num(m->d_h) + ", \"final_time\": " +
       num(m->d_tFinal) + ", \"steps\": " + num(double(m->d_Nt)) +
       ", \"dt\": " + num(m->d_dt) + ", \"restart\": " +
       (m->d_isRestartActive ? "true" : "false") + "}";
  s += ", \"restart\": {\"file\": " + q(d.getRestartDeck()->d_file) +
       ", \"step\": " + num(double(d.getRestartDeck()->d_step)) + "}";
  s += ", \"mesh\": {\"file\": " + q(d.getMeshDeck()->d_filename) +
       ", \"compute_mesh_size\": " +
       (d.getMeshDeck()->d_computeMeshSize ? "true" : "false") +
       ", \"keep_element_conn\": " +
       (d.getMeshDeck()->d_keepElementConn ? "true" : "false") + "}";
  s += ", \"material\": {\"type\": " + q(mat->d_materialType) +
       ", \"density\": " + num(mat->d_density) + ", \"E\": " +
       num(mat->d_matData.d_E) + ", \"G\": " + num(mat->d_matData.d_G) +
       ", \"K\": " + num(mat->d_matData.d_K) + ", \"nu\": " +
       num(mat->d_matData.d_nu) + ", \"Gc\": " + num(mat->d_matData.d_Gc) +
       ", \"KIc\": " + num(mat->d_matData.d_KIc) + ", \"plane_strain\": " +
       (mat->d_isPlaneStrain ? "true" : "false") +
       ", \"from_classical\": " +
       (mat->d_computeParamsFromElastic ? "true" : "false") +
       ", \"bond_type\": " + num(double(mat->d_bondPotentialType)) +
       ", \"check_sc\": " + num(mat->d_checkScFactor) +
       ", \"irreversible\": " +
       (mat->d_irreversibleBondBreak ? "true" : "false") +
       ", \"bond_params\": " + list(mat->d_bondPotentialParams) +
       ", \"influence_type\": " + num(double(mat->d_influenceFnType)) +
       ", \"influence_params\": " + list(mat->d_influenceFnParams) +
       ", \"critical_stretch\": " + num(mat->d_criticalStretch) +
       ", \"no_penetration\": " + (mat->d_applyContact ? "true" : "false") +
       "}";
  s += ", \"cracks\": [";
  const auto& cr = d.getFractureDeck()->d_cracks;
  for (size_t k = 0; k < cr.size(); k++)
    s += std::string(k ? ", " : "") + "{\"o\": " + num(cr[k].d_o) +
         ", \"pb\": [" + num(cr[k].d_pb[0]) + ", " + num(cr[k].d_pb[1]) +
         "], \"pt\": [" + num(cr[k].d_pt[0]) + ", " + num(cr[k].d_pt[1]) +
         "], \"theta\": " + num(cr[k].d_theta) + ", \"activation\": " +
         num(cr[k].d_activationTime) + "}";
  s += "], \"ubc\": [";
  for (size_t k = 0; k < l->d_uBCData.size(); k++)
    s += (k ? ", " : "") + bc(l->d_uBCData[k]);
  s += "], \"fbc\": [";
  for (size_t k = 0; k < l->d_fBCData.size(); k++)
    s += (k ? ", " : "") + bc(l->d_fBCData[k]);
  s += "], \"output\": {\"path\": " + q(o->d_path) + ", \"tags\": [";
  for (size_t k = 0; k < o->d_outTags.size(); k++)
    s += (k ? ", " : "") + q(o->d_outTags[k]);
  s += "], \"interval\": " + num(double(o->d_dtOut)) + ", \"compress\": " +
       q(o->d_compressType) + ", \"criteria\": " + q(o->d_outCriteria) +
       ", \"criteria_interval\": " + num(double(o->d_dtOutCriteria)) +
       ", \"criteria_params\": " + list(o->d_outCriteriaParams) +
       ", \"perform_fe_out\": " + (o->d_performFEOut ? "true" : "false") +
       "}";
  const auto* sd = d.getSolverDeck();
  s += ", \"solver\": {\"type\": " + q(sd->d_solverType) +
       ", \"max_iter\": " + num(sd->d_maxIters) + ", \"tol\": " +
       num(sd->d_tol) + ", \"perturbation\": " + num(sd->d_perturbation) +
       "}";
  s += ", \"postprocessing\": " +
       std::string(d.getPolicyDeck()->d_enablePostProcessing ? "true"
                                                             : "false");
  const auto* nf = d.getInteriorFlagsDeck();
  s += ", \"no_fail\": {\"active\": " +
       std::string(nf->d_noFailActive ? "true" : "false") +
       ", \"tol\": " + num(nf->d_noFailTol) + "}";
  const auto* ic = d.getInitialConditionDeck();
  s += ", \"ic\": {\"u_type\": " + q(ic->d_uICData.d_type) +
       ", \"u_params\": " + list(ic->d_uICData.d_params) +
       ", \"v_type\": " + q(ic->d_vICData.d_type) +
       ", \"v_params\": " + list(ic->d_vICData.d_params) + "}";
  s += ", \"gpu\": {\"precision\": " + q(d.getGpuDeck()->d_precision) +
       ", \"graph_chunk\": " + num(d.getGpuDeck()->d_graphChunk) + "}";
  s += "}";
  std::cout << s << std::endl;
}

// The mesh the deck names, as the model reads it (nlm_fd --dump-mesh): node
// coordinates, nodal volumes and element connectivity as JSON.
void dump_mesh(nlm::inp::Input& d) {
  nlm::fe::Mesh mesh(*d.getMeshDeck());
  std::string s = "{\"dim\": " + num(double(mesh.getDimension())) +
                  ", \"h\": " + num(mesh.getMeshSize()) + ", \"nodes\": [";
  const auto& x = mesh.getNodes();
  for (size_t k = 0; k < x.size(); k++) s += (k ? ", " : "") + num(x[k]);
  s += "], \"volumes\": " + list(mesh.getNodalVolumes()) +
       ", \"element_type\": " + num(double(mesh.getElementType())) +
       ", \"elements\": [";
  const auto& e = mesh.getElementConnectivities();
  for (size_t k = 0; k < e.size(); k++)
    s += (k ? ", " : "") + std::to_string(e[k]);
  s += "]}";
  std::cout << s << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
  std::string file;
  bool dump_only = false, dump_mesh_only = false;
  for (int k = 1; k < argc; k++) {
    if (!std::strcmp(argv[k], "--dump-deck")) dump_only = true;
    if (!std::strcmp(argv[k], "--dump-mesh")) dump_mesh_only = true;
    if ((!std::strcmp(argv[k], "-i") || !std::strcmp(argv[k], "--input-file")) &&
        k + 1 < argc)
      file = argv[++k];
    else if (!std::strncmp(argv[k], "--input-file=", 13))
      file = argv[k] + 13;
  }
  if (file.empty()) {
    std::cerr << argv[0] << " (" << nlm_version() << ") -i input.yaml"
              << std::endl;
    return 1;
  }
  if (dump_only || dump_mesh_only) {
    try {
      nlm::inp::Input deck(file);
      if (dump_only) dump(deck);
      else dump_mesh(deck);
    } catch (const nlm::inp::DeckError& e) {
      std::cerr << e.what() << std::endl;
      return 1;
    }
    return 0;
  }
  std::cout << argv[0] << " (" << nlm_version() << ")" << std::endl;
  const auto begin = std::chrono::steady_clock::now();
  try {
    nlm::inp::Input deck(file);
    if (deck.getSpatialDiscretization() != "finite_difference") {
      std::cerr << "Warning no model for the spatial discretization "
                   "specified!"
                << std::endl;
      return 1;
    }
    const std::string& td = deck.getModelDeck()->d_timeDiscretization;
    const std::string& mt = deck.getMaterialDeck()->d_materialType;
    if (td == "quasi_static") {
      if (mt == "ElasticState")
        nlm::model::QuasiStaticModel<nlm::material::pd::ElasticState> m(&deck);
      else if (mt == "RNPBond")
        nlm::model::QuasiStaticModel<nlm::material::pd::RNPBond> m(&deck);
      else
        throw nlm::inp::DeckError("Error: no quasi-static model for material " +
                                  mt);
    } else if (td == "central_difference" || td == "velocity_verlet") {
      if (mt == "RNPBond")
        nlm::model::FDModel<nlm::material::pd::RNPBond> m(&deck);
      else if (mt == "ElasticState")
        nlm::model::FDModel<nlm::material::pd::ElasticState> m(&deck);
      else
        throw nlm::inp::DeckError("Error: no explicit model for material " +
                                  mt);
    } else {
      std::cerr << "Warning no model for the spatial discretization "
                   "specified!"
                << std::endl;
      return 1;
    }
  } catch (const nlm::inp::DeckError& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return 2;
  }
  const double s = std::chrono::duration<double>(
                       std::chrono::steady_clock::now() - begin)
                       .count();
  std::cout << " Time elapsed = " << s << " sec \n";
  return 0;
}
