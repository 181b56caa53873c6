// Input deck reader (see input.h).  Section by section it follows
// inp::Input's setters in src/inp/input.cpp (line ranges cited per setter),
// reading the same keys with the same defaults and checks.
#include "input.h"

#include <cmath>
#include <iostream>

namespace nlm {
namespace inp {

namespace {

[[noreturn]] void error(const std::string& msg) { throw DeckError(msg); }

// definitelyGreaterThan (src/util/compare.cpp:25-28)
bool greater_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (a - b) > m * 1e-5;
}

}  // namespace

Input::Input(const std::string& filename) {
  yaml::Node config;
  try {
    config = yaml::load_file(filename);
  } catch (const yaml::Error& e) {
    error(std::string("Error: ") + e.what());
  }
  read(config);
}

Input Input::fromText(const std::string& text) {
  Input in;
  yaml::Node config;
  try {
    config = yaml::load(text);
  } catch (const yaml::Error& e) {
    error(std::string("Error: ") + e.what());
  }
  in.read(config);
  return in;
}

void Input::read(const yaml::Node& config) {
  try {
    // same order as inp::Input::Input (input.cpp:42-72)
    setModelDeck(config);
    setRestartDeck(config);
    setMeshDeck(config);
    setFractureDeck(config);
    setInteriorFlagsDeck(config);
    setInitialConditionDeck(config);
    setLoadingDeck(config);
    setMaterialDeck(config);
    setOutputDeck(config);
    setPolicyDeck(config);
    setSolverDeck(config);
    setGpuDeck(config);
  } catch (const yaml::Error& e) {
    error(std::string("Error: ") + e.what());
  }
}

// setModelDeck (input.cpp:124-188)
void Input::setModelDeck(const yaml::Node& config) {
  const auto& m = config["Model"];
  if (m["Dimension"])
    d_model.d_dim = m["Dimension"].as_size();
  else
    error("Error: Please specify the dimension.");

  if (m["Discretization_Type"]["Time"])
    d_model.d_timeDiscretization = m["Discretization_Type"]["Time"].str();
  if (m["Discretization_Type"]["Spatial"])
    d_model.d_spatialDiscretization =
        m["Discretization_Type"]["Spatial"].str();
  if (d_model.d_timeDiscretization == "central_difference" ||
      d_model.d_timeDiscretization == "velocity_verlet")
    d_model.d_simType = "explicit";

  if (m["Horizon"]) d_model.d_horizon = m["Horizon"].as_double();
  if (m["Horizon_h_Ratio"]) d_model.d_rh = int(m["Horizon_h_Ratio"].as_int());
  if (m["Mesh_Size"]) d_model.d_h = m["Mesh_Size"].as_double();

  if (!m["Horizon"]) {
    if (!m["Horizon_h_Ratio"] || !m["Mesh_Size"])
      error(
          "Error: Horizon is not provided. In this case Horizon_h_Ratio and "
          "Mesh_Size are necessary to compute horizon.");
    d_model.d_horizon = d_model.d_h * double(d_model.d_rh);
  }
  if (!m["Mesh_Size"])
    if (m["Horizon_h_Ratio"])
      d_model.d_h = d_model.d_horizon / double(d_model.d_rh);

  if (m["Final_Time"]) d_model.d_tFinal = m["Final_Time"].as_double();
  if (m["Time_Steps"]) d_model.d_Nt = m["Time_Steps"].as_int();
  if (std::abs(d_model.d_tFinal) < 1.0E-10 || d_model.d_Nt <= 0)
    error("Error: Check Final_Time and Time_Steps data.");
  d_model.d_dt = d_model.d_tFinal / double(d_model.d_Nt);

  if (config["Restart"]) d_model.d_isRestartActive = true;
}

// setRestartDeck (input.cpp:190-212)
void Input::setRestartDeck(const yaml::Node& config) {
  const auto& r = config["Restart"];
  if (!r) return;
  if (r["File"])
    d_restart.d_file = r["File"].str();
  else
    error("Error: Please specify the file for restart.");
  if (r["Step"])
    d_restart.d_step = r["Step"].as_size();
  else
    error("Error: Please specify the time step from which to restart the "
          "simulation.");
}

// setMeshDeck (input.cpp:214-254)
void Input::setMeshDeck(const yaml::Node& config) {
  d_mesh.d_dim = d_model.d_dim;
  d_mesh.d_spatialDiscretization = d_model.d_spatialDiscretization;
  if (config["Mesh"]["File"])
    d_mesh.d_filename = config["Mesh"]["File"].str();
  else
    error("Error: Please specify mesh filename.");
  if (d_model.d_h < 1.0E-12)
    d_mesh.d_computeMeshSize = true;
  else
    d_mesh.d_h = d_model.d_h;
  if (config["Mesh"]["Is_Centroid_Based_Discretization"])
    d_mesh.d_isCentroidBasedDiscretization =
        config["Mesh"]["Is_Centroid_Based_Discretization"].as_bool();
  if (config["Mesh"]["Keep_Element_Conn"])
    d_mesh.d_keepElementConn = config["Mesh"]["Keep_Element_Conn"].as_bool();
}

// setFractureDeck (input.cpp:290-355)
void Input::setFractureDeck(const yaml::Node& config) {
  const auto& c = config["Fracture"]["Cracks"];
  size_t ncracks = 0;
  if (c["Sets"]) ncracks = c["Sets"].as_size();
  if (ncracks == 0) return;
  if (d_model.d_dim != 2)
    error("Error: Pre-cracks can only be prescribed for material in "
          "dimension 2.");
  for (size_t s = 1; s <= ncracks; s++) {
    const auto& e = c["Set_" + std::to_string(s)];
    EdgeCrack crack;
    if (e["Orientation"]) crack.d_o = int(e["Orientation"].as_int());
    if (e["Line"]) {
      std::vector<double> locs = e["Line"].as_doubles();
      if (locs.size() != 4)
        error("Error: Input data in tag Line does not have enough elements.");
      crack.d_pb[0] = locs[0];
      crack.d_pb[1] = locs[1];
      crack.d_pt[0] = locs[2];
      crack.d_pt[1] = locs[3];
      const double dx = locs[2] - locs[0], dy = locs[3] - locs[1];
      crack.d_l = std::sqrt(dx * dx + dy * dy);
    } else {
      error("Error: Need Line data for complete specification of pre-crack.");
    }
    if (crack.d_o == 0) {
      if (e["Angle"])
        crack.d_theta = e["Angle"].as_double();
      else
        error("Error: Need Angle data for crack with arbitrary orientation.");
    }
    if (e["Activation_Time"])
      crack.d_activationTime = e["Activation_Time"].as_double();
    d_fracture.d_cracks.push_back(crack);
  }
}

// setInteriorFlagsDeck (input.cpp:357-401)
void Input::setInteriorFlagsDeck(const yaml::Node& config) {
  const auto& e = config["No_Fail_Region"];
  if (!e) return;
  d_interior.d_noFailActive = true;
  if (!e["Factor"]) error("Error: No_Fail_Region needs Factor.");
  d_interior.d_noFailTol = d_model.d_horizon * e["Factor"].as_double();
  if (e["Compute_And_Not_Store"])
    d_interior.d_computeAndNotStoreFlag = e["Compute_And_Not_Store"].as_bool();
  size_t num_regions = 0;
  if (e["Num_Regions"]) num_regions = e["Num_Regions"].as_size();
  for (size_t i = 1; i <= num_regions; i++) {
    const auto& ee = e["Region_" + std::to_string(i)];
    std::pair<std::string, std::vector<double>> data;
    if (ee["Type"])
      data.first = ee["Type"].str();
    else
      error("Error: Region type must be provided for no-fail region");
    data.second = ee["Parameters"].as_doubles();
    d_interior.d_noFailRegions.push_back(data);
  }
}

// setInitialConditionDeck (input.cpp:403-436)
void Input::setInitialConditionDeck(const yaml::Node& config) {
  for (int k = 0; k < 2; k++) {
    const auto& e = config["Initial_Condition"][k == 0 ? "Displacement"
                                                       : "Velocity"];
    if (!e) continue;
    ICData& d = k == 0 ? d_ic.d_uICData : d_ic.d_vICData;
    if (e["File"]) {
      d.d_file = e["File"].str();
      d.d_type = "file";
    }
    if (e["Type"]) d.d_type = e["Type"].str();
    d.d_params = e["Parameters"].as_doubles();
  }
}

// setLoadingDeck (input.cpp:438-612)
void Input::setLoadingDeck(const yaml::Node& config) {
  for (const std::string tag : {"Displacement_BC", "Force_BC"}) {
    const auto& t = config[tag];
    if (!t) continue;
    const long long nsets = t["Sets"].as_int();
    for (long long s = 1; s <= nsets; s++) {
      const auto& e = t["Set_" + std::to_string(s)];
      const auto& loc = e["Location"];
      BCData bc;
      if (loc["Line"]) {
        bc.d_regionType = "line";
        auto l = loc["Line"].as_doubles();
        if (l.size() != 2) error("Error: Check Line data in " + tag + ".");
        bc.d_x1 = l[0];
        bc.d_x2 = l[1];
      } else if (loc["Rectangle"]) {
        bc.d_regionType = "rectangle";
        auto l = loc["Rectangle"].as_doubles();
        if (l.size() != 4) error("Error: Check Rectangle data in " + tag + ".");
        bc.d_x1 = l[0];
        bc.d_x2 = l[2];
        bc.d_y1 = l[1];
        bc.d_y2 = l[3];
      } else if (loc["Angled_Rectangle"]) {
        bc.d_regionType = "angled_rectangle";
        auto l = loc["Angled_Rectangle"].as_doubles();
        if (l.size() != 4) error("Error: Check Rectangle data in " + tag + ".");
        bc.d_x1 = l[0];
        bc.d_x2 = l[2];
        bc.d_y1 = l[1];
        bc.d_y2 = l[3];
        bc.d_theta = loc["Angle"].as_double();
        double alpha = std::atan((bc.d_y2 - bc.d_y1) / (bc.d_x2 - bc.d_x1));
        if (greater_than(0.0, alpha)) alpha = alpha + M_PI;
        if (greater_than(bc.d_theta, alpha) ||
            std::abs(alpha - bc.d_theta) <= 1.0E-5)
          error("Error: Check angled rectangle. Either angle of diagonal is "
                "smaller than theta or it is very close to theta");
      } else if (loc["Point"] && loc["Radius"]) {
        bc.d_regionType = "circle";
        auto p = loc["Point"].as_doubles();
        if (p.size() != 2) error("Error: Check Point data in " + tag + ".");
        bc.d_x1 = p[0];
        bc.d_y1 = p[1];
        auto r = loc["Radius"].as_doubles();
        if (r.size() != 1) error("Error: Check Radii data in " + tag + ".");
        bc.d_r1 = r[0];
      } else if (loc["Point"] && loc["Radii"]) {
        bc.d_regionType = "torus";
        auto p = loc["Point"].as_doubles();
        if (p.size() != 2) error("Error: Check Point data in " + tag + ".");
        bc.d_x1 = p[0];
        bc.d_y1 = p[1];
        auto r = loc["Radii"].as_doubles();
        if (r.size() != 2) error("Error: Check Radii data in " + tag + ".");
        bc.d_r1 = r[0];
        bc.d_r2 = r[1];
      } else if (loc["Cuboid"]) {
        bc.d_regionType = "cuboid";
        auto l = loc["Cuboid"].as_doubles();
        if (l.size() != 6) error("Error: Check Cuboid data in " + tag + ".");
        bc.d_x1 = l[0];
        bc.d_x2 = l[3];
        bc.d_y1 = l[1];
        bc.d_y2 = l[4];
        bc.d_z1 = l[2];
        bc.d_z2 = l[5];
      } else if (loc["Displacement_from_pum"]) {
        bc.d_regionType = "displacement_from_pum";
      }
      for (const auto& d : e["Direction"].items())
        bc.d_direction.push_back(d.as_size());
      if (e["Direction"].is_scalar())
        bc.d_direction.push_back(e["Direction"].as_size());
      if (e["Time_Function"]) {
        bc.d_timeFnType = e["Time_Function"]["Type"].str();
        if (e["Time_Function"]["Parameters"])
          bc.d_timeFnParams = e["Time_Function"]["Parameters"].as_doubles();
      }
      if (e["Spatial_Function"]) {
        bc.d_spatialFnType = e["Spatial_Function"]["Type"].str();
        if (e["Spatial_Function"]["Parameters"])
          bc.d_spatialFnParams =
              e["Spatial_Function"]["Parameters"].as_doubles();
      }
      (tag == "Displacement_BC" ? d_loading.d_uBCData : d_loading.d_fBCData)
          .push_back(bc);
    }
  }
}

// setMaterialDeck (input.cpp:614-699)
void Input::setMaterialDeck(const yaml::Node& config) {
  const auto& e = config["Material"];
  if (e["Is_Plane_Strain"])
    d_material.d_isPlaneStrain = e["Is_Plane_Strain"].as_bool();
  if (e["Type"])
    d_material.d_materialType = e["Type"].str();
  else
    error("Error: Please specify the material type.");
  if (d_material.d_materialType == "PDBond") {
    std::cout << "Warning: PDBond is changed to RNPBond.\n";
    d_material.d_materialType = "RNPBond";
  }
  if (e["Compute_From_Classical"])
    d_material.d_computeParamsFromElastic =
        e["Compute_From_Classical"].as_bool();
  MatData& md = d_material.d_matData;
  if (e["E"]) md.d_E = e["E"].as_double();
  if (e["G"]) md.d_G = e["G"].as_double();
  if (e["K"]) md.d_K = e["K"].as_double();
  if (e["Lambda"]) md.d_lambda = e["Lambda"].as_double();
  if (e["Mu"]) md.d_mu = e["Mu"].as_double();
  if (e["Poisson_Ratio"]) md.d_nu = e["Poisson_Ratio"].as_double();
  if (e["Gc"]) md.d_Gc = e["Gc"].as_double();
  if (e["KIc"]) md.d_KIc = e["KIc"].as_double();
  if (const auto& f = e["Bond_Potential"]) {
    if (f["Type"]) d_material.d_bondPotentialType = f["Type"].as_size();
    if (f["Check_Sc_Factor"])
      d_material.d_checkScFactor = f["Check_Sc_Factor"].as_double();
    if (f["Irreversible_Bond_Fracture"])
      d_material.d_irreversibleBondBreak =
          f["Irreversible_Bond_Fracture"].as_bool();
    if (f["Parameters"])
      d_material.d_bondPotentialParams = f["Parameters"].as_doubles();
  }
  if (const auto& f = e["State_Potential"]) {
    if (f["Type"]) d_material.d_statePotentialType = f["Type"].as_size();
    if (f["Contribution_From_Broken_Bond"])
      d_material.d_stateContributionFromBrokenBond =
          f["Contribution_From_Broken_Bond"].as_bool();
    if (f["Parameters"])
      d_material.d_statePotentialParams = f["Parameters"].as_doubles();
  }
  if (const auto& f = e["Influence_Function"]) {
    if (f["Type"]) d_material.d_influenceFnType = f["Type"].as_size();
    if (f["Parameters"])
      d_material.d_influenceFnParams = f["Parameters"].as_doubles();
  }
  if (e["Density"])
    d_material.d_density = e["Density"].as_double();
  else
    error("Error: Please specify the density of the material.");
  if (e["No_Penetration"])
    d_material.d_applyContact = e["No_Penetration"].as_bool();
  if (e["Critical_Stretch"])
    d_material.d_criticalStretch = e["Critical_Stretch"].as_double();
}

// setOutputDeck (input.cpp:701-735)
void Input::setOutputDeck(const yaml::Node& config) {
  const auto& e = config["Output"];
  if (!e) return;
  if (e["File_Format"]) d_output.d_outFormat = e["File_Format"].str();
  // rw::writer::Writer::open (src/rw/writer.cpp:30-40) also knows "msh" (the
  // gmsh writer) and writes nothing for other names
  if (d_output.d_outFormat != "vtu" && d_output.d_outFormat != "legacy_vtk")
    throw DeckError("Error: Output File_Format " + d_output.d_outFormat +
                    " is not supported (vtu or legacy_vtk).");
  if (e["Path"]) d_output.d_path = e["Path"].str();
  if (e["Tags"]) d_output.d_outTags = e["Tags"].as_strings();
  if (e["Output_Interval"]) d_output.d_dtOut = e["Output_Interval"].as_size();
  d_output.d_dtOutOld = d_output.d_dtOut;
  d_output.d_dtOutCriteria = d_output.d_dtOut;
  if (e["Debug"]) d_output.d_debug = e["Debug"].as_size();
  if (e["Perform_FE_Out"])
    d_output.d_performFEOut = e["Perform_FE_Out"].as_bool();
  if (e["Compress_Type"]) d_output.d_compressType = e["Compress_Type"].str();
  if (const auto& c = e["Output_Criteria"]) {
    if (c["Type"]) d_output.d_outCriteria = c["Type"].str();
    if (c["New_Interval"])
      d_output.d_dtOutCriteria = c["New_Interval"].as_size();
    if (c["Parameters"])
      d_output.d_outCriteriaParams = c["Parameters"].as_doubles();
  }
  if (e["Bond_State"]) d_output.d_bondState = e["Bond_State"].as_bool();
}

// setPolicyDeck (input.cpp:737-746)
void Input::setPolicyDeck(const yaml::Node& config) {
  const auto& p = config["Policy"];
  if (p["Memory_Consumption_Flag"])
    d_policy.d_memControlFlag = int(p["Memory_Consumption_Flag"].as_int());
  if (p["Enable_PostProcessing"])
    d_policy.d_enablePostProcessing = p["Enable_PostProcessing"].as_bool();
}

// setSolverDeck (input.cpp:748-762)
void Input::setSolverDeck(const yaml::Node& config) {
  const auto& e = config["Solver"];
  if (!e) return;
  if (e["Type"]) d_solver.d_solverType = e["Type"].str();
  if (e["Max_Iteration"]) d_solver.d_maxIters = int(e["Max_Iteration"].as_int());
  if (e["Tolerance"]) d_solver.d_tol = e["Tolerance"].as_double();
  if (e["Perturbation"]) d_solver.d_perturbation = e["Perturbation"].as_double();
  if (e["Preconditioner"]) d_solver.d_preconditioner = e["Preconditioner"].str();
  if (e["Linear_Tolerance"])
    d_solver.d_linTol = e["Linear_Tolerance"].as_double();
  if (e["Linear_Max_Iteration"])
    d_solver.d_linMaxIter = int(e["Linear_Max_Iteration"].as_int());
  if (e["Relative_Tolerance"])
    d_solver.d_tolRel = e["Relative_Tolerance"].as_double();
}

void Input::setGpuDeck(const yaml::Node& config) {
  const auto& g = config["GPU"];
  if (!g) return;
  if (g["Precision"]) d_gpu.d_precision = g["Precision"].str();
  if (d_gpu.d_precision != "fp64_exact" && d_gpu.d_precision != "fp64" &&
      d_gpu.d_precision != "fp32_force")
    error("Error: GPU Precision must be fp64_exact, fp64 or fp32_force.");
  if (g["Device"]) d_gpu.d_device = int(g["Device"].as_int());
  if (g["Graph_Chunk"]) d_gpu.d_graphChunk = int(g["Graph_Chunk"].as_int());
  if (g["Autotune"]) d_gpu.d_autotune = g["Autotune"].as_bool();
}

}  // namespace inp
}  // namespace nlm
