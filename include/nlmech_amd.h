/*
 * nlmech_amd.h -- C-ABI of the MI355X peridynamics engine.
 *
 * The library replaces the explicit-dynamics hot path of NLMech's FDModel:
 * neighbor search, the RNPBond bond force with its break test, and the
 * central-difference update, computed by hand-written HIP kernels for gfx950.
 *
 * Conventions
 *   - Every entry point returns NLM_OK (0) or a negative NLM_ERR_* code; the
 *     message of the last failure is nlm_last_error(ctx).  Nothing in the
 *     library calls exit() (the reference's error convention is
 *     std::cerr + exit(1), e.g. src/material/pd/rnpBond.cpp:58-60).
 *   - Host arrays are owned by the caller, passed with explicit sizes, copied
 *     during the call and never retained.
 *   - Node vectors at this boundary are AoS Point3, double[3*n] with z = 0 in
 *     2D: the layout of the reference's std::vector<util::Point3>
 *     (src/util/point.h:29-516, src/data/DataManager.h:74-372), so a binding
 *     passes &u[0].d_x directly.
 *   - Neighbor lists are CSR with int64 row offsets and int32 columns, each row
 *     ascending by global id (the order of the reference's brute-force loop,
 *     src/geometry/neighbor.cpp:69-89).  Bond bits are addressed as
 *     (node i, local index jl) exactly as geometry::Fracture does
 *     (src/geometry/fracture.cpp:162-187).
 *   - One host thread per context; all device work is ordered on the
 *     context's HIP stream.  Contexts are independent.
 */
#ifndef NLMECH_AMD_H
#define NLMECH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLM_OK 0
#define NLM_ERR_INVALID (-1) /* bad argument or unsupported option */
#define NLM_ERR_HIP (-2)     /* HIP runtime failure */
#define NLM_ERR_STATE (-3)   /* call out of order (e.g. step before setup) */
#define NLM_ERR_MEMORY (-4)  /* allocation failed */

typedef struct nlm_ctx nlm_ctx;

/* Precision modes of the bond-force kernels. */
#define NLM_PREC_FP64_EXACT 0 /* reference operation order, no FMA, IEEE div */
#define NLM_PREC_FP64 1       /* FP64, break test as the reference, FMA elsewhere */
#define NLM_PREC_FP32_FORCE 2 /* FP64 state / stretch / break test, FP32 force */

/* Time integrators, inp::ModelDeck::d_timeDiscretization
 * (src/inp/input.cpp:137-147). */
#define NLM_TIME_CENTRAL_DIFFERENCE 0 /* FDModel::integrateCD, fDModel.cpp:377 */
#define NLM_TIME_VELOCITY_VERLET 1    /* FDModel::integrateVerlet, :446 */

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
int nlm_ctx_create(nlm_ctx** out, int device);
void nlm_ctx_destroy(nlm_ctx* ctx);
const char* nlm_last_error(const nlm_ctx* ctx);
const char* nlm_version(void);
/* Name of the GPU the context runs on (e.g. "gfx950"). */
int nlm_device_arch(const nlm_ctx* ctx, char* buf, int len);

/* ------------------------------------------------------------------------ */
/* Material: RNPBond (src/material/pd/rnpBond.cpp)                           */
/* ------------------------------------------------------------------------ */

/* Fields of inp::MaterialDeck used by RNPBond (src/inp/decks/materialDeck.h:
 * 226-288; MatData entries default to -1, :20-99) plus the model sizes. */
typedef struct {
  int dim;
  double horizon;
  double h; /* mesh size */
  double density;
  double E, K, Gc, KIc; /* < 0: not given */
  int plane_strain;
  int compute_from_classical;
  int n_bond_params;
  double bond_params[2];
  int inf_type; /* 0 constant, 1 linear, 2 gaussian (influenceFn.h) */
  int n_inf_params;
  double inf_params[2];
  double check_sc_factor;
  int irreversible;
  int no_penetration;
} nlm_material_deck;

/* Derived per-bond parameters: the state RNPBond keeps after its constructor
 * (rnpBond.cpp:23-97, computeParameters :99-189). */
typedef struct {
  int dim;
  double horizon, h, density;
  double C, beta, rbar, inv_factor, factor_sc;
  int irreversible;
  int inf_type;
  double inf_a, inf_b;
  double inf_moment;
  int contact;
  double contact_kn, contact_rc;
  double E, K, nu, lambda, G, Gc, KIc;
} nlm_rnp_params;

/* Replaces RNPBond::RNPBond + computeParameters (rnpBond.cpp:23-189). */
int nlm_rnp_params_from_deck(const nlm_material_deck* deck,
                             nlm_rnp_params* out);

/* ------------------------------------------------------------------------ */
/* Material: ElasticState (src/material/pd/ElasticState.cpp)                 */
/* ------------------------------------------------------------------------ */

/* Fields of inp::MaterialDeck used by ElasticState (materialDeck.h:226-288)
 * plus the critical stretch of the explicit model's bond breaking (the
 * reference ElasticState has no damage, ElasticState.cpp:273; here a bond
 * whose stretch (|Y|-|X|)/|X| is definitelyGreaterThan critical_stretch at a
 * force evaluation breaks irreversibly; <= 0 disables breaking). */
typedef struct {
  int dim;
  double horizon;
  double h;
  double density;
  double E, K, G; /* < 0: not given; 1D needs E, 2D/3D need K and G */
  int plane_strain;
  double critical_stretch;
} nlm_es_deck;

/* ElasticState constants after computeParameters (ElasticState.cpp:47-158):
 * nu = (3K - 2G) / (2 (3K + G)) (materialDeck.h:120), mu = G, factor2D =
 * (2 nu - 1) / (nu - 1) for 2D plane stress, else 1. */
typedef struct {
  int dim;
  double horizon, h, density;
  double E, K, G, nu, mu, factor2D;
  int plane_strain;
  double critical_stretch;
} nlm_es_params;

/* Replaces ElasticState::ElasticState + computeParameters
 * (ElasticState.cpp:17-158); errors where the reference exits. */
int nlm_es_params_from_deck(const nlm_es_deck* deck, nlm_es_params* out);

/* ------------------------------------------------------------------------ */
/* Neighbor search: geometry::Neighbor::Neighbor (neighbor.cpp:24-92)       */
/* ------------------------------------------------------------------------ */

/* Builds the neighbor lists of n nodes on the GPU with a cell list.  Node j
 * is a neighbor of i iff j != i and definitelyLessThan(|xi - xj|,
 * horizon + 1e-10) (neighbor.cpp:81-83); rows come out ascending.  The lists
 * stay on the device for nlm_fd_setup; *nnz_out receives the total count. */
int nlm_neighbor_build(nlm_ctx* ctx, int dim, int64_t n, const double* x,
                       double horizon, int64_t* nnz_out);
/* Copies the lists of the last nlm_neighbor_build (row_ptr[n+1], col[nnz]).
 * Replaces Neighbor::getNeighborsList (neighbor.cpp:106-112). */
int nlm_neighbor_get(nlm_ctx* ctx, int64_t* row_ptr, int32_t* col);

/* ------------------------------------------------------------------------ */
/* Explicit FD model: FDModel<RNPBond> (src/model/fd/fDModel.cpp)            */
/* ------------------------------------------------------------------------ */

/* Mesh, neighbor lists and model constants.  row_ptr/col may be NULL to use
 * the lists of the last nlm_neighbor_build on this context.  interior may be
 * NULL (every node interior: geometry::BaseInterior, interiorFlags.cpp:27-30);
 * otherwise interior[i] != 0 marks node i as breakable
 * (InteriorFlags::getInteriorFlag, interiorFlags.cpp:228-231). */
typedef struct {
  int dim; /* 1, 2 or 3; a 1D model (nodes on a line, y = z = 0) runs on the
              2D kernels with every y dof held at zero */
  int64_t n;
  const double* x;   /* [3n] reference coordinates, Mesh::getNodes */
  const double* vol; /* [n] nodal volumes, Mesh::getNodalVolumes */
  const int64_t* row_ptr;
  const int32_t* col;
  const uint8_t* interior;
  double dt; /* ModelDeck::d_dt = Final_Time / Time_Steps (input.cpp:184) */
  int precision;   /* NLM_PREC_* */
  int time_scheme; /* NLM_TIME_* */
  nlm_rnp_params mat; /* material == NLM_MATERIAL_RNP_BOND */
  int material;       /* NLM_MATERIAL_* (0: RNPBond) */
  nlm_es_params es;   /* material == NLM_MATERIAL_ELASTIC_STATE */
} nlm_fd_problem;

/* The model's material: FDModel<T>'s template argument.  The state-based
 * material computes, per force evaluation, the weighted-volume dilatation
 * (StateBasedHelperFunctions::dilatation, stateBasedHelperFunctions.cpp:
 * 26-83) and the ElasticState bond force of both bond directions, summed as
 * QuasiStaticModel::computeForces does (f_i += T_ij V_j, f_j -= T_ij V_i,
 * QuasiStaticModel.cpp:169-239); central difference only. */
#define NLM_MATERIAL_RNP_BOND 0
#define NLM_MATERIAL_ELASTIC_STATE 1

/* Region / function codes of inp::BCData (src/inp/decks/loadingDeck.h). */
#define NLM_REGION_RECTANGLE 0
#define NLM_REGION_ANGLED_RECTANGLE 1
#define NLM_REGION_CIRCLE 2
#define NLM_REGION_TORUS 3
#define NLM_REGION_LINE 4
#define NLM_REGION_CUBOID 5
#define NLM_TFN_CONSTANT 0
#define NLM_TFN_LINEAR 1
#define NLM_TFN_QUADRATIC 2
#define NLM_TFN_SIN 3
#define NLM_TFN_LINEAR_STEP 4
#define NLM_TFN_LINEAR_SLOW_FAST 5
#define NLM_SFN_CONSTANT 0
#define NLM_SFN_SIN_X 1
#define NLM_SFN_SIN_Y 2
#define NLM_SFN_SIN_Z 3
#define NLM_SFN_LINEAR_X 4
#define NLM_SFN_LINEAR_Y 5
#define NLM_SFN_LINEAR_Z 6
#define NLM_SFN_HAT_X 7
#define NLM_SFN_HAT_Y 8
#define NLM_SFN_HAT_Z 9
/* Force BCs only: FLoading's line_load (fLoading.cpp:172-228), for each
 * loaded direction d the factor floor((x_d - min_d)/h) / floor(|max_d -
 * min_d|/h) over the set's extent in d (floor((max_d - x_d)/h) when
 * space_params[0] == -1), h the mesh size given at setup. */
#define NLM_SFN_LINE_LOAD 10

typedef struct {
  int region;
  double x1, x2, y1, y2, z1, z2, r1, r2, theta;
  int ndir;
  int dir[3]; /* 1-based dofs */
  int time_fn;
  int n_time_params;
  double time_params[4];
  int space_fn;
  int n_space_params;
  double space_params[4];
} nlm_bc;

/* Edge crack of inp::FractureDeck (src/inp/decks/fractureDeck.h:22-104). */
typedef struct {
  int o; /* 1 horizontal, -1 vertical, 0 angled */
  double pb[2], pt[2];
  double theta;
  double activation_time; /* < 0: applied at setup (Fracture ctor) */
} nlm_crack;

/* Uploads mesh, lists and constants; allocates the device state (u = v = f
 * = 0, every bond intact).  Replaces the allocation part of
 * FDModel::initHObjects/init (fDModel.cpp:134-315). */
int nlm_fd_setup(nlm_ctx* ctx, const nlm_fd_problem* p);
/* Displacement BC sets: ULoading::ULoading node selection and fixity
 * (uLoading.cpp:18-164); applied by nlm_fd_init / every step
 * (ULoading::apply, :166-253).  At most 14 sets.  Clears node-list BCs. */
int nlm_fd_set_ubc(nlm_ctx* ctx, int nsets, const nlm_bc* sets);
/* Node-list displacement BCs (no counterpart in the reference's loading
 * deck; used for batched specimens): dof d of node nodes[k] is prescribed
 * when bit d of dofs[k] is set, with u = rates[3 k + d] t and v =
 * rates[3 k + d] (ULoading's "linear" time function per node).  Replaces
 * earlier node-list BCs; call after nlm_fd_set_ubc. */
int nlm_fd_set_node_ubc(nlm_ctx* ctx, int64_t count, const int64_t* nodes,
                        const uint8_t* dofs, const double* rates);
/* Force BC sets: FLoading (fLoading.cpp:18-279).  At most 8 sets. */
int nlm_fd_set_fbc(nlm_ctx* ctx, int nsets, const nlm_bc* sets);
/* Pre-cracks: cracks with activation_time < 0 break their crossing bonds now
 * (Fracture::Fracture -> computeFracturedBondFd, fracture.cpp:20-160); the
 * others are applied during stepping when the time passes their activation
 * time (Fracture::addCrack, :45-72). */
int nlm_fd_add_cracks(nlm_ctx* ctx, int ncracks, const nlm_crack* cracks);

/* State transfer, AoS double[3n]; any pointer may be NULL. */
int nlm_fd_set_state(nlm_ctx* ctx, const double* u, const double* v,
                     const double* f);
int nlm_fd_get_state(nlm_ctx* ctx, double* u, double* v, double* f);
/* Bond bits, one byte (0/1) per directed bond in CSR order. */
int nlm_fd_set_bonds(nlm_ctx* ctx, const uint8_t* bond_state);
int nlm_fd_get_bonds(nlm_ctx* ctx, uint8_t* bond_state);
/* Bond bits in the reference packing: ceil(n_i/8) bytes per node,
 * concatenated over nodes (Fracture::d_fracture, fracture.cpp:24-33). */
int nlm_fd_bond_bytes_size(nlm_ctx* ctx, int64_t* size);
int nlm_fd_get_bond_bytes(nlm_ctx* ctx, uint8_t* bytes);
/* Fixity bits per node (Mesh::d_fix, mesh.h:194-200). */
int nlm_fd_get_fixity(nlm_ctx* ctx, uint8_t* fix);
/* Model time and step count (FDModel::d_time, d_n). */
int nlm_fd_get_time(nlm_ctx* ctx, double* time, int64_t* step);
int nlm_fd_set_time(nlm_ctx* ctx, double time, int64_t step);

/* integrate() prologue (fDModel.cpp:318-333): displacement and force BCs at
 * the current time, then the internal forces (bond bits updated). */
int nlm_fd_init(nlm_ctx* ctx);
/* nsteps iterations of the integrate() loop (fDModel.cpp:344-373) with the
 * configured scheme: update, BCs at t+dt, internal forces, crack activation.
 * On return u, v, f, bond bits, time and step equal the reference's state
 * after the same iterations. */
int nlm_fd_step(nlm_ctx* ctx, int nsteps);
/* computeForces (fDModel.cpp:528-541): f += internal force of the current u
 * (force BCs are not added; bond bits updated). */
int nlm_fd_compute_forces(nlm_ctx* ctx);
/* Damage and energy fields of computePostProcFields (fDModel.cpp:652-778):
 * Z_i = max_j |S|/Sc(r), phi_i = 1 - sum unbroken V_j / sum V_j (float like
 * the reference), strain energy of node i.  Any pointer may be NULL. */
int nlm_fd_postproc(nlm_ctx* ctx, float* Z, float* phi, double* energy);
/* The same fields with no device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as
 * serving loops, harnesses that forbid hipMalloc and code near HIP graph
 * capture.  The device staging is a static array of the library, 16 MiB per
 * GPU, held until the process ends: at most NLM_NOALLOC_MAX_NODES nodes
 * (larger models get NLM_ERR_INVALID; use nlm_fd_postproc).  Results equal
 * nlm_fd_postproc's bit for bit.  Every context on one GPU shares that
 * GPU's staging array, so calls that overlap on one GPU (other host
 * threads' contexts) race. */
#define NLM_NOALLOC_MAX_NODES 1048576
int nlm_fd_postproc_noalloc(nlm_ctx* ctx, float* Z, float* phi,
                            double* energy);
/* External force density of every node at time t, 3 components per node in
 * reference order (z = 0 in 2D): FLoading::apply on a zero force vector
 * (fLoading.cpp:117-279), which computePostProcFields takes for the work
 * done u . f_ext (fDModel.cpp:655-662, 734-735). */
int nlm_fd_external_force(nlm_ctx* ctx, double t, double* f_ext);
/* Reaction force of computeForce (fDModel.cpp:548-580): for each node, the
 * sum of V_i F_ij over its bonds for which is_reaction_force(i, j) holds
 * (fDModel.cpp:585-605: node in max_y - 1.5 horizon < y < max_y - horizon,
 * bond crossing the line y = max_y - horizon + 1e-6 between min_x and max_x,
 * tested with util::geometry::doLinesIntersect as the reference writes it,
 * utilGeom.cpp:281-326), at the current state.  box = {min_x, max_x, max_y}
 * of the mesh bounding box (fe::Mesh::computeBBox), horizon = the model
 * deck's.  R: 3 per node (reference order), R_len: |R_i| (the per-node
 * Total_Reaction_Force).  Either output may be NULL.  Bond-based only. */
int nlm_fd_reaction_force(nlm_ctx* ctx, const double* box, double horizon,
                          double* R, double* R_len);

/* State-based material: dilatation theta_i of the last force evaluation and
 * the weighted volume m_i (VolumeCorrection::weightedVolume,
 * volumeCorrection.cpp:51-67), reference node order; either may be NULL. */
int nlm_fd_get_node_fields(nlm_ctx* ctx, double* theta, double* mw);
/* State-based material: the output fields of QuasiStaticModel::
 * computeForces(true) (QuasiStaticModel.cpp:206-229) at the current state,
 * reference node order.  strain_energy (in/out, float like the reference's
 * vector): the strain energy of every unbroken bond (ElasticState::getBondEF,
 * ElasticState.cpp:250-262) is added to it, as the reference adds at every
 * load step without clearing; strain, stress: 6 per node, (xx, yy, zz, yz,
 * xz, xy) = entries (0,0), (1,1), (2,2), (1,2), (0,2), (0,1) of
 * ElasticState::getStrain / getStress (:336-345, :432-471), the layout the
 * reference's .vtu writer uses (vtkWriter.cpp:213-233).  Any pointer may be
 * NULL. */
int nlm_fd_es_fields(nlm_ctx* ctx, float* strain_energy, double* strain,
                     double* stress);

/* Measurement: runs nlm_fd_step(nsteps) bracketed by HIP events recorded on
 * the context's stream; returns the elapsed milliseconds of the whole call
 * and the mean duration of the fused step kernel. */
int nlm_fd_timed_steps(nlm_ctx* ctx, int nsteps, double* ms_total,
                       double* ms_per_step);
/* nlm_fd_step with the fused steps launched from a captured HIP graph of
 * `chunk` steps (even) that is replayed; the step times advance on the
 * device with the host's accumulation (t += dt).  Central difference, no
 * pending timed cracks.  The graph is re-captured when chunk or the buffer
 * parity changes. */
int nlm_fd_step_graph(nlm_ctx* ctx, int nsteps, int chunk);
/* Timed variant: elapsed milliseconds of the call and the mean time of one
 * graph-launched step. */
int nlm_fd_timed_steps_graph(nlm_ctx* ctx, int nsteps, int chunk,
                             double* ms_total, double* ms_per_step);
/* Autotuner (no counterpart in the reference): times the fused step kernel
 * for each neighbor-slot chunk width (2, 4, 8 slots per wavefront iteration)
 * on the current state without changing it, and keeps the fastest.  Returns
 * the choice and its time per step in microseconds.  The FP64 exact mode
 * always uses 4. */
int nlm_fd_autotune(nlm_ctx* ctx, int* chosen, double* best_us);
/* Sets the chunk width by hand (2, 4 or 8). */
int nlm_fd_set_unroll(nlm_ctx* ctx, int unroll);
/* Joint autotuner (no counterpart in the reference) of the 2D bond-based
 * LDS-staged step kernel: for each tile (nodes per workgroup and halo block:
 * 128, 256, 512) the halos are rebuilt and each chunk width timed as in
 * nlm_fd_autotune; the fastest pair is kept.  us (optional, 9 entries,
 * tile-major: us[3 t + k] for tile {128, 256, 512}[t] and chunk {2, 4, 8}[k])
 * receives the times per step in microseconds, 0 where a pair was not run
 * (a tile whose halo exceeds the 64 KiB LDS budget; chunks other than 4 in
 * the FP64 exact mode).  Models without the staged kernel tune the chunk
 * width alone.  The tile changes the data movement only: FP64 modes give
 * the same bits at every tile; the FP32-force mode's block-relative FP32
 * positions take their origin from the tile, so its rounding (within that
 * mode's tolerance) depends on it. */
int nlm_fd_autotune_tiles(nlm_ctx* ctx, int* tile, int* chunk, double* us);
/* Sets the staged kernel's tile by hand (128, 256 or 512). */
int nlm_fd_set_tile(nlm_ctx* ctx, int tile);
/* Number of directed bonds of the configured model. */
int nlm_fd_num_bonds(nlm_ctx* ctx, int64_t* nnz);
/* Name of the fused step kernel launched for the configured model. */
const char* nlm_fd_kernel_name(nlm_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* Quasi-static model: QuasiStaticModel<T> (src/model/quasistatic/)          */
/* ------------------------------------------------------------------------ */

/* The quasi-static solver runs on a context set up with nlm_fd_setup (2D;
 * bond-based material, or the state-based material without bond breaking
 * with the finite-difference tangent) and its displacement / force BC sets.  Load step n
 * (the context's step count plus one, then advanced) is at t = n dt as
 * QuasiStaticModel::solver sets it: FLoading gives the body force b(t), ULoading the
 * prescribed displacements, and Newton iterations solve R(u) = f_int(u) + b
 * = 0 on the free dofs (QuasiStaticModel::solver, QuasiStaticModel.cpp:
 * 527-592; computeResidual :597-646).  The tangent is the analytic derivative
 * of the bond force (on the neighbor sparsity, 2x2 blocks); each Newton
 * system is solved by BiCGSTAB to a relative residual. */
#define NLM_QS_TANGENT_FULL 0      /* exact derivative, all 2x2 entries */
#define NLM_QS_TANGENT_COMPONENT 1 /* same-component entries only, the
                                      entries the reference's central-
                                      difference Jacobian keeps (:436-440) */
#define NLM_QS_TANGENT_FD 2        /* the reference's Jacobian itself:
                                      central differences of the force with
                                      eps = perturbation * h, same-component
                                      entries on the neighbor sparsity
                                      (assembly_jacobian_matrix_part,
                                      :363-454); the state-based material's
                                      tangent */
#define NLM_QS_PRECOND_NONE 0
#define NLM_QS_PRECOND_JACOBI 1  /* point Jacobi */
#define NLM_QS_PRECOND_BLOCKS 2  /* sparse approximate inverse with 2x2
                                    blocks on the neighbor sparsity, set by
                                    nlm_qs_set_precond or predicted by the
                                    graph network (nlm_gnn_*) */
#define NLM_QS_PRECOND_COARSE 16 /* flag: add the two-level coarse correction
                                    R A_c^-1 R^T (aggregates of 256 nodes,
                                    piecewise constant; A_c = R^T A_ff R
                                    inverted densely, at most 4096
                                    aggregates) to JACOBI or BLOCKS */

typedef struct {
  int nsteps;         /* load steps (ModelDeck Time_Steps) */
  double dt;          /* load-step increment of t (Final_Time / Time_Steps) */
  double tol_abs;     /* Newton stops when ||R||_2 <= tol_abs ... */
  double tol_rel;     /* ... or ||R||_2 <= tol_rel ||R_0|| (0: unused) */
  int max_newton;     /* SolverDeck Max_Iteration */
  int tangent;        /* NLM_QS_TANGENT_* */
  int precond;        /* NLM_QS_PRECOND_* */
  double lin_tol;     /* BiCGSTAB: ||r|| <= lin_tol ||b|| */
  int lin_max_iter;
  int predictor;      /* 1: linearized first iteration with the prescribed
                         increment as Dirichlet data (no force evaluation at
                         the jumped state); 0: the reference's order */
  int gnn_refresh;    /* preconditioner construction (network prediction,
                         coarse operator): 1 at every Newton iteration, 0 once
                         per load step, 2 once per call (the first load
                         step's tangent serves the whole call) */
  double perturbation; /* NLM_QS_TANGENT_FD: eps = perturbation * h
                          (SolverDeck Perturbation); <= 0 keeps the last
                          value (initially 1e-7) */
} nlm_qs_options;

typedef struct {
  int load_steps;
  int newton_iters;
  int linear_iters;
  int linear_failures;   /* solves that hit lin_max_iter or broke down */
  double residual;       /* ||R||_2 after the last load step */
  double residual0;      /* ||R_0|| of the last load step */
  double ms_total;
  double ms_linear;      /* BiCGSTAB iterations */
  double ms_precond;     /* preconditioner construction (incl. network) */
  double ms_tangent;     /* tangent assembly */
} nlm_qs_stats;

int nlm_qs_solve(nlm_ctx* ctx, const nlm_qs_options* opt, nlm_qs_stats* st);
/* Tangent A = -df_int/du at the current u and bond state (no force
 * evaluation); tangent = NLM_QS_TANGENT_*. */
int nlm_qs_assemble(nlm_ctx* ctx, int tangent);
/* The assembled blocks: blocks[4 k .. 4 k + 3] = A_{i,j} (row-major 2x2) for
 * directed bond k = (i, j) in the reference CSR order, diag[4 i ..] = A_ii.
 * Either pointer may be NULL. */
int nlm_qs_get_tangent(nlm_ctx* ctx, double* blocks, double* diag);
/* y = A_ff x on the free dofs (x, y AoS double[3n]; entries of prescribed
 * dofs of y are 0 and those of x are ignored). */
int nlm_qs_apply(nlm_ctx* ctx, const double* x, double* y);
/* Preconditioner blocks in the same layout as nlm_qs_get_tangent (stored
 * in FP32 on the device). */
int nlm_qs_set_precond(nlm_ctx* ctx, const double* blocks, const double* diag);
int nlm_qs_get_precond(nlm_ctx* ctx, double* blocks, double* diag);
/* z = R A_c^-1 R^T r: the coarse correction of the two-level preconditioner
 * on the assembled tangent (r, z AoS double[3n]). */
int nlm_qs_coarse_apply(nlm_ctx* ctx, const double* r, double* z);
/* BiCGSTAB launch mode (no counterpart in the reference): on (default),
 * each batch of 8 iterations is one replay of a captured HIP graph; off,
 * the same kernels are launched one by one.  Results are identical.  The
 * environment variable NLM_QS_GRAPH=0 makes off the default of new
 * contexts. */
int nlm_qs_set_graph(nlm_ctx* ctx, int on);
/* BiCGSTAB configuration (no counterpart in the reference): graph as
 * nlm_qs_set_graph; bond_scalar (default on): the products with the
 * analytic bond tangent (NLM_QS_TANGENT_FULL / _COMPONENT) read one scalar
 * per bond, g with -B_ij = g dx dx^T (dx the reference bond vector), and
 * form the block from the positions, instead of reading the 2x2 block --
 * 12 bytes per bond slot instead of 36.  The two forms agree to rounding
 * (each block entry within a few units in the last place); graph replays
 * and kernel-by-kernel launches of one form agree bit for bit. */
int nlm_qs_set_launch(nlm_ctx* ctx, int graph, int bond_scalar);
/* Solves A_ff x = b with the assembled tangent (b, x AoS double[3n]). */
int nlm_qs_linear_solve(nlm_ctx* ctx, const double* b, double* x,
                        int precond, double tol, int max_iter, int* iters,
                        double* relres);

/* ------------------------------------------------------------------------ */
/* Learned preconditioner (no counterpart in the reference)                  */
/* ------------------------------------------------------------------------ */

/* A graph network over the neighbor graph predicts the blocks of a sparse
 * approximate inverse M ~ A_ff^-1 on the tangent's sparsity (one 2x2 block
 * per directed bond and per node).  Inputs: the tangent blocks scaled by the
 * nodes' diagonal magnitudes, bond geometry, fixity.  Four message-passing
 * layers of width 64: m_ij = relu(g_ij Wc + h_i Wa + h_j Wb + bm), g <- m,
 * h_i <- relu(h_i Wd + mean_j(m_ij) Wf + bu); linear decoders give the
 * edge and node blocks.  Weights are BF16 copies of FP32 master weights;
 * every layer product runs on the matrix cores (BF16 inputs, FP32
 * accumulation), forward and backward.  Training minimizes
 * sum_k ||M r_k - e_k||^2 / ||e_k||^2 over residual / error pairs of the
 * tangent (A e_k = r_k) with Adam. */
#define NLM_GNN_LAYERS 4
#define NLM_GNN_WIDTH 64
/* Number of FP32 parameters of the network. */
int64_t nlm_gnn_num_params(void);
/* Attaches a network with deterministic initial weights to the context
 * (after nlm_fd_setup); NLM_QS_PRECOND_BLOCKS then uses its prediction. */
int nlm_gnn_create(nlm_ctx* ctx, uint64_t seed);
int nlm_gnn_get_weights(nlm_ctx* ctx, float* w);
int nlm_gnn_set_weights(nlm_ctx* ctx, const float* w);
/* Predicts the preconditioner blocks from the assembled tangent. */
int nlm_gnn_predict(nlm_ctx* ctx);
/* Loss and gradient for K pairs (r_k, e_k), AoS double[K][3n] each, on the
 * assembled tangent; the gradient is added to the context's accumulator and
 * optionally copied to grad (FP32, nlm_gnn_num_params entries). */
int nlm_gnn_loss_grad(nlm_ctx* ctx, int K, const double* r, const double* e,
                      double* loss, float* grad);
/* Adam step with the accumulated gradient (then cleared). */
int nlm_gnn_adam_step(nlm_ctx* ctx, double lr);
/* Clears the gradient accumulator (after loss evaluations that are not to
 * be stepped on). */
int nlm_gnn_zero_grad(nlm_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* NLMECH_AMD_H */
