/* hemocell_gpu.h -- C ABI of libhcgpu.so, the MI355X (gfx950) implementation of
 * HemoCell's IB-LBM time step.
 *
 * The reference drives the step from C++ (hemo::HemoCell::iterate(),
 * core/hemoCell.cpp:299-376) over Palabos lattices and per-block particle
 * vectors.  This header is the drop-in boundary that a host (the hemo::HemoCell
 * C++ mirror in hemocell_amd/csrc/hemocell.h, a ctypes/cgo/JNI binding, ...)
 * binds to.  Plain pointers and sizes only; all arrays are host memory owned by
 * the caller and copied at the call.  Every call returns hc_status; the text of
 * the last error is available from hc_last_error().
 *
 * Units: lattice units everywhere (the SI -> lattice conversion of
 * mechanics/constantConversion.cpp:36-90 is host-side: hc_lbm_base_parameters(),
 * hc_model_constants()).
 * Lattice node index n = z + nz*(y + ny*x) (patch/palabos.patch:244-250).
 * D3Q19 population order is Palabos's: c_0 = 0, c_1..9 =
 * (-1,0,0),(0,-1,0),(0,0,-1),(-1,-1,0),(-1,1,0),(-1,0,-1),(-1,0,1),(0,-1,-1),(0,-1,1),
 * c_{i+9} = -c_i.  Populations are exchanged shifted by the lattice weights
 * (f_i - t_i).
 */
#ifndef HEMOCELL_GPU_H
#define HEMOCELL_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hc_ctx hc_ctx;

typedef enum {
  HC_OK = 0,
  HC_INVALID_ARG = 1, /* bad size, index, order of calls */
  HC_OOM = 2,         /* device allocation failed */
  HC_DEVICE = 3,      /* HIP runtime / kernel error */
  HC_NUMERIC = 4,     /* NaN or Inf detected by hc_check_finite() */
  HC_STATE = 5        /* call not valid in the current state */
} hc_status;

typedef enum {
  HC_FP64 = 0,          /* FP64 populations and collision: the correctness mode */
  HC_FP32_SHIFTED = 1   /* FP32 populations (f_i - t_i) and FP32 collision; vertex
                           state, IBM weights and mechanics stay FP64 */
} hc_precision;

/* Node types of the lattice mask (one uint8 per node). */
enum {
  HC_NODE_FLUID = 0,      /* GuoExternalForceBGKdynamics */
  HC_NODE_BOUNCEBACK = 1, /* Palabos BounceBack (full-way, on-node); isBoundary() */
  HC_NODE_VELOCITY = 2    /* 2 + face: local regularized velocity boundary on domain
                             face 0..5 = x-,x+,y-,y+,z-,z+ (helper/hemocellInit.hh:58-92) */
};

/* Cell mechanics models (the mechanics/ plugins of the reference). */
enum {
  HC_MODEL_RBC = 0, /* RbcHighOrderModel  mechanics/rbcHighOrderModel.cpp:38-207 */
  HC_MODEL_PLT = 1, /* PltSimpleModel     mechanics/pltSimpleModel.cpp:44-208 */
  HC_MODEL_WBC = 2  /* WbcHighOrderModel  mechanics/wbcHighOrderModel.cpp:42-225 */
};

/* Mesh shapes (config/constant_defaults.h:80-84). */
enum {
  HC_SHAPE_WBC_SPHERE = 0,
  HC_SHAPE_RBC_FROM_SPHERE = 1,
  HC_SHAPE_ELLIPSOID_FROM_SPHERE = 6
};

/* Stiffness and geometry constants of one cell type, lattice units (what the
 * reference's CellMechanics::calculate_k* return, mechanics/cellMechanics.h:50-78,
 * wbcHighOrderModel.cpp:243-262). */
typedef struct {
  double k_link, k_bend, k_volume, k_area, eta_m;
  double k_inner_rigid, k_cytoskeleton, core_radius, radius; /* WBC only */
} hc_model_params;

/* SI inputs (dx [m], dt [s] (< 0: choose dt for tau = 1), nu_p [m^2/s], rho_p [kg/m^3],
 * kBT_p [J]) and the derived lattice quantities. */
typedef struct {
  double dx, dt, nu_p, rho_p, kBT_p;          /* in (dt is replaced when < 0) */
  double tau, nu_lbm, dm, df, f_limit, kBT_lbm; /* out */
} hc_units;

/* Material constants of a cell type as in its <MaterialModel> XML (SI where dimensional). */
typedef struct {
  double kLink, kBend, kVolume, kArea, eta_m;
  double kInnerRigid, kCytoskeleton, coreRadius, radius; /* WBC only */
} hc_material;

/* Replaces: Parameters::lbm_base_parameters (mechanics/constantConversion.cpp:36-59). Host only. */
hc_status hc_lbm_base_parameters(hc_units* u);
/* Replaces: CellMechanics::calculate_kLink/kBend/kVolume/kArea/etaM (mechanics/cellMechanics.h:50-78)
 * and WbcHighOrderModel::calculate_* (mechanics/wbcHighOrderModel.cpp:243-262). Host only. */
hc_status hc_model_constants(int model, const hc_material* m, int n_triangles, const hc_units* u,
                             hc_model_params* out);

/* ---- context -------------------------------------------------------------
 * Replaces: hemo::HemoCell::HemoCell (core/hemoCell.cpp:69-95) ownership of the
 * lattice and cell fields.  One context = one HIP stream on `device`. */
hc_status hc_create(int device, hc_ctx** out);
void hc_destroy(hc_ctx* ctx);
const char* hc_last_error(const hc_ctx* ctx);

/* ---- lattice ---------------------------------------------------------------
 * Replaces: new MultiBlockLattice3D<T,ForcedD3Q19>(..., new GuoExternalForceBGKdynamics(1/tau))
 * (core/hemoCell.cpp:455-469), defineDynamics(BounceBack) (helper/genericFunctions.cpp:144,154),
 * periodicity().toggle (core/hemoCell.cpp:162-176).  node_type may be NULL (all fluid). */
hc_status hc_lattice_init(hc_ctx* ctx, int nx, int ny, int nz, const uint8_t* node_type,
                          double tau, const int periodic[3], int precision);
/* Replaces: setBoundaryVelocity on a face (helper/hemocellInit.hh:75-77). */
hc_status hc_set_wall_velocity(hc_ctx* ctx, int face, const double u[3]);
/* Replaces: the driver's setExternalVector(lattice, bbox, force, g) after every
 * iterate (examples/pipeflow/pipeflow.cpp:144-146) -- a uniform body force. */
hc_status hc_set_body_force(hc_ctx* ctx, const double g[3]);
/* Replaces: HemoCell::latticeEquilibrium / initializeAtEquilibrium (core/hemoCell.cpp:129-133). */
hc_status hc_init_equilibrium(hc_ctx* ctx, double rho, const double u[3]);

/* ---- cell types and cells ------------------------------------------------
 * Replaces: HemoCell::addCellType<Model>(name, shape) (hemocell.h:122-128) with the
 * mesh built by hc_build_mesh and the constants of CommonCellConstants
 * (mechanics/commonCellConstants.cpp:70-409) derived here from the rest mesh. */
hc_status hc_build_mesh(int shape, double radius_lu, int min_triangles, double aspect_ratio,
                        int* nv, int* nt, double* vertices /* nullable, nv*3 */,
                        int32_t* triangles /* nullable, nt*3 */);
hc_status hc_add_cell_type(hc_ctx* ctx, int model, int nv, int nt, const int32_t* triangles,
                           const double* rest_vertices, int n_inner, const int32_t* inner_edges,
                           const hc_model_params* params, int timescale, int* type_id);
/* Replaces: readPositionsBloodCellField3D placement (io/readPositionsBloodCells.cpp:120-170):
 * vertex positions (lattice units, unwrapped) of ncells cells of one type, nv*3 per cell.
 * cell_ids may be NULL (sequential ids). */
hc_status hc_add_cells(hc_ctx* ctx, int type_id, int ncells, const double* xyz,
                       const int32_t* cell_ids);

/* ---- step controls -------------------------------------------------------
 * Replaces: setParticleVelocityUpdateTimeScaleSeparation, setMaterialTimeScaleSeparation,
 * setRepulsion (core/hemoCell.cpp:380-426) and param::f_limit (constantConversion.cpp:55-57). */
hc_status hc_set_velocity_timescale(hc_ctx* ctx, int timescale);
hc_status hc_set_material_timescale(hc_ctx* ctx, int type_id, int timescale);
hc_status hc_set_force_limit(hc_ctx* ctx, double f_limit);
/* The reference spreads in particle order (core/hemoCellParticleField.cpp:841-863), so
 * its runs repeat bit for bit; the GPU spread adds with atomics in no fixed order.  With
 * on != 0 (FP64 form only; needs a force limit, and the lattice), every force contribution
 * is rounded to a multiple of 2^(e-49), 2^e <= f_limit < 2^(e+1), before it is added: the
 * sums are then exact up to 8 force limits per node, so repeated runs, a run alone or in a
 * batch, and a run resumed from a checkpoint agree bit for bit.  A vertex force at the
 * limit keeps 49 bits; smaller ones lose relatively more (absolute rounding 2^(e-50)). */
hc_status hc_set_exact_spread(hc_ctx* ctx, int on);
hc_status hc_set_repulsion(hc_ctx* ctx, double k_rep, double cutoff_lu, int timescale);
/* Replaces: HemoCell::enableBoundaryParticles (core/hemoCell.cpp:428-436) with
 * populateBoundaryParticles / applyBoundaryRepulsionForce
 * (core/hemoCellParticleField.cpp:865-918): bounce-back nodes next to a non-bounce-back
 * node repel vertices of the 3x3x3 node neighbourhood of their bin closer than the
 * cut-off (lattice units) with k (r_c / r) r_hat, every `timescale` steps, before the
 * spread; the force is recomputed at each application and kept between. */
hc_status hc_set_boundary_repulsion(hc_ctx* ctx, double k_rep, double cutoff_lu, int timescale);
/* Replaces: HemoCell::setSystemPeriodicity after the lattice exists (core/hemoCell.cpp:
 * 168-180, which toggles the lattice's and the particle field's periodicity): the periodic
 * flags of the three axes; positions wrap and populations stream across the periodic faces
 * from the next step on. */
hc_status hc_set_periodicity(hc_ctx* ctx, const int periodic[3]);
/* Replaces: LeesEdwardsBC (helper/leesEdwardsBC.h:301-364, as set up in
 * cases/leesEdwards/leesEdwards.cpp:43-44,73): after every collide-stream the nodes of the
 * planes z = nz-1 and z = 0 relax once more towards the equilibrium of the momentum
 * (-+ (nz-1) shear_rate / 2, 0, 0), and their five populations that entered across the
 * plane are interpolated from the plane at x +- d, d = fmod(displacement_per_step x
 * iteration, nx).  The reference passes shear_rate x dt (seconds) as the displacement per
 * step.  The lattice must be periodic in x and z. */
hc_status hc_set_lees_edwards(hc_ctx* ctx, double shear_rate, double displacement_per_step);
/* Replaces: the interior viscosity of a cell type, HemoCellField (core/hemoCellField.cpp:
 * 99-110, INTERIOR_VISCOSITY build): bulk fluid nodes inside the membranes of type_id relax
 * with tau_inner (the reference's viscosityRatio (tau - 1/2) + 1/2) instead of tau.
 * tau_inner <= 0 turns it off for the type.  Type ids 0..14. */
hc_status hc_set_interior_viscosity(hc_ctx* ctx, int type_id, double tau_inner);
/* Replaces: HemoCell::setInteriorViscosityTimeScaleSeperation (core/hemoCell.cpp:404-410).
 * After the constitutive model of every step whose iteration is a multiple of entire_grid,
 * all classes are reset and each cell's inner nodes found by ray casting
 * (findInternalParticleGridPoints, core/hemoCellParticleField.cpp:776-808); of every
 * multiple of membrane, the support nodes of each vertex are re-classified against the
 * membrane normal (internalGridPointsMembrane, :745-774).  Both must be multiples of the
 * particle-velocity timescale (hc_step checks, as sanityCheck, hemoCell.cpp:614-620). */
hc_status hc_set_interior_viscosity_timescales(hc_ctx* ctx, int membrane, int entire_grid);
/* Replaces: the interiorViscosityField (helper/interiorViscosity.cpp:144-163), read for the
 * InteriorViscosity output and the checkpoint: per node (z + nz (y + ny x)) 0 or the
 * tau_inner of the type whose membrane holds it.  The setter restores such a field
 * (InteriorViscosityHelper::restore, interiorViscosity.cpp:107-133): every value must be 0
 * or the tau_inner of a type with interior viscosity. */
hc_status hc_get_interior_viscosity(hc_ctx* ctx, double* tau);
hc_status hc_set_interior_viscosity_field(hc_ctx* ctx, const double* tau);
/* Replaces: HemoCellStretch::applyForce (helper/hemoCellStretch.cpp:63-110): an extra
 * force added to the listed vertices (global vertex ids) before every step's spread. */
hc_status hc_set_extra_vertex_force(hc_ctx* ctx, int n, const int64_t* vertex_ids,
                                    const double* forces);

/* ---- the step --------------------------------------------------------------
 * Replaces: HemoCell::iterate() (core/hemoCell.cpp:299-376), nsteps times. Asynchronous
 * on the context stream; errors of the launches are reported. */
hc_status hc_step(hc_ctx* ctx, int nsteps);
hc_status hc_synchronize(hc_ctx* ctx);
/* nsteps of each of n contexts, stepped together: every step launches the collide-stream of
 * all contexts of one AA parity and boundary kind at once, then their velocity fields, then
 * the cells of one mesh at once, so many small cases cost a few launches per step.  Each
 * context computes what hc_step computes; a context whose step is not the plain fused one
 * (repulsion, interior viscosity, Lees-Edwards planes, a timescale above one) is stepped by
 * hc_step instead.  All contexts share one device and precision.  Asynchronous: each
 * context's stream is ordered after the batch.  No counterpart in the reference (one
 * HemoCell per process); used by the serving loop to batch independent cases. */
hc_status hc_step_batch(hc_ctx* const* ctxs, int n, int nsteps);

/* ---- request loop for small single-cell cases ------------------------------
 * No counterpart in the reference (one HemoCell per process, one case per run): a server
 * keeps each submitted case in its own context, admits queued cases in arrival order while
 * fewer than max_active run and their device footprints fit hbm_budget_bytes, and advances
 * every running case together (hc_step_batch) by a chunk per poll.  A shear case is a cubic
 * box with velocity walls at +-wall_u (x velocity) on its z faces, periodic in x and y
 * (helper/hemocellInit.hh:58-92); a stretch case has walls at rest on every face and pulls
 * the stretch_n lowest and highest vertices in x with -+stretch_force / stretch_n
 * (helper/hemoCellStretch.cpp:30-110).  The cell (a mesh from hc_build_mesh with constants
 * from hc_model_constants) is centred in the box; lattice units throughout. */
typedef struct hc_server hc_server;
enum { HC_CASE_SHEAR = 0, HC_CASE_STRETCH = 1 };
enum { HC_CASE_QUEUED = 0, HC_CASE_RUNNING = 1, HC_CASE_DONE = 2, HC_CASE_FAILED = 3 };
typedef struct {
  int kind;               /* HC_CASE_SHEAR or HC_CASE_STRETCH */
  int edge;               /* nodes per side of the cubic box */
  int64_t steps;
  int precision;          /* hc_precision */
  double tau;
  double f_limit;
  double wall_u;          /* shear: wall speed */
  double stretch_force;   /* stretch: total force per end */
  int stretch_n;          /* stretch: vertices per end */
  int model, nv, nt, n_inner;
  const double* vertices; /* nv*3, copied at submission */
  const int32_t* triangles;
  const int32_t* inner_edges;
  hc_model_params params;
} hc_case_spec;
hc_status hc_server_create(int device, int max_active, double hbm_budget_bytes, hc_server** out);
void hc_server_destroy(hc_server* s);
const char* hc_server_last_error(const hc_server* s);
hc_status hc_case_submit(hc_server* s, const hc_case_spec* spec, int64_t* case_id);
/* retire drained cases, admit, step every running case by min(chunk, its steps left)
 * (asynchronous; one batch per precision and step count); n_running (may be NULL)
 * receives the cases neither done nor failed.  A case whose set-up, batch, stream or
 * result read-back fails is marked HC_CASE_FAILED and its context released; the other
 * cases go on in the same pass, and the first error of the pass is returned
 * (hc_server_last_error). */
hc_status hc_server_poll(hc_server* s, int chunk, int* n_running);
/* Test hook, no counterpart in the reference: makes a case fail from the host at its
 * next poll, as a failed result read-back at retirement (HC_SERVER_INJECT_RETIRE) or as
 * an error reported by its stream (HC_SERVER_INJECT_STREAM_ERROR), so the failure paths
 * of hc_server_poll can be exercised without a device error. */
enum { HC_SERVER_INJECT_NONE = 0, HC_SERVER_INJECT_RETIRE = 1, HC_SERVER_INJECT_STREAM_ERROR = 2 };
hc_status hc_server_inject_failure(hc_server* s, int64_t case_id, int what);
/* HC_CASE_QUEUED / RUNNING / DONE / FAILED, -1 for an unknown id */
int hc_case_status(const hc_server* s, int64_t case_id);
/* a finished case's final vertex positions (nv*3) and stats {volume, area, largest
 * vertex-vertex distance}; either pointer may be NULL */
hc_status hc_case_result(hc_server* s, int64_t case_id, double* x, double* stats);
/* Non-blocking: 1 while work queued on the context is still running, 0 when idle, -1 on
 * error.  No counterpart in the reference (its iterate() is synchronous); used by hosts
 * that drive several contexts from one thread. */
int hc_busy(hc_ctx* ctx);
int64_t hc_iteration(const hc_ctx* ctx);
/* Picks the launch shapes of the per-cell kernels for this context's mesh mix (spread box
 * bytes, threads per RBC cell, side types' position layout, threads per WBC cell) by
 * timing steps_per_trial real steps per candidate, one parameter at a time; the steps
 * advance the simulation.  report (may be NULL) receives the timings and choices as JSON.
 * No counterpart in the reference (its kernels have no launch shape to choose). */
hc_status hc_autotune(hc_ctx* ctx, int steps_per_trial, char* report, int report_len);
/* Replaces: the iteration read back by HemoCellFields::load (core/hemoCellFields.cpp:245);
 * the material, velocity and repulsion timescales count from it. */
hc_status hc_set_iteration(hc_ctx* ctx, int64_t iter);

/* Single stages of iterate(), in its order (parity tests and partial drivers). */
hc_status hc_stage_spread(hc_ctx* ctx);          /* spreadParticleForce  hemoCellParticleField.cpp:841-863 */
hc_status hc_stage_collide_stream(hc_ctx* ctx);  /* lattice->collideAndStream()  hemoCell.cpp:317, with the
                                                    Lees-Edwards planes when set (its data processors) */
hc_status hc_stage_interpolate(hc_ctx* ctx);     /* interpolateFluidVelocity  hemoCellParticleField.cpp:819-839 */
hc_status hc_stage_advance(hc_ctx* ctx);         /* advanceParticles  hemoCellParticleField.cpp:566-588 */
hc_status hc_stage_mechanics(hc_ctx* ctx, int forced); /* applyConstitutiveModel  :633-675 */
hc_status hc_stage_repulsion(hc_ctx* ctx);       /* applyRepulsionForce  :677-743 */
hc_status hc_stage_boundary_repulsion(hc_ctx* ctx); /* applyBoundaryRepulsionForce  :889-918 */
/* pass 0: findInternalParticleGridPoints :776-808; pass 1: internalGridPointsMembrane
 * :745-774 with the support of the current positions */
hc_status hc_stage_interior_viscosity(hc_ctx* ctx, int pass);
hc_status hc_stage_lees_edwards(hc_ctx* ctx); /* LeesEdwardsBCGet/SetPopulations  leesEdwardsBC.h:101-298 */
hc_status hc_stage_end(hc_ctx* ctx);             /* setExternalVector reset + iter++  hemoCell.cpp:369-374 */

/* ---- state access (synchronous) ------------------------------------------ */
int64_t hc_num_nodes(const hc_ctx* ctx);
int64_t hc_num_vertices(const hc_ctx* ctx);
int hc_num_cells(const hc_ctx* ctx);
/* populations, natural (post-stream) layout, f[i*N + n], shifted, as double */
hc_status hc_get_populations(hc_ctx* ctx, double* f);
hc_status hc_set_populations(hc_ctx* ctx, const double* f);
/* IBM part of the force field (3*N, comp*N + n); the body force is not included */
hc_status hc_get_force_field(hc_ctx* ctx, double* F);
/* density and velocity u = j/rho + F/2 (F = body force + IBM force) per node; either may be NULL */
hc_status hc_get_fluid(hc_ctx* ctx, double* rho, double* u /* 3*N, n*3+c */);
/* vertex state, n*3+c; any pointer may be NULL */
hc_status hc_get_vertices(hc_ctx* ctx, double* x, double* v, double* f);
hc_status hc_set_vertices(hc_ctx* ctx, const double* x, const double* v, const double* f);
hc_status hc_get_repulsion_force(hc_ctx* ctx, double* frep);
/* Replaces: the particle removal of HemoCellParticleField::advanceParticles (core/
 * hemoCellParticleField.cpp:566-588) and the incomplete-cell handling of
 * applyConstitutiveModel (:633-675) and deleteIncompleteCells (:510-553).  A vertex whose
 * nearest node is a wall node after its advance is removed: it stops moving and spreads
 * nothing (and counts once in hc_wall_hits).  Its cell is incomplete from then on: no
 * mechanics, zero force on its other vertices at each material step, which keep drifting
 * with the fluid; at an interior-viscosity grid pass the whole cell is removed.
 * removed: one byte per vertex, 1 if removed (get) / to remove (set, for a restart);
 * incomplete: one byte per cell, 1 if any of its vertices was removed. */
hc_status hc_get_removed_vertices(hc_ctx* ctx, uint8_t* removed);
hc_status hc_set_removed_vertices(hc_ctx* ctx, const uint8_t* removed);
hc_status hc_get_incomplete_cells(hc_ctx* ctx, uint8_t* incomplete);
/* the wall repulsion force of every vertex, n*3+c */
hc_status hc_get_boundary_repulsion_force(hc_ctx* ctx, double* fbr);
/* Replaces: the force_repulsion member restored with the particle field
 * (HemoCellParticle::serializeValues_t, core/hemoCellParticle.h:45-63); it is kept between
 * repulsion updates when the repulsion timescale is above one. */
hc_status hc_set_repulsion_force(hc_ctx* ctx, const double* frep);
/* cell type id and cell id of every cell, in cell order (vertex blocks follow the same
 * order); either pointer may be NULL.  Replaces the celltype/cellId members of the
 * particles (core/hemoCellParticle.h:52-58). */
hc_status hc_get_cells(hc_ctx* ctx, int32_t* type_ids, int32_t* cell_ids);
/* nearest-node bin of every vertex, node index z + nz*(y + ny*x) or -1 outside the domain.
 * Replaces: HemoCellParticleField::update_pg (core/hemoCellParticleField.cpp:137-168). */
hc_status hc_get_vertex_bins(hc_ctx* ctx, int64_t* bins);
/* per-cell FP64 volume and area of the current shape (helper/cellInfo.cpp:39-120) */
hc_status hc_get_cell_volume_area(hc_ctx* ctx, double* volume, double* area);
/* Replaces: CellInformationFunctionals::allCellInformation (helper/cellInfo.cpp:170-266): per
 * cell (in cell order) volume, area, centroid[3] and mean velocity[3] (vertex means),
 * bounding box[6] (xmin, xmax, ymin, ymax, zmin, zmax) and the largest vertex-vertex
 * distance ("stretch"); lattice units; any pointer may be NULL. */
hc_status hc_get_cell_stats(hc_ctx* ctx, double* volume, double* area, double* centroid,
                            double* velocity, double* bbox, double* stretch);
/* No counterpart in the reference: hc_get_cell_stats for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).  Allocates nothing, on the device or the host: the statistics are staged
 * in a static device array sized for HC_CELL_STATS_NOALLOC_MAX cells (31.5 MB on each
 * device the library runs on, held until the process ends) and copied into
 * stats[HC_CELL_STATS_QUANTITIES * c + q], q = 0 volume, 1 area, 2-4 centroid, 5-7 mean
 * velocity, 8-13 bounding box (xmin, xmax, ymin, ymax, zmin, zmax), 14 largest
 * vertex-vertex distance, the values hc_get_cell_stats returns.  max_cells: the cells
 * stats holds.  HC_INVALID_ARG for more cells than HC_CELL_STATS_NOALLOC_MAX or than
 * max_cells; HC_STATE when cells were added since the last step or statistics call
 * (their device tables are built by those calls, which allocate).  Calls on contexts of
 * one device share the staging array: calls that overlap on a device race. */
#define HC_CELL_STATS_QUANTITIES 15
#define HC_CELL_STATS_NOALLOC_MAX 262144
hc_status hc_get_cell_stats_noalloc(hc_ctx* ctx, double* stats, int64_t max_cells);
/* Replaces: the per-term force vectors of HemoCellParticleField::separateForceVectors
 * (core/hemoCellParticleField.cpp:590-614) that writeOutput recomputes with
 * applyConstitutiveModel(true): the area, volume, bending, link, membrane-viscosity and
 * inner-link forces of every vertex at the current positions and velocities, n*3+c;
 * any pointer may be NULL.  Their sum is the mechanics force to round-off. */
hc_status hc_get_force_terms(hc_ctx* ctx, double* area, double* volume, double* bending, double* link,
                             double* visc, double* inner_link);
/* Replaces: FluidInfo::calculateVelocityStatistics / calculateForceStatistics
 * (helper/fluidInfo.cpp:102-141).  out = {min, max, mean, count} over the nodes that are
 * not bounce-back of |u| (what 0, u as hc_get_fluid reports it) or of |F| (what 1, the
 * vertex forces spread into a reset field, capped in place as the spread does). */
hc_status hc_get_fluid_stats(hc_ctx* ctx, int what, double out[4]);
/* vertices removed so far because their nearest node became a wall node during their
 * advance (hemoCellParticleField.cpp:566-588); each counts once (see hc_get_removed_vertices) */
int64_t hc_wall_hits(hc_ctx* ctx);
/* HC_NUMERIC if any population or vertex coordinate is NaN/Inf */
hc_status hc_check_finite(hc_ctx* ctx);

/* ---- timing (HIP events on the context stream) --------------------------- */
enum { HC_PROF_SPREAD = 0, HC_PROF_COLLIDE = 1, HC_PROF_INTERP = 2, HC_PROF_MECH = 3,
       HC_PROF_REPULSION = 4, HC_PROF_STEP = 5, HC_PROF_NSTAGES = 6 };
hc_status hc_profile_enable(hc_ctx* ctx, int enable);
/* total milliseconds and number of timed launches of a stage since enabling */
hc_status hc_profile_read(hc_ctx* ctx, int stage, double* total_ms, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* HEMOCELL_GPU_H */
