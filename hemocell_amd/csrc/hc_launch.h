// hc_launch.h -- host-callable launchers of the libhcgpu kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "hc_internal.h"

namespace hcg {

// lattice (hc_lbm.hip); R = float or double
struct CollideArgs {
  double omega;
  double g[3];
  double wall_u[6][3];
  // interior viscosity: per-node class (null when off) and the relaxation rate
  // of each class (class 0 = omega)
  const uint8_t* visc;
  double omega_cls[16];
};
template <typename R>
hipError_t launch_collide_stream(hipStream_t s, R* f, const R* Fcur, R* Fclear, const uint8_t* type,
                                 const LatticeDev& L, const CollideArgs& a, int odd, bool general,
                                 const uint32_t* vel_nodes, int n_vel, bool fuse_vel);
template <typename R>
hipError_t launch_init_equilibrium(hipStream_t s, R* f, const uint8_t* type, const LatticeDev& L,
                                   double rho, const double u[3]);
template <typename R>
hipError_t launch_natural_populations(hipStream_t s, const R* f, double* out, const LatticeDev& L, int phase);
template <typename R>
hipError_t launch_store_populations(hipStream_t s, R* f, const double* in, const LatticeDev& L);
template <typename R>
hipError_t launch_fluid_moments(hipStream_t s, const R* f, const R* Fcur, const uint8_t* type, const LatticeDev& L,
                                int phase, const CollideArgs& a, double* rho, double* u);
// per-block {min, max, sum, count} of |u| (what 0) or |Fstat| (what 1) over the
// non-bounce-back nodes; nblocks * 4 doubles in partials
template <typename R>
hipError_t launch_fluid_stats(hipStream_t s, const R* f, const R* Fcur, const uint8_t* type, const LatticeDev& L,
                              int phase, const CollideArgs& a, const R* Fstat, int what, double* partials,
                              int nblocks);
template <typename R>
hipError_t launch_fill_zero(hipStream_t s, R* p, int64_t n);
template <typename R>
hipError_t launch_to_double(hipStream_t s, const R* in, double* out, int64_t n);

// immersed boundary (hc_ibm.hip)
struct VertexDev {
  double* x;      // 3 components, stride vs
  double* v;
  double* f;
  double* frep;   // cell-cell repulsion force
  double* fbr;    // wall (boundary particle) repulsion force
  int64_t n;      // vertices
  int64_t vs;     // component stride
  // per vertex: bit 0 set once the vertex landed on a wall node and was removed
  // (HemoCellParticleField::advanceParticles); a cell with a removed vertex is
  // incomplete: no mechanics, zero force, its other vertices still drift
  uint8_t* flag;
};
constexpr uint8_t VFLAG_REMOVED = 1;
template <typename R>
hipError_t launch_spread(hipStream_t s, const VertexDev& V, R* Fibm, const uint8_t* type, const LatticeDev& L,
                         double f_limit, int use_repulsion);
template <typename R>
hipError_t launch_spread_cells(hipStream_t s, const VertexDev& V, R* Fibm, const uint8_t* type, const LatticeDev& L,
                               const int64_t* cell_base, const int32_t* cell_nv, int ncells, double f_limit,
                               int use_repulsion, int lds_bytes);
template <typename R>
hipError_t launch_velocity_field(hipStream_t s, const R* f, const R* Fcur, R* uf, const LatticeDev& L, int phase,
                                 bool general, const double g[3]);
template <typename R>
hipError_t launch_interp_advance(hipStream_t s, const VertexDev& V, const R* uf, const uint8_t* type,
                                 const LatticeDev& L, int do_interp, int do_advance, unsigned long long* wall_hits);
hipError_t launch_add_extra_force(hipStream_t s, const VertexDev& V, const int64_t* ids, const double* forces, int n);

// fused per-cell step (hc_cell_body.h): interpolation + advance, mechanics, next spread
enum { CELL_FLAG_INTERP = 1, CELL_FLAG_MECH = 2, CELL_FLAG_SPREAD = 4, CELL_FLAG_REPULSION = 8 };
template <typename R>
hipError_t launch_cell_step(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                            const int64_t* cell_base, const R* uf, const uint8_t* type, const LatticeDev& L,
                            R* Fnext, const double* extra, double* scratch, double f_limit, int flags, int box_bytes,
                            int lds_device, unsigned long long* wall_hits, int block, bool gpos = false);

// all cell types in one persistent launch (hc_cell_*.hip): a queue of (cell,
// type index) entries, the types with most vertices first; flags per type
constexpr int kMaxMixedTypes = 8;
constexpr int kMixedBlock = 256;
struct MixedArgs {
  int flags[kMaxMixedTypes];
  int box_nodes[kMaxMixedTypes];
};
template <typename R>
hipError_t launch_cell_step_mixed(hipStream_t s, const VertexDev& V, const TypeDev* d_types, const TypeDev* h_types,
                                  int ntypes, const int* flags, const int2* queue, int nqueue, unsigned int* head,
                                  int grid, const int64_t* cell_base, const R* uf, const uint8_t* type,
                                  const LatticeDev& L, R* Fnext, const double* extra, double* scratch, double f_limit,
                                  int box_bytes, int lds_device, unsigned long long* wall_hits);

// binning and cell-cell repulsion (hc_repulsion.hip)
struct RepulsionBufs {
  uint32_t *keys_in, *keys_out;  // nearest-node key per vertex, before / after the sort
  int32_t *ids_in, *ids_out;     // vertex ids, sorted alongside
  int2* bins;                    // [start, end) per node into ids_out; all zero between launches
  const int32_t* vcell;          // cell index of each vertex
  void* temp;                    // radix-sort scratch
  size_t temp_bytes;
};
size_t repulsion_temp_bytes(int64_t nv, int64_t N);
hipError_t launch_repulsion(hipStream_t s, const VertexDev& V, const LatticeDev& L, const RepulsionBufs& B,
                            double k_rep, double r_c);
// boundary particles (bounce-back nodes next to a non-bounce-back node) and the
// repulsion they exert on nearby vertices (hc_repulsion.hip)
hipError_t launch_boundary_flags(hipStream_t s, const uint8_t* type, const LatticeDev& L, uint8_t* flags);
hipError_t launch_boundary_repulsion(hipStream_t s, const VertexDev& V, const LatticeDev& L, const uint8_t* flags,
                                     double k, double r_c);
hipError_t launch_vertex_bins(hipStream_t s, const VertexDev& V, const LatticeDev& L, int64_t* out);

// One lattice of a batched launch: its buffers, geometry and collision
// constants, and the first block of its range (blocks of 256 nodes).
struct LatCase {
  void* f;
  const void* Fcur;
  void* Fclr;
  const uint8_t* type;
  LatticeDev L;
  CollideArgs a;
  uint32_t block0;
};
// ncase lattices of one parity and boundary kind in one launch of `blocks` blocks
template <typename R>
hipError_t launch_collide_multi(hipStream_t s, const LatCase* cases, int ncase, uint32_t blocks, int odd,
                                bool general, bool vel);
// the same for the velocity field: uf of each case from its post-stream populations
struct VelCase {
  const void* f;
  const void* Fcur;
  void* uf;
  LatticeDev L;
  double g[3];
  uint32_t block0;
};
template <typename R>
hipError_t launch_velocity_field_multi(hipStream_t s, const VelCase* cases, int ncase, uint32_t blocks, int phase,
                                       bool general);

// one cell of a batched cell-step launch (hc_cell.hip): everything the per-type
// launch passes, for the cell at vertex base `base` of its case
struct CellCase {
  VertexDev V;
  TypeDev T;
  int64_t base;
  const void* uf;
  const uint8_t* type;
  LatticeDev L;
  void* Fnext;
  const double* extra;
  double* scratch;
  double f_limit;
  int flags;
  unsigned long long* wall_hits;
};
// cells of one mesh (T gives the model, vertex count and viscosity for the LDS
// layout, identical over the cases) from several cases in one launch
template <typename R>
hipError_t launch_cell_step_multi(hipStream_t s, const TypeDev& T, const CellCase* cases, int ncells, int box_bytes,
                                  int lds_device, int block);

// Lees-Edwards planes after the collide-stream (hc_lbm.hip): d is the current
// displacement; scratch holds 2 nx ny x 19 populations
template <typename R>
hipError_t launch_lees_edwards(hipStream_t s, R* f, const uint8_t* type, const LatticeDev& L, int phase,
                               const CollideArgs& a, double d, double u_top, double u_bottom, R* scratch);

// interior viscosity (hc_visc.hip)
hipError_t launch_visc_clear(hipStream_t s, uint8_t* visc, int64_t n);
hipError_t launch_visc_grid(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                            const int64_t* cell_base, const LatticeDev& L, const uint8_t* type, int cls,
                            uint8_t* visc);
hipError_t launch_delete_incomplete(hipStream_t s, const VertexDev& V, const int64_t* cell_base, const int32_t* cell_nv,
                                    int ncells);
hipError_t launch_visc_membrane(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells,
                                int ncells, const int64_t* cell_base, const double* xold, const LatticeDev& L,
                                const uint8_t* type, int cls, uint32_t* stamp, uint8_t* visc, bool apply);

// mechanics (hc_mech.hip)
hipError_t launch_mechanics(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                            const int64_t* cell_base, double* bend_scratch, int lds_limit, int term = 63, double* out = nullptr);
hipError_t launch_cell_stats(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                             const int64_t* cell_base, double* out);
// the device address of the statistics staging array of hc_get_cell_stats_noalloc:
// looked up once per device (that device current) at context creation
hipError_t noalloc_scratch_setup(int device);
double* noalloc_scratch(int device);
hipError_t launch_check_finite(hipStream_t s, const double* p, int64_t n, unsigned int* bad);
hipError_t launch_check_finite_f(hipStream_t s, const float* p, int64_t n, unsigned int* bad);

}  // namespace hcg
