// hemocell.h -- the hemo::HemoCell host API over libhcgpu (MI355X).
//
// Mirrors the classes and calls a HemoCell driver uses (hemocell.h:68-253,
// config/config.h, mechanics/constantConversion.h, helper/cellInfo.h,
// helper/hemoCellStretch.h, helper/hemocellInit.hh), so that a driver such as
// examples/oneCellShear/oneCellShear.cpp or
// tests/validation/stretch_cell/test_stretch_cell.cpp compiles against it with
// its Palabos lattice replaced by hemo::Lattice.  Every step runs on the GPU
// through the C ABI of hemocell_gpu.h; this layer only reads configuration
// files, places cells and keeps the reference's bookkeeping (iteration counter,
// timescales, per-step external forces).
//
// Differences a driver sees:
//   * hemo::Lattice replaces plb::MultiBlockLattice3D: geometry is a node mask
//     (fluid, bounce-back, velocity wall) set with defineBounceBack() and the
//     boundary-condition helpers below, not with Palabos dynamics objects.
//   * writeOutput() writes the HDF5 and CSV files of the reference's layout
//     (positions, velocities, total, repulsion and per-term forces, ids,
//     triangles; fluid velocity, density and boundary); checkpoints are this
//     build's own binary format behind the reference's checkpoint.xml; load
//     balancing does not exist (one device holds the domain).
//   * Errors that the reference reports with a log line and exit(1) do the same
//     here; errors from the device return as hemo::HemoCellError.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "hemocell_gpu.h"

namespace hemo {

typedef double T;
typedef long plint;
typedef unsigned long pluint;

// ---------------------------------------------------------------- Array -----
// Fixed-size vector with element-wise arithmetic (helper/array.h).
template <typename _Tp, std::size_t _Nm>
struct Array : public std::array<_Tp, _Nm> {
  Array() : std::array<_Tp, _Nm>() { this->fill(_Tp()); }
  Array(std::initializer_list<_Tp> l) : std::array<_Tp, _Nm>() {
    this->fill(_Tp());
    std::size_t i = 0;
    for (const _Tp& v : l)
      if (i < _Nm) (*this)[i++] = v;
  }
  Array& operator+=(const Array& o) {
    for (std::size_t i = 0; i < _Nm; ++i) (*this)[i] += o[i];
    return *this;
  }
  Array& operator-=(const Array& o) {
    for (std::size_t i = 0; i < _Nm; ++i) (*this)[i] -= o[i];
    return *this;
  }
  Array operator+(const Array& o) const { Array r(*this); r += o; return r; }
  Array operator-(const Array& o) const { Array r(*this); r -= o; return r; }
  Array operator*(_Tp s) const { Array r(*this); for (auto& v : r) v *= s; return r; }
  Array operator/(_Tp s) const { Array r(*this); for (auto& v : r) v /= s; return r; }
};
template <typename _Tp, std::size_t _Nm>
Array<_Tp, _Nm> operator*(_Tp s, const Array<_Tp, _Nm>& a) { return a * s; }

// ------------------------------------------------------ constant_defaults ----
enum { WBC_SPHERE = 0, RBC_FROM_SPHERE = 1, ELLIPSOID_FROM_SPHERE = 6 };
enum {
  OUTPUT_POSITION = 1, OUTPUT_FORCE = 2, OUTPUT_FORCE_VOLUME = 3, OUTPUT_FORCE_BENDING = 4, OUTPUT_FORCE_LINK = 5,
  OUTPUT_FORCE_AREA = 6, OUTPUT_FORCE_VISC = 7, OUTPUT_TRIANGLES = 8, OUTPUT_CELL_ID = 9, OUTPUT_VERTEX_ID = 10,
  OUTPUT_VELOCITY = 11, OUTPUT_DENSITY = 12, OUTPUT_FORCE_REPULSION = 13, OUTPUT_SHEAR_RATE = 14,
  OUTPUT_STRAIN_RATE = 15, OUTPUT_SHEAR_STRESS = 16, OUTPUT_BOUNDARY = 17, OUTPUT_OMEGA = 18,
  OUTPUT_CELL_DENSITY = 19, OUTPUT_FORCE_INNER_LINK = 20, OUTPUT_INTERIOR_POINTS = 21
};

class HemoCellError : public std::runtime_error {
 public:
  explicit HemoCellError(const std::string& m) : std::runtime_error(m) {}
};

// File-system helpers (the reference's mkpath / file_exists, io/genericFunctions)
bool exists(const std::string& path);
void mkpath(const std::string& path);

// --------------------------------------------------------------- Config -----
// XML configuration access (config/config.h): cfg["domain"]["dx"].read<T>();
// a missing element throws std::invalid_argument (config/config.cpp:85).
struct XMLNode {
  std::string name, text;
  std::vector<std::unique_ptr<XMLNode>> children;
  const XMLNode* child(const std::string& n) const {
    for (const auto& c : children)
      if (c->name == n) return c.get();
    return nullptr;
  }
};

class XMLElement {
  const XMLNode* orig;

 public:
  explicit XMLElement(const XMLNode* o) : orig(o) {}
  XMLElement operator[](std::string name) const;
  template <typename V>
  V read() const {
    std::stringstream value(orig->text);
    V ret = V();
    if (!(value >> ret)) throw std::invalid_argument("cannot convert the value of XML element " + orig->name);
    return ret;
  }
  std::string getText() const { return orig->text; }
  const XMLNode* getOrig() const { return orig; }
};
template <>
inline bool XMLElement::read<bool>() const {
  std::stringstream value(orig->text);
  std::string w;
  value >> w;
  return w == "1" || w == "true" || w == "True" || w == "TRUE";
}
template <>
inline std::string XMLElement::read<std::string>() const {
  std::stringstream value(orig->text);
  std::string w;
  value >> w;
  return w;
}

class Config {
  std::unique_ptr<XMLNode> doc;

 public:
  bool checkpointed = false;
  explicit Config(std::string paramXmlFilename);
  XMLElement operator[](std::string name) const;
  // the whole document (above <hemocell> or <Checkpoint>)
  XMLElement document() const { return XMLElement(doc.get()); }
  // the <hemocell> element, also inside a checkpoint file
  const XMLNode* hemocellNode() const;
};

// XML text of a node and its children, indented two spaces per level
std::string toXml(const XMLNode& node, int indent = 0);

// ----------------------------------------------------------- Parameters -----
// SI -> lattice conversion (mechanics/constantConversion.{h,cpp}).
class Lattice;
class Parameters {
 public:
  static T dt, dx, dm, df;
  static T nu_p, rho_p, kBT_p;
  static T tau, re;
  static T nu_lbm, u_lbm_max, kBT_lbm, shearrate_lbm;
  static T f_limit;
  static T ef_lbm;
  static T pipe_radius;
  static void lbm_base_parameters(Config& cfg);
  static void lbm_pipe_parameters(Config& cfg, int nY);
  // the radius from the fluid nodes of the lattice's x = 0 plane, sqrt(area / pi)
  // (the reference's overload taking the voxelised flag matrix)
  static void lbm_pipe_parameters(Config& cfg, const Lattice& lattice);
  static void lbm_shear_parameters(Config& cfg, T nx);
  static void printParameters();
  static hc_units units();
};
typedef Parameters param;

// -------------------------------------------------------------- Lattice -----
struct Box3D {
  plint x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  Box3D() = default;
  Box3D(plint x0_, plint x1_, plint y0_, plint y1_, plint z0_, plint z1_)
      : x0(x0_), x1(x1_), y0(y0_), y1(y1_), z0(z0_), z1(z1_) {}
  plint getNx() const { return x1 - x0 + 1; }
  plint getNy() const { return y1 - y0 + 1; }
  plint getNz() const { return z1 - z0 + 1; }
};

class Periodicity {
  bool p[3] = {false, false, false};

 public:
  void toggle(int axis, bool periodic) { p[axis] = periodic; }
  void toggleAll(bool periodic) { p[0] = p[1] = p[2] = periodic; }
  bool get(int axis) const { return p[axis]; }
};

class HemoCell;

// The fluid lattice of a HemoCell run (replaces plb::MultiBlockLattice3D with
// GuoExternalForceBGKdynamics, core/hemoCell.cpp:455-469): its geometry is a
// node mask, its state lives on the GPU once HemoCell binds it.
class Lattice {
 public:
  Lattice(plint nx, plint ny, plint nz);
  plint getNx() const { return nx; }
  plint getNy() const { return ny; }
  plint getNz() const { return nz; }
  Box3D getBoundingBox() const { return Box3D(0, nx - 1, 0, ny - 1, 0, nz - 1); }
  Periodicity& periodicity() { return period; }
  void toggleInternalStatistics(bool) {}
  void initialize() {}
  // Fluid-only step (no particles), as drivers use for warm-up.
  void collideAndStream();
  uint8_t nodeType(plint x, plint y, plint z) const { return mask[index(x, y, z)]; }
  void setNodeType(plint x, plint y, plint z, uint8_t t) { mask[index(x, y, z)] = t; }
  bool isBoundary(plint x, plint y, plint z) const { return mask[index(x, y, z)] != HC_NODE_FLUID; }
  plint index(plint x, plint y, plint z) const { return z + nz * (y + ny * x); }

  T wall_u[6][3] = {{0}};
  std::vector<uint8_t> mask;
  HemoCell* owner = nullptr;  // set when a HemoCell binds the lattice
  // Lees-Edwards planes registered by LeesEdwardsBC::initialize()
  bool leesEdwards = false;
  T leesEdwardsShear = 0;
  double leesEdwardsDisplacement = 0;  // per step

 private:
  plint nx, ny, nz;
  Periodicity period;
};

// Lees-Edwards planes z = 0 and z = nz-1 (helper/leesEdwardsBC.h:301-364): the
// lattice becomes periodic on every axis; after each collide-stream the planes
// are driven at -+ (nz-1) shearRate / 2 and the populations crossing them are
// taken from the plane displaced by fmod(shearRate dt iter, nx) along x (the
// reference's rule: dt as passed, in seconds).  Construct and initialize() it
// before the lattice is first used (addCellType, latticeEquilibrium).
class LeesEdwardsBC {
 public:
  LeesEdwardsBC(Lattice& lattice, T shearRate, T dt, double** hemoCellLEcurrentDisplacement,
                plint dataProcessorLevel = 1);
  void initialize();
  // the host copy of the current displacement (the device derives its own from the iteration)
  void updateLECurDisplacement(unsigned int iter);
  Lattice& lattice;
  plint nx, ny, nz;
  double LEdisplacement;  // per step
  static double LEcurrentDisplacement;
  T dt, topVelocity, bottomVelocity;
  plint dataProcessorLevel;
};

// Full-way bounce-back on every node of the box (defineDynamics(..., new BounceBack),
// helper/genericFunctions.cpp:144,154).
void defineBounceBack(Lattice& lattice, Box3D box);

// createLocalBoundaryCondition3D() (helper/hemocellInit.hh:72-77): marks the
// nodes of a box that lie on the lattice's outer faces as regularized velocity
// nodes; a node on an edge or corner belongs to the first of its faces in the
// order x-, x+, y-, y+, z-, z+.
class OnLatticeBoundaryCondition3D {
 public:
  void setVelocityConditionOnBlockBoundaries(Lattice& lattice);
  void setVelocityConditionOnBlockBoundaries(Lattice& lattice, Box3D box);
};
inline OnLatticeBoundaryCondition3D* createLocalBoundaryCondition3D() { return new OnLatticeBoundaryCondition3D(); }

// Imposed velocity of the velocity nodes in the box (one value per domain face).
void setBoundaryVelocity(Lattice& lattice, Box3D box, Array<T, 3> u);
// A uniform force density on the whole lattice (the driver's body force); the
// next iterate() resets it, as the reference's force reset does.
void setExternalVector(Lattice& lattice, Box3D box, int forceBeginsAt, Array<T, 3> force);
// The reference's Couette set-up (helper/hemocellInit.hh:58-92).
void iniLatticeSquareCouette(Lattice& lattice, plint nx, plint ny, plint nz, OnLatticeBoundaryCondition3D& bc,
                             T shearRate);

// --------------------------------------------------- cell types / models -----
// Model descriptors: addCellType<Model>() selects the device mechanics kernel.
struct RbcHighOrderModel { static constexpr int model = HC_MODEL_RBC; };
struct PltSimpleModel { static constexpr int model = HC_MODEL_PLT; };
struct WbcHighOrderModel { static constexpr int model = HC_MODEL_WBC; };

// Rest-mesh metrics (helper/meshMetrics.h:127-222).
struct MeshMetrics {
  T volume = 0, area = 0, mean_edge = 0;
  T getVolume() const { return volume; }
  T getArea() const { return area; }
  T getSurface() const { return area; }  // meshMetrics.h:90, the total area
};

class HemoCellFields;

// One cell type (core/hemoCellField.h): its material configuration, rest mesh
// and mechanics constants.
class HemoCellField {
 public:
  std::string name;
  int ctype = 0;        // index in registration order
  int type_id = -1;     // libhcgpu cell-type id
  int constructType = 0;
  int numVertex = 0;
  int timescale = 1;
  T minimumDistanceFromSolid = 0.0;
  // interior viscosity: <MaterialModel><enableInteriorViscosity> and the tau inside
  // the membrane, <viscosityRatio> (tau - 1/2) + 1/2
  bool doInteriorViscosity = false;
  T interiorViscosityTau = 1.0;
  std::unique_ptr<Config> materialCfg;
  std::vector<T> rest;                          // nv*3, lattice units, centred
  std::vector<int32_t> triangles;               // nt*3
  std::vector<int32_t> inner_edges;             // ni*2
  std::vector<Array<plint, 3>> triangle_list;
  std::unique_ptr<MeshMetrics> meshmetric;
  hc_model_params params{};
  HemoCellFields& cellFields;
  std::vector<int> outputs{OUTPUT_POSITION};  // HDF5 datasets written by writeOutput
  HemoCellField(HemoCellFields& cf, const std::string& name, int ctype, int constructType, int model);
  void setOutputVariables(const std::vector<int>& o) { outputs = o; }
};

class HemoCellFields {
 public:
  explicit HemoCellFields(HemoCell& hc) : hemocell(hc) {}
  HemoCellField* operator[](const std::string& name);
  HemoCellField* operator[](pluint ctype) { return fields.at(ctype).get(); }
  pluint size() const { return fields.size(); }
  HemoCellField* addCellType(const std::string& name, int constructType, int model);
  HemoCell& hemocell;
  std::vector<std::unique_ptr<HemoCellField>> fields;
  int number_of_cells = 0;
  // cell id of each device cell, in device cell order
  std::vector<int32_t> cell_ids;
  std::vector<int> cell_types;
};

// ------------------------------------------------------------- HemoCell -----
class HemoCell {
 public:
  enum class MPIHandle { Internal, External };
  HemoCell(char* configFileName, int argc, char* argv[]) : HemoCell(configFileName, argc, argv, MPIHandle::Internal) {}
  HemoCell(char* configFileName, int argc, char* argv[], MPIHandle mpi_handle);
  ~HemoCell();
  HemoCell(const HemoCell&) = delete;
  HemoCell& operator=(const HemoCell&) = delete;

  void latticeEquilibrium(T rho, Array<T, 3> vel);
  void initializeCellfield();
  template <class Mechanics>
  void addCellType(std::string name, int constructType) {
    if (!cellfields) initializeCellfield();
    cellfields->addCellType(name, constructType, Mechanics::model);
  }
  void setOutputs(std::string name, std::vector<int> outputs) { (*cellfields)[name]->setOutputVariables(outputs); }
  void setFluidOutputs(std::vector<int> outputs) { fluidOutputs = outputs; }
  void setRepulsion(T repulsionConstant, T repulsionCutoff);
  // core/hemoCell.cpp:428-436: wall repulsion from the boundary particles (cut-off in
  // micrometres), applied every `timestep` steps
  void enableBoundaryParticles(T boundaryRepulsionConstant, T boundaryRepulsionCutoff, unsigned int timestep = 1);
  void setRepulsionTimeScaleSeperation(unsigned int separation);
  void setMaterialTimeScaleSeparation(std::string name, unsigned int separation);
  void setParticleVelocityUpdateTimeScaleSeparation(unsigned int separation);
  // membrane pass every `separation` steps, ray-casting pass every
  // `separation_entire_grid` steps (core/hemoCell.cpp:404-410)
  void setInteriorViscosityTimeScaleSeperation(unsigned int separation, unsigned int separation_entire_grid);
  void setInitialMinimumDistanceFromSolid(std::string name, T distance);
  void setSystemPeriodicity(unsigned int axis, bool bePeriodic);
  void loadParticles();
  // Checkpoints (core/hemoCell.cpp:197-219, core/hemoCellFields.cpp:240-319): the
  // directory <out>/<parameters><checkpointDirectory> (default <out>/checkpoint/)
  // holds checkpoint.xml (the configuration wrapped in <Checkpoint> with
  // <General><Iteration> and <OutDirectory>), lattice.dat (populations) and
  // particleField.dat (cells and vertex state); the previous set is renamed to
  // *.old first.  A run started from checkpoint.xml calls loadCheckPoint() in
  // place of loadParticles() and continues from the saved iteration.
  void loadCheckPoint();
  void saveCheckPoint();
  std::string checkpointDirectory();
  // HDF5 (per cell type and fluid) and CSV output of the current state in the
  // reference's file layout
  void writeOutput();
  std::string outputDirectory();
  void iterate();

  // Device handle (created by the constructor).  Exposed for drivers that
  // read state through the C ABI directly.
  hc_ctx* context() { return ctx; }
  // Bind the lattice to the device (called on first use; idempotent).
  void bindLattice();
  // Pending per-step vertex forces (HemoCellStretch::applyForce).
  void addVertexForces(const std::vector<int64_t>& ids, const std::vector<T>& forces);
  void check(hc_status s, const char* what) const;

  bool outputInSiUnits = true;
  // Lees-Edwards planes on (hemocell.h:146-147 of the reference)
  bool leesEdwardsBC = false;
  double* LEcurrentDisplacement = nullptr;
  bool repulsionEnabled = false;
  bool boundaryRepulsionEnabled = false;
  Lattice* lattice = nullptr;
  Config* cfg = nullptr;
  HemoCellFields* cellfields = nullptr;
  unsigned int iter = 0;
  int precision = HC_FP64;  // HC_FP32_SHIFTED with <hemocell><domain><precision>fp32 or HEMOCELL_PRECISION=fp32

  // body force set through setExternalVector() for the next iterate()
  Array<T, 3> pending_force;
  bool force_pending = false;

 private:
  hc_ctx* ctx = nullptr;
  bool bound = false;
  bool loadParticlesIsCalled = false;
  unsigned int repulsionTimescale = 1;
  T repulsionConstant = 0, repulsionCutoff = 0;
  T boundaryRepulsionConstant = 0, boundaryRepulsionCutoff = 0;
  unsigned int boundaryRepulsionTimescale = 1;
  std::vector<int64_t> extra_ids, pending_ids;
  std::vector<T> extra_f, pending_f;
  bool extra_set = false;
  std::vector<int> fluidOutputs;
  std::string outDir;
};

// ------------------------------------------------------------- fluidInfo -----
// helper/fluidInfo.h: minimum, maximum and mean of the fluid velocity magnitude,
// or of the spread IBM force magnitude, over the nodes that are not bounce-back
// (lattice units).  ncells is the number of nodes counted.
struct FluidStatistics {
  double min = 0, max = 0, avg = 0;
  pluint ncells = 0;
};
class FluidInfo {
 public:
  static FluidStatistics calculateVelocityStatistics(HemoCell* hemocell);
  // spreads the vertex forces into a reset field first (helper/fluidInfo.cpp:117-120)
  static FluidStatistics calculateForceStatistics(HemoCell* hemocell);
};

// --------------------------------------------------------- particleInfo -----
// helper/particleInfo.h: minimum, maximum and mean over all vertices of
// |force + repulsion force| or of |velocity| (lattice units); ncells is the
// number of vertices counted.
struct ParticleStatistics {
  double min = 0, max = 0, avg = 0;
  pluint ncells = 0;
};
class ParticleInfo {
 public:
  static ParticleStatistics calculateForceStatistics(HemoCell* hemocell);
  static ParticleStatistics calculateVelocityStatistics(HemoCell* hemocell);
};

// -------------------------------------------------------------- cellInfo -----
// Per-cell statistics (helper/cellInfo.h): lattice units, keyed by cell id.
struct CellInformation {
  T volume = 0, area = 0, stretch = 0;
  Array<T, 3> position, velocity;
  Array<T, 6> bbox;
  int cellType = 0;
  bool centerLocal = true;
};

class CellInformationFunctionals {
 public:
  static std::map<int, CellInformation> info_per_cell;
  static void calculateCellVolume(HemoCell* hemocell) { fill(hemocell); }
  static void calculateCellArea(HemoCell* hemocell) { fill(hemocell); }
  static void calculateCellPosition(HemoCell* hemocell) { fill(hemocell); }
  static void calculateCellStretch(HemoCell* hemocell) { fill(hemocell); }
  static void calculateCellBoundingBox(HemoCell* hemocell) { fill(hemocell); }
  static void calculateCellVelocity(HemoCell* hemocell) { fill(hemocell); }
  static void calculate_vol_pos_area(HemoCell* hemocell) { fill(hemocell); }
  static void clear_list() { info_per_cell.clear(); }
  static unsigned int getTotalNumberOfCells(HemoCell* hemocell);
  static unsigned int getNumberOfCellsFromType(HemoCell* hemocell, std::string type);

 private:
  static void fill(HemoCell* hemocell);
};

// ----------------------------------------------------------- cell stretch ----
// Stretching force on the n lowest / highest-x vertices of the single cell
// (helper/hemoCellStretch.cpp:30-110).
class HemoCellStretch {
 public:
  HemoCellStretch(HemoCellField& cellfield, unsigned int n_forced_lsps, T external_force);
  void applyForce();
  std::vector<plint> lower_lsps, upper_lsps;
  unsigned int n_forced_lsps = 0;
  T external_force = 0.0;  // per vertex
  T scale = 1.0;

 private:
  HemoCellField& cellfield;
};

// the reference's log streams
inline std::ostream& pcout = std::cout;
inline std::ostream& hlog = std::cout;

}  // namespace hemo
