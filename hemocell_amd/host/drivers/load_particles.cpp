// load_particles -- places cells from .pos files through hemo::HemoCell and
// writes what landed on the device (the io/readPositionsBloodCells.cpp path).
//
// usage: load_particles <config.xml> <nx> <ny> <nz> <walls> <out.bin> <KIND:name[@min_dist_um]>...
//   walls: none (periodic box) | bb-y (bounce-back planes y = 0 and y = ny-1,
//          periodic x and z) | vel-all (velocity nodes on every face, no periodicity)
//   KIND:  RBC (RBC_FROM_SPHERE, RbcHighOrderModel), PLT (ELLIPSOID_FROM_SPHERE,
//          PltSimpleModel) or WBC (WBC_SPHERE, WbcHighOrderModel); <name>.xml is
//          the material and <name>.pos the positions, as for addCellType(name).
// out.bin: int64 ncells, int64 nvertices, int32 cell id[ncells], int32 cell
//          type[ncells], float64 x[nvertices][3].
// With HEMOCELL_OUTPUT_AFTER=<n> the run steps n times and calls writeOutput()
// (all cell and fluid outputs) before writing out.bin.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "hemocell.h"

using hemo::T;

static int run(int argc, char* argv[]) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: %s <config.xml> <nx> <ny> <nz> <none|bb-y|vel-all> <out.bin> <KIND:name[@um]>...\n",
                 argv[0]);
    return 2;
  }
  const hemo::plint nx = std::atol(argv[2]), ny = std::atol(argv[3]), nz = std::atol(argv[4]);
  const std::string walls = argv[5];
  char* args[] = {argv[0], nullptr};
  hemo::HemoCell hemocell(argv[1], 0, args);
  hemo::param::lbm_base_parameters(*hemocell.cfg);
  hemocell.lattice = new hemo::Lattice(nx, ny, nz);
  hemo::Lattice& L = *hemocell.lattice;
  if (walls == "none") {
    L.periodicity().toggleAll(true);
  } else if (walls == "bb-y") {
    L.periodicity().toggleAll(true);
    L.periodicity().toggle(1, false);
    hemo::defineBounceBack(L, hemo::Box3D(0, nx - 1, 0, 0, 0, nz - 1));
    hemo::defineBounceBack(L, hemo::Box3D(0, nx - 1, ny - 1, ny - 1, 0, nz - 1));
  } else if (walls == "vel-all") {
    L.periodicity().toggleAll(false);
    hemo::OnLatticeBoundaryCondition3D bc;
    bc.setVelocityConditionOnBlockBoundaries(L);
  } else {
    std::fprintf(stderr, "unknown wall set-up %s\n", walls.c_str());
    return 2;
  }
  hemocell.latticeEquilibrium(1.0, hemo::Array<T, 3>({0., 0., 0.}));
  hemocell.initializeCellfield();
  for (int a = 7; a < argc; ++a) {
    std::string spec = argv[a];
    const size_t colon = spec.find(':');
    if (colon == std::string::npos) throw hemo::HemoCellError("cell type spec needs KIND:name");
    const std::string kind = spec.substr(0, colon);
    std::string name = spec.substr(colon + 1);
    T dist = -1;
    const size_t at = name.find('@');
    if (at != std::string::npos) {
      dist = std::atof(name.substr(at + 1).c_str());
      name = name.substr(0, at);
    }
    if (kind == "RBC") hemocell.addCellType<hemo::RbcHighOrderModel>(name, hemo::RBC_FROM_SPHERE);
    else if (kind == "PLT") hemocell.addCellType<hemo::PltSimpleModel>(name, hemo::ELLIPSOID_FROM_SPHERE);
    else if (kind == "WBC") hemocell.addCellType<hemo::WbcHighOrderModel>(name, hemo::WBC_SPHERE);
    else throw hemo::HemoCellError("unknown cell kind " + kind);
    if (dist >= 0) hemocell.setInitialMinimumDistanceFromSolid(name, dist);
  }
  hemocell.loadParticles();
  if (const char* steps = std::getenv("HEMOCELL_OUTPUT_AFTER")) {
    // step, then write the HDF5 / CSV output of the reference layout under ./tmp
    for (auto& fp : hemocell.cellfields->fields)
      hemocell.setOutputs(fp->name, {hemo::OUTPUT_POSITION, hemo::OUTPUT_VELOCITY, hemo::OUTPUT_FORCE,
                                     hemo::OUTPUT_TRIANGLES, hemo::OUTPUT_CELL_ID, hemo::OUTPUT_VERTEX_ID,
                                     hemo::OUTPUT_FORCE_AREA, hemo::OUTPUT_FORCE_VOLUME, hemo::OUTPUT_FORCE_BENDING,
                                     hemo::OUTPUT_FORCE_LINK, hemo::OUTPUT_FORCE_VISC,
                                     hemo::OUTPUT_FORCE_INNER_LINK});
    hemocell.setFluidOutputs({hemo::OUTPUT_VELOCITY, hemo::OUTPUT_DENSITY, hemo::OUTPUT_BOUNDARY,
                              hemo::OUTPUT_OMEGA, hemo::OUTPUT_INTERIOR_POINTS});
    const int n = std::atoi(steps);
    for (int k = 0; k < n; ++k) hemocell.iterate();
    hemocell.writeOutput();
  }
  const int64_t nc = hc_num_cells(hemocell.context());
  const int64_t nv = hc_num_vertices(hemocell.context());
  std::vector<double> x((size_t)nv * 3);
  if (nv) hemocell.check(hc_get_vertices(hemocell.context(), x.data(), nullptr, nullptr), "vertices");
  FILE* f = std::fopen(argv[6], "wb");
  if (!f) throw hemo::HemoCellError(std::string("cannot write ") + argv[6]);
  std::fwrite(&nc, 8, 1, f);
  std::fwrite(&nv, 8, 1, f);
  std::fwrite(hemocell.cellfields->cell_ids.data(), 4, (size_t)nc, f);
  std::vector<int32_t> types(hemocell.cellfields->cell_types.begin(), hemocell.cellfields->cell_types.end());
  std::fwrite(types.data(), 4, (size_t)nc, f);
  std::fwrite(x.data(), 8, x.size(), f);
  std::fclose(f);
  std::printf("{\"cells\": %lld, \"vertices\": %lld, \"listed\": %d}\n", (long long)nc, (long long)nv,
              hemocell.cellfields->number_of_cells);
  return 0;
}

int main(int argc, char* argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "load_particles: %s\n", e.what());
    return 1;
  }
}
