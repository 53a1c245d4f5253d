// pipeflow -- the reference's pipe-flow validation (tests/validation/pipeflow/
// test_pipeflow.cpp) against the host layer.  The reference voxelises
// examples/pipeflow/tube.stl (a cylinder of radius 5 and length 20 STL units,
// along x) with 50 nodes across its 10-unit height; here the tube is the
// analytic mask of the same geometry: a 101 x 51 x 51 box whose nodes with
// (y-25)^2 + (z-25)^2 < 25^2 are fluid, the others bounce-back, periodic in x.
//
// usage: pipeflow <config_pipeflow.xml> <tmax> <check_every>
//   the RBC and PLT materials and positions are read as RBC.xml/RBC.pos and
//   PLT.xml/PLT.pos from the working directory
// Like the reference: pipe parameters from the fluid cross-section, driving
// force 8 nu (u_max / 2) / R^2 along x, material and particle timescales from
// <ibm>, minimum distance from the wall 0.5 for the RBCs, 100 warm-up
// collide-streams of the fluid, then iterate() to tmax.  After iteration 100,
// every check_every iterations: the number of cells, the apparent viscosity
// 0.5 u_max / mean |u| (FluidInfo) and the mean vertex force in pN
// (ParticleInfo).  Prints one JSON line with the extremes of those.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "hemocell.h"

using hemo::T;

static int run(int argc, char* argv[]) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <config_pipeflow.xml> <tmax> <check_every>\n", argv[0]);
    return 2;
  }
  const unsigned tmax = (unsigned)std::atol(argv[2]);
  const unsigned every = std::max(1, std::atoi(argv[3]));
  char* args[] = {argv[0], nullptr};
  hemo::HemoCell hemocell(argv[1], 0, args);
  hemo::Config* cfg = hemocell.cfg;
  const hemo::plint nx = 101, ny = 51, nz = 51;
  hemocell.lattice = new hemo::Lattice(nx, ny, nz);
  for (hemo::plint y = 0; y < ny; ++y)
    for (hemo::plint z = 0; z < nz; ++z) {
      const double r2 = (y - 25.0) * (y - 25.0) + (z - 25.0) * (z - 25.0);
      if (r2 < 25.0 * 25.0) continue;
      hemo::defineBounceBack(*hemocell.lattice, hemo::Box3D(0, nx - 1, y, y, z, z));
    }
  hemo::param::lbm_pipe_parameters(*cfg, *hemocell.lattice);
  hemo::param::printParameters();
  hemocell.lattice->periodicity().toggleAll(false);
  hemocell.latticeEquilibrium(1., hemo::Array<T, 3>({0., 0., 0.}));
  const T force = 8 * hemo::param::nu_lbm * (hemo::param::u_lbm_max * 0.5) / hemo::param::pipe_radius /
                  hemo::param::pipe_radius;
  const hemo::Array<T, 3> driving({force, 0., 0.});
  hemocell.lattice->initialize();
  hemocell.initializeCellfield();
  hemocell.addCellType<hemo::RbcHighOrderModel>("RBC", hemo::RBC_FROM_SPHERE);
  hemocell.setMaterialTimeScaleSeparation("RBC", (*cfg)["ibm"]["stepMaterialEvery"].read<int>());
  hemocell.setInitialMinimumDistanceFromSolid("RBC", 0.5);
  hemocell.addCellType<hemo::PltSimpleModel>("PLT", hemo::ELLIPSOID_FROM_SPHERE);
  hemocell.setMaterialTimeScaleSeparation("PLT", (*cfg)["ibm"]["stepMaterialEvery"].read<int>());
  hemocell.setParticleVelocityUpdateTimeScaleSeparation((*cfg)["ibm"]["stepParticleEvery"].read<int>());
  hemocell.setSystemPeriodicity(0, true);
  hemocell.loadParticles();
  hemo::setExternalVector(*hemocell.lattice, hemocell.lattice->getBoundingBox(), 0, driving);
  for (int i = 0; i < 100; ++i) hemocell.lattice->collideAndStream();
  const unsigned cells0 = hemo::CellInformationFunctionals::getTotalNumberOfCells(&hemocell);
  double vmin = 1e300, vmax = -1e300, fmax = 0;
  unsigned cmin = cells0, cmax = cells0, checks = 0, cells_at_100 = cells0;
  long last_loss = -1;
  while (hemocell.iter < tmax) {
    hemocell.iterate();
    hemo::setExternalVector(*hemocell.lattice, hemocell.lattice->getBoundingBox(), 0, driving);
    // the number of cells after every iteration (test_pipeflow.cpp:84-88)
    const unsigned count = hemo::CellInformationFunctionals::getTotalNumberOfCells(&hemocell);
    if (count < cmin) last_loss = (long)hemocell.iter;
    cmin = std::min(cmin, count);
    cmax = std::max(cmax, count);
    if (hemocell.iter == 100) cells_at_100 = count;
    if (hemocell.iter < 100 || hemocell.iter % every) continue;
    const hemo::FluidStatistics fi = hemo::FluidInfo::calculateVelocityStatistics(&hemocell);
    const double viscosity = 0.5 * hemo::param::u_lbm_max / fi.avg;
    vmin = std::min(vmin, viscosity);
    vmax = std::max(vmax, viscosity);
    const hemo::ParticleStatistics pi = hemo::ParticleInfo::calculateForceStatistics(&hemocell);
    fmax = std::max(fmax, pi.avg * hemo::param::df * 1.0e12);
    ++checks;
  }
  std::printf("{\"cells_loaded\": %u, \"cells_min\": %u, \"cells_max\": %u, \"cells_at_100\": %u, "
              "\"last_loss_iteration\": %ld, \"removed_vertices\": %lld, \"viscosity_min\": %.17g, "
              "\"viscosity_max\": %.17g, \"mean_force_pN_max\": %.17g, \"checks\": %u, \"pipe_radius\": %.17g, "
              "\"iteration\": %u}\n",
              cells0, cmin, cmax, cells_at_100, last_loss, (long long)hc_wall_hits(hemocell.context()), vmin, vmax,
              fmax, checks, hemo::param::pipe_radius, hemocell.iter);
  return 0;
}

int main(int argc, char* argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "pipeflow: %s\n", e.what());
    return 1;
  }
}
