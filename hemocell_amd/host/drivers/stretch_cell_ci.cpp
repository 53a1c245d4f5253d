// stretch_cell_ci -- the reference's optical-tweezers example
// (examples/stretchCell/stretchCell.cpp:31-175), the case its continuous
// integration runs with scripts/ci/config-stretchCell.xml and checks with
// scripts/ci/stretchCell_sanity.sh, written against the hemo::HemoCell host API
// of this library.
//
// usage: stretch_cell_ci <config.xml>
// Run from a directory holding RBC.xml and RBC.pos, as the example is.  At
// iteration 1 and every <tmeas> iterations it prints one JSON line with the
// quantities the example logs: cell centre, bounding-box diameters, surface and
// volume (um^2, um^3, and in % of the rest mesh) and the largest diameter (um).
#include <cmath>
#include <cstdio>
#include <string>

#include "hemocell.h"

using hemo::T;

static int run(int argc, char* argv[]) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <config.xml>\n", argv[0]);
    return 2;
  }
  char* args[] = {(char*)"stretch_cell_ci", nullptr};
  hemo::HemoCell hemocell(argv[1], 0, args, hemo::HemoCell::MPIHandle::External);
  hemo::Config& cfg = *hemocell.cfg;
  hemo::param::lbm_base_parameters(cfg);

  // stretchCell.cpp:53-55: the force in pN and in lattice units
  const T stretch_force = cfg["parameters"]["stretchForce"].read<T>();
  const T stretch_force_lbm = stretch_force * 1e-12 / hemo::param::df;

  // stretchCell.cpp:57-61: a 26 x 13 x 13 um box
  const hemo::plint nz = 13 * (1e-6 / cfg["domain"]["dx"].read<T>());
  const hemo::plint nx = 2 * nz;
  const hemo::plint ny = nz;

  hemocell.lattice = new hemo::Lattice(nx, ny, nz);
  hemocell.lattice->toggleInternalStatistics(false);
  hemocell.lattice->periodicity().toggleAll(false);

  // every face a velocity boundary at rest (stretchCell.cpp:74-80)
  auto* boundaryCondition = hemo::createLocalBoundaryCondition3D();
  boundaryCondition->setVelocityConditionOnBlockBoundaries(*hemocell.lattice);
  hemo::setBoundaryVelocity(*hemocell.lattice, hemocell.lattice->getBoundingBox(), hemo::Array<T, 3>({0., 0., 0.}));

  hemocell.latticeEquilibrium(1., hemo::Array<T, 3>({0., 0., 0.}));
  hemocell.lattice->initialize();

  hemocell.initializeCellfield();
  hemocell.addCellType<hemo::RbcHighOrderModel>("RBC", hemo::RBC_FROM_SPHERE);
  if (cfg.checkpointed)
    hemocell.loadCheckPoint();
  else
    hemocell.loadParticles();

  auto* rbc_field = (*hemocell.cellfields)["RBC"];
  const unsigned n_forced_lsps = 1 + 6;
  hemo::HemoCellStretch cellStretch(*rbc_field, n_forced_lsps, stretch_force_lbm);

  // undeformed values in um (stretchCell.cpp:121-124)
  const T um = 1e-6 / hemo::param::dx;  // lattice units per um
  const T volume_eq = rbc_field->meshmetric->getVolume() / std::pow(um, 3);
  const T surface_eq = rbc_field->meshmetric->getSurface() / std::pow(um, 2);

  const unsigned tmax = cfg["sim"]["tmax"].read<unsigned>();
  const unsigned tmeas = cfg["sim"]["tmeas"].read<unsigned>();
  std::printf("{\"volume_eq_um3\": %.9g, \"surface_eq_um2\": %.9g, \"force_pN\": %g, \"lattice\": [%ld, %ld, %ld]}\n",
              volume_eq, surface_eq, stretch_force, (long)nx, (long)ny, (long)nz);
  while (hemocell.iter < tmax) {
    // applyForce() before every iterate() (stretchCell.cpp:146-149)
    cellStretch.applyForce();
    hemocell.iterate();
    if (hemocell.iter == 1 || hemocell.iter % tmeas == 0) {
      hemo::CellInformationFunctionals::calculateCellVolume(&hemocell);
      hemo::CellInformationFunctionals::calculateCellArea(&hemocell);
      hemo::CellInformationFunctionals::calculateCellPosition(&hemocell);
      hemo::CellInformationFunctionals::calculateCellStretch(&hemocell);
      hemo::CellInformationFunctionals::calculateCellBoundingBox(&hemocell);
      const hemo::CellInformation& ci = hemo::CellInformationFunctionals::info_per_cell[0];
      const T volume = ci.volume / std::pow(um, 3);
      const T surface = ci.area / std::pow(um, 2);
      const hemo::Array<T, 6> bb = ci.bbox / um;
      const hemo::Array<T, 3> pos = ci.position / um;
      const T largest = ci.stretch / um;
      std::printf("{\"iteration\": %u, \"centre_um\": [%.9g, %.9g, %.9g], \"diameters_um\": [%.9g, %.9g, %.9g], "
                  "\"surface_um2\": %.9g, \"surface_pct\": %.9g, \"volume_um3\": %.9g, \"volume_pct\": %.9g, "
                  "\"largest_diameter_um\": %.9g, \"wall_hits\": %lld, \"finite\": %d}\n",
                  hemocell.iter, pos[0], pos[1], pos[2], bb[1] - bb[0], bb[3] - bb[2], bb[5] - bb[4], surface,
                  surface / surface_eq * 100.0, volume, volume / volume_eq * 100.0, largest,
                  (long long)hc_wall_hits(hemocell.context()), hc_check_finite(hemocell.context()) == HC_OK);
      std::fflush(stdout);
      hemo::CellInformationFunctionals::clear_list();
    }
  }
  delete boundaryCondition;
  return 0;
}

int main(int argc, char* argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "stretch_cell_ci: %s\n", e.what());
    return 1;
  }
}
