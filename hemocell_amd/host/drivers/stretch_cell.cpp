// stretch_cell -- the reference's optical-tweezers validation
// (tests/validation/stretch_cell/test_stretch_cell.cpp) written against the
// hemo::HemoCell host API of this library.
//
// usage: stretch_cell <case directory> <force pN> [iterations, default 10000]
// The directory holds config_stretch_cell.xml, stretch_RBC.xml and
// stretch_RBC.pos.  Prints one JSON line with the final axial and transverse
// diameters (um), the volume ratio and the cell count.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "hemocell.h"

using hemo::T;
static const double to_micro_meter = 1e-6;
static const unsigned n_forced_lsps = 7;

static T axial_displacement(hemo::Array<T, 6> bbox) {
  auto bbox_in_si = bbox / (to_micro_meter / hemo::param::dx);
  return bbox_in_si[1] - bbox_in_si[0];
}
static T transverse_displacement(hemo::Array<T, 6> bbox) {
  auto bbox_in_si = bbox / (to_micro_meter / hemo::param::dx);
  return bbox_in_si[3] - bbox_in_si[2];
}

static int run(int argc, char* argv[]) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <case directory> <force pN> [iterations]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const double force = std::atof(argv[2]);
  const unsigned max_iteration = argc > 3 ? (unsigned)std::atoi(argv[3]) : 10000u;
  std::string inp = dir + "/config_stretch_cell.xml";
  char* args[] = {(char*)"stretch_cell", nullptr};

  hemo::HemoCell hemocell((char*)inp.c_str(), 0, args, hemo::HemoCell::MPIHandle::External);
  hemo::Config* cfg = hemocell.cfg;
  hemo::param::lbm_base_parameters(*cfg);
  hemo::param::ef_lbm = force * (1e-12 / hemo::param::df);

  hemo::plint nz = 13 * (to_micro_meter / (*cfg)["domain"]["dx"].read<T>());
  hemo::plint nx = 2 * nz;
  hemo::plint ny = nz;

  hemocell.lattice = new hemo::Lattice(nx, ny, nz);
  hemocell.lattice->toggleInternalStatistics(false);
  hemocell.lattice->periodicity().toggleAll(false);

  auto* boundaryCondition = hemo::createLocalBoundaryCondition3D();
  boundaryCondition->setVelocityConditionOnBlockBoundaries(*hemocell.lattice);
  hemo::setBoundaryVelocity(*hemocell.lattice, hemocell.lattice->getBoundingBox(), hemo::Array<T, 3>({0., 0., 0.}));

  hemocell.latticeEquilibrium(1., hemo::Array<T, 3>({0., 0., 0.}));
  hemocell.lattice->initialize();

  hemocell.initializeCellfield();
  const std::string name = dir + "/stretch_RBC";
  hemocell.addCellType<hemo::RbcHighOrderModel>(name, hemo::RBC_FROM_SPHERE);
  hemocell.loadParticles();

  auto cellfield = (*hemocell.cellfields)[name];
  hemo::HemoCellStretch cellStretch(*cellfield, n_forced_lsps, hemo::param::ef_lbm);

  const T initial_volume = cellfield->meshmetric->getVolume() / std::pow(to_micro_meter, 3);

  while (hemocell.iter < max_iteration) {
    cellStretch.applyForce();  // cell force is applied before each step
    hemocell.iterate();
    if (hemocell.iter == 1 || hemocell.iter % 1000 == 0) {
      hemo::CellInformationFunctionals::calculateCellBoundingBox(&hemocell);
      auto bbox = hemo::CellInformationFunctionals::info_per_cell[0].bbox;
      std::fprintf(stderr, "Iteration: %u\ttransverse:\t%.6f\taxial:\t%.6f\tstretch force: %g\n", hemocell.iter,
                   transverse_displacement(bbox), axial_displacement(bbox), force);
      hemo::CellInformationFunctionals::clear_list();
    }
  }

  const unsigned cell_count = hemo::CellInformationFunctionals::getTotalNumberOfCells(&hemocell);
  hemo::CellInformationFunctionals::calculateCellBoundingBox(&hemocell);
  auto bbox = hemo::CellInformationFunctionals::info_per_cell[0].bbox;
  hemo::CellInformationFunctionals::calculateCellVolume(&hemocell);
  T volume = hemo::CellInformationFunctionals::info_per_cell[0].volume;
  volume /= std::pow(to_micro_meter, 3);
  const int finite = hc_check_finite(hemocell.context()) == HC_OK;
  std::printf("{\"force_pN\": %g, \"iterations\": %u, \"cells\": %u, \"transverse_um\": %.6f, \"axial_um\": %.6f, "
              "\"volume_ratio\": %.6f, \"wall_hits\": %lld, \"finite\": %d}\n",
              force, hemocell.iter, cell_count, transverse_displacement(bbox), axial_displacement(bbox),
              volume / initial_volume, (long long)hc_wall_hits(hemocell.context()), finite);
  delete boundaryCondition;
  return 0;
}

int main(int argc, char* argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "stretch_cell: %s\n", e.what());
    return 1;
  }
}
