// lees_edwards -- the reference's Lees-Edwards shear case
// (cases/leesEdwards/leesEdwards.cpp) against the host layer, on an n^3 box.
//
// usage: lees_edwards <config.xml> <n> <warmup> <tmax> <out.bin> [RBC name]
//   config.xml: <domain> with dx, dt, shearrate, ... (param::lbm_shear_parameters)
//   name:       an RBC material <name>.xml and positions <name>.pos, loaded with
//               material timescale 20 and particle-velocity timescale 5 as the
//               reference case does
// out.bin: int64 iteration, int64 n, float64 populations[19][n^3]
// prints one JSON line: iteration, the host copy of the current displacement,
// cells, and the x-velocity averaged over each z plane
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "hemocell.h"

using hemo::T;

static int run(int argc, char* argv[]) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s <config.xml> <n> <warmup> <tmax> <out.bin> [RBC name]\n", argv[0]);
    return 2;
  }
  const hemo::plint n = std::atol(argv[2]);
  const int warmup = std::atoi(argv[3]);
  const unsigned int tmax = (unsigned)std::atol(argv[4]);
  char* args[] = {argv[0], nullptr};
  hemo::HemoCell hemocell(argv[1], 0, args);
  hemo::Config* cfg = hemocell.cfg;
  const double dt = (*cfg)["domain"]["dt"].read<double>();
  hemo::param::lbm_shear_parameters(*cfg, n);
  hemocell.lattice = new hemo::Lattice(n, n, n);
  hemo::LeesEdwardsBC LEbc(*hemocell.lattice, hemo::param::shearrate_lbm, dt, &hemocell.LEcurrentDisplacement);
  LEbc.initialize();
  hemocell.lattice->initialize();
  hemocell.initializeCellfield();
  const bool cells = argc > 6;
  if (cells) {
    hemocell.addCellType<hemo::RbcHighOrderModel>(argv[6], hemo::RBC_FROM_SPHERE);
    hemocell.setMaterialTimeScaleSeparation(argv[6], 20);
    hemocell.setParticleVelocityUpdateTimeScaleSeparation(5);
    hemocell.loadParticles();
  } else {
    hemocell.latticeEquilibrium(1.0, hemo::Array<T, 3>({0., 0., 0.}));
  }
  if (hemocell.iter == 0)
    for (int k = 0; k < warmup; ++k) hemocell.lattice->collideAndStream();
  while (hemocell.iter < tmax) {
    hemocell.iterate();
    LEbc.updateLECurDisplacement(hemocell.iter);
  }
  hc_ctx* ctx = hemocell.context();
  const int64_t N = hc_num_nodes(ctx), it = hemocell.iter;
  std::vector<double> pop((size_t)(19 * N)), rho((size_t)N), u((size_t)(3 * N));
  hemocell.check(hc_get_populations(ctx, pop.data()), "populations");
  hemocell.check(hc_get_fluid(ctx, rho.data(), u.data()), "fluid");
  FILE* o = std::fopen(argv[5], "wb");
  if (!o) throw hemo::HemoCellError(std::string("cannot write ") + argv[5]);
  const int64_t nn = n;
  std::fwrite(&it, 8, 1, o);
  std::fwrite(&nn, 8, 1, o);
  std::fwrite(pop.data(), 8, pop.size(), o);
  std::fclose(o);
  std::printf("{\"iteration\": %lld, \"displacement\": %.17g, \"lees_edwards\": %s, \"cells\": %d, \"ux\": [",
              (long long)it, *hemocell.LEcurrentDisplacement, hemocell.leesEdwardsBC ? "true" : "false",
              hc_num_cells(ctx));
  for (hemo::plint z = 0; z < n; ++z) {
    double s = 0;
    for (hemo::plint x = 0; x < n; ++x)
      for (hemo::plint y = 0; y < n; ++y) s += u[3 * (size_t)hemocell.lattice->index(x, y, z)];
    std::printf("%s%.17g", z ? ", " : "", s / (double)(n * n));
  }
  std::printf("]}\n");
  return 0;
}

int main(int argc, char* argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "lees_edwards: %s\n", e.what());
    return 1;
  }
}
