// checkpoint_run -- a sheared suspension in the shape of the reference's
// examples (examples/cube/cube.cpp:128-133, examples/capillary/capillary.cpp:
// 153-208): a fresh configuration loads particles, a checkpoint.xml resumes
// with loadCheckPoint(); the run iterates to tmax and saves a checkpoint every
// tcheckpoint steps.
//
// usage: checkpoint_run <config.xml> <nx> <ny> <nz> <shear_lu> <tmax> <tcheckpoint> <out.bin> <KIND:name>...
//   shear_lu: lattice shear rate of the Couette walls on the z faces
//             (iniLatticeSquareCouette), periodic in x and y
//   KIND:     RBC, PLT or WBC as for load_particles
//   HEMOCELL_TIMESCALES=m,v,r  material, particle-velocity and repulsion
//             timescale separations (default 1,1,1); with HEMOCELL_REPULSION=
//             k,cutoff_um the repulsion is on; HEMOCELL_INTERIOR_VISCOSITY=m,e
//             sets the interior-viscosity timescales (membrane, entire grid) of
//             the types whose material enables it
// out.bin: int64 iteration, int64 ncells, int64 nvertices, int32 cell id[ncells],
//          float64 x[nvertices][3], v[nvertices][3], f[nvertices][3],
//          float64 populations[19][nodes]
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "hemocell.h"

using hemo::T;

static int run(int argc, char* argv[]) {
  if (argc < 10) {
    std::fprintf(stderr,
                 "usage: %s <config.xml> <nx> <ny> <nz> <shear_lu> <tmax> <tcheckpoint> <out.bin> <KIND:name>...\n",
                 argv[0]);
    return 2;
  }
  const hemo::plint nx = std::atol(argv[2]), ny = std::atol(argv[3]), nz = std::atol(argv[4]);
  const T shear = std::atof(argv[5]);
  const unsigned int tmax = (unsigned)std::atol(argv[6]), tcheckpoint = (unsigned)std::atol(argv[7]);
  char* args[] = {argv[0], nullptr};
  hemo::HemoCell hemocell(argv[1], 0, args);
  hemo::Config* cfg = hemocell.cfg;
  hemo::param::lbm_base_parameters(*cfg);
  hemocell.lattice = new hemo::Lattice(nx, ny, nz);
  hemo::OnLatticeBoundaryCondition3D bc;
  hemo::iniLatticeSquareCouette(*hemocell.lattice, nx, ny, nz, bc, shear);
  hemocell.latticeEquilibrium(1.0, hemo::Array<T, 3>({0., 0., 0.}));
  hemocell.initializeCellfield();
  std::vector<std::string> names;
  for (int a = 9; a < argc; ++a) {
    const std::string spec = argv[a];
    const size_t colon = spec.find(':');
    if (colon == std::string::npos) throw hemo::HemoCellError("cell type spec needs KIND:name");
    const std::string kind = spec.substr(0, colon), name = spec.substr(colon + 1);
    if (kind == "RBC") hemocell.addCellType<hemo::RbcHighOrderModel>(name, hemo::RBC_FROM_SPHERE);
    else if (kind == "PLT") hemocell.addCellType<hemo::PltSimpleModel>(name, hemo::ELLIPSOID_FROM_SPHERE);
    else if (kind == "WBC") hemocell.addCellType<hemo::WbcHighOrderModel>(name, hemo::WBC_SPHERE);
    else throw hemo::HemoCellError("unknown cell kind " + kind);
    names.push_back(name);
  }
  unsigned ts[3] = {1, 1, 1};
  if (const char* e = std::getenv("HEMOCELL_TIMESCALES")) {
    std::stringstream ss(e);
    char sep;
    ss >> ts[0] >> sep >> ts[1] >> sep >> ts[2];
  }
  for (const auto& n : names) hemocell.setMaterialTimeScaleSeparation(n, ts[0]);
  hemocell.setParticleVelocityUpdateTimeScaleSeparation(ts[1]);
  if (const char* e = std::getenv("HEMOCELL_REPULSION")) {
    std::stringstream ss(e);
    T k = 0, cutoff = 0;
    char sep;
    ss >> k >> sep >> cutoff;
    hemocell.setRepulsion(k, cutoff);
    hemocell.setRepulsionTimeScaleSeperation(ts[2]);
  }

  if (const char* e = std::getenv("HEMOCELL_INTERIOR_VISCOSITY")) {
    std::stringstream ss(e);
    unsigned m = 1, g = 1;
    char sep;
    ss >> m >> sep >> g;
    hemocell.setInteriorViscosityTimeScaleSeperation(m, g);
  }

  if (!cfg->checkpointed) {
    hemocell.loadParticles();
  } else {
    hemocell.loadCheckPoint();
  }
  const unsigned start = hemocell.iter;
  while (hemocell.iter < tmax) {
    hemocell.iterate();
    if (tcheckpoint && hemocell.iter % tcheckpoint == 0) hemocell.saveCheckPoint();
  }

  hc_ctx* ctx = hemocell.context();
  const int64_t it = hemocell.iter, nc = hc_num_cells(ctx), nv = hc_num_vertices(ctx), N = hc_num_nodes(ctx);
  std::vector<double> x((size_t)nv * 3), v((size_t)nv * 3), f((size_t)nv * 3), pop((size_t)(19 * N));
  if (nv) hemocell.check(hc_get_vertices(ctx, x.data(), v.data(), f.data()), "vertices");
  hemocell.check(hc_get_populations(ctx, pop.data()), "populations");
  FILE* o = std::fopen(argv[8], "wb");
  if (!o) throw hemo::HemoCellError(std::string("cannot write ") + argv[8]);
  std::fwrite(&it, 8, 1, o);
  std::fwrite(&nc, 8, 1, o);
  std::fwrite(&nv, 8, 1, o);
  std::fwrite(hemocell.cellfields->cell_ids.data(), 4, (size_t)nc, o);
  std::fwrite(x.data(), 8, x.size(), o);
  std::fwrite(v.data(), 8, v.size(), o);
  std::fwrite(f.data(), 8, f.size(), o);
  std::fwrite(pop.data(), 8, pop.size(), o);
  std::fclose(o);
  const hemo::FluidStatistics fs = hemo::FluidInfo::calculateVelocityStatistics(&hemocell);
  std::vector<double> tau_iv((size_t)N);
  hemocell.check(hc_get_interior_viscosity(ctx, tau_iv.data()), "interior viscosity");
  long inner = 0;
  for (double t : tau_iv) inner += t != 0.0;
  std::printf("{\"start\": %u, \"iteration\": %lld, \"cells\": %lld, \"vertices\": %lld, \"out\": \"%s\", "
              "\"u_max\": %.17g, \"u_mean\": %.17g, \"fluid_nodes\": %lu, \"inner_nodes\": %ld}\n",
              start, (long long)it, (long long)nc, (long long)nv, hemocell.outputDirectory().c_str(), fs.max, fs.avg,
              (unsigned long)fs.ncells, inner);
  return 0;
}

int main(int argc, char* argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "checkpoint_run: %s\n", e.what());
    return 1;
  }
}
