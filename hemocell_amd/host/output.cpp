// output.cpp -- HemoCell::writeOutput(): HDF5 and CSV files in the reference's
// layout, so its post-processing scripts read them unchanged.
//
//   <out>/hdf5/<iter>/<Type>.<iter>.p.0.h5   per cell type
//        (io/ParticleHdf5IO.cpp:39-229, io/hemoCellParticleFieldOutputFunctions.cpp)
//   <out>/hdf5/<iter>/Fluid.<iter>.p.0.h5    fluid field with a one-node envelope
//        (io/FluidHdf5IO.hh:74-211)
//   <out>/csv/<Type>.<iter>.csv              per-cell information (io/writeCellInfoCSV.cpp:30-78)
//
// <iter> is the iteration zero-padded to 12 digits; <out> is
// <parameters><outputDirectory> of the configuration (default ./tmp), with _0,
// _1, ... appended when it already exists (config/config.cpp:91-130).  One
// domain means one "atomic block": p.0, processorId 0, one processor.
#include <hdf5.h>
#include <hdf5_hl.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iomanip>

#include "hemocell.h"

namespace hemo {

bool exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

void mkpath(const std::string& p) {
  std::string cur;
  for (size_t i = 0; i < p.size(); ++i) {
    cur += p[i];
    if (p[i] == '/' || i + 1 == p.size())
      if (!cur.empty() && cur != "/" && cur != "./") mkdir(cur.c_str(), 0777);
  }
}

namespace {

std::string zero_pad(long num, int w = 12) {
  std::ostringstream ss;
  ss << std::setw(w) << std::setfill('0') << num;
  return ss.str();
}

// 2-D dataset, chunked by up to 1000 rows and deflated (level 7), as the reference writes
template <typename V>
void dataset2(hid_t file, const std::string& name, hid_t type, const std::vector<V>& data, hsize_t rows,
              hsize_t cols) {
  hsize_t dims[2] = {rows, cols};
  hsize_t chunk[2] = {std::max<hsize_t>(std::min<hsize_t>(1000, rows), 1), std::max<hsize_t>(cols, 1)};
  hid_t sid = H5Screate_simple(2, dims, nullptr);
  hid_t plist = H5Pcreate(H5P_DATASET_CREATE);
  H5Pset_chunk(plist, 2, chunk);
  H5Pset_deflate(plist, 7);
  hid_t did = H5Dcreate2(file, name.c_str(), type, sid, H5P_DEFAULT, plist, H5P_DEFAULT);
  if (rows * cols > 0) H5Dwrite(did, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.data());
  H5Dclose(did);
  H5Pclose(plist);
  H5Sclose(sid);
}

// 4-D fluid dataset (z, y, x, components), chunked per axis up to 1000 (io/FluidHdf5IO.hh:116-121)
void dataset4(hid_t file, const std::string& name, const std::vector<float>& data, const hsize_t dim[4]) {
  hsize_t chunk[4] = {std::min<hsize_t>(1000, dim[0]), std::min<hsize_t>(1000, dim[1]),
                      std::min<hsize_t>(1000, dim[2]), dim[3]};
  hid_t sid = H5Screate_simple(4, dim, nullptr);
  hid_t plist = H5Pcreate(H5P_DATASET_CREATE);
  H5Pset_chunk(plist, 4, chunk);
  H5Pset_deflate(plist, 7);
  hid_t did = H5Dcreate2(file, name.c_str(), H5T_NATIVE_FLOAT, sid, H5P_DEFAULT, plist, H5P_DEFAULT);
  H5Dwrite(did, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.data());
  H5Dclose(did);
  H5Pclose(plist);
  H5Sclose(sid);
}

hid_t open_file(const std::string& name, double dx, double dt, long iter) {
  hid_t f = H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) throw HemoCellError("cannot create " + name);
  int id = 0;
  H5LTset_attribute_double(f, "/", "dx", &dx, 1);
  H5LTset_attribute_double(f, "/", "dt", &dt, 1);
  H5LTset_attribute_long(f, "/", "iteration", &iter, 1);
  H5LTset_attribute_int(f, "/", "processorId", &id, 1);
  return f;
}

}  // namespace

std::string HemoCell::outputDirectory() {
  if (!outDir.empty()) return outDir;
  std::string d;
  try {
    d = (*cfg)["parameters"]["outputDirectory"].read<std::string>();
    while (!d.empty() && d.back() == '/') d.pop_back();
    if (!d.empty() && d[0] != '/') d = "./" + d;
  } catch (std::invalid_argument&) {
    d = "./tmp";
  }
  if (exists(d))
    for (int i = 0;; i++)
      if (!exists(d + "_" + std::to_string(i))) {
        d = d + "_" + std::to_string(i);
        break;
      }
  outDir = d + "/";
  mkpath(outDir + "hdf5/");
  return outDir;
}

// core/hemoCell.cpp:221-287: forces are recomputed for the output (and the
// repulsion when enabled); nothing of the step state changes.
void HemoCell::writeOutput() {
  bindLattice();
  const std::string out = outputDirectory();
  const long it = (long)iter;
  std::cout << "(HemoCell) (Output) writing output at timestep " << iter << " (" << param::dt * iter << " s)."
            << std::endl;
  check(hc_stage_mechanics(ctx, 1), "writeOutput");
  if (repulsionEnabled) check(hc_stage_repulsion(ctx), "writeOutput");
  if (boundaryRepulsionEnabled) check(hc_stage_boundary_repulsion(ctx), "writeOutput");
  mkpath(out + "hdf5/" + zero_pad(it) + "/");
  mkpath(out + "csv/");
  const double dx = param::dx, dt = param::dt, df = param::df;
  const bool si = outputInSiUnits;

  // ---- cells ----
  const int64_t nv = hc_num_vertices(ctx);
  std::vector<double> x((size_t)nv * 3), v((size_t)nv * 3), f((size_t)nv * 3), fr;
  std::vector<std::vector<double>> terms;  // area, volume, bending, link, visc, inner link (on first use)
  // force_repulsion: the cell-cell and the wall terms (hemoCellParticleField.cpp:690, :911)
  auto repulsion_force = [&](int64_t n) {
    std::vector<double> r((size_t)n * 3, 0.0);
    if (n && repulsionEnabled) check(hc_get_repulsion_force(ctx, r.data()), "writeOutput");
    if (n && boundaryRepulsionEnabled) {
      std::vector<double> b((size_t)n * 3);
      check(hc_get_boundary_repulsion_force(ctx, b.data()), "writeOutput");
      for (size_t k = 0; k < r.size(); ++k) r[k] = r[k] + b[k];
    }
    return r;
  };
  if (nv) check(hc_get_vertices(ctx, x.data(), v.data(), f.data()), "writeOutput");
  if (cellfields) {
    // vertex base of every device cell
    std::vector<int64_t> base(cellfields->cell_types.size());
    int64_t b = 0;
    for (size_t c = 0; c < base.size(); ++c) {
      base[c] = b;
      b += (*cellfields)[(pluint)cellfields->cell_types[c]]->numVertex;
    }
    // Cells that lost a vertex to a wall are incomplete: the reference deletes
    // them before it writes the positions of their type
    // (hemoCellParticleFieldOutputFunctions.cpp:51-52, deleteIncompleteCells) and
    // writes no incomplete cell.  Deleting = removing every vertex of the cell.
    std::vector<uint8_t> incomplete(base.size(), 0);
    if (!base.empty()) check(hc_get_incomplete_cells(ctx, incomplete.data()), "writeOutput");
    bool drop = false;
    std::vector<uint8_t> removed((size_t)nv, 0);
    for (size_t c = 0; c < base.size(); ++c) {
      if (!incomplete[c]) continue;
      const HemoCellField& F = *(*cellfields)[(pluint)cellfields->cell_types[c]];
      if (std::find(F.outputs.begin(), F.outputs.end(), (int)OUTPUT_POSITION) == F.outputs.end()) continue;
      if (!drop) check(hc_get_removed_vertices(ctx, removed.data()), "writeOutput");
      drop = true;
      for (int k = 0; k < F.numVertex; ++k) removed[(size_t)(base[c] + k)] = 1;
    }
    if (drop) check(hc_set_removed_vertices(ctx, removed.data()), "writeOutput");
    for (auto& fp : cellfields->fields) {
      HemoCellField& F = *fp;
      if (F.outputs.empty()) continue;
      std::vector<size_t> cells;
      for (size_t c = 0; c < base.size(); ++c)
        if (cellfields->cell_types[c] == F.ctype && !incomplete[c]) cells.push_back(c);
      const hsize_t rows = (hsize_t)cells.size() * F.numVertex;
      hid_t file = open_file(out + "hdf5/" + zero_pad(it) + "/" + F.name + "." + zero_pad(it) + ".p.0.h5", dx, dt, it);
      long nproc = 1;
      H5LTset_attribute_long(file, "/", "numberOfProcessors", &nproc, 1);
      auto vec3 = [&](const std::vector<double>& src, double scale, const char* name) {
        std::vector<float> o;
        o.reserve(rows * 3);
        for (size_t c : cells)
          for (int k = 0; k < F.numVertex; ++k)
            for (int q = 0; q < 3; ++q) o.push_back((float)(src[3 * (size_t)(base[c] + k) + q] * scale));
        dataset2(file, name, H5T_NATIVE_FLOAT, o, rows, 3);
      };
      for (int o : F.outputs) {
        if (o == OUTPUT_POSITION) {
          vec3(x, si ? dx : 1.0, "Position");
          long n = (long)rows;
          H5LTset_attribute_long(file, "/", "numberOfParticles", &n, 1);
        } else if (o == OUTPUT_VELOCITY) {
          vec3(v, si ? dx / dt : 1.0, "Velocity");
        } else if (o == OUTPUT_FORCE) {
          // force_total = force + force_repulsion (hemoCellParticleField.cpp:596)
          if (fr.empty()) fr = repulsion_force(nv);
          std::vector<double> tot(f);
          for (size_t k = 0; k < tot.size(); ++k) tot[k] += fr[k];
          vec3(tot, si ? df : 1.0, "Total force");
        } else if (o == OUTPUT_FORCE_REPULSION) {
          if (fr.empty()) fr = repulsion_force(nv);
          vec3(fr, si ? df : 1.0, "Repulsion force");
        } else if (o == OUTPUT_FORCE_AREA || o == OUTPUT_FORCE_VOLUME || o == OUTPUT_FORCE_BENDING ||
                   o == OUTPUT_FORCE_LINK || o == OUTPUT_FORCE_VISC || o == OUTPUT_FORCE_INNER_LINK) {
          // separateForceVectors + applyConstitutiveModel(true): every term on its own
          // (hemoCellParticleField.cpp:590-614, hemoCellParticleFieldOutputFunctions.cpp:113-288)
          if (terms.empty()) {
            terms.assign(6, std::vector<double>((size_t)nv * 3, 0.0));
            if (nv)
              check(hc_get_force_terms(ctx, terms[0].data(), terms[1].data(), terms[2].data(), terms[3].data(),
                                       terms[4].data(), terms[5].data()),
                    "writeOutput");
          }
          const int k = o == OUTPUT_FORCE_AREA ? 0 : o == OUTPUT_FORCE_VOLUME ? 1 : o == OUTPUT_FORCE_BENDING ? 2
                        : o == OUTPUT_FORCE_LINK ? 3 : o == OUTPUT_FORCE_VISC ? 4 : 5;
          static const char* names[6] = {"Area force", "Volume force", "Bending force", "Link force",
                                         "Viscous force", "Inner link force"};
          vec3(terms[(size_t)k], si ? df : 1.0, names[k]);
        } else if (o == OUTPUT_CELL_ID || o == OUTPUT_VERTEX_ID) {
          std::vector<float> ids;
          for (size_t c : cells)
            for (int k = 0; k < F.numVertex; ++k)
              ids.push_back((float)(o == OUTPUT_CELL_ID ? cellfields->cell_ids[c] : k));
          dataset2(file, o == OUTPUT_CELL_ID ? "Cell Id" : "Vertex Id", H5T_NATIVE_FLOAT, ids, rows, 1);
        } else if (o == OUTPUT_TRIANGLES) {
          std::vector<int> tri;
          int counter = 0;
          for (size_t c = 0; c < cells.size(); ++c) {
            for (size_t t = 0; t < F.triangles.size(); ++t) tri.push_back(F.triangles[t] + counter);
            counter += F.numVertex;
          }
          const hsize_t nt = (hsize_t)tri.size() / 3;
          dataset2(file, "Triangles", H5T_NATIVE_INT, tri, nt, 3);
          long n = (long)nt;
          H5LTset_attribute_long(file, "/", "numberOfTriangles", &n, 1);
        }
      }
      H5Fclose(file);
    }
  }

  // ---- fluid ----
  if (!fluidOutputs.empty() && lattice) {
    const plint nx = lattice->getNx(), ny = lattice->getNy(), nz = lattice->getNz();
    const size_t N = (size_t)(nx * ny * nz);
    std::vector<double> rho(N), u(3 * N);
    check(hc_get_fluid(ctx, rho.data(), u.data()), "writeOutput");
    const hsize_t Nx = nx + 2, Ny = ny + 2, Nz = nz + 2;  // one-node envelope on each side
    hid_t file = open_file(out + "hdf5/" + zero_pad(it) + "/Fluid." + zero_pad(it) + ".p.0.h5", dx, dt, it);
    int ncells = (int)(Nx * Ny * Nz);
    int sub[3] = {(int)Nz, (int)Ny, (int)Nx};
    float dxyz[3] = {1.f, 1.f, 1.f};
    float rel[3] = {-1.5f, -1.5f, -1.5f};
    if (si)
      for (int q = 0; q < 3; ++q) {
        rel[q] *= (float)dx;
        dxyz[q] = (float)dx;
      }
    H5LTset_attribute_int(file, "/", "numberOfCells", &ncells, 1);
    H5LTset_attribute_int(file, "/", "subdomainSize", sub, 3);
    H5LTset_attribute_float(file, "/", "relativePosition", rel, 3);
    H5LTset_attribute_float(file, "/", "dxdydz", dxyz, 3);
    // envelope nodes take the value of their periodic image, or of the nearest
    // domain node across a non-periodic face
    auto src = [&](plint c, plint n, int axis) {
      if (c >= 0 && c < n) return c;
      if (lattice->periodicity().get(axis)) return (c + n) % n;
      return c < 0 ? (plint)0 : n - 1;
    };
    // relaxation time inside membranes (0 elsewhere), for Omega and InteriorPoints
    std::vector<double> tau_iv;
    for (int o : fluidOutputs)
      if ((o == OUTPUT_OMEGA || o == OUTPUT_INTERIOR_POINTS) && tau_iv.empty()) {
        tau_iv.resize(N);
        check(hc_get_interior_viscosity(ctx, tau_iv.data()), "writeOutput");
      }
    for (int o : fluidOutputs) {
      const bool vel = o == OUTPUT_VELOCITY, den = o == OUTPUT_DENSITY, bnd = o == OUTPUT_BOUNDARY;
      const bool omg = o == OUTPUT_OMEGA, inp = o == OUTPUT_INTERIOR_POINTS;
      if (!vel && !den && !bnd && !omg && !inp) continue;  // shear, strain, force, cell density: not written
      const hsize_t comp = vel ? 3 : 1;
      std::vector<float> d(Nx * Ny * Nz * comp);
      size_t n = 0;
      for (plint z = -1; z <= nz; ++z)
        for (plint y = -1; y <= ny; ++y)
          for (plint xx = -1; xx <= nx; ++xx) {
            const plint sx = src(xx, nx, 0), sy = src(y, ny, 1), sz = src(z, nz, 2);
            const size_t node = (size_t)lattice->index(sx, sy, sz);
            if (vel)
              for (int q = 0; q < 3; ++q) d[n++] = (float)(u[3 * node + q] * (si ? dx / dt : 1.0));
            else if (den)
              d[n++] = (float)(rho[node] * (si ? df / (dx * dx) : 1.0));
            else if (omg)
              // FluidHdf5IO.hh:354-370, scaled as the reference does in SI output;
              // bounce-back nodes have no relaxation rate
              d[n++] = lattice->isBoundary(sx, sy, sz)
                           ? 0.f
                           : (float)((1.0 / (tau_iv[node] > 0.0 ? tau_iv[node] : param::tau)) *
                                     (si ? df / (dx * dx) : 1.0));
            else if (inp)
              d[n++] = (float)tau_iv[node];  // FluidHdf5IO.hh:333-352
            else
              d[n++] = lattice->isBoundary(sx, sy, sz) ? 1.f : 0.f;
          }
      const hsize_t dim[4] = {Nz, Ny, Nx, comp};
      const char* dname = vel ? "Velocity" : den ? "Density" : omg ? "Omega" : inp ? "InteriorPoints" : "Boundary";
      dataset4(file, dname, d, dim);
    }
    H5Fclose(file);
  }

  // ---- csv: per-cell information ----
  if (cellfields && !cellfields->fields.empty()) {
    CellInformationFunctionals::clear_list();
    CellInformationFunctionals::calculate_vol_pos_area(this);
    std::vector<std::ofstream> csv(cellfields->size());
    for (pluint i = 0; i < cellfields->size(); ++i) {
      csv[i].open(out + "csv/" + (*cellfields)[i]->name + "." + zero_pad(it) + ".csv", std::ofstream::trunc);
      csv[i] << "X,Y,Z,area,volume,atomic_block,cellId,baseCellId,velocity_x,velocity_y,velocity_z" << std::endl;
    }
    const int ncell_total = std::max(1, cellfields->number_of_cells);
    for (auto& kv : CellInformationFunctionals::info_per_cell) {
      CellInformation ci = kv.second;
      if (si) {
        ci.position = ci.position * dx;
        ci.area *= dx * dx;
        ci.velocity = ci.velocity * (dx / dt);
        ci.volume *= dx * dx * dx;
      }
      const int base_id = ((kv.first % ncell_total) + ncell_total) % ncell_total;
      std::ofstream& o = csv[(size_t)ci.cellType];
      o << ci.position[0] << "," << ci.position[1] << "," << ci.position[2] << "," << ci.area << "," << ci.volume
        << ",0," << kv.first << "," << base_id << "," << ci.velocity[0] << "," << ci.velocity[1] << ","
        << ci.velocity[2] << std::endl;
    }
    CellInformationFunctionals::clear_list();
  }
}

}  // namespace hemo
