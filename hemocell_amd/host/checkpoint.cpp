// checkpoint.cpp -- HemoCell::saveCheckPoint() / loadCheckPoint().
//
// The reference saves the Palabos lattice and particle field with parallelIO
// (core/hemoCellFields.cpp:270-319) and reads them back in
// HemoCellFields::load (core/hemoCellFields.cpp:240-268).  Here the state that
// a step boundary carries is:
//   lattice.dat        header + the 19 populations per node (natural layout,
//                      f[i*N + n], as double: hc_get_populations)
//   particleField.dat  header + the cell types (name, vertices per cell), per
//                      cell its type and id, then per vertex position,
//                      velocity, force and repulsion force (the fields of
//                      HemoCellParticle::serializeValues_t,
//                      core/hemoCellParticle.h:45-63), then one byte per
//                      vertex, 1 for a vertex removed by wall contact (the
//                      reference saves only the particles that exist)
//   internalViscosity.dat  (with interior viscosity on) header + the tau inside
//                      membranes per node, 0 elsewhere (InteriorViscosityHelper::
//                      checkpoint / restore, helper/interiorViscosity.cpp:82-133)
//   checkpoint.xml     <Checkpoint><hemocell>...</hemocell><General>
//                      <Iteration>N</Iteration><OutDirectory>d</OutDirectory>
//                      </General></Checkpoint>
// Both binary files are little-endian and open with an 8-byte tag and the
// iteration, so a lattice and a particle field of different steps are refused.
// The IBM force field is not stored: the next step spreads it again from the
// vertex forces, as the reference's next iterate() does.
#include <cstdio>
#include <cstring>
#include <fstream>

#include "hemocell.h"

namespace hemo {

namespace {

const char kLatticeTag[8] = {'H', 'C', 'C', 'K', 'L', 'A', 'T', '1'};
const char kParticleTag[8] = {'H', 'C', 'C', 'K', 'P', 'A', 'R', '1'};
const char kViscosityTag[8] = {'H', 'C', 'C', 'K', 'I', 'V', 'F', '1'};

std::string escape(const std::string& t) {
  std::string r;
  for (char c : t) {
    if (c == '&') r += "&amp;";
    else if (c == '<') r += "&lt;";
    else if (c == '>') r += "&gt;";
    else r += c;
  }
  return r;
}

// core/hemoCellFields.cpp:278-291: the previous checkpoint survives as *.old
void rename_to_old(const std::string& p) {
  if (exists(p)) std::rename(p.c_str(), (p + ".old").c_str());
}

template <typename V>
void put(std::ofstream& o, const V& v) {
  o.write(reinterpret_cast<const char*>(&v), sizeof(V));
}
template <typename V>
void put_vec(std::ofstream& o, const std::vector<V>& v) {
  if (!v.empty()) o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(sizeof(V) * v.size()));
}
template <typename V>
V get(std::ifstream& i, const std::string& file) {
  V v;
  if (!i.read(reinterpret_cast<char*>(&v), sizeof(V))) throw HemoCellError("loadCheckPoint: " + file + " is truncated");
  return v;
}
template <typename V>
void get_vec(std::ifstream& i, std::vector<V>& v, const std::string& file) {
  if (!v.empty() && !i.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(sizeof(V) * v.size())))
    throw HemoCellError("loadCheckPoint: " + file + " is truncated");
}
void check_tag(std::ifstream& i, const char* tag, const std::string& file) {
  char t[8];
  if (!i.read(t, 8) || std::memcmp(t, tag, 8) != 0)
    throw HemoCellError("loadCheckPoint: " + file + " is not a checkpoint file of this format");
}

}  // namespace

std::string toXml(const XMLNode& n, int indent) {
  const std::string pad((size_t)indent * 2, ' ');
  std::string r = pad + "<" + n.name + ">";
  if (n.children.empty()) return r + escape(n.text) + "</" + n.name + ">\n";
  r += "\n";
  if (!n.text.empty()) r += pad + "  " + escape(n.text) + "\n";
  for (const auto& c : n.children) r += toXml(*c, indent + 1);
  return r + pad + "</" + n.name + ">\n";
}

// config/config.cpp:168-173
std::string HemoCell::checkpointDirectory() {
  std::string name = "checkpoint";
  try {
    name = (*cfg)["parameters"]["checkpointDirectory"].read<std::string>();
    while (!name.empty() && name.back() == '/') name.pop_back();
  } catch (std::invalid_argument&) {
  }
  return outputDirectory() + name + "/";
}

// core/hemoCell.cpp:209-218, core/hemoCellFields.cpp:270-319
void HemoCell::saveCheckPoint() {
  hlog << "(HemoCell) (Saving Functions) Saving Checkpoint at timestep " << iter << std::endl;
  bindLattice();
  check(hc_synchronize(ctx), "saveCheckPoint");
  const std::string dir = checkpointDirectory();
  mkpath(dir);
  rename_to_old(dir + "lattice.dat");
  rename_to_old(dir + "particleField.dat");
  rename_to_old(dir + "checkpoint.xml");
  const int64_t it = (int64_t)iter;
  bool iv = false;
  if (cellfields)
    for (auto& F : cellfields->fields) iv = iv || F->doInteriorViscosity;
  if (iv) rename_to_old(dir + "internalViscosity.dat");

  // checkpoint.xml
  const XMLNode* hc = cfg->hemocellNode();
  if (!hc) throw HemoCellError("saveCheckPoint: the configuration has no <hemocell> element");
  {
    std::ofstream x(dir + "checkpoint.xml");
    x << "<?xml version=\"1.0\" ?>\n<Checkpoint>\n" << toXml(*hc, 1) << "  <General>\n"
      << "    <Iteration>" << it << "</Iteration>\n"
      << "    <OutDirectory>" << escape(outputDirectory()) << "</OutDirectory>\n"
      << "  </General>\n</Checkpoint>\n";
    if (!x) throw HemoCellError("saveCheckPoint: cannot write " + dir + "checkpoint.xml");
  }

  // lattice.dat
  {
    const int64_t N = hc_num_nodes(ctx);
    std::vector<double> f((size_t)(19 * N));
    check(hc_get_populations(ctx, f.data()), "saveCheckPoint: populations");
    std::ofstream o(dir + "lattice.dat", std::ios::binary);
    o.write(kLatticeTag, 8);
    put<int64_t>(o, lattice->getNx());
    put<int64_t>(o, lattice->getNy());
    put<int64_t>(o, lattice->getNz());
    put<int64_t>(o, it);
    put<int32_t>(o, precision);
    put<int32_t>(o, 19);
    put_vec(o, f);
    if (!o) throw HemoCellError("saveCheckPoint: cannot write " + dir + "lattice.dat");
  }

  // particleField.dat
  {
    const int nc = hc_num_cells(ctx);
    const int64_t nv = hc_num_vertices(ctx);
    std::vector<int32_t> tid((size_t)nc), cid((size_t)nc);
    check(hc_get_cells(ctx, tid.data(), cid.data()), "saveCheckPoint: cells");
    std::vector<double> x((size_t)(3 * nv)), v((size_t)(3 * nv)), fv((size_t)(3 * nv)), fr((size_t)(3 * nv), 0.0);
    check(hc_get_vertices(ctx, x.data(), v.data(), fv.data()), "saveCheckPoint: vertices");
    if (repulsionEnabled && nv > 0) check(hc_get_repulsion_force(ctx, fr.data()), "saveCheckPoint: repulsion force");
    // cell types by registration order (the reference's celltype)
    std::vector<int32_t> ctype((size_t)nc);
    const int ntypes = cellfields ? (int)cellfields->fields.size() : 0;
    for (int c = 0; c < nc; ++c) {
      ctype[(size_t)c] = -1;
      for (int t = 0; t < ntypes; ++t)
        if (cellfields->fields[(size_t)t]->type_id == tid[(size_t)c]) ctype[(size_t)c] = t;
      if (ctype[(size_t)c] < 0) throw HemoCellError("saveCheckPoint: a cell of an unregistered type");
    }
    std::ofstream o(dir + "particleField.dat", std::ios::binary);
    o.write(kParticleTag, 8);
    put<int64_t>(o, it);
    put<int32_t>(o, ntypes);
    for (int t = 0; t < ntypes; ++t) {
      const HemoCellField& F = *cellfields->fields[(size_t)t];
      put<int32_t>(o, F.numVertex);
      put<int32_t>(o, (int32_t)F.name.size());
      o.write(F.name.data(), (std::streamsize)F.name.size());
    }
    put<int32_t>(o, nc);
    put<int64_t>(o, nv);
    put_vec(o, ctype);
    put_vec(o, cid);
    put_vec(o, x);
    put_vec(o, v);
    put_vec(o, fv);
    put_vec(o, fr);
    std::vector<uint8_t> removed((size_t)nv, 0);
    if (nv > 0) check(hc_get_removed_vertices(ctx, removed.data()), "saveCheckPoint: removed vertices");
    put_vec(o, removed);
    if (!o) throw HemoCellError("saveCheckPoint: cannot write " + dir + "particleField.dat");
  }

  // internalViscosity.dat
  if (iv) {
    const int64_t N = hc_num_nodes(ctx);
    std::vector<double> tau((size_t)N);
    check(hc_get_interior_viscosity(ctx, tau.data()), "saveCheckPoint: interior viscosity");
    std::ofstream o(dir + "internalViscosity.dat", std::ios::binary);
    o.write(kViscosityTag, 8);
    put<int64_t>(o, it);
    put<int64_t>(o, N);
    put_vec(o, tau);
    if (!o) throw HemoCellError("saveCheckPoint: cannot write " + dir + "internalViscosity.dat");
  }
}

// core/hemoCell.cpp:197-207, core/hemoCellFields.cpp:240-268
void HemoCell::loadCheckPoint() {
  hlog << "(HemoCell) (Saving Functions) Loading Checkpoint" << std::endl;
  if (!cellfields) throw HemoCellError("loadCheckPoint: initializeCellfield and addCellType first");
  int64_t it_xml = -1;
  if (cfg->checkpointed) {
    XMLElement g = cfg->document()["Checkpoint"]["General"];
    it_xml = g["Iteration"].read<int64_t>();
    std::string d = g["OutDirectory"].read<std::string>();
    if (!d.empty() && d.back() != '/') d += "/";
    outDir = d;
    mkpath(outDir + "hdf5/");
  }
  const std::string dir = checkpointDirectory();
  bindLattice();
  if (hc_num_cells(ctx) != 0) throw HemoCellError("loadCheckPoint: cells are already loaded");

  int64_t it_lat = 0;
  {
    const std::string file = dir + "lattice.dat";
    std::ifstream i(file, std::ios::binary);
    if (!i.is_open()) throw HemoCellError("loadCheckPoint: cannot open " + file);
    check_tag(i, kLatticeTag, file);
    const int64_t nx = get<int64_t>(i, file), ny = get<int64_t>(i, file), nz = get<int64_t>(i, file);
    it_lat = get<int64_t>(i, file);
    (void)get<int32_t>(i, file);  // precision of the run that wrote it
    const int32_t q = get<int32_t>(i, file);
    if (nx != lattice->getNx() || ny != lattice->getNy() || nz != lattice->getNz() || q != 19)
      throw HemoCellError("loadCheckPoint: " + file + " holds a lattice of another size");
    std::vector<double> f((size_t)(19 * nx * ny * nz));
    get_vec(i, f, file);
    check(hc_set_populations(ctx, f.data()), "loadCheckPoint: populations");
  }
  {
    const std::string file = dir + "particleField.dat";
    std::ifstream i(file, std::ios::binary);
    if (!i.is_open()) throw HemoCellError("loadCheckPoint: cannot open " + file);
    check_tag(i, kParticleTag, file);
    const int64_t it_par = get<int64_t>(i, file);
    if (it_par != it_lat) throw HemoCellError("loadCheckPoint: lattice.dat and particleField.dat are of different steps");
    const int32_t ntypes = get<int32_t>(i, file);
    if (ntypes != (int32_t)cellfields->fields.size())
      throw HemoCellError("loadCheckPoint: the run registers another number of cell types than the checkpoint");
    for (int t = 0; t < ntypes; ++t) {
      const int32_t nvt = get<int32_t>(i, file);
      const int32_t len = get<int32_t>(i, file);
      if (len < 0 || len > 4096) throw HemoCellError("loadCheckPoint: " + file + " is corrupt");
      std::string name((size_t)len, ' ');
      if (len && !i.read(&name[0], len)) throw HemoCellError("loadCheckPoint: " + file + " is truncated");
      const HemoCellField& F = *cellfields->fields[(size_t)t];
      if (name != F.name || nvt != F.numVertex)
        throw HemoCellError("loadCheckPoint: cell type " + std::to_string(t) + " is " + F.name +
                            " here and " + name + " in the checkpoint");
    }
    const int32_t nc = get<int32_t>(i, file);
    const int64_t nv = get<int64_t>(i, file);
    if (nc < 0 || nv < 0) throw HemoCellError("loadCheckPoint: " + file + " is corrupt");
    std::vector<int32_t> ctype((size_t)nc), cid((size_t)nc);
    get_vec(i, ctype, file);
    get_vec(i, cid, file);
    int64_t sum = 0;
    for (int c = 0; c < nc; ++c) {
      if (ctype[(size_t)c] < 0 || ctype[(size_t)c] >= ntypes) throw HemoCellError("loadCheckPoint: " + file + " is corrupt");
      sum += cellfields->fields[(size_t)ctype[(size_t)c]]->numVertex;
    }
    if (sum != nv) throw HemoCellError("loadCheckPoint: " + file + " is corrupt");
    std::vector<double> x((size_t)(3 * nv)), v((size_t)(3 * nv)), fv((size_t)(3 * nv)), fr((size_t)(3 * nv));
    get_vec(i, x, file);
    get_vec(i, v, file);
    get_vec(i, fv, file);
    get_vec(i, fr, file);
    std::vector<uint8_t> removed((size_t)nv, 0);
    if (i.peek() != std::char_traits<char>::eof()) get_vec(i, removed, file);  // absent in older files
    // cells go back in their saved order, one hc_add_cells per run of one type
    int64_t off = 0;
    for (int c = 0; c < nc;) {
      int e = c;
      while (e < nc && ctype[(size_t)e] == ctype[(size_t)c]) ++e;
      const HemoCellField& F = *cellfields->fields[(size_t)ctype[(size_t)c]];
      check(hc_add_cells(ctx, F.type_id, e - c, x.data() + 3 * off, cid.data() + c), "loadCheckPoint: cells");
      for (int k = c; k < e; ++k) {
        cellfields->cell_ids.push_back(cid[(size_t)k]);
        cellfields->cell_types.push_back(ctype[(size_t)k]);
      }
      off += (int64_t)(e - c) * F.numVertex;
      c = e;
    }
    cellfields->number_of_cells = nc;
    if (nv > 0) {
      check(hc_set_vertices(ctx, x.data(), v.data(), fv.data()), "loadCheckPoint: vertices");
      check(hc_set_repulsion_force(ctx, fr.data()), "loadCheckPoint: repulsion force");
      check(hc_set_removed_vertices(ctx, removed.data()), "loadCheckPoint: removed vertices");
    }
  }
  bool iv = false;
  for (auto& F : cellfields->fields) iv = iv || F->doInteriorViscosity;
  if (iv) {
    // helper/interiorViscosity.cpp:107-133: the field must exist when interior viscosity is on
    const std::string file = dir + "internalViscosity.dat";
    std::ifstream i(file, std::ios::binary);
    if (!i.is_open()) throw HemoCellError("loadCheckPoint: interior viscosity is on but " + file + " does not exist");
    check_tag(i, kViscosityTag, file);
    const int64_t it_iv = get<int64_t>(i, file), N = get<int64_t>(i, file);
    if (it_iv != it_lat) throw HemoCellError("loadCheckPoint: internalViscosity.dat is of another step");
    if (N != hc_num_nodes(ctx)) throw HemoCellError("loadCheckPoint: " + file + " holds a lattice of another size");
    std::vector<double> tau((size_t)N);
    get_vec(i, tau, file);
    check(hc_set_interior_viscosity_field(ctx, tau.data()), "loadCheckPoint: interior viscosity");
  }
  const int64_t it = it_xml >= 0 ? it_xml : it_lat;
  check(hc_set_iteration(ctx, it), "loadCheckPoint: iteration");
  iter = (unsigned int)it;
  loadParticlesIsCalled = true;
}

}  // namespace hemo
