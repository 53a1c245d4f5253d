// hemocell.cpp -- hemo::HemoCell host layer over libhcgpu (see hemocell.h).
#include "hemocell.h"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <unordered_map>

namespace hemo {

static const T kPi = 3.14159265358979323846;

// ============================================================ XML config ====
namespace {

// A small XML reader for HemoCell configuration files: elements, text,
// comments, processing instructions and attributes (ignored).
class XmlReader {
 public:
  explicit XmlReader(const std::string& s) : s_(s) {}
  std::unique_ptr<XMLNode> document() {
    auto doc = std::make_unique<XMLNode>();
    doc->name = "#document";
    while (true) {
      skip_misc();
      if (i_ >= s_.size()) break;
      if (s_[i_] != '<') throw std::invalid_argument("XML: text outside the root element");
      doc->children.push_back(element());
    }
    return doc;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;

  bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
  void skip_ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  void skip_until(const char* end) {
    size_t k = s_.find(end, i_);
    if (k == std::string::npos) throw std::invalid_argument(std::string("XML: unterminated construct, missing ") + end);
    i_ = k + std::strlen(end);
  }
  // comments, <?...?> and <!DOCTYPE ...> between elements
  void skip_misc() {
    while (true) {
      skip_ws();
      if (starts("<!--")) skip_until("-->");
      else if (starts("<?")) skip_until("?>");
      else if (starts("<!")) skip_until(">");
      else return;
    }
  }
  static std::string decode(const std::string& t) {
    std::string r;
    for (size_t k = 0; k < t.size(); ++k) {
      if (t[k] == '&') {
        static const char* ent[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"}, {"&quot;", "\""}, {"&apos;", "'"}};
        bool hit = false;
        for (auto& e : ent)
          if (t.compare(k, std::strlen(e[0]), e[0]) == 0) {
            r += e[1];
            k += std::strlen(e[0]) - 1;
            hit = true;
            break;
          }
        if (hit) continue;
      }
      r += t[k];
    }
    return r;
  }
  std::unique_ptr<XMLNode> element() {
    auto n = std::make_unique<XMLNode>();
    ++i_;  // '<'
    size_t b = i_;
    while (i_ < s_.size() && !std::isspace((unsigned char)s_[i_]) && s_[i_] != '>' && s_[i_] != '/') ++i_;
    n->name = s_.substr(b, i_ - b);
    if (n->name.empty()) throw std::invalid_argument("XML: empty element name");
    // attributes are skipped (quoted values may hold '>')
    while (i_ < s_.size() && s_[i_] != '>' && !(s_[i_] == '/' && i_ + 1 < s_.size() && s_[i_ + 1] == '>')) {
      if (s_[i_] == '"' || s_[i_] == '\'') {
        const char q = s_[i_++];
        while (i_ < s_.size() && s_[i_] != q) ++i_;
      }
      ++i_;
    }
    if (i_ >= s_.size()) throw std::invalid_argument("XML: unterminated tag <" + n->name);
    if (s_[i_] == '/') {
      i_ += 2;
      return n;
    }
    ++i_;  // '>'
    std::string text;
    while (true) {
      if (i_ >= s_.size()) throw std::invalid_argument("XML: element <" + n->name + "> not closed");
      if (starts("<!--")) {
        skip_until("-->");
      } else if (starts("<![CDATA[")) {
        size_t k = s_.find("]]>", i_);
        if (k == std::string::npos) throw std::invalid_argument("XML: unterminated CDATA");
        text += s_.substr(i_ + 9, k - i_ - 9);
        i_ = k + 3;
      } else if (starts("<?")) {
        skip_until("?>");
      } else if (starts("</")) {
        i_ += 2;
        size_t e = s_.find('>', i_);
        if (e == std::string::npos) throw std::invalid_argument("XML: unterminated closing tag");
        std::string close = s_.substr(i_, e - i_);
        while (!close.empty() && std::isspace((unsigned char)close.back())) close.pop_back();
        if (close != n->name) throw std::invalid_argument("XML: <" + n->name + "> closed by </" + close + ">");
        i_ = e + 1;
        break;
      } else if (s_[i_] == '<') {
        n->children.push_back(element());
      } else {
        size_t e = s_.find('<', i_);
        if (e == std::string::npos) e = s_.size();
        text += s_.substr(i_, e - i_);
        i_ = e;
      }
    }
    // trim
    size_t a = 0, z = text.size();
    while (a < z && std::isspace((unsigned char)text[a])) ++a;
    while (z > a && std::isspace((unsigned char)text[z - 1])) --z;
    n->text = decode(text.substr(a, z - a));
    return n;
  }
};

bool file_exists(const std::string& p) {
  std::ifstream f(p);
  return f.good();
}

}  // namespace

XMLElement XMLElement::operator[](std::string name) const {
  const XMLNode* c = orig ? orig->child(name) : nullptr;
  if (!c) throw std::invalid_argument("XML child " + name + " does not exist.");
  return XMLElement(c);
}

Config::Config(std::string path) {
  if (!file_exists(path)) {
    std::cout << path + " is not an existing config file, exiting ..." << std::endl;
    std::exit(1);
  }
  std::ifstream f(path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  doc = XmlReader(text).document();
  // a checkpoint file wraps the configuration in <Checkpoint> (config/config.cpp:52-58)
  checkpointed = !doc->children.empty() && doc->children.front()->name == "Checkpoint";
}

const XMLNode* Config::hemocellNode() const {
  if (checkpointed) {
    const XMLNode* c = doc->child("Checkpoint");
    return c ? c->child("hemocell") : nullptr;
  }
  return doc->child("hemocell");
}

XMLElement Config::operator[](std::string name) const {
  if (checkpointed) return XMLElement(doc->child("Checkpoint"))["hemocell"][name];
  const XMLNode* root = doc->child("hemocell");
  if (!root) throw std::invalid_argument("XML child hemocell does not exist.");
  return XMLElement(root)[name];
}

// ============================================================ Parameters ====
T Parameters::dt = 0, Parameters::dx = 0, Parameters::dm = 0, Parameters::df = 0;
T Parameters::nu_p = 0, Parameters::rho_p = 0, Parameters::kBT_p = 0;
T Parameters::tau = 0, Parameters::re = 0;
T Parameters::nu_lbm = 0, Parameters::u_lbm_max = 0, Parameters::kBT_lbm = 0, Parameters::shearrate_lbm = 0;
T Parameters::f_limit = 0;
T Parameters::ef_lbm = 0;
T Parameters::pipe_radius = 0;

hc_units Parameters::units() {
  hc_units u{};
  u.dx = dx;
  u.dt = dt;
  u.nu_p = nu_p;
  u.rho_p = rho_p;
  u.kBT_p = kBT_p;
  u.tau = tau;
  u.nu_lbm = nu_lbm;
  u.dm = dm;
  u.df = df;
  u.f_limit = f_limit;
  u.kBT_lbm = kBT_lbm;
  return u;
}

// constantConversion.cpp:36-59
void Parameters::lbm_base_parameters(Config& cfg) {
  hc_units u{};
  u.dt = cfg["domain"]["dt"].read<T>();
  u.dx = cfg["domain"]["dx"].read<T>();
  u.nu_p = cfg["domain"]["nuP"].read<T>();
  u.rho_p = cfg["domain"]["rhoP"].read<T>();
  u.kBT_p = cfg["domain"]["kBT"].read<T>();
  if (hc_lbm_base_parameters(&u) != HC_OK) throw HemoCellError("invalid <domain> parameters (dx, dt, nuP, rhoP, kBT)");
  if (u.dt != cfg["domain"]["dt"].read<T>()) std::cout << "(HemoCell) dt is set to *auto*. Tau will be set to 1!" << std::endl;
  dt = u.dt;
  dx = u.dx;
  nu_p = u.nu_p;
  rho_p = u.rho_p;
  kBT_p = u.kBT_p;
  tau = u.tau;
  nu_lbm = u.nu_lbm;
  dm = u.dm;
  df = u.df;
  f_limit = u.f_limit;
  kBT_lbm = u.kBT_lbm;
}

// constantConversion.cpp:78-86
void Parameters::lbm_pipe_parameters(Config& cfg, int nY) {
  lbm_base_parameters(cfg);
  re = cfg["domain"]["Re"].read<T>();
  pipe_radius = nY;
  u_lbm_max = re * nu_lbm / (pipe_radius * 2);
}

// constantConversion.cpp:61-72
void Parameters::lbm_pipe_parameters(Config& cfg, const Lattice& L) {
  lbm_base_parameters(cfg);
  re = cfg["domain"]["Re"].read<T>();
  T area = 0;
  for (plint y = 0; y < L.getNy(); ++y)
    for (plint z = 0; z < L.getNz(); ++z) area += L.isBoundary(0, y, z) ? 0 : 1;
  pipe_radius = std::sqrt(area / M_PI);
  hlog << "(Parameters) The channel has a calculated radius of " << pipe_radius
       << " LU, assuming a perfect circle cross-section." << std::endl;
  u_lbm_max = re * nu_lbm / (pipe_radius * 2);
}

// constantConversion.cpp:88-94
void Parameters::lbm_shear_parameters(Config& cfg, T nx) {
  lbm_base_parameters(cfg);
  const T shearrate_p = cfg["domain"]["shearrate"].read<T>();
  re = (nx * (shearrate_p * (nx * 0.5))) / nu_p;
  shearrate_lbm = shearrate_p * dt;
  u_lbm_max = shearrate_lbm;
}

void Parameters::printParameters() {
  std::cout << "(HemoCell) System parameters:\n\t dx: \t" << dx << "\n\t dt: \t" << dt << "\n\t dm: \t" << dm
            << "\n\t dN: \t" << df << "\n\t tau: \t" << tau << "\n\t nu_lbm: \t" << nu_lbm << "\n\t u_lb_max: \t"
            << u_lbm_max << "\n\t f_limit: \t" << f_limit << std::endl;
}

// =============================================================== Lattice ====
namespace {
// velocity nodes keyed by node index -> imposed velocity; mask value 2 marks
// them until the lattice is bound and each gets its face
std::unordered_map<const Lattice*, std::unordered_map<plint, Array<T, 3>>>& velocity_nodes() {
  static std::unordered_map<const Lattice*, std::unordered_map<plint, Array<T, 3>>> m;
  return m;
}
struct LatticeForce {
  Array<T, 3> g;
  bool set = false;
};
std::unordered_map<const Lattice*, LatticeForce>& lattice_forces() {
  static std::unordered_map<const Lattice*, LatticeForce> m;
  return m;
}

Box3D clip(const Lattice& L, Box3D b) {
  b.x0 = std::max<plint>(b.x0, 0);
  b.y0 = std::max<plint>(b.y0, 0);
  b.z0 = std::max<plint>(b.z0, 0);
  b.x1 = std::min<plint>(b.x1, L.getNx() - 1);
  b.y1 = std::min<plint>(b.y1, L.getNy() - 1);
  b.z1 = std::min<plint>(b.z1, L.getNz() - 1);
  return b;
}

// the domain face a velocity node belongs to: the first of x-, x+, y-, y+,
// z-, z+ that it lies on along a non-periodic axis, -1 for none
int node_face(Lattice& L, plint x, plint y, plint z) {
  const plint c[3] = {x, y, z};
  const plint n[3] = {L.getNx(), L.getNy(), L.getNz()};
  for (int a = 0; a < 3; ++a) {
    if (L.periodicity().get(a)) continue;
    if (c[a] == 0) return 2 * a;
    if (c[a] == n[a] - 1) return 2 * a + 1;
  }
  return -1;
}
}  // namespace

Lattice::Lattice(plint nx_, plint ny_, plint nz_) : nx(nx_), ny(ny_), nz(nz_) {
  if (nx < 1 || ny < 1 || nz < 1) throw HemoCellError("lattice sizes must be >= 1");
  mask.assign((size_t)(nx * ny * nz), HC_NODE_FLUID);
}

void Lattice::collideAndStream() {
  if (!owner) throw HemoCellError("Lattice::collideAndStream: the lattice is not attached to a HemoCell");
  owner->bindLattice();
  LatticeForce& lf = lattice_forces()[this];
  owner->check(hc_set_body_force(owner->context(), lf.g.data()), "hc_set_body_force");
  owner->check(hc_stage_collide_stream(owner->context()), "collideAndStream");
}

void defineBounceBack(Lattice& L, Box3D box) {
  box = clip(L, box);
  for (plint x = box.x0; x <= box.x1; ++x)
    for (plint y = box.y0; y <= box.y1; ++y)
      for (plint z = box.z0; z <= box.z1; ++z) {
        velocity_nodes()[&L].erase(L.index(x, y, z));
        L.setNodeType(x, y, z, HC_NODE_BOUNCEBACK);
      }
}

void OnLatticeBoundaryCondition3D::setVelocityConditionOnBlockBoundaries(Lattice& L) {
  setVelocityConditionOnBlockBoundaries(L, L.getBoundingBox());
}

void OnLatticeBoundaryCondition3D::setVelocityConditionOnBlockBoundaries(Lattice& L, Box3D box) {
  box = clip(L, box);
  auto& vn = velocity_nodes()[&L];
  for (plint x = box.x0; x <= box.x1; ++x)
    for (plint y = box.y0; y <= box.y1; ++y)
      for (plint z = box.z0; z <= box.z1; ++z) {
        const bool on_face = x == 0 || y == 0 || z == 0 || x == L.getNx() - 1 || y == L.getNy() - 1 ||
                             z == L.getNz() - 1;
        if (!on_face) continue;
        L.setNodeType(x, y, z, HC_NODE_VELOCITY);
        vn.emplace(L.index(x, y, z), Array<T, 3>{0.0, 0.0, 0.0});
      }
}

void setBoundaryVelocity(Lattice& L, Box3D box, Array<T, 3> u) {
  box = clip(L, box);
  auto& vn = velocity_nodes()[&L];
  for (plint x = box.x0; x <= box.x1; ++x)
    for (plint y = box.y0; y <= box.y1; ++y)
      for (plint z = box.z0; z <= box.z1; ++z) {
        auto it = vn.find(L.index(x, y, z));
        if (it != vn.end()) it->second = u;
      }
  if (L.owner && L.owner->context() && hc_num_nodes(L.owner->context()) > 0) {
    // already on the device: update the faces of the nodes in the box
    for (plint x = box.x0; x <= box.x1; ++x)
      for (plint y = box.y0; y <= box.y1; ++y)
        for (plint z = box.z0; z <= box.z1; ++z)
          if (vn.count(L.index(x, y, z))) {
            const int f = node_face(L, x, y, z);
            if (f >= 0) {
              for (int c = 0; c < 3; ++c) L.wall_u[f][c] = u[c];
              L.owner->check(hc_set_wall_velocity(L.owner->context(), f, u.data()), "setBoundaryVelocity");
            }
          }
  }
}

void setExternalVector(Lattice& L, Box3D box, int, Array<T, 3> force) {
  const Box3D all = L.getBoundingBox();
  if (box.x0 > all.x0 || box.y0 > all.y0 || box.z0 > all.z0 || box.x1 < all.x1 || box.y1 < all.y1 || box.z1 < all.z1)
    throw HemoCellError("setExternalVector: only a force on the whole lattice is supported");
  LatticeForce& lf = lattice_forces()[&L];
  lf.g = force;
  lf.set = true;
}

// helper/hemocellInit.hh:58-92
void iniLatticeSquareCouette(Lattice& lattice, plint nx, plint ny, plint nz, OnLatticeBoundaryCondition3D& bc,
                             T shearRate) {
  const Box3D top(0, nx - 1, 0, ny - 1, nz - 1, nz - 1);
  const Box3D bottom(0, nx - 1, 0, ny - 1, 0, 0);
  lattice.periodicity().toggle(0, true);
  lattice.periodicity().toggle(1, true);
  lattice.periodicity().toggle(2, false);
  bc.setVelocityConditionOnBlockBoundaries(lattice, top);
  bc.setVelocityConditionOnBlockBoundaries(lattice, bottom);
  const T vHalf = (nz - 1) * shearRate * 0.5;
  setBoundaryVelocity(lattice, top, Array<T, 3>{-vHalf, 0.0, 0.0});
  setBoundaryVelocity(lattice, bottom, Array<T, 3>{vHalf, 0.0, 0.0});
  setExternalVector(lattice, lattice.getBoundingBox(), 0, Array<T, 3>{0.0, 0.0, 0.0});
  lattice.initialize();
}

// ============================================================ cell types ====
namespace {
// vertex pairs opposite through the mesh centre, greedy in vertex order
std::vector<int32_t> antipodal_pairs(const std::vector<T>& v, size_t count) {
  const size_t nv = v.size() / 3;
  T m[3] = {0, 0, 0};
  for (size_t i = 0; i < nv; ++i)
    for (int c = 0; c < 3; ++c) m[c] += v[3 * i + c];
  for (int c = 0; c < 3; ++c) m[c] /= (T)nv;
  std::vector<T> w(v.size());
  for (size_t i = 0; i < nv; ++i)
    for (int c = 0; c < 3; ++c) w[3 * i + c] = v[3 * i + c] - m[c];
  std::vector<char> used(nv, 0);
  std::vector<int32_t> out;
  for (size_t i = 0; i < nv && out.size() / 2 < count; ++i) {
    if (used[i]) continue;
    size_t best = 0;
    T bd = INFINITY;
    for (size_t j = 0; j < nv; ++j) {
      T d2 = 0;
      for (int c = 0; c < 3; ++c) d2 += (w[3 * j + c] + w[3 * i + c]) * (w[3 * j + c] + w[3 * i + c]);
      const T d = std::sqrt(d2);
      if (d < bd) {
        bd = d;
        best = j;
      }
    }
    const T ni = std::sqrt(w[3 * i] * w[3 * i] + w[3 * i + 1] * w[3 * i + 1] + w[3 * i + 2] * w[3 * i + 2]);
    if (best != i && !used[best] && bd < 1e-6 * ni + 1e-9) {
      out.push_back((int32_t)i);
      out.push_back((int32_t)best);
      used[i] = used[best] = 1;
    }
  }
  return out;
}

T read_or(Config& cfg, const char* key, T dflt) {
  try {
    return cfg["MaterialModel"][key].read<T>();
  } catch (std::invalid_argument&) {
    return dflt;
  }
}
}  // namespace

HemoCellField::HemoCellField(HemoCellFields& cf, const std::string& name_, int ctype_, int constructType_, int model)
    : name(name_), ctype(ctype_), constructType(constructType_), cellFields(cf) {
  if (constructType != WBC_SPHERE && constructType != RBC_FROM_SPHERE && constructType != ELLIPSOID_FROM_SPHERE) {
    std::cout << "(HemoCell) (AddCellType) construct type " << constructType << " is not supported" << std::endl;
    std::exit(1);
  }
  materialCfg = std::make_unique<Config>(name + ".xml");
  Config& m = *materialCfg;
  if (param::dx <= 0) throw HemoCellError("addCellType: call param::lbm_base_parameters before adding cell types");
  T aspect = 0.3;
  if (constructType == ELLIPSOID_FROM_SPHERE) aspect = m["MaterialModel"]["aspectRatio"].read<T>();
  bool has_stl = true;
  try {
    (void)m["MaterialModel"]["StlFile"];
  } catch (std::invalid_argument&) {
    has_stl = false;
  }
  if (has_stl) {
    std::cout << "(HemoCell) (AddCellType) STL cell meshes are not supported; use minNumTriangles" << std::endl;
    std::exit(1);
  }
  const T radius_lu = m["MaterialModel"]["radius"].read<T>() / param::dx;
  const int min_tri = (int)m["MaterialModel"]["minNumTriangles"].read<T>();
  int nv = 0, nt = 0;
  if (hc_build_mesh(constructType, radius_lu, min_tri, aspect, &nv, &nt, nullptr, nullptr) != HC_OK)
    throw HemoCellError("mesh construction failed for " + name);
  rest.resize((size_t)nv * 3);
  triangles.resize((size_t)nt * 3);
  hc_build_mesh(constructType, radius_lu, min_tri, aspect, &nv, &nt, rest.data(), triangles.data());
  numVertex = nv;
  for (int t = 0; t < nt; ++t) triangle_list.push_back({triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]});
  // rest metrics (helper/meshMetrics.h:176-222)
  meshmetric = std::make_unique<MeshMetrics>();
  for (int t = 0; t < nt; ++t) {
    const T* a = &rest[3 * triangles[3 * t]];
    const T* b = &rest[3 * triangles[3 * t + 1]];
    const T* c = &rest[3 * triangles[3 * t + 2]];
    const T e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const T n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    meshmetric->area += 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    meshmetric->volume += (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                           a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
  }
  // inner links: as many as the material file lists, regenerated in this mesh's
  // numbering (the file's vertex ids follow Palabos's numbering)
  size_t n_inner = 0;
  try {
    n_inner = m["MaterialModel"]["InnerEdges"].getOrig()->children.size();
  } catch (std::invalid_argument&) {
  }
  if (model == HC_MODEL_WBC && n_inner == 0) n_inner = (size_t)nv;  // every antipodal pair
  if (n_inner) inner_edges = antipodal_pairs(rest, n_inner);
  hc_material mat{};
  mat.kLink = m["MaterialModel"]["kLink"].read<T>();
  mat.kBend = m["MaterialModel"]["kBend"].read<T>();
  mat.kVolume = m["MaterialModel"]["kVolume"].read<T>();
  mat.kArea = m["MaterialModel"]["kArea"].read<T>();
  mat.eta_m = m["MaterialModel"]["eta_m"].read<T>();
  mat.kInnerRigid = read_or(m, "kInnerRigid", 0.0);
  mat.kCytoskeleton = read_or(m, "kCytoskeleton", 0.0);
  mat.coreRadius = read_or(m, "coreRadius", 0.0);
  mat.radius = read_or(m, "radius", 0.0);
  const hc_units u = param::units();
  if (hc_model_constants(model, &mat, nt, &u, &params) != HC_OK)
    throw HemoCellError("model constants failed for " + name);
  hc_ctx* ctx = cf.hemocell.context();
  cf.hemocell.check(hc_add_cell_type(ctx, model, nv, nt, triangles.data(), rest.data(), (int)inner_edges.size() / 2,
                                     inner_edges.empty() ? nullptr : inner_edges.data(), &params, 1, &type_id),
                    "addCellType");
  // interior viscosity (core/hemoCellField.cpp:99-110): tau inside the membrane
  // from the viscosity ratio, on when the material enables it
  interiorViscosityTau = read_or(m, "viscosityRatio", 1.0) * (param::tau - 0.5) + 0.5;
  try {
    doInteriorViscosity = m["MaterialModel"]["enableInteriorViscosity"].read<bool>();
  } catch (std::invalid_argument&) {
  }
  if (doInteriorViscosity) {
    std::cout << "(HemoCell) (AddCellType) (" << name << ") Enabling interior viscosity" << std::endl;
    cf.hemocell.check(hc_set_interior_viscosity(ctx, type_id, interiorViscosityTau), "interior viscosity");
  }
}

HemoCellField* HemoCellFields::operator[](const std::string& name) {
  for (auto& f : fields)
    if (f->name == name) return f.get();
  throw std::invalid_argument("no cell type named " + name);
}

HemoCellField* HemoCellFields::addCellType(const std::string& name, int constructType, int model) {
  if (hemocell.lattice == nullptr) throw HemoCellError("addCellType: set HemoCell::lattice first");
  hemocell.bindLattice();
  fields.push_back(std::make_unique<HemoCellField>(*this, name, (int)fields.size(), constructType, model));
  return fields.back().get();
}

// ============================================================== HemoCell ====
HemoCell::HemoCell(char* configFileName, int, char*[], MPIHandle) {
  cfg = new Config(configFileName);
  try {
    std::string p = (*cfg)["domain"]["precision"].read<std::string>();
    if (p == "fp32" || p == "FP32") precision = HC_FP32_SHIFTED;
  } catch (std::invalid_argument&) {
  }
  if (const char* e = std::getenv("HEMOCELL_PRECISION")) {
    if (std::string(e) == "fp32") precision = HC_FP32_SHIFTED;
    if (std::string(e) == "fp64") precision = HC_FP64;
  }
  int device = 0;
  if (const char* d = std::getenv("HEMOCELL_DEVICE")) device = std::atoi(d);
  if (hc_create(device, &ctx) != HC_OK) throw HemoCellError("no usable GPU for HemoCell");
}

HemoCell::~HemoCell() {
  if (lattice) {
    velocity_nodes().erase(lattice);
    lattice_forces().erase(lattice);
  }
  delete cellfields;
  delete cfg;
  if (ctx) hc_destroy(ctx);
}

void HemoCell::check(hc_status s, const char* what) const {
  if (s != HC_OK) throw HemoCellError(std::string(what) + ": " + hc_last_error(ctx));
}

void HemoCell::bindLattice() {
  if (bound) return;
  if (!lattice) throw HemoCellError("HemoCell::lattice is not set");
  if (param::tau <= 0.5) throw HemoCellError("call param::lbm_base_parameters (or a variant) before using the lattice");
  lattice->owner = this;
  Lattice& L = *lattice;
  std::vector<uint8_t> mask = L.mask;
  bool face_set[6] = {false, false, false, false, false, false};
  T face_u[6][3] = {{0}};
  for (auto& kv : velocity_nodes()[&L]) {
    const plint n = kv.first;
    const plint z = n % L.getNz(), y = (n / L.getNz()) % L.getNy(), x = n / (L.getNy() * L.getNz());
    const int f = node_face(L, x, y, z);
    if (f < 0) throw HemoCellError("a velocity boundary node lies on no non-periodic domain face");
    mask[n] = (uint8_t)(HC_NODE_VELOCITY + f);
    for (int c = 0; c < 3; ++c) {
      if (face_set[f] && face_u[f][c] != kv.second[c])
        throw HemoCellError("velocity nodes of one domain face need one velocity");
      face_u[f][c] = kv.second[c];
    }
    face_set[f] = true;
  }
  const int periodic[3] = {L.periodicity().get(0), L.periodicity().get(1), L.periodicity().get(2)};
  check(hc_lattice_init(ctx, (int)L.getNx(), (int)L.getNy(), (int)L.getNz(), mask.data(), param::tau, periodic,
                        precision),
        "lattice");
  for (int f = 0; f < 6; ++f)
    if (face_set[f]) {
      check(hc_set_wall_velocity(ctx, f, face_u[f]), "wall velocity");
      for (int c = 0; c < 3; ++c) L.wall_u[f][c] = face_u[f][c];
    }
  check(hc_set_force_limit(ctx, param::f_limit), "force limit");
  if (repulsionEnabled) check(hc_set_repulsion(ctx, repulsionConstant, repulsionCutoff, (int)repulsionTimescale), "repulsion");
  if (boundaryRepulsionEnabled)
    check(hc_set_boundary_repulsion(ctx, boundaryRepulsionConstant, boundaryRepulsionCutoff,
                                    (int)boundaryRepulsionTimescale),
          "boundary repulsion");
  if (L.leesEdwards) {
    check(hc_set_lees_edwards(ctx, L.leesEdwardsShear, L.leesEdwardsDisplacement), "Lees-Edwards");
    leesEdwardsBC = true;
  }
  bound = true;
}

// helper/leesEdwardsBC.h:301-364
double LeesEdwardsBC::LEcurrentDisplacement = 0.0;

LeesEdwardsBC::LeesEdwardsBC(Lattice& l, T shearRate, T dt_, double** hemoCellLEcurrentDisplacement,
                             plint dataProcessorLevel_)
    : lattice(l), nx(l.getNx()), ny(l.getNy()), nz(l.getNz()), LEdisplacement(shearRate * dt_), dt(dt_),
      dataProcessorLevel(dataProcessorLevel_) {
  const T vHalf = (nz - 1) * shearRate * 0.5;
  topVelocity = -vHalf;
  bottomVelocity = vHalf;
  if (hemoCellLEcurrentDisplacement) *hemoCellLEcurrentDisplacement = &LEcurrentDisplacement;
  lattice.leesEdwardsShear = shearRate;
  lattice.leesEdwardsDisplacement = LEdisplacement;
}

void LeesEdwardsBC::initialize() {
  lattice.periodicity().toggleAll(true);
  lattice.leesEdwards = true;
  if (lattice.owner && lattice.owner->context() && hc_num_nodes(lattice.owner->context()) > 0)
    throw HemoCellError("LeesEdwardsBC::initialize: the lattice is already bound; initialize it before first use");
}

void LeesEdwardsBC::updateLECurDisplacement(unsigned int iter) {
  LEcurrentDisplacement = std::fmod(LEdisplacement * iter, (double)nx);
}

// core/hemoCell.cpp:129-133
void HemoCell::latticeEquilibrium(T rho, Array<T, 3> vel) {
  bindLattice();
  check(hc_init_equilibrium(ctx, rho, vel.data()), "latticeEquilibrium");
}

void HemoCell::initializeCellfield() {
  if (!cellfields) cellfields = new HemoCellFields(*this);
}

// core/hemoCell.cpp:404-410
void HemoCell::setInteriorViscosityTimeScaleSeperation(unsigned int separation, unsigned int separation_entire_grid) {
  std::cout << "(HemoCell) (Interior Viscosity Timescale Seperation) Setting seperation to " << separation
            << " timesteps" << std::endl;
  std::cout << "(HemoCell) (Interior Viscosity Timescale Seperation) Setting entire grid refresh to "
            << separation_entire_grid << " timesteps" << std::endl;
  check(hc_set_interior_viscosity_timescales(ctx, (int)separation, (int)separation_entire_grid),
        "setInteriorViscosityTimeScaleSeperation");
}

// core/hemoCell.cpp:419-426: the cut-off is given in micrometres
void HemoCell::setRepulsion(T k, T cutoff_um) {
  repulsionConstant = k;
  repulsionCutoff = cutoff_um * (1e-6 / param::dx);
  repulsionEnabled = true;
  if (bound) check(hc_set_repulsion(ctx, repulsionConstant, repulsionCutoff, (int)repulsionTimescale), "setRepulsion");
}

void HemoCell::enableBoundaryParticles(T k, T cutoff_um, unsigned int timestep) {
  hlog << "(HemoCell) (Repulsion) Setting boundary repulsion constant to " << k << ". boundary repulsionCutoff to"
       << cutoff_um << " µm" << std::endl;
  boundaryRepulsionConstant = k;
  boundaryRepulsionCutoff = cutoff_um * (1e-6 / param::dx);
  boundaryRepulsionTimescale = std::max(1u, timestep);
  boundaryRepulsionEnabled = true;
  if (bound)
    check(hc_set_boundary_repulsion(ctx, boundaryRepulsionConstant, boundaryRepulsionCutoff,
                                    (int)boundaryRepulsionTimescale),
          "enableBoundaryParticles");
}

void HemoCell::setRepulsionTimeScaleSeperation(unsigned int separation) {
  repulsionTimescale = std::max(1u, separation);
  if (bound && repulsionEnabled)
    check(hc_set_repulsion(ctx, repulsionConstant, repulsionCutoff, (int)repulsionTimescale), "setRepulsion");
}

void HemoCell::setMaterialTimeScaleSeparation(std::string name, unsigned int separation) {
  HemoCellField* f = (*cellfields)[name];
  f->timescale = (int)std::max(1u, separation);
  check(hc_set_material_timescale(ctx, f->type_id, f->timescale), "setMaterialTimeScaleSeparation");
}

void HemoCell::setParticleVelocityUpdateTimeScaleSeparation(unsigned int separation) {
  check(hc_set_velocity_timescale(ctx, (int)std::max(1u, separation)), "setParticleVelocityUpdateTimeScaleSeparation");
}

void HemoCell::setInitialMinimumDistanceFromSolid(std::string name, T distance) {
  if (loadParticlesIsCalled)
    std::cout << "(HemoCell) (Set Distance) WARNING: this function is called after the particles are loaded, so it has "
                 "no effect!"
              << std::endl;
  (*cellfields)[name]->minimumDistanceFromSolid = distance;
}

// core/hemoCell.cpp:168-180
void HemoCell::setSystemPeriodicity(unsigned int axis, bool bePeriodic) {
  if (!lattice || axis > 2) throw HemoCellError("setSystemPeriodicity: no lattice or a bad axis");
  lattice->periodicity().toggle((int)axis, bePeriodic);
  if (bound) {
    const int p[3] = {lattice->periodicity().get(0), lattice->periodicity().get(1), lattice->periodicity().get(2)};
    check(hc_set_periodicity(ctx, p), "setSystemPeriodicity");
  }
}

// io/readPositionsBloodCells.cpp:36-361: <name>.pos holds the cell count, then
// per cell its centre (um) and three rotation angles (degrees, applied about
// x, then y, then z around the mesh's bounding-box centre).  Vertices whose
// nearest node is a wall, or lies within the type's minimum distance from a
// wall, are dropped together with their cell (nodes beyond a periodic face by
// their periodic image).  Cell ids run over all types in
// registration order, one per line.
void HemoCell::loadParticles() {
  if (!cellfields) throw HemoCellError("loadParticles: initializeCellfield and addCellType first");
  bindLattice();
  loadParticlesIsCalled = true;
  Lattice& L = *lattice;
  const T posRatio = 1e-6 / param::dx;
  int cellid = 0;
  int total = 0;
  for (auto& fp : cellfields->fields) {
    HemoCellField& F = *fp;
    std::ifstream in(F.name + ".pos");
    if (!in.is_open()) {
      std::cout << "*** WARNING! particle positions input file " << F.name << ".pos does not exist!" << std::endl;
      continue;
    }
    int np = 0;
    in >> np;
    total += np;
    // mesh centred on its bounding box
    const int nv = F.numVertex;
    T lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < nv; ++k)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], F.rest[3 * k + c]);
        hi[c] = std::max(hi[c], F.rest[3 * k + c]);
      }
    std::vector<T> base(F.rest);
    for (int k = 0; k < nv; ++k)
      for (int c = 0; c < 3; ++c) base[3 * k + c] -= (lo[c] + hi[c]) / 2.0;
    const int deny = (int)((F.minimumDistanceFromSolid * 1e-6) / param::dx);
    std::vector<T> xyz;
    std::vector<int32_t> ids;
    for (int i = 0; i < np; ++i) {
      T p[3], ang[3];
      if (!(in >> p[0] >> p[1] >> p[2] >> ang[0] >> ang[1] >> ang[2]))
        throw HemoCellError("malformed line in " + F.name + ".pos");
      for (int c = 0; c < 3; ++c) ang[c] = ang[c] * (kPi / 180.0) * -1.0;
      const int id = cellid++;
      // meshRotation: about the bounding-box centre, X then Y then Z
      T blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int k = 0; k < nv; ++k)
        for (int c = 0; c < 3; ++c) {
          blo[c] = std::min(blo[c], base[3 * k + c]);
          bhi[c] = std::max(bhi[c], base[3 * k + c]);
        }
      const T ctr[3] = {(bhi[0] + blo[0]) * 0.5, (bhi[1] + blo[1]) * 0.5, (bhi[2] + blo[2]) * 0.5};
      const T al = ang[0], be = ang[1], ga = ang[2];
      T a[3][3] = {{1.0, 0.0, 0.0}, {0.0, std::cos(al), std::sin(al)}, {0.0, -std::sin(al), std::cos(al)}};
      T b[3][3] = {{std::cos(be), 0.0, -std::sin(be)}, {0.0, 1.0, 0.0}, {std::sin(be), 0.0, std::cos(be)}};
      T cm[3][3];
      for (int r = 0; r < 3; r++)
        for (int s = 0; s < 3; s++) {
          cm[r][s] = 0.0;
          for (int k = 0; k < 3; k++) cm[r][s] += a[k][s] * b[r][k];
        }
      T bz[3][3] = {{std::cos(ga), std::sin(ga), 0.0}, {-std::sin(ga), std::cos(ga), 0.0}, {0.0, 0.0, 1.0}};
      for (int r = 0; r < 3; r++)
        for (int s = 0; s < 3; s++) {
          a[r][s] = 0.0;
          for (int k = 0; k < 3; k++) a[r][s] += cm[k][s] * bz[r][k];
        }
      std::vector<T> cell((size_t)nv * 3);
      bool complete = true;
      for (int k = 0; k < nv && complete; ++k) {
        T q[3], rq[3];
        for (int c = 0; c < 3; ++c) q[c] = base[3 * k + c] + -1.0 * ctr[c];
        for (int r = 0; r < 3; ++r) {
          rq[r] = 0.0;
          for (int s = 0; s < 3; ++s) rq[r] += a[r][s] * q[s];
        }
        T v[3];
        for (int c = 0; c < 3; ++c) v[c] = p[c] * posRatio + (rq[c] + ctr[c]);
        // nearest node, truncated as int(v + 0.5); on a periodic axis a node
        // beyond the domain is its periodic image (the reference reads it from
        // the lattice envelope, which holds that image)
        const plint nx = (plint)(v[0] + 0.5), ny = (plint)(v[1] + 0.5), nz = (plint)(v[2] + 0.5);
        auto image = [&](plint c, int axis, plint n) -> plint {
          if (c >= 0 && c < n) return c;
          return L.periodicity().get(axis) ? ((c % n) + n) % n : -1;
        };
        auto boundary_at = [&](plint x, plint y, plint z) {
          const plint X = image(x, 0, L.getNx()), Y = image(y, 1, L.getNy()), Z = image(z, 2, L.getNz());
          return X >= 0 && Y >= 0 && Z >= 0 && L.isBoundary(X, Y, Z);
        };
        if (boundary_at(nx, ny, nz)) complete = false;
        for (int dx = -deny; dx <= deny && complete; ++dx)
          for (int dy = -deny; dy <= deny && complete; ++dy)
            for (int dz = -deny; dz <= deny && complete; ++dz)
              if (boundary_at(nx + dx, ny + dy, nz + dz)) complete = false;
        for (int c = 0; c < 3; ++c) cell[3 * k + c] = v[c];
      }
      if (!complete) continue;  // deleteIncompleteCells
      xyz.insert(xyz.end(), cell.begin(), cell.end());
      ids.push_back(id);
    }
    const int n = (int)ids.size();
    if (n) check(hc_add_cells(ctx, F.type_id, n, xyz.data(), ids.data()), "loadParticles");
    for (int k = 0; k < n; ++k) {
      cellfields->cell_ids.push_back(ids[k]);
      cellfields->cell_types.push_back(F.ctype);
    }
  }
  cellfields->number_of_cells = total;
}

void HemoCell::addVertexForces(const std::vector<int64_t>& ids, const std::vector<T>& forces) {
  pending_ids.insert(pending_ids.end(), ids.begin(), ids.end());
  pending_f.insert(pending_f.end(), forces.begin(), forces.end());
}

// core/hemoCell.cpp:299-376; the forces a driver set before this call (vertex
// forces, the lattice force) act on this step only.
void HemoCell::iterate() {
  bindLattice();
  if (pending_ids != extra_ids || pending_f != extra_f) {
    check(hc_set_extra_vertex_force(ctx, (int)pending_ids.size(), pending_ids.data(), pending_f.data()),
          "vertex forces");
    extra_ids = pending_ids;
    extra_f = pending_f;
  }
  LatticeForce& lf = lattice_forces()[lattice];
  check(hc_set_body_force(ctx, lf.g.data()), "body force");
  check(hc_step(ctx, 1), "iterate");
  // setExternalVector(..., 0) at the end of the step
  lf.g = Array<T, 3>{0.0, 0.0, 0.0};
  pending_ids.clear();
  pending_f.clear();
  ++iter;
}

// ============================================================= fluidInfo ====
// helper/fluidInfo.cpp:102-141 (the reference's force gatherer never counts its
// nodes, so its mean divides by zero; here both means divide by the node count)
namespace {
FluidStatistics fluid_statistics(HemoCell* h, int what) {
  h->bindLattice();
  double out[4];
  h->check(hc_get_fluid_stats(h->context(), what, out), "FluidInfo");
  FluidStatistics r;
  r.min = out[0];
  r.max = out[1];
  r.avg = out[2];
  r.ncells = (pluint)out[3];
  return r;
}
}  // namespace

FluidStatistics FluidInfo::calculateVelocityStatistics(HemoCell* hemocell) { return fluid_statistics(hemocell, 0); }
FluidStatistics FluidInfo::calculateForceStatistics(HemoCell* hemocell) { return fluid_statistics(hemocell, 1); }

// ========================================================== particleInfo ====
// helper/particleInfo.cpp:57-150: every vertex of every cell
namespace {
ParticleStatistics particle_statistics(HemoCell* h, bool force) {
  h->bindLattice();
  hc_ctx* ctx = h->context();
  const int64_t nv = hc_num_vertices(ctx);
  ParticleStatistics r;
  if (nv == 0) return r;
  std::vector<double> x((size_t)(3 * nv)), v((size_t)(3 * nv)), f((size_t)(3 * nv)), fr((size_t)(3 * nv), 0.0);
  h->check(hc_get_vertices(ctx, x.data(), v.data(), f.data()), "ParticleInfo");
  if (force && h->repulsionEnabled) h->check(hc_get_repulsion_force(ctx, fr.data()), "ParticleInfo");
  double sum = 0;
  for (int64_t i = 0; i < nv; ++i) {
    double q[3];
    for (int c = 0; c < 3; ++c) q[c] = force ? f[(size_t)(3 * i + c)] + fr[(size_t)(3 * i + c)] : v[(size_t)(3 * i + c)];
    const double m = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    if (i == 0 || m < r.min) r.min = m;
    if (i == 0 || m > r.max) r.max = m;
    sum += m;
  }
  r.avg = sum / (double)nv;
  r.ncells = (pluint)nv;
  return r;
}
}  // namespace

ParticleStatistics ParticleInfo::calculateForceStatistics(HemoCell* hemocell) {
  return particle_statistics(hemocell, true);
}
ParticleStatistics ParticleInfo::calculateVelocityStatistics(HemoCell* hemocell) {
  return particle_statistics(hemocell, false);
}

// ============================================================== cellInfo ====
std::map<int, CellInformation> CellInformationFunctionals::info_per_cell;

// deleteIncompleteCells (core/hemoCellParticleField.cpp:510-553), which the
// cellInfo functions call first (helper/cellInfo.cpp:271-319): every vertex of a
// cell that lost one to a wall is removed.  Returns one flag per cell.
static std::vector<uint8_t> delete_incomplete_cells(HemoCell* h) {
  hc_ctx* ctx = h->context();
  const int nc = hc_num_cells(ctx);
  std::vector<uint8_t> inc((size_t)nc, 0);
  if (nc == 0) return inc;
  h->check(hc_get_incomplete_cells(ctx, inc.data()), "deleteIncompleteCells");
  if (std::find(inc.begin(), inc.end(), 1) == inc.end()) return inc;
  const int64_t nv = hc_num_vertices(ctx);
  std::vector<uint8_t> rm((size_t)nv);
  h->check(hc_get_removed_vertices(ctx, rm.data()), "deleteIncompleteCells");
  int64_t b = 0;
  for (int c = 0; c < nc; ++c) {
    const int n = (*h->cellfields)[(pluint)h->cellfields->cell_types.at(c)]->numVertex;
    if (inc[(size_t)c])
      for (int k = 0; k < n; ++k) rm[(size_t)(b + k)] = 1;
    b += n;
  }
  h->check(hc_set_removed_vertices(ctx, rm.data()), "deleteIncompleteCells");
  return inc;
}

void CellInformationFunctionals::fill(HemoCell* h) {
  hc_ctx* ctx = h->context();
  const int nc = hc_num_cells(ctx);
  const std::vector<uint8_t> gone = h->cellfields ? delete_incomplete_cells(h) : std::vector<uint8_t>((size_t)nc, 0);
  std::vector<T> vol(nc), area(nc), pos(3 * nc), vel(3 * nc), bbox(6 * nc), stretch(nc);
  h->check(hc_get_cell_stats(ctx, vol.data(), area.data(), pos.data(), vel.data(), bbox.data(), stretch.data()),
           "cell information");
  for (int c = 0; c < nc; ++c) {
    if (gone[(size_t)c]) continue;
    const int id = h->cellfields ? h->cellfields->cell_ids.at(c) : c;
    CellInformation& ci = info_per_cell[id];
    ci.volume = vol[c];
    ci.area = area[c];
    ci.stretch = stretch[c];
    for (int k = 0; k < 3; ++k) {
      ci.position[k] = pos[3 * c + k];
      ci.velocity[k] = vel[3 * c + k];
    }
    for (int k = 0; k < 6; ++k) ci.bbox[k] = bbox[6 * c + k];
    ci.cellType = h->cellfields ? h->cellfields->cell_types.at(c) : 0;
  }
}

unsigned int CellInformationFunctionals::getTotalNumberOfCells(HemoCell* hemocell) {
  if (!hemocell->cellfields) return (unsigned)hc_num_cells(hemocell->context());
  const std::vector<uint8_t> gone = delete_incomplete_cells(hemocell);
  return (unsigned)std::count(gone.begin(), gone.end(), 0);
}

unsigned int CellInformationFunctionals::getNumberOfCellsFromType(HemoCell* hemocell, std::string type) {
  const int ct = (*hemocell->cellfields)[type]->ctype;
  const std::vector<uint8_t> gone = delete_incomplete_cells(hemocell);
  unsigned n = 0;
  for (size_t c = 0; c < gone.size(); ++c) n += !gone[c] && hemocell->cellfields->cell_types[c] == ct;
  return n;
}

// ========================================================== cell stretch ====
// helper/hemoCellStretch.cpp:30-110: the single cell's vertices sorted on x by
// a bubble sort (stable), the n first and n last forced.
HemoCellStretch::HemoCellStretch(HemoCellField& cf, unsigned int n, T force) : cellfield(cf) {
  HemoCell& h = cf.cellFields.hemocell;
  if (hc_num_cells(h.context()) != 1) {
    std::cout << "(HemoCellStretch) Refusing to run with more or less than 1 cell" << std::endl;
    std::exit(1);
  }
  n_forced_lsps = n;
  external_force = force / n;
  const int64_t nv = hc_num_vertices(h.context());
  std::vector<T> x((size_t)nv * 3);
  h.check(hc_get_vertices(h.context(), x.data(), nullptr, nullptr), "HemoCellStretch");
  std::vector<int64_t> found((size_t)nv);
  for (int64_t i = 0; i < nv; ++i) found[(size_t)i] = i;
  for (size_t i = 0; i + 1 < found.size(); i++)
    for (size_t j = 1; j < found.size() - i; j++)
      if (x[3 * found[j - 1]] > x[3 * found[j]]) std::swap(found[j - 1], found[j]);
  for (unsigned int i = 0; i < n; i++) {
    lower_lsps.push_back(found[i]);
    upper_lsps.push_back(found[found.size() - 1 - i]);
  }
}

void HemoCellStretch::applyForce() {
  if (cellfield.timescale != 1) {
    std::cout << "Refusing to stretch with particle update timestep larger than 1" << std::endl;
    std::exit(1);
  }
  std::vector<int64_t> ids;
  std::vector<T> f;
  const T ex = external_force * scale;
  for (plint v : lower_lsps) {
    ids.push_back(v);
    f.insert(f.end(), {-ex, 0.0, 0.0});
  }
  for (plint v : upper_lsps) {
    ids.push_back(v);
    f.insert(f.end(), {ex, 0.0, 0.0});
  }
  cellfield.cellFields.hemocell.addVertexForces(ids, f);
}

}  // namespace hemo
