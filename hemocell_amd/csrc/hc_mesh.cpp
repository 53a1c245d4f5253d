// hc_mesh.cpp -- cell rest meshes and the per-type constant tables (host C++).
//
// Mesh construction follows helper/meshGeneratingFunctions.hh:32-271 and
// helper/meshGeneratingFunctions.h:69-94: an icosahedron (WBC sphere, RBC) or an
// octahedron (platelet ellipsoid) refined by midpoint subdivision while it has
// fewer triangles than requested, mapped to the Evans-Fung RBC shape or an
// ellipsoid, de-duplicated into an indexed mesh and inflated by 1e-3 lattice
// units along the vertex normals.  The constant tables follow
// mechanics/commonCellConstants.cpp:70-409 and helper/meshMetrics.h:127-184.
#include <cstring>
#include <array>
#include <cmath>
#include <map>

#include "hc_internal.h"

namespace hcg {
namespace {

using P3 = std::array<double, 3>;
using Tri = std::array<P3, 3>;

double len3(const P3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
P3 diff(const P3& a, const P3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
P3 crs(const P3& a, const P3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dt3(const P3& a, const P3& b) {
  double r = 0.0;
  for (int i = 0; i < 3; ++i) r += a[i] * b[i];
  return r;
}
double norm3(const P3& a) {
  double r = 0.0;
  for (int i = 0; i < 3; ++i) r += a[i] * a[i];
  return std::sqrt(r);
}

// Unit icosahedron, vertex order and face list as in the reference
// (meshGeneratingFunctions.hh:42-97); kept alike because they fix the mesh numbering.
std::vector<Tri> icosahedron() {
  const double t = -0.8506508084, o = -0.5257311121;
  const P3 v[13] = {{0, 0, 0},
                    {t, o, 0},  {-t, o, 0}, {-t, -o, 0}, {t, -o, 0},
                    {o, 0, t},  {o, 0, -t}, {-o, 0, -t}, {-o, 0, t},
                    {0, t, o},  {0, -t, o}, {0, -t, -o}, {0, t, -o}};
  const int f[20][3] = {{5, 8, 9},  {5, 10, 8}, {6, 12, 7}, {6, 7, 11}, {1, 4, 5},
                        {1, 6, 4},  {3, 2, 8},  {3, 7, 2},  {9, 12, 1}, {9, 2, 12},
                        {10, 4, 11}, {10, 11, 3}, {9, 1, 5}, {12, 6, 1}, {5, 4, 10},
                        {6, 11, 4}, {8, 2, 9},  {7, 12, 2}, {8, 10, 3}, {7, 3, 11}};
  std::vector<Tri> out;
  for (auto& q : f) out.push_back({v[q[0]], v[q[1]], v[q[2]]});
  return out;
}

// Unit octahedron with outward (counter-clockwise) faces.
std::vector<Tri> octahedron() {
  const P3 px{1, 0, 0}, mx{-1, 0, 0}, py{0, 1, 0}, my{0, -1, 0}, pz{0, 0, 1}, mz{0, 0, -1};
  return {{px, py, pz}, {py, mx, pz}, {mx, my, pz}, {my, px, pz},
          {py, px, mz}, {mx, py, mz}, {my, mx, mz}, {px, my, mz}};
}

void refine(std::vector<Tri>& tris, int min_triangles) {
  size_t size;
  while ((size = tris.size()) < (size_t)min_triangles) {
    for (size_t i = 0; i < size; ++i) {
      const P3 a = tris[i][0], b = tris[i][1], c = tris[i][2];
      P3 d, e, f;
      for (int k = 0; k < 3; ++k) {
        d[k] = 0.5 * (a[k] + b[k]);
        e[k] = 0.5 * (b[k] + c[k]);
        f[k] = 0.5 * (c[k] + a[k]);
      }
      double nd = len3(d), ne = len3(e), nf = len3(f);
      for (int k = 0; k < 3; ++k) {
        d[k] /= nd;
        e[k] /= ne;
        f[k] /= nf;
      }
      tris[i] = {d, e, f};
      tris.push_back({a, d, f});
      tris.push_back({d, b, e});
      tris.push_back({f, e, c});
    }
  }
}

// z-x-z Euler rotation (phi, theta, psi).
void rotate(std::vector<Tri>& tris, double phi, double theta, double psi) {
  const double cf = std::cos(phi), sf = std::sin(phi), ct = std::cos(theta), st = std::sin(theta),
               cp = std::cos(psi), sp = std::sin(psi);
  const double m[3][3] = {{cp * cf - ct * sf * sp, cp * sf + ct * cf * sp, sp * st},
                          {-sp * cf - ct * sf * cp, -sp * sf + ct * cf * cp, cp * st},
                          {st * sf, -st * cf, ct}};
  for (auto& t : tris)
    for (auto& p : t) {
      P3 q;
      for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += m[i][j] * p[j];
        q[i] = s;
      }
      p = q;
    }
}

int sgn(double v) { return (0.0 < v) - (v < 0.0); }

// Evans-Fung biconcave map of a unit-sphere point (meshGeneratingFunctions.hh:154-168).
P3 rbc_point(P3 p) {
  double r2 = p[0] * p[0] + p[1] * p[1];
  int s = sgn(p[2]);
  if (1 - r2 < 0) r2 = 1;
  const double C0 = 0.054322, C2 = 1.001279, C4 = -0.561381;
  p[2] = s * 1.0 * std::sqrt(1 - r2) * (C0 + C2 * r2 + C4 * r2 * r2);
  return p;
}

P3 ellipsoid_point(P3 p, double R, double aspect) {
  double r2 = p[0] * p[0] + p[1] * p[1];
  int s = sgn(p[2]);
  if (1 - r2 < 0) r2 = 1;
  p[0] *= R;
  p[1] *= R;
  p[2] = s * aspect * R * std::sqrt(1 - r2);
  return p;
}

}  // namespace

bool build_mesh(int shape, double radius, int min_tri, double aspect, std::vector<double>& verts,
                std::vector<int32_t>& tris_out, std::string& err) {
  const double half = 3.14159265358979323846 / 2.0;
  std::vector<Tri> tris;
  if (!(radius > 0.0) || min_tri < 8) {
    err = "build_mesh: radius must be > 0 and min_triangles >= 8";
    return false;
  }
  if (shape == HC_SHAPE_WBC_SPHERE) {
    tris = icosahedron();
    refine(tris, min_tri);
    for (auto& t : tris)
      for (auto& p : t)
        for (auto& c : p) c *= radius;
  } else if (shape == HC_SHAPE_RBC_FROM_SPHERE) {
    tris = icosahedron();
    refine(tris, min_tri);
    rotate(tris, half, half, 0.0);
    for (auto& t : tris)
      for (auto& p : t) {
        p = rbc_point(p);
        for (auto& c : p) c *= radius;
      }
    rotate(tris, half, half, 0.0);
  } else if (shape == HC_SHAPE_ELLIPSOID_FROM_SPHERE) {
    tris = octahedron();
    refine(tris, min_tri);
    rotate(tris, half, half, 0.0);
    for (auto& t : tris)
      for (auto& p : t) p = ellipsoid_point(p, radius, aspect);
    rotate(tris, half, half, 0.0);
  } else {
    err = "build_mesh: unsupported shape " + std::to_string(shape);
    return false;
  }
  // indexed mesh: vertices numbered by first appearance (exact coordinates)
  std::map<P3, int> index;
  std::vector<P3> vv;
  tris_out.clear();
  for (auto& t : tris)
    for (auto& p : t) {
      auto it = index.find(p);
      int id;
      if (it == index.end()) {
        id = (int)vv.size();
        index.emplace(p, id);
        vv.push_back(p);
      } else {
        id = it->second;
      }
      tris_out.push_back(id);
    }
  // inflate by 1e-3 along the normalised sum of unit triangle normals
  std::vector<P3> acc(vv.size(), P3{0, 0, 0});
  for (size_t t = 0; t < tris_out.size() / 3; ++t) {
    const P3 &a = vv[tris_out[3 * t]], &b = vv[tris_out[3 * t + 1]], &c = vv[tris_out[3 * t + 2]];
    P3 n = crs(diff(b, a), diff(c, a));
    double l = len3(n);
    for (auto& q : n) q /= l;
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 3; ++q) acc[tris_out[3 * t + k]][q] += n[q];
  }
  verts.resize(3 * vv.size());
  for (size_t i = 0; i < vv.size(); ++i) {
    double l = len3(acc[i]);
    for (int q = 0; q < 3; ++q) verts[3 * i + q] = vv[i][q] + 1e-3 * (acc[i][q] / l);
  }
  return true;
}

bool build_type_tables(TypeHost& T, std::string& err) {
  const int nv = T.nv, nt = T.nt;
  auto X = [&](int v) { return P3{T.rest[3 * v], T.rest[3 * v + 1], T.rest[3 * v + 2]}; };
  for (int i = 0; i < 3 * nt; ++i)
    if (T.tri[i] < 0 || T.tri[i] >= nv) {
      err = "triangle references a vertex out of range";
      return false;
    }
  // triangles around each vertex (ascending)
  std::vector<std::vector<int>> vt(nv);
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) vt[T.tri[3 * t + k]].push_back(t);
  auto shared = [&](int a, int b, int out[2]) -> bool {
    int n = 0;
    for (int t : vt[a]) {
      const int32_t* q = &T.tri[3 * t];
      if (q[0] == b || q[1] == b || q[2] == b) {
        if (n < 2) out[n] = t;
        ++n;
      }
    }
    return n == 2;
  };
  auto tri_unit_normal = [&](int t) {
    P3 a = X(T.tri[3 * t]), b = X(T.tri[3 * t + 1]), c = X(T.tri[3 * t + 2]);
    P3 n = crs(diff(b, a), diff(c, a));
    double l = norm3(n);
    if (l != 0.0) for (auto& q : n) q /= l;
    else n = {0, 0, 0};
    return n;
  };
  T.edge.clear();
  for (int t = 0; t < nt; ++t) {
    const int32_t* q = &T.tri[3 * t];
    for (int k = 0; k < 3; ++k) {
      int a = q[k], b = q[(k + 1) % 3];
      if (a < b) {
        T.edge.push_back(a);
        T.edge.push_back(b);
      }
    }
  }
  T.ne = (int)T.edge.size() / 2;
  const int ne = T.ne;
  T.edge_len_eq.assign(ne, 0);
  T.edge_angle_eq.assign(ne, 0);
  T.edge_tris.assign(2 * ne, -1);
  T.edge_outer.assign(2 * ne, -1);
  for (int e = 0; e < ne; ++e) {
    int a = T.edge[2 * e], b = T.edge[2 * e + 1];
    P3 ev = diff(X(b), X(a));
    T.edge_len_eq[e] = norm3(ev);
    int adj[2];
    if (!shared(a, b, adj)) {
      err = "mesh is not closed: an edge is not shared by exactly two triangles";
      return false;
    }
    // first the triangle holding the directed edge b->a, then the one holding
    // a->b: with outward triangles the platelet's dihedral bending force
    // (pltSimpleModel.cpp:160-178) then opposes a fold of the edge
    for (int k = 0; k < 3; ++k)
      if (T.tri[3 * adj[0] + k] == a && T.tri[3 * adj[0] + (k + 1) % 3] == b) {
        std::swap(adj[0], adj[1]);
        break;
      }
    T.edge_tris[2 * e] = adj[0];
    T.edge_tris[2 * e + 1] = adj[1];
    P3 n1 = tri_unit_normal(adj[0]), n2 = tri_unit_normal(adj[1]);
    double l = norm3(ev);
    P3 uv{ev[0] / l, ev[1] / l, ev[2] / l};
    T.edge_angle_eq[e] = std::atan2(dt3(crs(n1, n2), uv), dt3(n1, n2));
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < 3; ++k) {
        int w = T.tri[3 * adj[s] + k];
        if (w != a && w != b) T.edge_outer[2 * e + s] = w;
      }
  }
  T.ni = (int)T.inner.size() / 2;
  T.inner_len_eq.assign(T.ni, 0);
  for (int k = 0; k < T.ni; ++k) {
    int a = T.inner[2 * k], b = T.inner[2 * k + 1];
    if (a < 0 || a >= nv || b < 0 || b >= nv) {
      err = "inner edge references a vertex out of range";
      return false;
    }
    T.inner_len_eq[k] = norm3(diff(X(b), X(a)));
  }
  T.tri_area_eq.assign(nt, 0);
  for (int t = 0; t < nt; ++t) {
    P3 a = X(T.tri[3 * t]), b = X(T.tri[3 * t + 1]), c = X(T.tri[3 * t + 2]);
    T.tri_area_eq[t] = 0.5 * norm3(crs(diff(b, a), diff(c, a)));
  }
  double vol = 0.0;
  for (int v = 0; v < nv; ++v)
    for (int t : vt[v]) {
      P3 a = X(T.tri[3 * t]), b = X(T.tri[3 * t + 1]), c = X(T.tri[3 * t + 2]);
      vol += dt3(a, crs(b, c)) / 6.0 / 3.0;
    }
  T.volume_eq = vol;
  double s = 0.0;
  for (double a : T.tri_area_eq) s += a;
  T.area_mean_eq = s / nt;
  s = 0.0;
  for (double l : T.edge_len_eq) s += l;
  T.edge_mean_eq = s / ne;
  s = 0.0;
  for (double a : T.edge_angle_eq) s += a;
  T.angle_mean_eq = s / ne;
  // ordered 1-rings
  T.ring.assign(6 * nv, -1);
  T.nring.assign(nv, 0);
  for (int e = 0; e < ne; ++e) {
    int a = T.edge[2 * e], b = T.edge[2 * e + 1];
    for (int pass = 0; pass < 2; ++pass) {
      int u = pass ? b : a, w = pass ? a : b;
      int k = 0;
      while (k < 6 && T.ring[6 * u + k] != -1) ++k;
      if (k == 6) {
        err = "vertex with more than 6 neighbours (the mechanics tables hold at most 6)";
        return false;
      }
      T.ring[6 * u + k] = w;
    }
  }
  for (int v = 0; v < nv; ++v) {
    int c = 0;
    while (c < 6 && T.ring[6 * v + c] != -1) ++c;
    T.nring[v] = c;
    if (c < 3) {
      err = "vertex with fewer than 3 neighbours";
      return false;
    }
    int cur = T.ring[6 * v];
    for (int k = 1; k < c; ++k) {
      int adj[2];
      shared(v, cur, adj);
      int next = -1;
      for (int sd = 0; sd < 2 && next < 0; ++sd) {
        const int32_t* q = &T.tri[3 * adj[sd]];
        for (int m = 0; m < 3; ++m)
          if (q[m] == v && q[(m + 1) % 3] == cur) next = q[(m + 2) % 3];
      }
      cur = next;
      T.ring[6 * v + k] = cur;
    }
  }
  T.patch_dist_eq.assign(nv, 0);
  for (int v = 0; v < nv; ++v) {
    const int c = T.nring[v];
    const int32_t* R = &T.ring[6 * v];
    P3 sum{0, 0, 0};
    for (int j = 0; j < c; ++j)
      for (int q = 0; q < 3; ++q) sum[q] += X(R[j])[q];
    P3 x = X(v), dev;
    for (int q = 0; q < 3; ++q) dev[q] = sum[q] / (unsigned)c - x[q];
    P3 pn{0, 0, 0};
    for (int j = 0; j < c; ++j) {
      P3 n = crs(diff(X(R[j]), x), diff(X(R[(j + 1) % c]), x));
      double l = norm3(n);
      for (int q = 0; q < 3; ++q) pn[q] += n[q] / l;
    }
    double l = norm3(pn);
    for (auto& q : pn) q /= l;
    T.patch_dist_eq[v] = dt3(pn, dev);
  }
  // gather lists in reference accumulation order
  auto csr = [&](std::vector<std::vector<int32_t>>& lists, std::vector<int32_t>& off, std::vector<int32_t>& flat) {
    off.assign(nv + 1, 0);
    flat.clear();
    for (int v = 0; v < nv; ++v) {
      off[v] = (int32_t)flat.size();
      flat.insert(flat.end(), lists[v].begin(), lists[v].end());
    }
    off[nv] = (int32_t)flat.size();
  };
  std::vector<std::vector<int32_t>> L(nv);
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) L[T.tri[3 * t + k]].push_back(4 * t + k);
  csr(L, T.vtri_off, T.vtri);
  for (auto& l : L) l.clear();
  for (int i = 0; i < nv; ++i) {
    L[i].push_back(i);
    for (int j = 0; j < T.nring[i]; ++j) L[T.ring[6 * i + j]].push_back(i);
  }
  for (auto& l : L) std::sort(l.begin(), l.end());
  csr(L, T.vbend_off, T.vbend);
  for (auto& l : L) l.clear();
  for (int e = 0; e < ne; ++e) {
    L[T.edge[2 * e]].push_back(4 * e + 0);
    L[T.edge[2 * e + 1]].push_back(4 * e + 1);
    if (T.model == HC_MODEL_PLT) {
      L[T.edge_outer[2 * e]].push_back(4 * e + 2);
      L[T.edge_outer[2 * e + 1]].push_back(4 * e + 2);
    }
  }
  csr(L, T.vedge_off, T.vedge);
  for (auto& l : L) l.clear();
  for (int k = 0; k < T.ni; ++k) {
    L[T.inner[2 * k]].push_back(2 * k + 0);
    L[T.inner[2 * k + 1]].push_back(2 * k + 1);
  }
  csr(L, T.vinner_off, T.vinner);
  T.vtri4.clear();
  for (int32_t e : T.vtri) {
    const int t = e >> 2;
    T.vtri4.insert(T.vtri4.end(), {t, T.tri[3 * t], T.tri[3 * t + 1], T.tri[3 * t + 2]});
  }
  T.vedge4.clear();
  for (int32_t e : T.vedge) {
    const int ed = e >> 2;
    T.vedge4.insert(T.vedge4.end(), {e, T.edge[2 * ed], T.edge[2 * ed + 1], 0});
  }
  T.padded = true;
  for (int v = 0; v < nv; ++v)
    if (T.vtri_off[v + 1] - T.vtri_off[v] > kPadW || T.vedge_off[v + 1] - T.vedge_off[v] > kPadW ||
        T.vbend_off[v + 1] - T.vbend_off[v] > kPadB)
      T.padded = false;
  T.vtri4p.clear();
  T.vedge4p.clear();
  T.vbendp.clear();
  if (T.padded) {
    T.vtri4p.assign((size_t)nv * kPadW * 4, -1);
    T.vedge4p.assign((size_t)nv * kPadW * 4, -1);
    T.vbendp.assign((size_t)nv * kPadB, -1);
    T.vtri_aeq.assign((size_t)nv * kPadW, 1.0);
    T.vedge_l0.assign((size_t)nv * kPadW, 1.0);
    T.vbend_n.assign((size_t)nv * kPadB, 1.0);
    for (int v = 0; v < nv; ++v) {
      for (int e = T.vtri_off[v], j = 0; e < T.vtri_off[v + 1]; ++e, ++j) {
        std::copy(&T.vtri4[4 * (size_t)e], &T.vtri4[4 * (size_t)e] + 4, &T.vtri4p[4 * ((size_t)v * kPadW + j)]);
        T.vtri_aeq[(size_t)v * kPadW + j] = T.tri_area_eq[T.vtri4[4 * (size_t)e]];
      }
      for (int e = T.vedge_off[v], j = 0; e < T.vedge_off[v + 1]; ++e, ++j) {
        std::copy(&T.vedge4[4 * (size_t)e], &T.vedge4[4 * (size_t)e] + 4, &T.vedge4p[4 * ((size_t)v * kPadW + j)]);
        T.vedge_l0[(size_t)v * kPadW + j] = T.edge_len_eq[T.vedge4[4 * (size_t)e] >> 2];
      }
      for (int e = T.vbend_off[v], j = 0; e < T.vbend_off[v + 1]; ++e, ++j) {
        T.vbendp[(size_t)v * kPadB + j] = T.vbend[e];
        T.vbend_n[(size_t)v * kPadB + j] = (double)(unsigned)T.nring[T.vbend[e]];
      }
    }
  }
  T.cslot.clear();
  T.cbend.clear();
  T.cring.clear();
  if (T.padded && T.model == HC_MODEL_RBC && nv < 0xFFFF) {
    auto fbits = [](double x) {
      const float f = (float)x;
      uint32_t u;
      std::memcpy(&u, &f, 4);
      return u;
    };
    T.cslot.assign((size_t)nv * kPadW * 4, 0);
    T.cbend.assign((size_t)nv * kPadB, 0xFFFFFFFFu);
    T.cring.assign((size_t)nv * 4, 0);
    for (int v = 0; v < nv; ++v) {
      for (int j = 0; j < kPadW; ++j) {
        uint32_t c0 = 0xFFFF, c1 = 0xFFFF, c2 = 0xFFFF, other = 0xFFFF;
        double aeq = 1.0, l0 = 1.0;
        const int et = T.vtri_off[v] + j;
        if (et < T.vtri_off[v + 1]) {
          const int32_t* q = &T.vtri4[4 * (size_t)et];
          c0 = (uint32_t)q[1];
          c1 = (uint32_t)q[2];
          c2 = (uint32_t)q[3];
          aeq = T.tri_area_eq[q[0]];
        }
        const int ee = T.vedge_off[v] + j;
        if (ee < T.vedge_off[v + 1]) {
          const int code = T.vedge[ee], ed = code >> 2, kind = code & 3;
          other = (uint32_t)(kind == 0 ? T.edge[2 * ed + 1] : T.edge[2 * ed]);
          l0 = T.edge_len_eq[ed];
        }
        uint32_t* r = &T.cslot[4 * ((size_t)v * kPadW + j)];
        r[0] = c0 | (c1 << 16);
        r[1] = c2 | (other << 16);
        r[2] = fbits(aeq);
        r[3] = fbits(l0);
      }
      for (int e = T.vbend_off[v], j = 0; e < T.vbend_off[v + 1]; ++e, ++j)
        T.cbend[(size_t)v * kPadB + j] = (uint32_t)T.vbend[e] | ((uint32_t)T.nring[T.vbend[e]] << 16);
      uint32_t rr[6];
      for (int j = 0; j < 6; ++j) rr[j] = j < T.nring[v] ? (uint32_t)T.ring[6 * v + j] : 0xFFFFu;
      T.cring[4 * (size_t)v + 0] = rr[0] | (rr[1] << 16);
      T.cring[4 * (size_t)v + 1] = rr[2] | (rr[3] << 16);
      T.cring[4 * (size_t)v + 2] = rr[4] | (rr[5] << 16);
      T.cring[4 * (size_t)v + 3] = fbits(T.patch_dist_eq[v]);
    }
  }
  T.rpar.clear();
  T.rnb.clear();
  if (!T.cring.empty()) {
    std::vector<float> par((size_t)nv * 12, 1.0f);
    std::vector<uint32_t> nb((size_t)nv, 0);
    bool ok = true;
    for (int v = 0; v < nv && ok; ++v) {
      const int nr = T.nring[v];
      const int32_t* R = &T.ring[6 * (size_t)v];
      ok = nr >= 3 && nr <= 6;
      auto slot = [&](int32_t a) {
        for (int j = 0; j < nr; ++j)
          if (R[j] == a) return j;
        return -1;
      };
      unsigned seen_t = 0, seen_e = 0;
      for (int e = T.vtri_off[v]; ok && e < T.vtri_off[v + 1]; ++e) {
        const int32_t* q = &T.vtri4[4 * (size_t)e];
        int r = 0;
        while (r < 3 && q[1 + r] != v) ++r;
        if (r == 3) {
          ok = false;
          break;
        }
        const int32_t a = q[1 + (r + 1) % 3], b = q[1 + (r + 2) % 3];
        const int j = slot(a);
        if (j < 0 || R[(j + 1) % nr] != b || (seen_t >> j & 1u)) {
          ok = false;
          break;
        }
        seen_t |= 1u << j;
        par[12 * (size_t)v + j] = (float)T.tri_area_eq[q[0]];
      }
      for (int e = T.vedge_off[v]; ok && e < T.vedge_off[v + 1]; ++e) {
        const int code = T.vedge[e], ed = code >> 2, kind = code & 3;
        const int32_t other = kind == 0 ? T.edge[2 * ed + 1] : (kind == 1 ? T.edge[2 * ed] : -1);
        const int j = slot(other);
        if (j < 0 || (seen_e >> j & 1u)) {
          ok = false;
          break;
        }
        seen_e |= 1u << j;
        par[12 * (size_t)v + 6 + j] = (float)T.edge_len_eq[ed];
      }
      const unsigned all = (1u << nr) - 1u;
      ok = ok && seen_t == all && seen_e == all && T.vbend_off[v + 1] - T.vbend_off[v] == nr + 1;
      unsigned seen_b = 0;
      bool self = false;
      for (int e = T.vbend_off[v]; ok && e < T.vbend_off[v + 1]; ++e) {
        const int32_t i = T.vbend[e];
        if (i == v) {
          self = true;
          continue;
        }
        const int j = slot(i);
        if (j < 0) ok = false;
        else seen_b |= 1u << j;
      }
      ok = ok && self && seen_b == all;
      if (!ok) break;
      uint32_t w = (uint32_t)nr;
      for (int j = 0; j < nr; ++j) {
        const int n = T.nring[R[j]];
        if (n < 1 || n > 7) ok = false;
        w |= (uint32_t)(n & 7) << (4 + 3 * j);
      }
      nb[v] = w;
    }
    if (ok) {
      T.rpar.swap(par);
      T.rnb.swap(nb);
    }
  }
  return true;
}

}  // namespace hcg
