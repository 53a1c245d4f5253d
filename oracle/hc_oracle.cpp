// hc_oracle.cpp -- FP64 CPU restatement of HemoCell's IB-LBM time step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is built into
// oracle/libhc_oracle.so and may be loaded only by tests/, by the smoke() check
// and by bench.py's cpu_baseline leg.  The product (libhcgpu.so) never links,
// loads or calls it.
//
// What it restates (reference paths relative to the HemoCell source tree):
//   * HemoCell::iterate() stage order                    core/hemoCell.cpp:299-376
//   * phi2 kernel + weight normalisation                 core/immersedBoundaryMethod.h:37-41,62-138
//   * force spreading with FORCE_LIMIT cap               core/hemoCellParticleField.cpp:841-863
//   * velocity interpolation                             core/hemoCellParticleField.cpp:819-839
//   * Euler advance + wall tagging                       core/hemoCellParticleField.cpp:566-588,
//                                                        core/hemoCellParticle.h:188-203
//   * force reset before mechanics                       core/hemoCellParticleField.cpp:633-675
//   * RBC / PLT / WBC constitutive models                mechanics/rbcHighOrderModel.cpp:38-207,
//                                                        mechanics/pltSimpleModel.cpp:44-208,
//                                                        mechanics/wbcHighOrderModel.cpp:42-225
//   * rest-geometry tables                               mechanics/commonCellConstants.cpp:70-409
//   * rest volume                                        helper/meshMetrics.h:127-184
//   * cell-cell repulsion                                core/hemoCellParticleField.cpp:677-743
//   * wall (boundary particle) repulsion                 core/hemoCellParticleField.cpp:865-918
//   * Palabos D3Q19 Guo-forced BGK, bounce-back, local regularized velocity
//     boundary (Palabos v2.3.0 is not vendored; restated from its published
//     scheme; call sites core/hemoCell.cpp:317,455-469, helper/hemocellInit.hh:58-92).
//
// Lattice storage: natural (post-stream) populations f[i*N + n] shifted by the
// lattice weights (f_i - t_i), node index n = z + nz*(y + ny*x), streaming by
// a two-lattice push (an independent formulation from the product's in-place
// AA scheme).  Every function is written for clarity, in the reference's
// arithmetic order, not for speed (OpenMP over independent nodes/cells only,
// so results do not depend on the thread count).
//
// Parity status: the HemoCell-owned parts (IBM, advance, mechanics, tables)
// are pinned by tests/test_oracle_*.py against closed forms and invariants;
// the Palabos arithmetic is pinned by closed forms (rest state, conservation,
// Guo momentum gain, Poiseuille/Couette/shear-wave) but its exact op order is
// "parity unpinned" (Palabos source absent).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <stdexcept>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- D3Q19 ----
// Palabos D3Q19 ordering: c_0 = 0, c_1..9 listed below, c_{i+9} = -c_i.
const int C[19][3] = {
    {0, 0, 0},
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
    {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
    {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
inline double W(int i) {
  if (i == 0) return 1.0 / 3.0;
  int k = (i - 1) % 9;           // position inside each half
  return k < 3 ? 1.0 / 18.0 : 1.0 / 36.0;
}
inline int OPP(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }
const double INVCS2 = 3.0;
const double CS2 = 1.0 / 3.0;

enum NodeType : uint8_t { FLUID = 0, BOUNCEBACK = 1, VELWALL0 = 2 };  // 2..7: velocity faces 0..5
inline bool is_boundary(uint8_t t) { return t != FLUID; }

// rhoBar and j from partial sums over the x = -1, 0, +1 planes (Palabos D3Q19 layout).
inline void rhobar_j(const double f[19], double& rhoBar, double j[3]) {
  double xm = f[1] + f[4] + f[5] + f[6] + f[7];
  double x0 = f[0] + f[2] + f[3] + f[8] + f[9] + f[11] + f[12] + f[17] + f[18];
  double xp = f[10] + f[13] + f[14] + f[15] + f[16];
  double ym = f[2] + f[4] + f[8] + f[9] + f[14];
  double yp = f[5] + f[11] + f[13] + f[17] + f[18];
  double zm = f[3] + f[6] + f[8] + f[16] + f[18];
  double zp = f[7] + f[9] + f[12] + f[15] + f[17];
  rhoBar = xm + x0 + xp;
  j[0] = xp - xm;
  j[1] = yp - ym;
  j[2] = zp - zm;
}

// Second-order BGK equilibrium in shifted form: t_i (rhoBar + 3 c.j + 3/2 (3 (c.j)^2 - j^2)/rho).
inline double feq(int i, double rhoBar, double invRho, const double j[3], double jSqr) {
  double c_j = C[i][0] * j[0];
  c_j += C[i][1] * j[1];
  c_j += C[i][2] * j[2];
  return W(i) * (rhoBar + INVCS2 * c_j + INVCS2 / 2.0 * invRho * (INVCS2 * c_j * c_j - jSqr));
}

// Guo-forced BGK collision of one fluid node (in place).  F is the force
// (body force + IBM) at the node; u = j/rho + F/2, j_eq = rho u,
// source (1 - omega/2) t_i [3 (c_i - u) + 9 (c_i.u) c_i] . (rho F).
inline void collide_guo(double f[19], const double F[3], double omega) {
  double rhoBar, j[3];
  rhobar_j(f, rhoBar, j);
  double rho = rhoBar + 1.0;
  double invRho = 1.0 / rho;
  double u[3], jeq[3];
  for (int d = 0; d < 3; ++d) u[d] = j[d] * invRho + F[d] / 2.0;
  for (int d = 0; d < 3; ++d) jeq[d] = rho * u[d];
  double jSqr = jeq[0] * jeq[0] + jeq[1] * jeq[1] + jeq[2] * jeq[2];
  for (int i = 0; i < 19; ++i) {
    f[i] *= 1.0 - omega;
    f[i] += omega * feq(i, rhoBar, invRho, jeq, jSqr);
  }
  double mu = rho * (1.0 - omega / 2.0);
  for (int i = 0; i < 19; ++i) {
    double c_u = 0.0;
    for (int d = 0; d < 3; ++d) c_u += C[i][d] * u[d];
    c_u *= INVCS2 * INVCS2;
    double ft = 0.0;
    for (int d = 0; d < 3; ++d) ft += ((C[i][d] - u[d]) * INVCS2 + c_u * C[i][d]) * F[d];
    ft *= W(i);
    ft *= mu;
    f[i] += ft;
  }
}

// Local regularized velocity boundary: rebuild all populations of a wall node
// from the known ones (on-wall and those arriving from the interior) as
// f_eq(rhoBar, rho u_w) + f_neq(Pi_neq) with Pi_neq from the bounce-back of
// the non-equilibrium part.  axis/orient: wall normal axis and the sign of
// c[axis] of the known "normal" populations (-1 on a low face, +1 on a high face).
inline void velocity_bc_complete(double f[19], int axis, int orient, const double uw[3]) {
  int onWall[19], nOn = 0, normal[19], nNo = 0;
  for (int i = 0; i < 19; ++i) {
    if (C[i][axis] == 0) onWall[nOn++] = i;
    else if (C[i][axis] == orient) normal[nNo++] = i;
  }
  double rhoOnWall = 0.0, rhoNormal = 0.0;
  for (int k = 0; k < nOn; ++k) rhoOnWall += f[onWall[k]];
  for (int k = 0; k < nNo; ++k) rhoNormal += f[normal[k]];
  double velNormal = orient * uw[axis];
  double rhoBar = (2.0 * rhoNormal + rhoOnWall - velNormal) / (1.0 + velNormal);
  double rho = rhoBar + 1.0;
  double j[3] = {rho * uw[0], rho * uw[1], rho * uw[2]};
  double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
  double invRho = 1.0 / rho;
  double fNeq[19] = {0};
  for (int k = 0; k < nOn; ++k) {
    int i = onWall[k];
    fNeq[i] = f[i] - feq(i, rhoBar, invRho, j, jSqr);
  }
  for (int k = 0; k < nNo; ++k) {
    int i = normal[k];
    fNeq[i] = f[i] - feq(i, rhoBar, invRho, j, jSqr);
  }
  double Pi[6];
  int iPi = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      double s = 0.0;
      for (int k = 0; k < nOn; ++k) {
        int i = onWall[k];
        s += C[i][a] * C[i][b] * fNeq[i];
      }
      for (int k = 0; k < nNo; ++k) {
        int i = normal[k];
        s += 2.0 * C[i][a] * C[i][b] * fNeq[i];
      }
      Pi[iPi++] = s;
    }
  for (int i = 0; i < 19; ++i) {
    double fn = 0.0;
    int p = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b) {
        double toAdd = C[i][a] * C[i][b];
        if (a == b) toAdd -= CS2;
        else toAdd *= 2.0;
        toAdd *= Pi[p++];
        fn += toAdd;
      }
    fn *= W(i) * INVCS2 * INVCS2 / 2.0;
    f[i] = feq(i, rhoBar, invRho, j, jSqr) + fn;
  }
}

// ------------------------------------------------------------ geometry ----
struct V3 { double x[3]; };
inline V3 sub(const double* a, const double* b) { return V3{{a[0] - b[0], a[1] - b[1], a[2] - b[2]}}; }
inline V3 cross(const V3& a, const V3& b) {
  return V3{{a.x[1] * b.x[2] - a.x[2] * b.x[1], a.x[2] * b.x[0] - a.x[0] * b.x[2],
             a.x[0] * b.x[1] - a.x[1] * b.x[0]}};
}
inline double dot(const V3& a, const V3& b) {
  double r = 0.0;
  for (int i = 0; i < 3; ++i) r += a.x[i] * b.x[i];
  return r;
}
inline double norm(const V3& a) {
  double r = 0.0;
  for (int i = 0; i < 3; ++i) r += a.x[i] * a.x[i];
  return std::sqrt(r);
}
inline void area_and_unit_normal(const double* v0, const double* v1, const double* v2, double& area, V3& n) {
  n = cross(sub(v1, v0), sub(v2, v0));
  double nn = norm(n);
  if (nn != 0.0) {
    area = 0.5 * nn;
    for (int i = 0; i < 3; ++i) n.x[i] /= nn;
  } else {
    area = 0.0;
    n = V3{{0, 0, 0}};
  }
}
inline V3 unit_normal(const double* v0, const double* v1, const double* v2) {
  V3 n = cross(sub(v1, v0), sub(v2, v0));
  double nn = norm(n);
  if (nn != 0.0) {
    for (int i = 0; i < 3; ++i) n.x[i] /= nn;
  } else {
    n = V3{{0, 0, 0}};
  }
  return n;
}
inline double angle_between_faces(const V3& n1, const V3& n2, const V3& e) {
  return std::atan2(dot(cross(n1, n2), e), dot(n1, n2));
}

const double MaxCellVolumetricChange = 0.01;
const double MaxCellSurfaceAreaChange = 0.09;
const double MaxCellBendingAngle = 0.0555;
const double MaxPLTBendingAngle = 2.467;
const double MaxCellPersistenceLength = 9.0;
const double FORCE_LIMIT = 50.0;  // pN; the viscosity cap compares the raw number / 4

enum Model { RBC = 0, PLT = 1, WBC = 2 };

// Rest-geometry tables of one cell type (commonCellConstants.cpp:70-409).
struct CellType {
  int model = RBC;
  int nv = 0, nt = 0, ne = 0;
  std::vector<int> tri;            // nt*3
  std::vector<double> rest;        // nv*3
  std::vector<int> edge;           // ne*2, i<j in triangle order
  std::vector<double> edge_len_eq, edge_angle_eq, tri_area_eq, patch_dist_eq;
  std::vector<int> edge_tris;      // ne*2 (ascending triangle ids)
  std::vector<int> edge_outer;     // ne*2
  std::vector<int> ring;           // nv*6 ordered 1-ring
  std::vector<int> nring;          // nv
  std::vector<int> inner;          // ni*2
  std::vector<double> inner_len_eq;
  double volume_eq = 0, area_mean_eq = 0, edge_mean_eq = 0, angle_mean_eq = 0;
  // stiffness, lattice units
  double k_link = 0, k_bend = 0, k_volume = 0, k_area = 0, eta_m = 0;
  double k_inner_rigid = 0, k_cytoskeleton = 0, core_radius = 0, radius = 0;
  int timescale = 1;
};

// Triangles incident on each vertex, ascending.
std::vector<std::vector<int>> vertex_triangles(const CellType& ct) {
  std::vector<std::vector<int>> vt(ct.nv);
  for (int t = 0; t < ct.nt; ++t)
    for (int k = 0; k < 3; ++k) vt[ct.tri[3 * t + k]].push_back(t);
  return vt;
}

// The two triangles that share edge (a,b), ascending triangle id.
void adjacent_triangles(const CellType& ct, const std::vector<std::vector<int>>& vt, int a, int b, int out[2]) {
  int n = 0;
  for (int t : vt[a]) {
    const int* T = &ct.tri[3 * t];
    if (T[0] == b || T[1] == b || T[2] == b) {
      if (n < 2) out[n] = t;
      ++n;
    }
  }
  if (n != 2) throw std::runtime_error("mesh is not a closed 2-manifold (edge not shared by exactly 2 triangles)");
}

// Whether triangle t holds the directed edge a->b in its cyclic vertex order.
bool directed_edge(const CellType& ct, int t, int a, int b) {
  const int* T = &ct.tri[3 * t];
  for (int k = 0; k < 3; ++k)
    if (T[k] == a && T[(k + 1) % 3] == b) return true;
  return false;
}

void build_tables(CellType& ct) {
  const double* X = ct.rest.data();
  auto P = [&](int v) { return X + 3 * v; };
  // unique edges: keep (a,b) only when a<b, in triangle order
  ct.edge.clear();
  for (int t = 0; t < ct.nt; ++t) {
    const int* T = &ct.tri[3 * t];
    if (T[0] < T[1]) { ct.edge.push_back(T[0]); ct.edge.push_back(T[1]); }
    if (T[1] < T[2]) { ct.edge.push_back(T[1]); ct.edge.push_back(T[2]); }
    if (T[2] < T[0]) { ct.edge.push_back(T[2]); ct.edge.push_back(T[0]); }
  }
  ct.ne = (int)ct.edge.size() / 2;
  auto vt = vertex_triangles(ct);
  ct.edge_len_eq.resize(ct.ne);
  ct.edge_angle_eq.resize(ct.ne);
  ct.edge_tris.resize(2 * ct.ne);
  ct.edge_outer.resize(2 * ct.ne);
  for (int e = 0; e < ct.ne; ++e) {
    int a = ct.edge[2 * e], b = ct.edge[2 * e + 1];
    ct.edge_len_eq[e] = norm(sub(P(b), P(a)));
    int adj[2];
    adjacent_triangles(ct, vt, a, b, adj);
    // the pair is ordered by orientation: the first triangle holds the directed
    // edge b->a, the second a->b (the one the edge list took the edge from).
    // With outward triangles this makes the platelet's dihedral bending force
    // (pltSimpleModel.cpp:160-178) oppose a fold of the edge; Palabos's
    // getAdjacentTriangleIds order is not available to check against.
    if (directed_edge(ct, adj[0], a, b)) std::swap(adj[0], adj[1]);
    ct.edge_tris[2 * e] = adj[0];
    ct.edge_tris[2 * e + 1] = adj[1];
    const int* T0 = &ct.tri[3 * adj[0]];
    const int* T1 = &ct.tri[3 * adj[1]];
    V3 n1 = unit_normal(P(T0[0]), P(T0[1]), P(T0[2]));
    V3 n2 = unit_normal(P(T1[0]), P(T1[1]), P(T1[2]));
    V3 ev = sub(P(b), P(a));
    double len = norm(ev);
    V3 uv{{ev.x[0] / len, ev.x[1] / len, ev.x[2] / len}};
    ct.edge_angle_eq[e] = angle_between_faces(n1, n2, uv);
    for (int k = 0; k < 3; ++k) {
      if (T0[k] != a && T0[k] != b) ct.edge_outer[2 * e] = T0[k];
      if (T1[k] != a && T1[k] != b) ct.edge_outer[2 * e + 1] = T1[k];
    }
  }
  int ni = (int)ct.inner.size() / 2;
  ct.inner_len_eq.resize(ni);
  for (int k = 0; k < ni; ++k) ct.inner_len_eq[k] = norm(sub(P(ct.inner[2 * k + 1]), P(ct.inner[2 * k])));
  ct.tri_area_eq.resize(ct.nt);
  for (int t = 0; t < ct.nt; ++t) {
    const int* T = &ct.tri[3 * t];
    ct.tri_area_eq[t] = 0.5 * norm(cross(sub(P(T[1]), P(T[0])), sub(P(T[2]), P(T[0]))));
  }
  // rest volume: every triangle visited once per corner, 1/6/3 each (meshMetrics.h:176-184)
  double vol = 0.0;
  for (int v = 0; v < ct.nv; ++v)
    for (int t : vt[v]) {
      const int* T = &ct.tri[3 * t];
      V3 v0{{P(T[0])[0], P(T[0])[1], P(T[0])[2]}};
      V3 v1{{P(T[1])[0], P(T[1])[1], P(T[1])[2]}};
      V3 v2{{P(T[2])[0], P(T[2])[1], P(T[2])[2]}};
      double t6 = dot(v0, cross(v1, v2));
      vol += t6 / 6.0 / 3.0;
    }
  ct.volume_eq = vol;
  double s = 0.0;
  for (double a : ct.tri_area_eq) s += a;
  ct.area_mean_eq = s / ct.nt;
  s = 0.0;
  for (double l : ct.edge_len_eq) s += l;
  ct.edge_mean_eq = s / ct.ne;
  s = 0.0;
  for (double a : ct.edge_angle_eq) s += a;
  ct.angle_mean_eq = s / ct.ne;
  // 1-ring: first-come from the edge list, then re-ordered around the vertex
  ct.ring.assign(6 * ct.nv, -1);
  ct.nring.assign(ct.nv, 0);
  for (int e = 0; e < ct.ne; ++e) {
    int a = ct.edge[2 * e], b = ct.edge[2 * e + 1];
    for (int k = 0; k < 6; ++k) if (ct.ring[6 * a + k] == -1) { ct.ring[6 * a + k] = b; break; }
    for (int k = 0; k < 6; ++k) if (ct.ring[6 * b + k] == -1) { ct.ring[6 * b + k] = a; break; }
  }
  for (int v = 0; v < ct.nv; ++v)
    for (int k = 0; k < 6; ++k)
      if (ct.ring[6 * v + k] != -1) ct.nring[v]++;
  for (int v = 0; v < ct.nv; ++v) {
    int nvtx = ct.ring[6 * v];
    for (int n = 1; n < ct.nring[v]; ++n) {
      int adj[2];
      adjacent_triangles(ct, vt, v, nvtx, adj);
      int next = -1;
      for (int s2 = 0; s2 < 2 && next < 0; ++s2) {
        const int* T = &ct.tri[3 * adj[s2]];
        if ((T[0] == v && T[1] == nvtx) || (T[1] == v && T[2] == nvtx) || (T[2] == v && T[0] == nvtx)) {
          for (int k = 0; k < 3; ++k)
            if (T[k] != v && T[k] != nvtx) next = T[k];
        }
      }
      nvtx = next;
      ct.ring[6 * v + n] = nvtx;
    }
  }
  ct.patch_dist_eq.resize(ct.nv);
  for (int v = 0; v < ct.nv; ++v) {
    int nr = ct.nring[v];
    const int* R = &ct.ring[6 * v];
    double sum[3] = {0, 0, 0};
    for (int j = 0; j < nr; ++j)
      for (int c = 0; c < 3; ++c) sum[c] += P(R[j])[c];
    V3 dev;
    for (int c = 0; c < 3; ++c) dev.x[c] = sum[c] / (unsigned)nr - P(v)[c];
    V3 pn{{0, 0, 0}};
    for (int j = 0; j < nr - 1; ++j) {
      V3 tn = cross(sub(P(R[j]), P(v)), sub(P(R[j + 1]), P(v)));
      double l = norm(tn);
      for (int c = 0; c < 3; ++c) tn.x[c] /= l;
      for (int c = 0; c < 3; ++c) pn.x[c] += tn.x[c];
    }
    V3 tn = cross(sub(P(R[nr - 1]), P(v)), sub(P(R[0]), P(v)));
    double l = norm(tn);
    for (int c = 0; c < 3; ++c) tn.x[c] /= l;
    for (int c = 0; c < 3; ++c) pn.x[c] += tn.x[c];
    l = norm(pn);
    for (int c = 0; c < 3; ++c) pn.x[c] /= l;
    ct.patch_dist_eq[v] = dot(pn, dev);
  }
}

// ------------------------------------------------------------- oracle ----
struct Kernel {
  int n = 0;
  long idx[27];
  double w[27];
};

struct Oracle {
  int nx, ny, nz;
  long N;
  int periodic[3];
  double omega;
  std::vector<uint8_t> type;
  std::vector<double> f, ftmp;      // 19*N natural layout, shifted
  std::vector<double> Fibm;         // 3*N (SoA: comp*N + n), IBM part of the force field
  double g[3] = {0, 0, 0};          // uniform body force
  double wall_u[6][3] = {{0}};
  double f_limit = 1e300;
  // vertices (SoA over all cells, cells contiguous, types in registration order)
  std::vector<double> x, v, force, frep;  // 3*NV each, AoS per vertex (x[3*i+c])
  std::vector<int> vcell;                 // cell index of each vertex
  // 1 for a vertex removed by wall contact (advanceParticles,
  // hemoCellParticleField.cpp:566-588: the particle is deleted); its cell is
  // incomplete from then on
  std::vector<uint8_t> removed;
  std::vector<CellType> types;
  std::vector<int> cell_type, cell_base;  // per cell
  std::vector<Kernel> kern;               // per vertex, built by the spread
  long iter = 0;
  int pv_timescale = 1;
  int rep_timescale = 1;
  bool repulsion = false;
  double rep_const = 0, rep_cutoff = 0;
  long wall_hits = 0;
  // wall repulsion: boundary particles and the force they exert (kept per vertex)
  bool brep = false;
  double brep_const = 0, brep_cutoff = 0;
  int brep_timescale = 1;
  std::vector<long> bparticles;  // node indices, ascending x, y, z
  std::vector<double> fbr;       // 3*NV
  // Lees-Edwards planes (z = 0 and z = nz-1): plane velocities and the
  // displacement per step
  bool le = false;
  double le_top_u = 0, le_bottom_u = 0, le_disp = 0;
  // interior viscosity: tau inside the membranes per type (0 = off), node classes
  std::vector<double> iv_tau;
  std::vector<uint8_t> visc;     // N: 0 background, 1 + type inside
  int iv_timescale = 1, iv_grid_timescale = 1;

  long idx(int X, int Y, int Z) const { return (long)Z + (long)nz * ((long)Y + (long)ny * X); }
  int dim(int a) const { return a == 0 ? nx : (a == 1 ? ny : nz); }

  // node coordinate along axis a wrapped (periodic) or -1 when outside the domain
  int wrap(int a, long c) const {
    int n = dim(a);
    if (c >= 0 && c < n) return (int)c;
    if (!periodic[a]) return -1;
    long m = c % n;
    if (m < 0) m += n;
    return (int)m;
  }

  void collide_node(long n, double fl[19]) const {
    uint8_t t = type[n];
    if (t == BOUNCEBACK) {
      double tmp[19];
      for (int i = 0; i < 19; ++i) tmp[i] = fl[OPP(i)];
      for (int i = 0; i < 19; ++i) fl[i] = tmp[i];
      return;
    }
    if (t >= VELWALL0) {
      int face = t - VELWALL0;
      velocity_bc_complete(fl, face / 2, (face % 2) ? 1 : -1, wall_u[face]);
    }
    double F[3] = {g[0] + Fibm[n], g[1] + Fibm[N + n], g[2] + Fibm[2 * N + n]};
    const double ti = (!visc.empty() && visc[n]) ? ivt(visc[n] - 1) : 0.0;
    const double om = ti > 0.0 ? 1.0 / ti : omega;
    collide_guo(fl, F, om);
  }

  void collide_and_stream() {
    std::fill(ftmp.begin(), ftmp.end(), 0.0);
#pragma omp parallel for schedule(static)
    for (int X = 0; X < nx; ++X)
      for (int Y = 0; Y < ny; ++Y)
        for (int Z = 0; Z < nz; ++Z) {
          long n = idx(X, Y, Z);
          double fl[19];
          for (int i = 0; i < 19; ++i) fl[i] = f[(long)i * N + n];
          collide_node(n, fl);
          for (int i = 0; i < 19; ++i) {
            int a = wrap(0, X + C[i][0]), b = wrap(1, Y + C[i][1]), c = wrap(2, Z + C[i][2]);
            if (a < 0 || b < 0 || c < 0) continue;  // leaves the domain
            ftmp[(long)i * N + idx(a, b, c)] = fl[i];
          }
        }
    f.swap(ftmp);
  }

  void velocity_at(long n, double u[3]) const {
    double fl[19];
    for (int i = 0; i < 19; ++i) fl[i] = f[(long)i * N + n];
    double rhoBar, j[3];
    rhobar_j(fl, rhoBar, j);
    double invRho = 1.0 / (rhoBar + 1.0);
    double F[3] = {g[0] + Fibm[n], g[1] + Fibm[N + n], g[2] + Fibm[2 * N + n]};
    for (int d = 0; d < 3; ++d) u[d] = j[d] * invRho + F[d] / 2.0;
  }

  // phi2 support of a vertex: 3x3x3 candidates in x->y->z order around the
  // nearest node, skipping nodes outside a non-periodic domain, zero weights
  // and wall nodes, then normalised to sum 1.
  void make_kernel(const double* p, Kernel& k) const {
    k.n = 0;
    long ctr[3];
    for (int a = 0; a < 3; ++a) ctr[a] = (long)std::floor(p[a] + 0.5);
    double total = 0.0;
    for (int dx = -1; dx < 2; ++dx)
      for (int dy = -1; dy < 2; ++dy)
        for (int dz = -1; dz < 2; ++dz) {
          long nd[3] = {ctr[0] + dx, ctr[1] + dy, ctr[2] + dz};
          int w3[3];
          bool out = false;
          for (int a = 0; a < 3; ++a) {
            w3[a] = wrap(a, nd[a]);
            if (w3[a] < 0) out = true;
          }
          if (out) continue;
          double phi[3];
          for (int a = 0; a < 3; ++a) phi[a] = p[a] - (double)nd[a];
          double w = 1.0;
          {
            double q0 = std::max(1.0 - std::fabs(phi[0]), 0.0);
            double q1 = std::max(1.0 - std::fabs(phi[1]), 0.0);
            double q2 = std::max(1.0 - std::fabs(phi[2]), 0.0);
            w = q0 * q1 * q2;
          }
          if (w == 0.0) continue;
          long n = idx(w3[0], w3[1], w3[2]);
          if (is_boundary(type[n])) continue;
          total += w;
          k.idx[k.n] = n;
          k.w[k.n] = w;
          k.n++;
        }
    double coeff = 1.0 / total;
    for (int i = 0; i < k.n; ++i) k.w[i] *= coeff;
  }

  long NV() const { return (long)vcell.size(); }

  void spread() {
    kern.resize(NV());
    for (long i = 0; i < NV(); ++i) {
      Kernel& k = kern[i];
      if (removed[i]) {  // no particle, no contribution
        k.n = 0;
        continue;
      }
      make_kernel(&x[3 * i], k);
      double* fv = &force[3 * i];
      double fm = std::sqrt(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2]);
      if (fm > f_limit) {
        double s = f_limit / fm;
        for (int c = 0; c < 3; ++c) fv[c] *= s;
      }
      double r[3];
      for (int c = 0; c < 3; ++c) r[c] = brep && (long)fbr.size() == 3 * NV() ? frep[3 * i + c] + fbr[3 * i + c]
                                                                              : frep[3 * i + c];
      for (int j = 0; j < k.n; ++j)
        for (int c = 0; c < 3; ++c) Fibm[(long)c * N + k.idx[j]] += (r[c] + fv[c]) * k.w[j];
    }
  }

  // The support is recomputed from the positions, which are unchanged since the
  // spread of this step (the reference stores it on the particle instead).
  void interpolate() {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < NV(); ++i) {
      if (removed[i]) continue;
      Kernel k;
      make_kernel(&x[3 * i], k);
      double vel[3] = {0, 0, 0};
      for (int j = 0; j < k.n; ++j) {
        double u[3];
        velocity_at(k.idx[j], u);
        for (int c = 0; c < 3; ++c) vel[c] += u[c] * k.w[j];
      }
      for (int c = 0; c < 3; ++c) v[3 * i + c] = vel[c];
    }
  }

  // A vertex whose nearest node is a wall node after its advance is removed
  // (hemoCellParticleField.cpp:566-588; the nearest node by floor(x + 0.5)).
  void advance() {
    for (long i = 0; i < NV(); ++i) {
      if (removed[i]) continue;
      for (int c = 0; c < 3; ++c) x[3 * i + c] += v[3 * i + c];
      int w3[3];
      bool out = false;
      for (int a = 0; a < 3; ++a) {
        w3[a] = wrap(a, (long)std::floor(x[3 * i + a] + 0.5));
        if (w3[a] < 0) out = true;
      }
      if (!out && is_boundary(type[idx(w3[0], w3[1], w3[2])])) {
        wall_hits++;
        removed[i] = 1;
      }
    }
  }

  bool incomplete(int cell) const {
    const long b = cell_base[cell];
    for (int k = 0; k < types[cell_type[cell]].nv; ++k)
      if (removed[b + k]) return true;
    return false;
  }

  // deleteIncompleteCells (hemoCellParticleField.cpp:510-553): every vertex of
  // an incomplete cell goes
  void delete_incomplete_cells() {
    for (int c = 0; c < (int)cell_type.size(); ++c)
      if (incomplete(c)) {
        const long b = cell_base[c];
        for (int k = 0; k < types[cell_type[c]].nv; ++k) removed[b + k] = 1;
      }
  }

  void mechanics_cell(int cell) {
    const CellType& ct = types[cell_type[cell]];
    long b = cell_base[cell];
    auto P = [&](int k) { return &x[3 * (b + k)]; };
    auto Vv = [&](int k) { return &v[3 * (b + k)]; };
    auto Fp = [&](int k) { return &force[3 * (b + k)]; };
    auto add = [&](int k, const V3& a) { double* q = Fp(k); for (int c = 0; c < 3; ++c) q[c] += a.x[c]; };
    auto subf = [&](int k, const V3& a) { double* q = Fp(k); for (int c = 0; c < 3; ++c) q[c] -= a.x[c]; };

    double volume = 0.0;
    std::vector<double> tarea(ct.nt);
    std::vector<V3> tnorm(ct.nt);
    // rbcHighOrderModel.cpp:48-100 sums the triple products of the absolute corner
    // positions; here they are taken relative to the cell's vertex 0 (the same
    // volume for a closed mesh), which removes the cancellation that absolute
    // coordinates far from the origin cause
    const double* r0 = P(0);
    for (int t = 0; t < ct.nt; ++t) {
      const int* T = &ct.tri[3 * t];
      const double *v0 = P(T[0]), *v1 = P(T[1]), *v2 = P(T[2]);
      const double w0[3] = {v0[0] - r0[0], v0[1] - r0[1], v0[2] - r0[2]};
      const double w1[3] = {v1[0] - r0[0], v1[1] - r0[1], v1[2] - r0[2]};
      const double w2[3] = {v2[0] - r0[0], v2[1] - r0[1], v2[2] - r0[2]};
      const double v210 = w2[0] * w1[1] * w0[2];
      const double v120 = w1[0] * w2[1] * w0[2];
      const double v201 = w2[0] * w0[1] * w1[2];
      const double v021 = w0[0] * w2[1] * w1[2];
      const double v102 = w1[0] * w0[1] * w2[2];
      const double v012 = w0[0] * w1[1] * w2[2];
      volume += (-v210 + v120 + v201 - v021 - v102 + v012);
      double area;
      V3 n;
      area_and_unit_normal(v0, v1, v2, area, n);
      const double ar = (area - ct.tri_area_eq[t]) / ct.tri_area_eq[t];
      const double afm = ct.k_area * (ar + ar / std::fabs(MaxCellSurfaceAreaChange - ar * ar));
      double cen[3];
      for (int c = 0; c < 3; ++c) cen[c] = (v0[c] + v1[c] + v2[c]) / 3.0;
      V3 a0{{(cen[0] - v0[0]) * afm, (cen[1] - v0[1]) * afm, (cen[2] - v0[2]) * afm}};
      V3 a1{{(cen[0] - v1[0]) * afm, (cen[1] - v1[1]) * afm, (cen[2] - v1[2]) * afm}};
      V3 a2{{(cen[0] - v2[0]) * afm, (cen[1] - v2[1]) * afm, (cen[2] - v2[2]) * afm}};
      add(T[0], a0);
      add(T[1], a1);
      add(T[2], a2);
      tarea[t] = area;
      tnorm[t] = n;
    }
    volume *= (1.0 / 6.0);
    const double vfrac = (volume - ct.volume_eq) / ct.volume_eq;
    const double vforce = -ct.k_volume * vfrac / std::fabs(MaxCellVolumetricChange - vfrac * vfrac);
    for (int t = 0; t < ct.nt; ++t) {
      const int* T = &ct.tri[3 * t];
      double s = tarea[t] / ct.area_mean_eq;
      V3 lvf{{(vforce * tnorm[t].x[0]) * s, (vforce * tnorm[t].x[1]) * s, (vforce * tnorm[t].x[2]) * s}};
      add(T[0], lvf);
      add(T[1], lvf);
      add(T[2], lvf);
    }
    if (ct.model == RBC || ct.model == WBC) {
      for (int i = 0; i < ct.nv; ++i) {
        int nr = ct.nring[i];
        const int* R = &ct.ring[6 * i];
        double sum[3] = {0, 0, 0};
        for (int j = 0; j < nr; ++j)
          for (int c = 0; c < 3; ++c) sum[c] += P(R[j])[c];
        V3 dev;
        for (int c = 0; c < 3; ++c) dev.x[c] = sum[c] / (unsigned)nr - P(i)[c];
        V3 pn{{0, 0, 0}};
        for (int j = 0; j < nr - 1; ++j) {
          V3 tn = cross(sub(P(R[j]), P(i)), sub(P(R[j + 1]), P(i)));
          double l = norm(tn);
          for (int c = 0; c < 3; ++c) tn.x[c] /= l;
          for (int c = 0; c < 3; ++c) pn.x[c] += tn.x[c];
        }
        V3 tn = cross(sub(P(R[nr - 1]), P(i)), sub(P(R[0]), P(i)));
        double l = norm(tn);
        for (int c = 0; c < 3; ++c) tn.x[c] /= l;
        for (int c = 0; c < 3; ++c) pn.x[c] += tn.x[c];
        l = norm(pn);
        for (int c = 0; c < 3; ++c) pn.x[c] /= l;
        const double ndev = dot(pn, dev);
        const double dDev = (ndev - ct.patch_dist_eq[i]) / ct.edge_mean_eq;
        const double s = ct.k_bend * (dDev + dDev / std::fabs(MaxCellBendingAngle - dDev * dDev));
        V3 bf{{pn.x[0] * s, pn.x[1] * s, pn.x[2] * s}};
        add(i, bf);
        V3 nbf{{-bf.x[0] / (unsigned)nr, -bf.x[1] / (unsigned)nr, -bf.x[2] / (unsigned)nr}};
        for (int j = 0; j < nr; ++j) add(R[j], nbf);
      }
    }
    const bool visc = (ct.model != RBC) || (ct.eta_m != 0.0);
    for (int e = 0; e < ct.ne; ++e) {
      int a = ct.edge[2 * e], bb = ct.edge[2 * e + 1];
      V3 ev = sub(P(bb), P(a));
      double len = (ct.model == PLT) ? std::sqrt(ev.x[0] * ev.x[0] + ev.x[1] * ev.x[1] + ev.x[2] * ev.x[2]) : norm(ev);
      V3 uv{{ev.x[0] / len, ev.x[1] / len, ev.x[2] / len}};
      const double ef = (len - ct.edge_len_eq[e]) / ct.edge_len_eq[e];
      const double fs = ct.k_link * (ef + ef / std::fabs(MaxCellPersistenceLength - ef * ef));
      V3 fl{{uv.x[0] * fs, uv.x[1] * fs, uv.x[2] * fs}};
      add(a, fl);
      subf(bb, fl);
      if (visc) {
        V3 rv = sub(Vv(bb), Vv(a));
        double pr = dot(rv, uv);
        V3 fv{{(pr * uv.x[0]) * ct.eta_m, (pr * uv.x[1]) * ct.eta_m, (pr * uv.x[2]) * ct.eta_m}};
        double m = norm(fv);
        if (m > FORCE_LIMIT / 4.0) {
          double s = (FORCE_LIMIT / 4.0) / m;
          for (int c = 0; c < 3; ++c) fv.x[c] *= s;
        }
        add(a, fv);
        subf(bb, fv);
      }
      if (ct.model == PLT) {
        const int* T0 = &ct.tri[3 * ct.edge_tris[2 * e]];
        const int* T1 = &ct.tri[3 * ct.edge_tris[2 * e + 1]];
        V3 n1 = unit_normal(P(T0[0]), P(T0[1]), P(T0[2]));
        V3 n2 = unit_normal(P(T1[0]), P(T1[1]), P(T1[2]));
        double ang = angle_between_faces(n1, n2, uv);
        double af = ang - ct.edge_angle_eq[e];
        double fm = ct.k_bend * (af + af / std::fabs(MaxPLTBendingAngle - af * af));
        V3 bfv{{((n1.x[0] + n2.x[0]) * fm) * 0.5, ((n1.x[1] + n2.x[1]) * fm) * 0.5, ((n1.x[2] + n2.x[2]) * fm) * 0.5}};
        add(a, bfv);
        add(bb, bfv);
        subf(ct.edge_outer[2 * e], bfv);
        subf(ct.edge_outer[2 * e + 1], bfv);
      }
    }
    int ni = (int)ct.inner.size() / 2;
    for (int k = 0; k < ni; ++k) {
      int a = ct.inner[2 * k], bb = ct.inner[2 * k + 1];
      V3 ev = sub(P(bb), P(a));
      if (ct.model == PLT) {
        double len = std::sqrt(ev.x[0] * ev.x[0] + ev.x[1] * ev.x[1] + ev.x[2] * ev.x[2]);
        V3 uv{{ev.x[0] / len, ev.x[1] / len, ev.x[2] / len}};
        double ef = (len - ct.inner_len_eq[k]) / ct.inner_len_eq[k];
        double fs = ct.k_link * 5.0 * ef;
        V3 fl{{uv.x[0] * fs, uv.x[1] * fs, uv.x[2] * fs}};
        add(a, fl);
        subf(bb, fl);
      } else if (ct.model == WBC) {
        double len = norm(ev);
        V3 uv{{ev.x[0] / len, ev.x[1] / len, ev.x[2] / len}};
        if (len < 2 * ct.radius) {
          double s = 1.0 - (len / (2 * ct.radius));
          V3 fl{{(uv.x[0] * s) * ct.k_cytoskeleton, (uv.x[1] * s) * ct.k_cytoskeleton, (uv.x[2] * s) * ct.k_cytoskeleton}};
          subf(a, fl);
          add(bb, fl);
        }
        if (len < 2 * ct.core_radius) {
          double s = 1 - (len / (2 * ct.core_radius));
          V3 fl{{(uv.x[0] * s) * ct.k_inner_rigid, (uv.x[1] * s) * ct.k_inner_rigid, (uv.x[2] * s) * ct.k_inner_rigid}};
          subf(a, fl);
          add(bb, fl);
        }
      }
    }
  }

  void mechanics(bool forced) {
    int nc = (int)cell_type.size();
    std::vector<char> due(types.size());
    for (size_t t = 0; t < types.size(); ++t) due[t] = forced || (iter % types[t].timescale == 0);
    for (int c = 0; c < nc; ++c) {
      if (!due[cell_type[c]]) continue;
      long b = cell_base[c];
      for (int k = 0; k < types[cell_type[c]].nv; ++k)
        for (int q = 0; q < 3; ++q) force[3 * (b + k) + q] = 0.0;
    }
    // applyConstitutiveModel (hemoCellParticleField.cpp:633-675): the forces of
    // every vertex of a due type are reset, the model then runs on the complete
    // cells only
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < nc; ++c)
      if (due[cell_type[c]] && !incomplete(c)) mechanics_cell(c);
  }

  // Cell-cell repulsion over nearest-node bins (hemoCellParticleField.cpp:677-743):
  // every node l is paired with itself and 13 forward neighbours (the
  // reference's half stencil); inside one bin each pair is visited in both
  // orders, as in the reference.  Bins hold vertices in vertex order; the
  // domain wraps on periodic axes and the separation uses the minimum image.
  void repulsion_force() {
    long nv = NV();
    for (long i = 0; i < 3 * nv; ++i) frep[i] = 0.0;
    std::vector<std::vector<long>> bin(N);
    for (long i = 0; i < nv; ++i) {
      if (removed[i]) continue;
      int w3[3];
      bool out = false;
      for (int a = 0; a < 3; ++a) {
        w3[a] = wrap(a, (long)std::floor(x[3 * i + a] + 0.5));
        if (w3[a] < 0) out = true;
      }
      if (!out) bin[idx(w3[0], w3[1], w3[2])].push_back(i);
    }
    static const int OFF[14][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                                   {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},
                                   {1, 0, 1}, {1, 1, -1}, {1, 1, 0}, {1, 1, 1}, {0, 1, -1}};
    for (int X = 0; X < nx; ++X)
      for (int Y = 0; Y < ny; ++Y)
        for (int Z = 0; Z < nz; ++Z) {
          long l = idx(X, Y, Z);
          if (bin[l].empty()) continue;
          for (int o = 0; o < 14; ++o) {
            int a = wrap(0, X + OFF[o][0]), b = wrap(1, Y + OFF[o][1]), c = wrap(2, Z + OFF[o][2]);
            if (a < 0 || b < 0 || c < 0) continue;
            long m = idx(a, b, c);
            for (long i : bin[l])
              for (long j : bin[m]) {
                if (i == j || vcell[i] == vcell[j]) continue;
                double dv[3];
                for (int q = 0; q < 3; ++q) {
                  double d = x[3 * i + q] - x[3 * j + q];
                  if (periodic[q]) {
                    int n = dim(q);
                    d -= n * std::floor(d / n + 0.5);
                  }
                  dv[q] = d;
                }
                double dist = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
                if (dist < rep_cutoff) {
                  double s = rep_const * (1 / (dist / rep_cutoff));
                  for (int q = 0; q < 3; ++q) {
                    double r = s * (dv[q] / dist);
                    frep[3 * i + q] += r;
                    frep[3 * j + q] -= r;
                  }
                }
              }
          }
        }
  }

  // populateBoundaryParticles (hemoCellParticleField.cpp:865-887): bounce-back nodes
  // with a node that is not bounce-back among their 26 neighbours, in x, y, z order
  // (periodic axes wrap; the reference's scan stops one node short of the far
  // faces of its block, which here would drop a wall plane, so every node counts)
  void populate_boundary() {
    bparticles.clear();
    for (int X = 0; X < nx; ++X)
      for (int Y = 0; Y < ny; ++Y)
        for (int Z = 0; Z < nz; ++Z) {
          if (type[idx(X, Y, Z)] != BOUNCEBACK) continue;
          bool found = false;
          for (int a = -1; a <= 1 && !found; ++a)
            for (int b = -1; b <= 1 && !found; ++b)
              for (int c = -1; c <= 1 && !found; ++c) {
                int xx = wrap(0, X + a), yy = wrap(1, Y + b), zz = wrap(2, Z + c);
                if (xx < 0 || yy < 0 || zz < 0) continue;
                if (type[idx(xx, yy, zz)] != BOUNCEBACK) found = true;
              }
          if (found) bparticles.push_back(idx(X, Y, Z));
        }
  }

  // applyBoundaryRepulsionForce (hemoCellParticleField.cpp:889-918): every boundary
  // particle, in order, adds k (r_c / r) r_hat to each vertex binned in its 3x3x3
  // neighbourhood closer than r_c; r from the node to the vertex, minimum image on
  // periodic axes.  The force is recomputed here at each application.
  void boundary_repulsion_force() {
    long nv = NV();
    fbr.assign(3 * nv, 0.0);
    std::vector<std::vector<long>> bin(N);
    for (long i = 0; i < nv; ++i) {
      if (removed[i]) continue;
      int w3[3];
      bool out = false;
      for (int a = 0; a < 3; ++a) {
        w3[a] = wrap(a, (long)std::floor(x[3 * i + a] + 0.5));
        if (w3[a] < 0) out = true;
      }
      if (!out) bin[idx(w3[0], w3[1], w3[2])].push_back(i);
    }
    for (long b : bparticles) {
      const int BZ = (int)(b % nz), BY = (int)((b / nz) % ny), BX = (int)(b / ((long)ny * nz));
      const int P[3] = {BX, BY, BZ};
      for (int a = -1; a <= 1; ++a)
        for (int bb = -1; bb <= 1; ++bb)
          for (int c = -1; c <= 1; ++c) {
            int xx = wrap(0, BX + a), yy = wrap(1, BY + bb), zz = wrap(2, BZ + c);
            if (xx < 0 || yy < 0 || zz < 0) continue;
            for (long i : bin[idx(xx, yy, zz)]) {
              double dv[3];
              for (int q = 0; q < 3; ++q) {
                double d = x[3 * i + q] - P[q];
                if (periodic[q]) {
                  int n = dim(q);
                  d -= n * std::floor(d / n + 0.5);
                }
                dv[q] = d;
              }
              double dist = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
              if (dist < brep_cutoff) {
                double s = brep_const * (1 / (dist / brep_cutoff));
                for (int q = 0; q < 3; ++q) fbr[3 * i + q] = fbr[3 * i + q] + s * (dv[q] / dist);
              }
            }
          }
    }
  }

  // ---- interior viscosity (the reference's INTERIOR_VISCOSITY build) ----
  // Fluid nodes inside a membrane relax with the type's tau_inner
  // (hemoCellField.cpp:99-110); visc holds 0 (background) or 1 + type per node.
  double ivt(int t) const { return t < (int)iv_tau.size() ? iv_tau[t] : 0.0; }
  bool iv_on() const {
    for (double t : iv_tau)
      if (t > 0.0) return true;
    return false;
  }

  // Moller-Trumbore test of the ray from lattice site s along (-40,-40,-40)
  // against one triangle: 1 when it crosses at t > 1e-7, t rounded to float
  // (helper/mollerTrumbore.h:26-78).
  static int ray_crosses(const double* v0, const double* v1, const double* v2, const long s[3]) {
    const double EPS = 0.0000001;
    const double ray[3] = {-40.0, -40.0, -40.0};
    double e1[3], e2[3], sv[3];
    for (int c = 0; c < 3; ++c) {
      e1[c] = v1[c] - v0[c];
      e2[c] = v2[c] - v0[c];
    }
    const double pv[3] = {ray[1] * e2[2] - ray[2] * e2[1], ray[2] * e2[0] - ray[0] * e2[2],
                          ray[0] * e2[1] - ray[1] * e2[0]};
    const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det > -EPS && det < EPS) return 0;
    const double inv = 1 / det;
    for (int c = 0; c < 3; ++c) sv[c] = (double)s[c] - v0[c];
    const double u = inv * (sv[0] * pv[0] + sv[1] * pv[1] + sv[2] * pv[2]);
    if (u < 0.0 || u > 1.0) return 0;
    const double qv[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2],
                          sv[0] * e1[1] - sv[1] * e1[0]};
    const double v = inv * (ray[0] * qv[0] + ray[1] * qv[1] + ray[2] * qv[2]);
    if (v < 0.0 || u + v > 1.0) return 0;
    const float t = (float)(inv * (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]));
    return t > EPS ? 1 : 0;
  }

  // Entire-grid pass (hemoCellParticleField.cpp:776-808): every node back to the
  // background, then each node of a cell's bounding box (corners truncated to
  // integers, octree.h:107-150; clamped to the domain on non-periodic axes) whose
  // ray crosses the cell's triangles an odd number of times is inner.  The
  // reference culls triangles with an octree first; the cull only drops
  // triangles the ray cannot reach, so every triangle is tested here.  Only
  // bulk fluid nodes change.
  void iv_entire_grid() {
    std::fill(visc.begin(), visc.end(), (uint8_t)0);
    const int nc = (int)cell_type.size();
    for (int c = 0; c < nc; ++c) {
      const int t = cell_type[c];
      if (!(ivt(t) > 0.0) || incomplete(c)) continue;
      const CellType& ct = types[t];
      const long b = cell_base[c];
      double lo[3], hi[3];
      for (int a = 0; a < 3; ++a) lo[a] = hi[a] = x[3 * b + a];
      for (int k = 1; k < ct.nv; ++k)
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(lo[a], x[3 * (b + k) + a]);
          hi[a] = std::max(hi[a], x[3 * (b + k) + a]);
        }
      long a0[3], a1[3];
      for (int a = 0; a < 3; ++a) {
        a0[a] = (long)lo[a];
        a1[a] = (long)hi[a];
        if (!periodic[a]) {
          a0[a] = std::max(a0[a], 0L);
          a1[a] = std::min(a1[a], (long)dim(a) - 1);
        }
      }
      for (long X = a0[0]; X <= a1[0]; ++X)
        for (long Y = a0[1]; Y <= a1[1]; ++Y)
          for (long Z = a0[2]; Z <= a1[2]; ++Z) {
            const long s[3] = {X, Y, Z};
            int cnt = 0;
            for (int q = 0; q < ct.nt; ++q) {
              const int* T = &ct.tri[3 * q];
              cnt += ray_crosses(&x[3 * (b + T[0])], &x[3 * (b + T[1])], &x[3 * (b + T[2])], s);
            }
            if (!(cnt % 2)) continue;
            const long n = idx(wrap(0, X), wrap(1, Y), wrap(2, Z));
            if (type[n] == FLUID) visc[n] = (uint8_t)(1 + t);
          }
    }
  }

  // Membrane pass (hemoCellParticleField.cpp:745-774): each vertex, in vertex
  // order, visits the support nodes of its position before this step's advance
  // (the kernel coordinates of the interpolation); a node within edge_mean_eq
  // of the new position is inner when (node - x) . normal < 0, else background.
  // normal = sum over the vertex's triangles, ascending, of unit normal x
  // area / area_mean_eq at the new positions (rbcHighOrderModel.cpp:114-121;
  // only the RBC model computes it, so the other models' normal is zero).
  void iv_membrane(const std::vector<double>& xold) {
    std::vector<std::vector<std::vector<int>>> vt(types.size());
    for (long i = 0; i < NV(); ++i) {
      const int c = vcell[i], t = cell_type[c];
      if (!(ivt(t) > 0.0) || removed[i]) continue;
      const CellType& ct = types[t];
      const long b = cell_base[c];
      const int k = (int)(i - b);
      double nrm[3] = {0.0, 0.0, 0.0};
      if (ct.model == RBC) {
        if (vt[t].empty()) vt[t] = vertex_triangles(ct);
        for (int q : vt[t][k]) {
          const int* T = &ct.tri[3 * q];
          double area;
          V3 n;
          area_and_unit_normal(&x[3 * (b + T[0])], &x[3 * (b + T[1])], &x[3 * (b + T[2])], area, n);
          const double s = area / ct.area_mean_eq;
          for (int a = 0; a < 3; ++a) nrm[a] += n.x[a] * s;
        }
      }
      long nd[27][3], nid[27];
      const int ns = support_nodes(&xold[3 * i], nd, nid);
      for (int j = 0; j < ns; ++j) {
        double lp[3];
        for (int a = 0; a < 3; ++a) lp[a] = (double)nd[j][a] - x[3 * i + a];
        const double len = std::sqrt(lp[0] * lp[0] + lp[1] * lp[1] + lp[2] * lp[2]);
        if (len > ct.edge_mean_eq) continue;
        const double d = lp[0] * nrm[0] + lp[1] * nrm[1] + lp[2] * nrm[2];
        visc[nid[j]] = d < 0.0 ? (uint8_t)(1 + t) : (uint8_t)0;
      }
    }
  }

  // the nodes make_kernel keeps for position p: unwrapped coordinates and index
  int support_nodes(const double* p, long nd[27][3], long nid[27]) const {
    int m = 0;
    long ctr[3];
    for (int a = 0; a < 3; ++a) ctr[a] = (long)std::floor(p[a] + 0.5);
    for (int dx = -1; dx < 2; ++dx)
      for (int dy = -1; dy < 2; ++dy)
        for (int dz = -1; dz < 2; ++dz) {
          const long q[3] = {ctr[0] + dx, ctr[1] + dy, ctr[2] + dz};
          int w3[3];
          bool out = false;
          for (int a = 0; a < 3; ++a) {
            w3[a] = wrap(a, q[a]);
            if (w3[a] < 0) out = true;
          }
          if (out) continue;
          double w = 1.0;
          for (int a = 0; a < 3; ++a) w *= std::max(1.0 - std::fabs(p[a] - (double)q[a]), 0.0);
          if (w == 0.0) continue;
          const long n = idx(w3[0], w3[1], w3[2]);
          if (is_boundary(type[n])) continue;
          for (int a = 0; a < 3; ++a) nd[m][a] = q[a];
          nid[m++] = n;
        }
    return m;
  }

  // Lees-Edwards planes after the collide-and-stream (helper/leesEdwardsBC.h:
  // 101-215, then :275-298).  Each node of the top (z = nz-1) and bottom (z = 0)
  // plane keeps a copy of its populations relaxed once more, by BGK with its
  // own omega, towards the equilibrium of its rhoBar and the momentum
  // (u_plane, 0, 0); the five populations that entered across the plane are
  // then replaced by the linear interpolation, at the x displaced by the
  // current displacement d, of the plane's un-relaxed populations (the two
  // diagonal ones exchanged); all copies are written back at the end.
  void lees_edwards() {
    const double d = std::fmod(le_disp * (double)iter, (double)nx);
    const double g = std::fmod(d, 1.0);
    auto mod = [](long a, long b) { return (a % b + b) % b; };
    std::vector<double> out((size_t)2 * nx * ny * 19);
    for (int side = 0; side < 2; ++side) {  // 0 top, 1 bottom
      const int Z = side == 0 ? nz - 1 : 0;
      const double uw = side == 0 ? le_top_u : le_bottom_u;
      for (int X = 0; X < nx; ++X)
        for (int Y = 0; Y < ny; ++Y) {
          const long n = idx(X, Y, Z);
          double* o = &out[(((size_t)side * nx + X) * ny + Y) * 19];
          double fl[19];
          for (int i = 0; i < 19; ++i) fl[i] = f[(long)i * N + n];
          if (type[n] != FLUID) {  // wall nodes keep their own dynamics
            for (int i = 0; i < 19; ++i) o[i] = fl[i];
            continue;
          }
          double rhoBar, j[3];
          rhobar_j(fl, rhoBar, j);
          const double je[3] = {uw, 0.0, 0.0};
          const double jSqr = je[0] * je[0] + je[1] * je[1] + je[2] * je[2];
          const double invRho = 1.0 / (rhoBar + 1.0);
          const double ti = (!visc.empty() && visc[n]) ? ivt(visc[n] - 1) : 0.0;
          const double om = ti > 0.0 ? 1.0 / ti : omega;
          for (int i = 0; i < 19; ++i) {
            o[i] = fl[i] * (1.0 - om);
            o[i] += om * feq(i, rhoBar, invRho, je, jSqr);
          }
          long s1, s2;
          if (side == 0) {
            s1 = mod((long)std::ceil(d + X), nx);
            s2 = mod((long)std::floor(d + X), nx);
          } else {
            s1 = mod((long)std::floor(-d + X), nx);
            s2 = mod((long)std::ceil(-d + X), nx);
          }
          auto P = [&](long sx, int i) { return f[(long)i * N + idx((int)sx, Y, Z)]; };
          auto lerp = [&](int i) { return g * P(s1, i) + (1 - g) * P(s2, i); };
          if (side == 0) {
            o[3] = lerp(3);
            o[6] = lerp(16);
            o[8] = lerp(8);
            o[16] = lerp(6);
            o[18] = lerp(18);
          } else {
            o[7] = lerp(15);
            o[9] = lerp(9);
            o[12] = lerp(12);
            o[15] = lerp(7);
            o[17] = lerp(17);
          }
        }
    }
    for (int side = 0; side < 2; ++side) {
      const int Z = side == 0 ? nz - 1 : 0;
      for (int X = 0; X < nx; ++X)
        for (int Y = 0; Y < ny; ++Y) {
          const double* o = &out[(((size_t)side * nx + X) * ny + Y) * 19];
          for (int i = 0; i < 19; ++i) f[(long)i * N + idx(X, Y, Z)] = o[i];
        }
    }
  }

  void reset_force_field() { std::fill(Fibm.begin(), Fibm.end(), 0.0); }

  void step() {
    if (repulsion && iter % rep_timescale == 0) repulsion_force();
    if (brep && iter % brep_timescale == 0) boundary_repulsion_force();
    spread();
    collide_and_stream();
    if (le) lees_edwards();
    if (iter % pv_timescale == 0) interpolate();
    const bool iv = iv_on();
    std::vector<double> xold;
    if (iv && iter % iv_timescale == 0) xold = x;
    advance();
    mechanics(false);
    // hemoCell.cpp:347-357: after the constitutive model
    if (iv && iter % iv_grid_timescale == 0) {
      delete_incomplete_cells();  // hemoCell.cpp:347-351
      iv_entire_grid();
    }
    if (iv && iter % iv_timescale == 0) iv_membrane(xold);
    reset_force_field();
    iter++;
  }
};

}  // namespace

// =================================================================== C ABI ==
extern "C" {

void* ho_create(int nx, int ny, int nz, const uint8_t* types, double tau, const int* periodic) {
  Oracle* o = new Oracle();
  o->nx = nx; o->ny = ny; o->nz = nz;
  o->N = (long)nx * ny * nz;
  for (int a = 0; a < 3; ++a) o->periodic[a] = periodic ? periodic[a] : 1;
  o->omega = 1.0 / tau;
  o->type.assign(o->N, FLUID);
  if (types) std::memcpy(o->type.data(), types, o->N);
  o->f.assign(19 * o->N, 0.0);
  o->ftmp.assign(19 * o->N, 0.0);
  o->Fibm.assign(3 * o->N, 0.0);
  return o;
}
void ho_destroy(void* h) { delete (Oracle*)h; }
void ho_set_body_force(void* h, const double* g) { Oracle* o = (Oracle*)h; for (int c = 0; c < 3; ++c) o->g[c] = g[c]; }
void ho_set_wall_velocity(void* h, int face, const double* u) { Oracle* o = (Oracle*)h; for (int c = 0; c < 3; ++c) o->wall_u[face][c] = u[c]; }
void ho_set_f_limit(void* h, double fl) { ((Oracle*)h)->f_limit = fl; }
void ho_set_timescales(void* h, int pv) { ((Oracle*)h)->pv_timescale = pv; }
void ho_set_repulsion(void* h, double k, double cutoff, int timescale) {
  Oracle* o = (Oracle*)h;
  o->repulsion = true; o->rep_const = k; o->rep_cutoff = cutoff; o->rep_timescale = timescale;
}
void ho_set_iter(void* h, long it) { ((Oracle*)h)->iter = it; }

// Equilibrium at density rho and velocity u on fluid and velocity-wall nodes; 0 on bounce-back nodes.
void ho_init_equilibrium(void* h, double rho, const double* u) {
  Oracle* o = (Oracle*)h;
  double rhoBar = rho - 1.0, invRho = 1.0 / rho;
  double j[3] = {rho * u[0], rho * u[1], rho * u[2]};
  double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
  for (long n = 0; n < o->N; ++n)
    for (int i = 0; i < 19; ++i)
      o->f[(long)i * o->N + n] = (o->type[n] == BOUNCEBACK) ? 0.0 : feq(i, rhoBar, invRho, j, jSqr);
}

int ho_add_cell_type(void* h, int model, int nv, int nt, const int* tri, const double* rest, int ninner,
                     const int* inner, const double* params, int timescale) {
  Oracle* o = (Oracle*)h;
  CellType ct;
  ct.model = model; ct.nv = nv; ct.nt = nt;
  ct.tri.assign(tri, tri + 3 * nt);
  ct.rest.assign(rest, rest + 3 * nv);
  if (ninner > 0) ct.inner.assign(inner, inner + 2 * ninner);
  ct.k_link = params[0]; ct.k_bend = params[1]; ct.k_volume = params[2]; ct.k_area = params[3];
  ct.eta_m = params[4]; ct.k_inner_rigid = params[5]; ct.k_cytoskeleton = params[6];
  ct.core_radius = params[7]; ct.radius = params[8];
  ct.timescale = timescale > 0 ? timescale : 1;
  try { build_tables(ct); } catch (...) { return -1; }
  o->types.push_back(ct);
  return (int)o->types.size() - 1;
}

int ho_add_cells(void* h, int t, int ncells, const double* xyz) {
  Oracle* o = (Oracle*)h;
  const CellType& ct = o->types[t];
  for (int c = 0; c < ncells; ++c) {
    int cell = (int)o->cell_type.size();
    o->cell_type.push_back(t);
    o->cell_base.push_back((int)o->NV());
    for (int k = 0; k < ct.nv; ++k) {
      for (int q = 0; q < 3; ++q) {
        o->x.push_back(xyz[3 * ((long)c * ct.nv + k) + q]);
        o->v.push_back(0.0);
        o->force.push_back(0.0);
        o->frep.push_back(0.0);
      }
      o->vcell.push_back(cell);
      o->removed.push_back(0);
    }
    if (o->brep) o->fbr.resize(3 * o->NV(), 0.0);
  }
  return 0;
}

// Table export for cross-checking the product's constant builder.
void ho_type_sizes(void* h, int t, int* out4) {
  const CellType& ct = ((Oracle*)h)->types[t];
  out4[0] = ct.nv; out4[1] = ct.nt; out4[2] = ct.ne; out4[3] = (int)ct.inner.size() / 2;
}
void ho_type_tables(void* h, int t, int* edge, int* edge_tris, int* edge_outer, int* ring, int* nring,
                    double* edge_len, double* edge_angle, double* tri_area, double* patch_dist,
                    double* inner_len, double* scalars4) {
  const CellType& ct = ((Oracle*)h)->types[t];
  std::copy(ct.edge.begin(), ct.edge.end(), edge);
  std::copy(ct.edge_tris.begin(), ct.edge_tris.end(), edge_tris);
  std::copy(ct.edge_outer.begin(), ct.edge_outer.end(), edge_outer);
  std::copy(ct.ring.begin(), ct.ring.end(), ring);
  std::copy(ct.nring.begin(), ct.nring.end(), nring);
  std::copy(ct.edge_len_eq.begin(), ct.edge_len_eq.end(), edge_len);
  std::copy(ct.edge_angle_eq.begin(), ct.edge_angle_eq.end(), edge_angle);
  std::copy(ct.tri_area_eq.begin(), ct.tri_area_eq.end(), tri_area);
  std::copy(ct.patch_dist_eq.begin(), ct.patch_dist_eq.end(), patch_dist);
  std::copy(ct.inner_len_eq.begin(), ct.inner_len_eq.end(), inner_len);
  scalars4[0] = ct.volume_eq; scalars4[1] = ct.area_mean_eq; scalars4[2] = ct.edge_mean_eq; scalars4[3] = ct.angle_mean_eq;
}

int ho_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
long ho_num_vertices(void* h) { return ((Oracle*)h)->NV(); }
long ho_wall_hits(void* h) { return ((Oracle*)h)->wall_hits; }
void ho_get_removed(void* h, uint8_t* out) {
  Oracle* o = (Oracle*)h;
  std::copy(o->removed.begin(), o->removed.end(), out);
}
void ho_set_removed(void* h, const uint8_t* in) {
  Oracle* o = (Oracle*)h;
  for (long i = 0; i < o->NV(); ++i) o->removed[i] = in[i] ? 1 : 0;
}

void ho_get_populations(void* h, double* f) { Oracle* o = (Oracle*)h; std::copy(o->f.begin(), o->f.end(), f); }
void ho_set_populations(void* h, const double* f) { Oracle* o = (Oracle*)h; std::copy(f, f + 19 * o->N, o->f.begin()); }
void ho_get_force_field(void* h, double* F) { Oracle* o = (Oracle*)h; std::copy(o->Fibm.begin(), o->Fibm.end(), F); }
void ho_set_force_field(void* h, const double* F) { Oracle* o = (Oracle*)h; std::copy(F, F + 3 * o->N, o->Fibm.begin()); }

// Density and velocity (u = j/rho + F/2 with the current force field) of every
// node; a velocity-boundary node reports its imposed velocity and the density of
// the mass balance over its known populations (the boundary dynamics' moments).
void ho_get_fluid(void* h, double* rho, double* u) {
  Oracle* o = (Oracle*)h;
  for (long n = 0; n < o->N; ++n) {
    double fl[19];
    for (int i = 0; i < 19; ++i) fl[i] = o->f[(long)i * o->N + n];
    double rhoBar, j[3], uu[3];
    rhobar_j(fl, rhoBar, j);
    o->velocity_at(n, uu);
    if (o->type[n] >= VELWALL0) {
      const int face = o->type[n] - VELWALL0, axis = face / 2, orient = (face % 2) ? 1 : -1;
      double on = 0.0, nrm = 0.0;
      for (int i = 0; i < 19; ++i) {
        if (C[i][axis] == 0) on += fl[i];
        else if (C[i][axis] == orient) nrm += fl[i];
      }
      const double vn = orient * o->wall_u[face][axis];
      rhoBar = (2.0 * nrm + on - vn) / (1.0 + vn);
      for (int c = 0; c < 3; ++c) uu[c] = o->wall_u[face][c];
    }
    if (rho) rho[n] = rhoBar + 1.0;
    if (u) for (int c = 0; c < 3; ++c) u[3 * n + c] = uu[c];
  }
}

void ho_get_vertices(void* h, double* x, double* v, double* f) {
  Oracle* o = (Oracle*)h;
  if (x) std::copy(o->x.begin(), o->x.end(), x);
  if (v) std::copy(o->v.begin(), o->v.end(), v);
  if (f) std::copy(o->force.begin(), o->force.end(), f);
}
void ho_set_vertices(void* h, const double* x, const double* v, const double* f) {
  Oracle* o = (Oracle*)h;
  if (x) std::copy(x, x + o->x.size(), o->x.begin());
  if (v) std::copy(v, v + o->v.size(), o->v.begin());
  if (f) std::copy(f, f + o->force.size(), o->force.begin());
}
// Nearest-node bin of every vertex as update_pg forms it (hemoCellParticleField.cpp:150-163),
// -1 for a vertex outside the domain.
void ho_vertex_bins(void* h, long* out) {
  Oracle* o = (Oracle*)h;
  for (long i = 0; i < o->NV(); ++i) {
    int w3[3];
    bool out_of = false;
    for (int a = 0; a < 3; ++a) {
      w3[a] = o->wrap(a, (long)std::floor(o->x[3 * i + a] + 0.5));
      if (w3[a] < 0) out_of = true;
    }
    out[i] = out_of ? -1 : o->idx(w3[0], w3[1], w3[2]);
  }
}
void ho_set_boundary_repulsion(void* h, double k, double cutoff, int timescale) {
  Oracle* o = (Oracle*)h;
  o->brep = true;
  o->brep_const = k;
  o->brep_cutoff = cutoff;
  o->brep_timescale = timescale;
  o->populate_boundary();
  o->fbr.assign(3 * o->NV(), 0.0);
}
void ho_get_boundary_repulsion(void* h, double* fr) {
  Oracle* o = (Oracle*)h;
  for (long i = 0; i < 3 * o->NV(); ++i) fr[i] = i < (long)o->fbr.size() ? o->fbr[i] : 0.0;
}
void ho_set_interior_viscosity(void* h, int t, double tau_inner) {
  Oracle* o = (Oracle*)h;
  if (o->iv_tau.size() < o->types.size()) o->iv_tau.resize(o->types.size(), 0.0);
  o->iv_tau[t] = tau_inner > 0.0 ? tau_inner : 0.0;
  if (o->visc.empty()) o->visc.assign(o->N, 0);
}
void ho_set_interior_viscosity_timescales(void* h, int membrane, int grid) {
  ((Oracle*)h)->iv_timescale = membrane;
  ((Oracle*)h)->iv_grid_timescale = grid;
}
// per node: 0, or the tau_inner of the type whose membrane holds it
void ho_get_interior_viscosity(void* h, double* tau) {
  Oracle* o = (Oracle*)h;
  for (long n = 0; n < o->N; ++n) tau[n] = (!o->visc.empty() && o->visc[n]) ? o->ivt(o->visc[n] - 1) : 0.0;
}
void ho_set_interior_viscosity_classes(void* h, const uint8_t* cls) {
  Oracle* o = (Oracle*)h;
  o->visc.assign(cls, cls + o->N);
}
// pass 0: entire grid; pass 1: membrane with the support at the current positions
void ho_stage_interior_viscosity(void* h, int pass) {
  Oracle* o = (Oracle*)h;
  if (pass == 0) o->iv_entire_grid();
  else o->iv_membrane(o->x);
}
void ho_set_lees_edwards(void* h, double shear_rate, double displacement_per_step) {
  Oracle* o = (Oracle*)h;
  o->le = true;
  const double vhalf = (o->nz - 1) * shear_rate * 0.5;
  o->le_top_u = -vhalf;
  o->le_bottom_u = vhalf;
  o->le_disp = displacement_per_step;
}
void ho_stage_lees_edwards(void* h) { ((Oracle*)h)->lees_edwards(); }
int ho_ray_crosses(const double* v0, const double* v1, const double* v2, const long* s) {
  return Oracle::ray_crosses(v0, v1, v2, s);
}
void ho_stage_boundary_repulsion(void* h) { ((Oracle*)h)->boundary_repulsion_force(); }
void ho_get_repulsion(void* h, double* fr) { Oracle* o = (Oracle*)h; std::copy(o->frep.begin(), o->frep.end(), fr); }

// Individual stages, in iterate() order, for stage-level parity tests.
void ho_stage_repulsion(void* h) { ((Oracle*)h)->repulsion_force(); }
void ho_stage_spread(void* h) { ((Oracle*)h)->spread(); }
void ho_stage_collide_stream(void* h) {
  Oracle* o = (Oracle*)h;
  o->collide_and_stream();
  if (o->le) o->lees_edwards();
}
void ho_stage_interpolate(void* h) { ((Oracle*)h)->interpolate(); }
void ho_stage_advance(void* h) { ((Oracle*)h)->advance(); }
void ho_stage_mechanics(void* h, int forced) { ((Oracle*)h)->mechanics(forced != 0); }
void ho_stage_reset_force(void* h) { ((Oracle*)h)->reset_force_field(); }
void ho_step(void* h, int n) { Oracle* o = (Oracle*)h; for (int s = 0; s < n; ++s) o->step(); }

// phi2 weights of one position (tests of the kernel itself).
int ho_ibm_kernel(void* h, const double* p, long* idx, double* w) {
  Kernel k;
  ((Oracle*)h)->make_kernel(p, k);
  for (int i = 0; i < k.n; ++i) { idx[i] = k.idx[i]; w[i] = k.w[i]; }
  return k.n;
}

// One Guo-BGK collision of a single population vector (unit tests).
void ho_collide_one(double* f, const double* F, double omega) { collide_guo(f, F, omega); }
void ho_velocity_bc_one(double* f, int axis, int orient, const double* uw) { velocity_bc_complete(f, axis, orient, uw); }
double ho_feq(int i, double rhoBar, const double* j) {
  double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
  return feq(i, rhoBar, 1.0 / (rhoBar + 1.0), j, jSqr);
}
}
