// hc_mech_dev.h -- device helpers of the membrane mechanics, shared by the
// mechanics kernels and the fused per-cell step kernel.
#pragma once
#include <cmath>

#include "hc_launch.h"

namespace hcg {
namespace {

constexpr double MaxCellVolumetricChange = 0.01;   // config/constant_defaults.h:157-173
constexpr double MaxCellSurfaceAreaChange = 0.09;
constexpr double MaxCellBendingAngle = 0.0555;
constexpr double MaxPLTBendingAngle = 2.467;
constexpr double MaxCellPersistenceLength = 9.0;
constexpr double FORCE_LIMIT = 50.0;               // config/constant_defaults.h:73-75

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 crossp(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dotp(D3 a, D3 b) {
  double r = 0.0;
  r += a.x * b.x;
  r += a.y * b.y;
  r += a.z * b.z;
  return r;
}
__device__ __forceinline__ double normp(D3 a) {
  double r = 0.0;
  r += a.x * a.x;
  r += a.y * a.y;
  r += a.z * a.z;
  return sqrt(r);
}

// Positions of one cell's vertices: three coordinate arrays, in LDS when the
// cell is staged there, otherwise in the global vertex planes.
struct CellView {
  const double *px, *py, *pz;
  __device__ __forceinline__ D3 pos(int k) const { return {px[k], py[k], pz[k]}; }
};

// area and unit normal of triangle t (array.h:271-285)
__device__ __forceinline__ void tri_geom(const CellView& c, const TypeDev& T, int t, D3& v0, D3& v1, D3& v2,
                                         double& area, D3& n) {
  v0 = c.pos(T.tri[3 * t]);
  v1 = c.pos(T.tri[3 * t + 1]);
  v2 = c.pos(T.tri[3 * t + 2]);
  n = crossp(v1 - v0, v2 - v0);
  const double nn = normp(n);
  if (nn != 0.0) {
    area = 0.5 * nn;
    n.x /= nn;
    n.y /= nn;
    n.z /= nn;
  } else {
    area = 0.0;
    n = {0.0, 0.0, 0.0};
  }
}

__device__ __forceinline__ D3 unit_normal(D3 v0, D3 v1, D3 v2) {
  D3 n = crossp(v1 - v0, v2 - v0);
  const double nn = normp(n);
  if (nn != 0.0) {
    n.x /= nn;
    n.y /= nn;
    n.z /= nn;
  } else {
    n = {0.0, 0.0, 0.0};
  }
  return n;
}

// bending force of vertex i from its ordered 1-ring (rbcHighOrderModel.cpp:127-166)
__device__ D3 bending_force(const CellView& c, const TypeDev& T, int i) {
  const int nr = T.nring[i];
  // the ring (at most 6 entries, stored 6 wide) is loaded at once
  int R[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) R[j] = T.ring[6 * i + j];
  const double patch_eq = T.patch_dist_eq[i];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= nr) break;
    D3 p = c.pos(R[j]);
    s0 += p.x;
    s1 += p.y;
    s2 += p.z;
  }
  const D3 xi = c.pos(i);
  const double inv_n = 1.0 / (double)(unsigned)nr;
  const D3 dev = {s0 * inv_n - xi.x, s1 * inv_n - xi.y, s2 * inv_n - xi.z};
  D3 pn = {0.0, 0.0, 0.0};
  D3 prev = c.pos(R[0]);
  const D3 first = prev;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= nr) break;
    const D3 nxt = (j + 1 < nr) ? c.pos(R[j + 1 < 6 ? j + 1 : 0]) : first;
    D3 tn = crossp(prev - xi, nxt - xi);
    const double il = 1.0 / normp(tn);
    pn.x += tn.x * il;
    pn.y += tn.y * il;
    pn.z += tn.z * il;
    prev = nxt;
  }
  const double il = 1.0 / normp(pn);
  pn.x *= il;
  pn.y *= il;
  pn.z *= il;
  const double ndev = dotp(pn, dev);
  const double dDev = (ndev - patch_eq) / T.edge_mean_eq;
  const double s = T.k_bend * (dDev + dDev / fabs(MaxCellBendingAngle - dDev * dDev));
  return {pn.x * s, pn.y * s, pn.z * s};
}

// Force terms, as the reference's models add them to force_area, force_volume,
// force_bending, force_link, force_visc and force_inner_link
// (rbcHighOrderModel.cpp, pltSimpleModel.cpp, wbcHighOrderModel.cpp).
enum { TERM_AREA = 1, TERM_VOLUME = 2, TERM_BENDING = 4, TERM_LINK = 8, TERM_VISC = 16, TERM_INNER = 32,
       TERM_ALL = 63 };

// Force on vertex k of one cell: every term gathered in the reference's order
// (area, volume, bending, edges, inner links), or only the terms in TERMS.
// bx/by/bz: per-vertex bending forces of the cell; vx/vy/vz: vertex velocities
// (membrane viscosity).
template <int MODEL, int TERMS = TERM_ALL>
__device__ __forceinline__ D3 vertex_force(const CellView& c, const TypeDev& T, int k, double vforce, const double* bx,
                                           const double* by, const double* bz, const double* vx, const double* vy,
                                           const double* vz, bool visc) {
  double F0 = 0.0, F1 = 0.0, F2 = 0.0;
  const D3 xk = c.pos(k);
  // area and volume terms, triangle order: each incident triangle is evaluated
  // once; the two sums are kept apart and added in the reference's order
  // (all area terms, then all volume terms) up to the grouping of that sum
  {
    double A0 = 0.0, A1 = 0.0, A2 = 0.0, B0 = 0.0, B1 = 0.0, B2 = 0.0;
    const double inv_area_mean = 1.0 / T.area_mean_eq;
    auto tri_term = [&](const int4 q, const double a_eq) {
      const D3 v0 = c.pos(q.y), v1 = c.pos(q.z), v2 = c.pos(q.w);
      D3 n = crossp(v1 - v0, v2 - v0);
      const double nn = normp(n);
      double area = 0.0;
      if (nn != 0.0) {
        area = 0.5 * nn;
        const double inn = 1.0 / nn;
        n.x *= inn;
        n.y *= inn;
        n.z *= inn;
      } else {
        n = {0.0, 0.0, 0.0};
      }
      const double ar = (area - a_eq) / a_eq;
      const double afm = T.k_area * (ar + ar / fabs(MaxCellSurfaceAreaChange - ar * ar));
      constexpr double third = 1.0 / 3.0;
      const double c0 = (v0.x + v1.x + v2.x) * third, c1 = (v0.y + v1.y + v2.y) * third,
                   c2 = (v0.z + v1.z + v2.z) * third;
      A0 += (c0 - xk.x) * afm;
      A1 += (c1 - xk.y) * afm;
      A2 += (c2 - xk.z) * afm;
      const double sc = area * inv_area_mean;
      B0 += (vforce * n.x) * sc;
      B1 += (vforce * n.y) * sc;
      B2 += (vforce * n.z) * sc;
    };
    if (T.vtri4p) {
      int4 q[kPadW];
      double aeq[kPadW];
#pragma unroll
      for (int j = 0; j < kPadW; ++j) {
        q[j] = T.vtri4p[kPadW * k + j];
        aeq[j] = T.vtri_aeq[kPadW * k + j];
      }
#pragma unroll
      for (int j = 0; j < kPadW; ++j)
        if (q[j].x >= 0) tri_term(q[j], aeq[j]);
    } else {
      for (int e = T.vtri_off[k]; e < T.vtri_off[k + 1]; ++e) tri_term(T.vtri4[e], T.tri_area_eq[T.vtri4[e].x]);
    }
    if ((TERMS & TERM_AREA) && (TERMS & TERM_VOLUME)) {
      F0 = A0 + B0;
      F1 = A1 + B1;
      F2 = A2 + B2;
    } else if (TERMS & TERM_AREA) {
      F0 = A0;
      F1 = A1;
      F2 = A2;
    } else if (TERMS & TERM_VOLUME) {
      F0 = B0;
      F1 = B1;
      F2 = B2;
    }
  }
  // bending: own force and -F_i/n_i of every vertex i whose 1-ring holds k, ascending i
  if (MODEL != HC_MODEL_PLT && (TERMS & TERM_BENDING)) {
    auto bend_term = [&](const int i, const double dn) {
      const double b0 = bx[i], b1 = by[i], b2 = bz[i];
      if (i == k) {
        F0 += b0;
        F1 += b1;
        F2 += b2;
      } else {
        const double in = 1.0 / dn;
        F0 += -b0 * in;
        F1 += -b1 * in;
        F2 += -b2 * in;
      }
    };
    if (T.vbendp) {
      int b[kPadB];
      double dn[kPadB];
#pragma unroll
      for (int j = 0; j < kPadB; ++j) {
        b[j] = T.vbendp[kPadB * k + j];
        dn[j] = T.vbend_n[kPadB * k + j];
      }
#pragma unroll
      for (int j = 0; j < kPadB; ++j)
        if (b[j] >= 0) bend_term(b[j], dn[j]);
    } else {
      for (int e = T.vbend_off[k]; e < T.vbend_off[k + 1]; ++e)
        bend_term(T.vbend[e], (double)(unsigned)T.nring[T.vbend[e]]);
    }
  }
  // edges: link, membrane viscosity, (PLT) dihedral bending; edge order
  auto edge_term = [&](const int4 q, const double l0) {
    const int ed = q.x >> 2, kind = q.x & 3;
    const int a = q.y, b = q.z;
    const D3 pa = c.pos(a), pb = c.pos(b);
    const D3 ev = pb - pa;
    const double len = normp(ev);
    const double il = 1.0 / len;
    const D3 uv = {ev.x * il, ev.y * il, ev.z * il};
    if (kind != 2) {
      const double sgn = kind == 0 ? 1.0 : -1.0;
      if (TERMS & TERM_LINK) {
        const double ef = (len - l0) / l0;
        const double fs = T.k_link * (ef + ef / fabs(MaxCellPersistenceLength - ef * ef));
        const D3 fl = {uv.x * fs, uv.y * fs, uv.z * fs};
        if (sgn > 0) { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
        else { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
      }
      if ((TERMS & TERM_VISC) && visc) {
        const D3 rv = {vx[b] - vx[a], vy[b] - vy[a], vz[b] - vz[a]};
        const double pr = dotp(rv, uv);
        D3 fv = {(pr * uv.x) * T.eta_m, (pr * uv.y) * T.eta_m, (pr * uv.z) * T.eta_m};
        const double m = normp(fv);
        if (m > FORCE_LIMIT / 4.0) {
          const double s = (FORCE_LIMIT / 4.0) / m;
          fv.x *= s;
          fv.y *= s;
          fv.z *= s;
        }
        if (sgn > 0) { F0 += fv.x; F1 += fv.y; F2 += fv.z; }
        else { F0 -= fv.x; F1 -= fv.y; F2 -= fv.z; }
      }
    }
    if (MODEL == HC_MODEL_PLT && (TERMS & TERM_BENDING)) {
      const int32_t* T0 = T.tri + 3 * T.edge_tris[2 * ed];
      const int32_t* T1 = T.tri + 3 * T.edge_tris[2 * ed + 1];
      const D3 n1 = unit_normal(c.pos(T0[0]), c.pos(T0[1]), c.pos(T0[2]));
      const D3 n2 = unit_normal(c.pos(T1[0]), c.pos(T1[1]), c.pos(T1[2]));
      const double ang = atan2(dotp(crossp(n1, n2), uv), dotp(n1, n2));
      const double af = ang - T.edge_angle_eq[ed];
      const double fm = T.k_bend * (af + af / fabs(MaxPLTBendingAngle - af * af));
      const D3 bf = {((n1.x + n2.x) * fm) * 0.5, ((n1.y + n2.y) * fm) * 0.5, ((n1.z + n2.z) * fm) * 0.5};
      if (kind != 2) { F0 += bf.x; F1 += bf.y; F2 += bf.z; }
      else { F0 -= bf.x; F1 -= bf.y; F2 -= bf.z; }
    }
  };
  if (MODEL != HC_MODEL_PLT && T.vedge4p) {
    int4 q[kPadW];
    double l0[kPadW];
#pragma unroll
    for (int j = 0; j < kPadW; ++j) {
      q[j] = T.vedge4p[kPadW * k + j];
      l0[j] = T.vedge_l0[kPadW * k + j];
    }
#pragma unroll
    for (int j = 0; j < kPadW; ++j)
      if (q[j].x >= 0) edge_term(q[j], l0[j]);
  } else {
    for (int e = T.vedge_off[k]; e < T.vedge_off[k + 1]; ++e) edge_term(T.vedge4[e], T.edge_len_eq[T.vedge4[e].x >> 2]);
  }
  // inner links, inner-edge order
  for (int e = T.vinner_off[k]; (TERMS & TERM_INNER) && e < T.vinner_off[k + 1]; ++e) {
    const int ie = T.vinner[e] >> 1, end = T.vinner[e] & 1;
    const D3 pa = c.pos(T.inner[2 * ie]), pb = c.pos(T.inner[2 * ie + 1]);
    const D3 ev = pb - pa;
    const double len = normp(ev);
    const D3 uv = {ev.x / len, ev.y / len, ev.z / len};
    if (MODEL == HC_MODEL_PLT) {
      const double ef = (len - T.inner_len_eq[ie]) / T.inner_len_eq[ie];
      const double fs = T.k_link * 5.0 * ef;
      const D3 fl = {uv.x * fs, uv.y * fs, uv.z * fs};
      if (end == 0) { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
      else { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
    } else if (MODEL == HC_MODEL_WBC) {
      if (len < 2 * T.radius) {
        const double s = 1.0 - (len / (2 * T.radius));
        const D3 fl = {(uv.x * s) * T.k_cyto, (uv.y * s) * T.k_cyto, (uv.z * s) * T.k_cyto};
        if (end == 0) { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
        else { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
      }
      if (len < 2 * T.core_radius) {
        const double s = 1 - (len / (2 * T.core_radius));
        const D3 fl = {(uv.x * s) * T.k_inner_rigid, (uv.y * s) * T.k_inner_rigid, (uv.z * s) * T.k_inner_rigid};
        if (end == 0) { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
        else { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
      }
    }
  }
  return {F0, F1, F2};
}

// Sum of the signed triple products of the cell's triangles handled by thread tid
// (the cell volume is 1/6 of the sum over all threads).
template <int BS>
__device__ __forceinline__ double volume_partial(const CellView& c, const TypeDev& T, int tid) {
  double part = 0.0;
  // corners relative to the cell's vertex 0: the same volume for a closed mesh,
  // without the cancellation of triple products of coordinates hundreds of
  // spacings from the origin
  const D3 r = c.pos(0);
  for (int t = tid; t < T.nt; t += BS) {
    const D3 v0 = c.pos(T.tri[3 * t]) - r, v1 = c.pos(T.tri[3 * t + 1]) - r, v2 = c.pos(T.tri[3 * t + 2]) - r;
    const double v210 = v2.x * v1.y * v0.z;
    const double v120 = v1.x * v2.y * v0.z;
    const double v201 = v2.x * v0.y * v1.z;
    const double v021 = v0.x * v2.y * v1.z;
    const double v102 = v1.x * v0.y * v2.z;
    const double v012 = v0.x * v1.y * v2.z;
    part += (-v210 + v120 + v201 - v021 - v102 + v012);
  }
  return part;
}

}  // namespace
}  // namespace hcg
