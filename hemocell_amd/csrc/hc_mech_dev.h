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

// A 3-vector of the mechanics scalar M: double in the FP64 form; float for the
// FP32-shifted form's cell kernels, which take positions relative to the cell.
template <typename M>
struct V3T {
  M x, y, z;
};
using D3 = V3T<double>;
template <typename M>
__device__ __forceinline__ V3T<M> operator-(V3T<M> a, V3T<M> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename M>
__device__ __forceinline__ V3T<M> crossp(V3T<M> a, V3T<M> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename M>
__device__ __forceinline__ M dotp(V3T<M> a, V3T<M> b) {
  M r = M(0);
  r += a.x * b.x;
  r += a.y * b.y;
  r += a.z * b.z;
  return r;
}
// square root, quotient and reciprocal: IEEE double in FP64; in FP32 the
// hardware's one-ulp square root and reciprocal
__device__ __forceinline__ double msqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float msqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double mdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float mdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double mrcp(double b) { return 1.0 / b; }
__device__ __forceinline__ float mrcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double mabs(double x) { return fabs(x); }
__device__ __forceinline__ float mabs(float x) { return fabsf(x); }
__device__ __forceinline__ double matan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float matan2(float y, float x) { return atan2f(y, x); }
template <typename M>
__device__ __forceinline__ M normp(V3T<M> a) {
  M r = M(0);
  r += a.x * a.x;
  r += a.y * a.y;
  r += a.z * a.z;
  return msqrt(r);
}

// Positions of one cell's vertices: three coordinate arrays, in LDS when the
// cell is staged there, otherwise in the global vertex planes.  With M = float
// they are offsets from a point of the cell (the mechanics only differences them).
template <typename M>
struct CellViewT {
  const M *px, *py, *pz;
  __device__ __forceinline__ V3T<M> pos(int k) const { return {px[k], py[k], pz[k]}; }
};
using CellView = CellViewT<double>;

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

template <typename M>
__device__ __forceinline__ V3T<M> unit_normal(V3T<M> v0, V3T<M> v1, V3T<M> v2) {
  V3T<M> n = crossp(v1 - v0, v2 - v0);
  const M nn = normp(n);
  if (nn != M(0)) {
    n.x = mdiv(n.x, nn);
    n.y = mdiv(n.y, nn);
    n.z = mdiv(n.z, nn);
  } else {
    n = {M(0), M(0), M(0)};
  }
  return n;
}

// bending force of vertex i from its ordered 1-ring (rbcHighOrderModel.cpp:127-166)
template <typename M>
__device__ V3T<M> bending_force(const CellViewT<M>& c, const TypeDev& T, int i) {
  const int nr = T.nring[i];
  // the ring (at most 6 entries, stored 6 wide) is loaded at once
  int R[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) R[j] = T.ring[6 * i + j];
  const M patch_eq = (M)T.patch_dist_eq[i];
  M s0 = M(0), s1 = M(0), s2 = M(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= nr) break;
    V3T<M> p = c.pos(R[j]);
    s0 += p.x;
    s1 += p.y;
    s2 += p.z;
  }
  const V3T<M> xi = c.pos(i);
  const M inv_n = mrcp((M)(unsigned)nr);
  const V3T<M> dev = {s0 * inv_n - xi.x, s1 * inv_n - xi.y, s2 * inv_n - xi.z};
  V3T<M> pn = {M(0), M(0), M(0)};
  V3T<M> prev = c.pos(R[0]);
  const V3T<M> first = prev;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= nr) break;
    const V3T<M> nxt = (j + 1 < nr) ? c.pos(R[j + 1 < 6 ? j + 1 : 0]) : first;
    V3T<M> tn = crossp(prev - xi, nxt - xi);
    const M il = mrcp(normp(tn));
    pn.x += tn.x * il;
    pn.y += tn.y * il;
    pn.z += tn.z * il;
    prev = nxt;
  }
  const M il = mrcp(normp(pn));
  pn.x *= il;
  pn.y *= il;
  pn.z *= il;
  const M ndev = dotp(pn, dev);
  const M dDev = mdiv(ndev - patch_eq, (M)T.edge_mean_eq);
  const M s = (M)T.k_bend * (dDev + mdiv(dDev, mabs(M(MaxCellBendingAngle) - dDev * dDev)));
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
template <int MODEL, int TERMS = TERM_ALL, typename M = double>
__device__ __forceinline__ V3T<M> vertex_force(const CellViewT<M>& c, const TypeDev& T, int k, M vforce,
                                               const M* bx, const M* by, const M* bz, const M* vx, const M* vy,
                                               const M* vz, bool visc) {
  M F0 = M(0), F1 = M(0), F2 = M(0);
  const V3T<M> xk = c.pos(k);
  // area and volume terms, triangle order: each incident triangle is evaluated
  // once; the two sums are kept apart and added in the reference's order
  // (all area terms, then all volume terms) up to the grouping of that sum
  {
    M A0 = M(0), A1 = M(0), A2 = M(0), B0 = M(0), B1 = M(0), B2 = M(0);
    const M inv_area_mean = mrcp((M)T.area_mean_eq);
    auto tri_term = [&](const int4 q, const M a_eq) {
      const V3T<M> v0 = c.pos(q.y), v1 = c.pos(q.z), v2 = c.pos(q.w);
      V3T<M> n = crossp(v1 - v0, v2 - v0);
      const M nn = normp(n);
      M area = M(0);
      if (nn != M(0)) {
        area = M(0.5) * nn;
        const M inn = mrcp(nn);
        n.x *= inn;
        n.y *= inn;
        n.z *= inn;
      } else {
        n = {M(0), M(0), M(0)};
      }
      const M ar = mdiv(area - a_eq, a_eq);
      const M afm = (M)T.k_area * (ar + mdiv(ar, mabs(M(MaxCellSurfaceAreaChange) - ar * ar)));
      constexpr M third = M(1.0 / 3.0);
      const M c0 = (v0.x + v1.x + v2.x) * third, c1 = (v0.y + v1.y + v2.y) * third,
              c2 = (v0.z + v1.z + v2.z) * third;
      A0 += (c0 - xk.x) * afm;
      A1 += (c1 - xk.y) * afm;
      A2 += (c2 - xk.z) * afm;
      const M sc = area * inv_area_mean;
      B0 += (vforce * n.x) * sc;
      B1 += (vforce * n.y) * sc;
      B2 += (vforce * n.z) * sc;
    };
    if (T.vtri4p) {
      int4 q[kPadW];
      M aeq[kPadW];
#pragma unroll
      for (int j = 0; j < kPadW; ++j) {
        q[j] = T.vtri4p[kPadW * k + j];
        aeq[j] = (M)T.vtri_aeq[kPadW * k + j];
      }
#pragma unroll
      for (int j = 0; j < kPadW; ++j)
        if (q[j].x >= 0) tri_term(q[j], aeq[j]);
    } else {
      for (int e = T.vtri_off[k]; e < T.vtri_off[k + 1]; ++e) tri_term(T.vtri4[e], (M)T.tri_area_eq[T.vtri4[e].x]);
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
    auto bend_term = [&](const int i, const M dn) {
      const M b0 = bx[i], b1 = by[i], b2 = bz[i];
      if (i == k) {
        F0 += b0;
        F1 += b1;
        F2 += b2;
      } else {
        const M in = mrcp(dn);
        F0 += -b0 * in;
        F1 += -b1 * in;
        F2 += -b2 * in;
      }
    };
    if (T.vbendp) {
      int b[kPadB];
      M dn[kPadB];
#pragma unroll
      for (int j = 0; j < kPadB; ++j) {
        b[j] = T.vbendp[kPadB * k + j];
        dn[j] = (M)T.vbend_n[kPadB * k + j];
      }
#pragma unroll
      for (int j = 0; j < kPadB; ++j)
        if (b[j] >= 0) bend_term(b[j], dn[j]);
    } else {
      for (int e = T.vbend_off[k]; e < T.vbend_off[k + 1]; ++e)
        bend_term(T.vbend[e], (M)(unsigned)T.nring[T.vbend[e]]);
    }
  }
  // edges: link, membrane viscosity, (PLT) dihedral bending; edge order
  auto edge_term = [&](const int4 q, const M l0) {
    const int ed = q.x >> 2, kind = q.x & 3;
    const int a = q.y, b = q.z;
    const V3T<M> pa = c.pos(a), pb = c.pos(b);
    const V3T<M> ev = pb - pa;
    const M len = normp(ev);
    const M il = mrcp(len);
    const V3T<M> uv = {ev.x * il, ev.y * il, ev.z * il};
    if (kind != 2) {
      const M sgn = kind == 0 ? M(1) : M(-1);
      if (TERMS & TERM_LINK) {
        const M ef = mdiv(len - l0, l0);
        const M fs = (M)T.k_link * (ef + mdiv(ef, mabs(M(MaxCellPersistenceLength) - ef * ef)));
        const V3T<M> fl = {uv.x * fs, uv.y * fs, uv.z * fs};
        if (sgn > 0) { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
        else { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
      }
      if ((TERMS & TERM_VISC) && visc) {
        const V3T<M> rv = {vx[b] - vx[a], vy[b] - vy[a], vz[b] - vz[a]};
        const M pr = dotp(rv, uv);
        const M eta = (M)T.eta_m;
        V3T<M> fv = {(pr * uv.x) * eta, (pr * uv.y) * eta, (pr * uv.z) * eta};
        const M m = normp(fv);
        if (m > M(FORCE_LIMIT / 4.0)) {
          const M s = mdiv(M(FORCE_LIMIT / 4.0), m);
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
      const V3T<M> n1 = unit_normal(c.pos(T0[0]), c.pos(T0[1]), c.pos(T0[2]));
      const V3T<M> n2 = unit_normal(c.pos(T1[0]), c.pos(T1[1]), c.pos(T1[2]));
      const M ang = matan2(dotp(crossp(n1, n2), uv), dotp(n1, n2));
      const M af = ang - (M)T.edge_angle_eq[ed];
      const M fm = (M)T.k_bend * (af + mdiv(af, mabs(M(MaxPLTBendingAngle) - af * af)));
      const V3T<M> bf = {((n1.x + n2.x) * fm) * M(0.5), ((n1.y + n2.y) * fm) * M(0.5),
                         ((n1.z + n2.z) * fm) * M(0.5)};
      if (kind != 2) { F0 += bf.x; F1 += bf.y; F2 += bf.z; }
      else { F0 -= bf.x; F1 -= bf.y; F2 -= bf.z; }
    }
  };
  if (MODEL != HC_MODEL_PLT && T.vedge4p) {
    int4 q[kPadW];
    M l0[kPadW];
#pragma unroll
    for (int j = 0; j < kPadW; ++j) {
      q[j] = T.vedge4p[kPadW * k + j];
      l0[j] = (M)T.vedge_l0[kPadW * k + j];
    }
#pragma unroll
    for (int j = 0; j < kPadW; ++j)
      if (q[j].x >= 0) edge_term(q[j], l0[j]);
  } else {
    for (int e = T.vedge_off[k]; e < T.vedge_off[k + 1]; ++e)
      edge_term(T.vedge4[e], (M)T.edge_len_eq[T.vedge4[e].x >> 2]);
  }
  // inner links, inner-edge order
  for (int e = T.vinner_off[k]; (TERMS & TERM_INNER) && e < T.vinner_off[k + 1]; ++e) {
    const int ie = T.vinner[e] >> 1, end = T.vinner[e] & 1;
    const V3T<M> pa = c.pos(T.inner[2 * ie]), pb = c.pos(T.inner[2 * ie + 1]);
    const V3T<M> ev = pb - pa;
    const M len = normp(ev);
    const V3T<M> uv = {mdiv(ev.x, len), mdiv(ev.y, len), mdiv(ev.z, len)};
    if (MODEL == HC_MODEL_PLT) {
      const M l0 = (M)T.inner_len_eq[ie];
      const M ef = mdiv(len - l0, l0);
      const M fs = (M)T.k_link * M(5.0) * ef;
      const V3T<M> fl = {uv.x * fs, uv.y * fs, uv.z * fs};
      if (end == 0) { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
      else { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
    } else if (MODEL == HC_MODEL_WBC) {
      const M rad = (M)T.radius, core = (M)T.core_radius;
      if (len < 2 * rad) {
        const M s = M(1) - mdiv(len, 2 * rad);
        const M kc = (M)T.k_cyto;
        const V3T<M> fl = {(uv.x * s) * kc, (uv.y * s) * kc, (uv.z * s) * kc};
        if (end == 0) { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
        else { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
      }
      if (len < 2 * core) {
        const M s = 1 - mdiv(len, 2 * core);
        const M kr = (M)T.k_inner_rigid;
        const V3T<M> fl = {(uv.x * s) * kr, (uv.y * s) * kr, (uv.z * s) * kr};
        if (end == 0) { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
        else { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
      }
    }
  }
  return {F0, F1, F2};
}

// The FP32 RBC kernels' bending and vertex forces over the compact tables
// (TypeHost::cslot / cbend / cring): the terms, their order and their operations
// are those of bending_force and vertex_force<HC_MODEL_RBC, TERM_ALL, float>; an
// edge is taken from the vertex's side (its force on the second end, -F(b - a),
// is F(a - b) in exact arithmetic; the multiply-adds the compiler fuses differ),
// so one uint4 per slot carries a triangle and an edge.
__device__ __forceinline__ int ring_entry(const uint4& r, int j) {
  const uint32_t w = j < 2 ? r.x : (j < 4 ? r.y : r.z);
  return (int)((w >> (16 * (j & 1))) & 0xFFFFu);
}

__device__ V3T<float> bending_force_c(const CellViewT<float>& c, const TypeDev& T, int i) {
  using M = float;
  const uint4 rq = T.cring[i];
  int R[6];
  int nr = 0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    R[j] = ring_entry(rq, j);
    nr += R[j] != 0xFFFF;
  }
  const M patch_eq = __uint_as_float(rq.w);
  M s0 = M(0), s1 = M(0), s2 = M(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= nr) break;
    V3T<M> p = c.pos(R[j]);
    s0 += p.x;
    s1 += p.y;
    s2 += p.z;
  }
  const V3T<M> xi = c.pos(i);
  const M inv_n = mrcp((M)(unsigned)nr);
  const V3T<M> dev = {s0 * inv_n - xi.x, s1 * inv_n - xi.y, s2 * inv_n - xi.z};
  V3T<M> pn = {M(0), M(0), M(0)};
  V3T<M> prev = c.pos(R[0]);
  const V3T<M> first = prev;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= nr) break;
    const V3T<M> nxt = (j + 1 < nr) ? c.pos(R[j + 1 < 6 ? j + 1 : 0]) : first;
    V3T<M> tn = crossp(prev - xi, nxt - xi);
    const M il = mrcp(normp(tn));
    pn.x += tn.x * il;
    pn.y += tn.y * il;
    pn.z += tn.z * il;
    prev = nxt;
  }
  const M il = mrcp(normp(pn));
  pn.x *= il;
  pn.y *= il;
  pn.z *= il;
  const M ndev = dotp(pn, dev);
  const M dDev = mdiv(ndev - patch_eq, (M)T.edge_mean_eq);
  const M s = (M)T.k_bend * (dDev + mdiv(dDev, mabs(M(MaxCellBendingAngle) - dDev * dDev)));
  return {pn.x * s, pn.y * s, pn.z * s};
}

__device__ __forceinline__ V3T<float> vertex_force_rbc_c(const CellViewT<float>& c, const TypeDev& T, int k,
                                                          float vforce, const float* bx, const float* by,
                                                          const float* bz, const float* vx, const float* vy,
                                                          const float* vz, bool visc) {
  using M = float;
  uint4 q[kPadW];
  uint32_t bq[kPadB];
#pragma unroll
  for (int j = 0; j < kPadW; ++j) q[j] = T.cslot[kPadW * k + j];
#pragma unroll
  for (int j = 0; j < kPadB; ++j) bq[j] = T.cbend[kPadB * k + j];
  M F0 = M(0), F1 = M(0), F2 = M(0);
  const V3T<M> xk = c.pos(k);
  {
    M A0 = M(0), A1 = M(0), A2 = M(0), B0 = M(0), B1 = M(0), B2 = M(0);
    const M inv_area_mean = mrcp((M)T.area_mean_eq);
#pragma unroll
    for (int j = 0; j < kPadW; ++j) {
      if ((q[j].x & 0xFFFFu) == 0xFFFFu) continue;
      const V3T<M> v0 = c.pos((int)(q[j].x & 0xFFFFu)), v1 = c.pos((int)(q[j].x >> 16)),
                   v2 = c.pos((int)(q[j].y & 0xFFFFu));
      const M a_eq = __uint_as_float(q[j].z);
      V3T<M> n = crossp(v1 - v0, v2 - v0);
      const M nn = normp(n);
      M area = M(0);
      if (nn != M(0)) {
        area = M(0.5) * nn;
        const M inn = mrcp(nn);
        n.x *= inn;
        n.y *= inn;
        n.z *= inn;
      } else {
        n = {M(0), M(0), M(0)};
      }
      const M ar = mdiv(area - a_eq, a_eq);
      const M afm = (M)T.k_area * (ar + mdiv(ar, mabs(M(MaxCellSurfaceAreaChange) - ar * ar)));
      constexpr M third = M(1.0 / 3.0);
      const M c0 = (v0.x + v1.x + v2.x) * third, c1 = (v0.y + v1.y + v2.y) * third,
              c2 = (v0.z + v1.z + v2.z) * third;
      A0 += (c0 - xk.x) * afm;
      A1 += (c1 - xk.y) * afm;
      A2 += (c2 - xk.z) * afm;
      const M sc = area * inv_area_mean;
      B0 += (vforce * n.x) * sc;
      B1 += (vforce * n.y) * sc;
      B2 += (vforce * n.z) * sc;
    }
    F0 = A0 + B0;
    F1 = A1 + B1;
    F2 = A2 + B2;
  }
#pragma unroll
  for (int j = 0; j < kPadB; ++j) {
    if (bq[j] == 0xFFFFFFFFu) continue;
    const int i = (int)(bq[j] & 0xFFFFu);
    const M b0 = bx[i], b1 = by[i], b2 = bz[i];
    if (i == k) {
      F0 += b0;
      F1 += b1;
      F2 += b2;
    } else {
      const M in = mrcp((M)(bq[j] >> 16));
      F0 += -b0 * in;
      F1 += -b1 * in;
      F2 += -b2 * in;
    }
  }
#pragma unroll
  for (int j = 0; j < kPadW; ++j) {
    const int o = (int)(q[j].y >> 16);
    if (o == 0xFFFF) continue;
    const M l0 = __uint_as_float(q[j].w);
    const V3T<M> ev = c.pos(o) - xk;
    const M len = normp(ev);
    const M il = mrcp(len);
    const V3T<M> uv = {ev.x * il, ev.y * il, ev.z * il};
    const M ef = mdiv(len - l0, l0);
    const M fs = (M)T.k_link * (ef + mdiv(ef, mabs(M(MaxCellPersistenceLength) - ef * ef)));
    F0 += uv.x * fs;
    F1 += uv.y * fs;
    F2 += uv.z * fs;
    if (visc) {
      const V3T<M> rv = {vx[o] - vx[k], vy[o] - vy[k], vz[o] - vz[k]};
      const M pr = dotp(rv, uv);
      const M eta = (M)T.eta_m;
      V3T<M> fv = {(pr * uv.x) * eta, (pr * uv.y) * eta, (pr * uv.z) * eta};
      const M m = normp(fv);
      if (m > M(FORCE_LIMIT / 4.0)) {
        const M s = mdiv(M(FORCE_LIMIT / 4.0), m);
        fv.x *= s;
        fv.y *= s;
        fv.z *= s;
      }
      F0 += fv.x;
      F1 += fv.y;
      F2 += fv.z;
    }
  }
  return {F0, F1, F2};
}

// The FP32 RBC kernels' vertex force over the ring-ordered tables (TypeHost::rpar
// / rnb): the terms of vertex_force<HC_MODEL_RBC>, summed around the 1-ring.
// Triangle j is (k, ring[j], ring[j+1]) and edge j is (k, ring[j]), so each
// neighbour's position is read once and enters as its offset d_j from vertex k:
// the triangle's centroid offset is (d_j + d_j+1) / 3 and its unit normal times
// its area is half the cross product.  The sums run in ring order (own bending
// force first) rather than in the reference's triangle, contributor and edge
// order: the same force up to FP32 rounding.
__device__ __forceinline__ V3T<float> vertex_force_rbc_r(const CellViewT<float>& c, const TypeDev& T, int k,
                                                          float vforce, const float* bx, const float* by,
                                                          const float* bz, const float* vx, const float* vy,
                                                          const float* vz, bool visc) {
  using M = float;
  const uint4 rq = T.cring[k];
  const uint32_t nbw = T.rnb[k];
  const float4 q0 = T.rpar[3 * k], q1 = T.rpar[3 * k + 1], q2 = T.rpar[3 * k + 2];
  const M aeq[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
  const M l0[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
  const int nr = (int)(nbw & 7u);  // 3..6
  const V3T<M> xk = c.pos(k);
  int R[6];
  V3T<M> d[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    R[j] = ring_entry(rq, j);
    if (j < 3 || j < nr) d[j] = c.pos(R[j]) - xk;
    else d[j] = {M(0), M(0), M(0)};
  }
  M F0, F1, F2;
  {
    M A0 = M(0), A1 = M(0), A2 = M(0), B0 = M(0), B1 = M(0), B2 = M(0);
    const M half_vf = M(0.5) * vforce * mrcp((M)T.area_mean_eq);
    constexpr M third = M(1.0 / 3.0);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      if (j >= 3 && j >= nr) break;
      const V3T<M> a = d[j];
      V3T<M> b = d[j < 5 ? j + 1 : 0];
      if (j + 1 >= nr) b = d[0];
      const V3T<M> cr = crossp(a, b);
      const M area = M(0.5) * normp(cr);
      const M ar = mdiv(area - aeq[j], aeq[j]);
      const M afm = (M)T.k_area * (ar + mdiv(ar, mabs(M(MaxCellSurfaceAreaChange) - ar * ar)));
      A0 += ((a.x + b.x) * third) * afm;
      A1 += ((a.y + b.y) * third) * afm;
      A2 += ((a.z + b.z) * third) * afm;
      B0 += cr.x * half_vf;
      B1 += cr.y * half_vf;
      B2 += cr.z * half_vf;
    }
    F0 = A0 + B0;
    F1 = A1 + B1;
    F2 = A2 + B2;
  }
  F0 += bx[k];
  F1 += by[k];
  F2 += bz[k];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= 3 && j >= nr) break;
    const M in = mrcp((M)((nbw >> (4 + 3 * j)) & 7u));
    F0 += -bx[R[j]] * in;
    F1 += -by[R[j]] * in;
    F2 += -bz[R[j]] * in;
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= 3 && j >= nr) break;
    const V3T<M> ev = d[j];
    const M len = normp(ev);
    const M il = mrcp(len);
    const V3T<M> uv = {ev.x * il, ev.y * il, ev.z * il};
    const M ef = mdiv(len - l0[j], l0[j]);
    const M fs = (M)T.k_link * (ef + mdiv(ef, mabs(M(MaxCellPersistenceLength) - ef * ef)));
    F0 += uv.x * fs;
    F1 += uv.y * fs;
    F2 += uv.z * fs;
    if (visc) {
      const int o = R[j];
      const V3T<M> rv = {vx[o] - vx[k], vy[o] - vy[k], vz[o] - vz[k]};
      const M pr = dotp(rv, uv);
      const M eta = (M)T.eta_m;
      V3T<M> fv = {(pr * uv.x) * eta, (pr * uv.y) * eta, (pr * uv.z) * eta};
      const M m = normp(fv);
      if (m > M(FORCE_LIMIT / 4.0)) {
        const M s = mdiv(M(FORCE_LIMIT / 4.0), m);
        fv.x *= s;
        fv.y *= s;
        fv.z *= s;
      }
      F0 += fv.x;
      F1 += fv.y;
      F2 += fv.z;
    }
  }
  return {F0, F1, F2};
}

// Signed triple product of triangle t's corners relative to r (0 for t >= nt);
// the cell volume is 1/6 of the sum over the triangles.
template <typename M>
__device__ __forceinline__ double volume_term(const CellViewT<M>& c, const TypeDev& T, int t, const D3& r) {
  if (t >= T.nt) return 0.0;
  const V3T<M> a = c.pos(T.tri[3 * t]), b = c.pos(T.tri[3 * t + 1]), e = c.pos(T.tri[3 * t + 2]);
  const D3 v0 = D3{(double)a.x, (double)a.y, (double)a.z} - r, v1 = D3{(double)b.x, (double)b.y, (double)b.z} - r,
           v2 = D3{(double)e.x, (double)e.y, (double)e.z} - r;
  const double v210 = v2.x * v1.y * v0.z;
  const double v120 = v1.x * v2.y * v0.z;
  const double v201 = v2.x * v0.y * v1.z;
  const double v021 = v0.x * v2.y * v1.z;
  const double v102 = v1.x * v0.y * v2.z;
  const double v012 = v0.x * v1.y * v2.z;
  return -v210 + v120 + v201 - v021 - v102 + v012;
}

// Sum of the signed triple products of the cell's triangles handled by thread tid
// (the cell volume is 1/6 of the sum over all threads).
template <int BS, typename M>
__device__ __forceinline__ double volume_partial(const CellViewT<M>& c, const TypeDev& T, int tid) {
  double part = 0.0;
  // corners relative to the cell's vertex 0: the same volume for a closed mesh,
  // without the cancellation of triple products of coordinates hundreds of
  // spacings from the origin (FP64 whatever the position type)
  const V3T<M> rm = c.pos(0);
  const D3 r = {(double)rm.x, (double)rm.y, (double)rm.z};
  for (int t = tid; t < T.nt; t += BS) {
    const V3T<M> a = c.pos(T.tri[3 * t]), b = c.pos(T.tri[3 * t + 1]), e = c.pos(T.tri[3 * t + 2]);
    const D3 v0 = D3{(double)a.x, (double)a.y, (double)a.z} - r, v1 = D3{(double)b.x, (double)b.y, (double)b.z} - r,
             v2 = D3{(double)e.x, (double)e.y, (double)e.z} - r;
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
