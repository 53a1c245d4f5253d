// hc_mech.hip -- membrane mechanics of RBC, PLT and WBC cells on MI355X.
//
// Replaces (reference paths relative to the HemoCell tree):
//   RbcHighOrderModel::ParticleMechanics  mechanics/rbcHighOrderModel.cpp:38-207
//   PltSimpleModel::ParticleMechanics     mechanics/pltSimpleModel.cpp:44-208
//   WbcHighOrderModel::ParticleMechanics  mechanics/wbcHighOrderModel.cpp:42-225
//   the force reset of applyConstitutiveModel  core/hemoCellParticleField.cpp:633-675
//
// One workgroup per cell.  The reference scatters every term into sv.force
// (area and volume per triangle, bending per vertex and its 1-ring, link and
// viscosity per edge, inner links); here each vertex *gathers* its terms from
// precomputed incidence lists that visit them in exactly the reference's order,
// so every per-vertex sum is formed in the same order without atomics.  Only the
// cell volume is a workgroup tree reduction (FP64).  Per-triangle and per-edge
// terms are recomputed by each incident vertex (cheap, no storage); the
// per-vertex bending force, needed by the whole 1-ring, is computed once into a
// scratch array and read back after a barrier.
#include <algorithm>
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
  const int32_t* R = T.ring + 6 * i;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int j = 0; j < nr; ++j) {
    D3 p = c.pos(R[j]);
    s0 += p.x;
    s1 += p.y;
    s2 += p.z;
  }
  const D3 xi = c.pos(i);
  const double dn = (double)(unsigned)nr;
  const D3 dev = {s0 / dn - xi.x, s1 / dn - xi.y, s2 / dn - xi.z};
  D3 pn = {0.0, 0.0, 0.0};
  D3 prev = c.pos(R[0]);
  const D3 first = prev;
  for (int j = 0; j < nr; ++j) {
    const D3 nxt = (j + 1 < nr) ? c.pos(R[j + 1]) : first;
    D3 tn = crossp(prev - xi, nxt - xi);
    const double l = normp(tn);
    pn.x += tn.x / l;
    pn.y += tn.y / l;
    pn.z += tn.z / l;
    prev = nxt;
  }
  const double l = normp(pn);
  pn.x /= l;
  pn.y /= l;
  pn.z /= l;
  const double ndev = dotp(pn, dev);
  const double dDev = (ndev - T.patch_dist_eq[i]) / T.edge_mean_eq;
  const double s = T.k_bend * (dDev + dDev / fabs(MaxCellBendingAngle - dDev * dDev));
  return {pn.x * s, pn.y * s, pn.z * s};
}

// STAGED: the cell's positions, bending forces and (with membrane viscosity)
// velocities live in LDS for the whole kernel (9 x 8 bytes per vertex at most).
template <int MODEL, int BS, bool STAGED>
__global__ __launch_bounds__(BS) void k_mechanics(VertexDev V, TypeDev T, const int32_t* __restrict__ cells,
                                                  const int64_t* __restrict__ cell_base, double* __restrict__ bend) {
  extern __shared__ double sm[];
  __shared__ double red[BS];
  __shared__ double vforce_sh;
  const int cell = cells[blockIdx.x];
  const int64_t base = cell_base[cell];
  const int tid = threadIdx.x;
  const int64_t vs = V.vs;
  const int nv = T.nv;
  const bool visc = (MODEL != HC_MODEL_RBC) || (T.eta_m != 0.0);
  const double *vx, *vy, *vz;
  double *bx, *by, *bz;
  CellView c;
  if (STAGED) {
    for (int k = tid; k < nv; k += BS) {
      sm[k] = V.x[base + k];
      sm[nv + k] = V.x[vs + base + k];
      sm[2 * nv + k] = V.x[2 * vs + base + k];
      if (visc) {
        sm[6 * nv + k] = V.v[base + k];
        sm[7 * nv + k] = V.v[vs + base + k];
        sm[8 * nv + k] = V.v[2 * vs + base + k];
      }
    }
    c = CellView{sm, sm + nv, sm + 2 * nv};
    bx = sm + 3 * nv;
    by = sm + 4 * nv;
    bz = sm + 5 * nv;
    vx = sm + 6 * nv;
    vy = sm + 7 * nv;
    vz = sm + 8 * nv;
    __syncthreads();
  } else {
    c = CellView{V.x + base, V.x + vs + base, V.x + 2 * vs + base};
    bx = bend + base;
    by = bend + vs + base;
    bz = bend + 2 * vs + base;
    vx = V.v + base;
    vy = V.v + vs + base;
    vz = V.v + 2 * vs + base;
  }

  // ---- cell volume (sum of signed triple products, x 1/6) --------------------
  double part = 0.0;
  for (int t = tid; t < T.nt; t += BS) {
    const D3 v0 = c.pos(T.tri[3 * t]), v1 = c.pos(T.tri[3 * t + 1]), v2 = c.pos(T.tri[3 * t + 2]);
    const double v210 = v2.x * v1.y * v0.z;
    const double v120 = v1.x * v2.y * v0.z;
    const double v201 = v2.x * v0.y * v1.z;
    const double v021 = v0.x * v2.y * v1.z;
    const double v102 = v1.x * v0.y * v2.z;
    const double v012 = v0.x * v1.y * v2.z;
    part += (-v210 + v120 + v201 - v021 - v102 + v012);
  }
  red[tid] = part;
  __syncthreads();
  for (int h = BS / 2; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) {
    const double volume = red[0] * (1.0 / 6.0);
    const double vf = (volume - T.volume_eq) / T.volume_eq;
    vforce_sh = -T.k_volume * vf / fabs(MaxCellVolumetricChange - vf * vf);
  }
  // ---- per-vertex bending force into scratch (RBC, WBC) ----------------------
  if (MODEL != HC_MODEL_PLT) {
    for (int i = tid; i < T.nv; i += BS) {
      const D3 b = bending_force(c, T, i);
      bx[i] = b.x;
      by[i] = b.y;
      bz[i] = b.z;
    }
  }
  __syncthreads();
  const double vforce = vforce_sh;

  // ---- gather every term of vertex k in the reference's order ----------------
  for (int k = tid; k < T.nv; k += BS) {
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
    const D3 xk = c.pos(k);
    // area terms, triangle order
    for (int e = T.vtri_off[k]; e < T.vtri_off[k + 1]; ++e) {
      const int t = T.vtri[e] >> 2;
      D3 v0, v1, v2, n;
      double area;
      tri_geom(c, T, t, v0, v1, v2, area, n);
      const double ar = (area - T.tri_area_eq[t]) / T.tri_area_eq[t];
      const double afm = T.k_area * (ar + ar / fabs(MaxCellSurfaceAreaChange - ar * ar));
      const double c0 = (v0.x + v1.x + v2.x) / 3.0, c1 = (v0.y + v1.y + v2.y) / 3.0, c2 = (v0.z + v1.z + v2.z) / 3.0;
      F0 += (c0 - xk.x) * afm;
      F1 += (c1 - xk.y) * afm;
      F2 += (c2 - xk.z) * afm;
    }
    // volume terms, triangle order
    for (int e = T.vtri_off[k]; e < T.vtri_off[k + 1]; ++e) {
      const int t = T.vtri[e] >> 2;
      D3 v0, v1, v2, n;
      double area;
      tri_geom(c, T, t, v0, v1, v2, area, n);
      const double s = area / T.area_mean_eq;
      F0 += (vforce * n.x) * s;
      F1 += (vforce * n.y) * s;
      F2 += (vforce * n.z) * s;
    }
    // bending: own force and -F_i/n_i of every vertex i whose 1-ring holds k, ascending i
    if (MODEL != HC_MODEL_PLT) {
      for (int e = T.vbend_off[k]; e < T.vbend_off[k + 1]; ++e) {
        const int i = T.vbend[e];
        const double b0 = bx[i], b1 = by[i], b2 = bz[i];
        if (i == k) {
          F0 += b0;
          F1 += b1;
          F2 += b2;
        } else {
          const double dn = (double)(unsigned)T.nring[i];
          F0 += -b0 / dn;
          F1 += -b1 / dn;
          F2 += -b2 / dn;
        }
      }
    }
    // edges: link, membrane viscosity, (PLT) dihedral bending; edge order
    for (int e = T.vedge_off[k]; e < T.vedge_off[k + 1]; ++e) {
      const int ed = T.vedge[e] >> 2, kind = T.vedge[e] & 3;
      const int a = T.edge[2 * ed], b = T.edge[2 * ed + 1];
      const D3 pa = c.pos(a), pb = c.pos(b);
      const D3 ev = pb - pa;
      const double len = normp(ev);
      const D3 uv = {ev.x / len, ev.y / len, ev.z / len};
      if (kind != 2) {
        const double sgn = kind == 0 ? 1.0 : -1.0;
        const double ef = (len - T.edge_len_eq[ed]) / T.edge_len_eq[ed];
        const double fs = T.k_link * (ef + ef / fabs(MaxCellPersistenceLength - ef * ef));
        const D3 fl = {uv.x * fs, uv.y * fs, uv.z * fs};
        if (sgn > 0) { F0 += fl.x; F1 += fl.y; F2 += fl.z; }
        else { F0 -= fl.x; F1 -= fl.y; F2 -= fl.z; }
        if (visc) {
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
      if (MODEL == HC_MODEL_PLT) {
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
    }
    // inner links, inner-edge order
    for (int e = T.vinner_off[k]; e < T.vinner_off[k + 1]; ++e) {
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
    const int64_t q = base + k;
    V.f[q] = F0;
    V.f[vs + q] = F1;
    V.f[2 * vs + q] = F2;
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_volume_area(VertexDev V, TypeDev T, const int32_t* __restrict__ cells,
                                                    const int64_t* __restrict__ cell_base, double* vol, double* area) {
  __shared__ double rv[BS], ra[BS];
  const int cell = cells[blockIdx.x];
  const int64_t base = cell_base[cell];
  const CellView c{V.x + base, V.x + V.vs + base, V.x + 2 * V.vs + base};
  const int tid = threadIdx.x;
  double pv = 0.0, pa = 0.0;
  for (int t = tid; t < T.nt; t += BS) {
    D3 v0, v1, v2, n;
    double a;
    tri_geom(c, T, t, v0, v1, v2, a, n);
    pv += dotp(v0, crossp(v1, v2));
    pa += a;
  }
  rv[tid] = pv;
  ra[tid] = pa;
  __syncthreads();
  for (int h = BS / 2; h > 0; h >>= 1) {
    if (tid < h) {
      rv[tid] += rv[tid + h];
      ra[tid] += ra[tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    vol[cell] = rv[0] / 6.0;
    area[cell] = ra[0];
  }
}

__global__ void k_check_finite(const double* p, int64_t n, unsigned int* bad) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * 256;
  bool b = false;
  for (; i < n; i += step) b |= !isfinite(p[i]);
  if (b) atomicOr(bad, 1u);
}
__global__ void k_check_finite_f(const float* p, int64_t n, unsigned int* bad) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * 256;
  bool b = false;
  for (; i < n; i += step) b |= !isfinite(p[i]);
  if (b) atomicOr(bad, 1u);
}

}  // namespace

hipError_t launch_mechanics(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                            const int64_t* cell_base, double* bend, int lds_limit) {
  if (ncells <= 0) return hipSuccess;
  dim3 g((unsigned)ncells);
  const bool visc = (T.model != HC_MODEL_RBC) || (T.eta_m != 0.0);
  const size_t lds = sizeof(double) * (size_t)T.nv * (visc ? 9 : 6);
  const bool staged = lds <= (size_t)std::min(lds_limit, 65536);
  const size_t dyn = staged ? lds : 0;
#define HC_MECH(M, B)                                                                                           \
  if (staged) hipLaunchKernelGGL((k_mechanics<M, B, true>), g, dim3(B), dyn, s, V, T, cells, cell_base, bend);   \
  else hipLaunchKernelGGL((k_mechanics<M, B, false>), g, dim3(B), 0, s, V, T, cells, cell_base, bend);
  switch (T.model) {
    case HC_MODEL_RBC:
      HC_MECH(HC_MODEL_RBC, 256)
      break;
    case HC_MODEL_PLT:
      HC_MECH(HC_MODEL_PLT, 64)
      break;
    case HC_MODEL_WBC:
      HC_MECH(HC_MODEL_WBC, 256)
      break;
    default:
      return hipErrorInvalidValue;
  }
#undef HC_MECH
  return hipGetLastError();
}

hipError_t launch_cell_volume_area(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells,
                                   int ncells, const int64_t* cell_base, double* vol, double* area) {
  if (ncells <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_volume_area<256>), dim3((unsigned)ncells), dim3(256), 0, s, V, T, cells, cell_base, vol, area);
  return hipGetLastError();
}

hipError_t launch_check_finite(hipStream_t s, const double* p, int64_t n, unsigned int* bad) {
  if (n <= 0) return hipSuccess;
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_check_finite, dim3((unsigned)b), dim3(256), 0, s, p, n, bad);
  return hipGetLastError();
}
hipError_t launch_check_finite_f(hipStream_t s, const float* p, int64_t n, unsigned int* bad) {
  if (n <= 0) return hipSuccess;
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_check_finite_f, dim3((unsigned)b), dim3(256), 0, s, p, n, bad);
  return hipGetLastError();
}

}  // namespace hcg
