// hc_visc.hip -- interior viscosity: which fluid nodes lie inside a membrane.
//
// Replaces (reference paths relative to the HemoCell tree, INTERIOR_VISCOSITY
// build):
//   HemoCellParticleField::findInternalParticleGridPoints  core/hemoCellParticleField.cpp:776-808
//   OctreeStructCell::findInnerNodes / findCrossings       helper/octree.h:107-150, helper/octree.cpp:163-182
//   MollerTrumbore                                         helper/mollerTrumbore.h:26-78
//   HemoCellParticleField::internalGridPointsMembrane      core/hemoCellParticleField.cpp:745-774
//
// The reference keeps a Dynamics object per node and swaps in a BGK clone with
// omega = 1/tau_inner for inner nodes.  Here a byte per node holds the class
// (0 background, 1 + type inside a membrane of that type) and the collide
// kernel looks the relaxation rate up in a 16-entry table.
//
// Entire-grid pass: one workgroup per cell.  The cell's positions are staged
// in LDS; the bounding box (corners truncated to integers, clamped on
// non-periodic axes) is split over the threads, and each node counts the
// crossings of the ray from it along (-40,-40,-40) with every triangle of the
// cell (Moller-Trumbore, same operations in the same order, this file being
// built without floating-point contraction).  The reference culls triangles
// with an octree first; here two culls drop only triangles the ray cannot
// reach, so the count is the same: a triangle whose projection along the ray,
// (x - z, y - z), keeps clear of the node's (with a margin far above the test's
// rounding), and one whose corners all lie above the node on some axis (the
// ray decreases in every coordinate).  The nodes are visited line by line
// along the ray direction: all nodes of a line share (x - z, y - z), so one
// thread lists the line's candidate triangles once (a uniform loop of LDS
// broadcasts) and tests only those for each node of the line.
//
// Membrane pass: the reference visits particles in order and the last write to
// a node wins.  Two kernels reproduce that without ordering the threads: the
// first takes atomicMax of ((vertex + 1) << 1 | inside) into a per-node stamp,
// the second lets the vertex that owns the maximum write the node's class and
// clear the stamp, so the stamps are zero again afterwards.
#include "hc_launch.h"
#include "hc_ibm_dev.h"
#include "hc_mech_dev.h"

namespace hcg {
namespace {

constexpr int VB = 256;

__device__ __forceinline__ int64_t wrap_node(int64_t c, int n) {
  int64_t m = c % n;
  return m < 0 ? m + n : m;
}

// helper/mollerTrumbore.h:26-78, ray (-40,-40,-40) from the integer site s
__device__ __forceinline__ int ray_crosses(double v0x, double v0y, double v0z, double v1x, double v1y, double v1z,
                                           double v2x, double v2y, double v2z, double sx, double sy, double sz) {
  const double EPS = 0.0000001;
  const double r = -40.0;
  const double e1x = v1x - v0x, e1y = v1y - v0y, e1z = v1z - v0z;
  const double e2x = v2x - v0x, e2y = v2y - v0y, e2z = v2z - v0z;
  const double px = r * e2z - r * e2y, py = r * e2x - r * e2z, pz = r * e2y - r * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (det > -EPS && det < EPS) return 0;
  const double inv = 1 / det;
  const double tx = sx - v0x, ty = sy - v0y, tz = sz - v0z;
  const double u = inv * (tx * px + ty * py + tz * pz);
  if (u < 0.0 || u > 1.0) return 0;
  const double qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
  const double v = inv * (r * qx + r * qy + r * qz);
  if (v < 0.0 || u + v > 1.0) return 0;
  const float t = (float)(inv * (e2x * qx + e2y * qy + e2z * qz));
  return t > EPS ? 1 : 0;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// node (X, Y, Z) of the bounding box: inner when the ray crosses the cell an
// odd number of times.  With pb, triangles whose projection along the ray's
// direction, (x - z, y - z), keeps clear of the node's are skipped first.
__device__ __forceinline__ void classify_node(int64_t X, int64_t Y, int64_t Z, const TypeDev& T, const double* px,
                                              const double* py, const double* pz, const float4* pb,
                                              const LatticeDev& L, const uint8_t* __restrict__ type, uint8_t cls,
                                              uint8_t* __restrict__ visc) {
  const int64_t n = wrap_node(Z, L.nz) + (int64_t)L.nz * (wrap_node(Y, L.ny) + (int64_t)L.ny * wrap_node(X, L.nx));
  const bool fluid = type == nullptr || type[n] == HC_NODE_FLUID;
  const double sx = (double)X, sy = (double)Y, sz = (double)Z;
  const float us = (float)(X - Z), vs = (float)(Y - Z);
  int cnt = 0;
  for (int t = 0; t < T.nt; ++t) {
    if (pb) {
      const float4 b = pb[t];
      if (us < b.x || us > b.y || vs < b.z || vs > b.w) continue;
    }
    const int i0 = T.tri[3 * t], i1 = T.tri[3 * t + 1], i2 = T.tri[3 * t + 2];
    const double ax = px[i0], ay = py[i0], az = pz[i0];
    const double bx = px[i1], by = py[i1], bz = pz[i1];
    const double cx = px[i2], cy = py[i2], cz = pz[i2];
    if (!fluid || fmin(fmin(ax, bx), cx) > sx || fmin(fmin(ay, by), cy) > sy || fmin(fmin(az, bz), cz) > sz)
      continue;
    cnt += ray_crosses(ax, ay, az, bx, by, bz, cx, cy, cz, sx, sy, sz);
  }
  if (fluid && (cnt & 1)) visc[n] = cls;
}

// float bounds that enclose a double value with room for the rounding of
// Moller-Trumbore near a triangle's edge
__device__ __forceinline__ float lower_f(double a) {
  const float f = (float)a;
  return f - (fabsf(f) * 1e-6f + 1e-2f);
}
__device__ __forceinline__ float upper_f(double a) {
  const float f = (float)a;
  return f + (fabsf(f) * 1e-6f + 1e-2f);
}

constexpr int CAND_MAX = 32;  // triangles listed per line before it falls back

// LINES: the box's nodes are visited line by line along the ray direction
// (fixed x - z and y - z), the triangles that can cross a line being listed
// once per line; the dynamic LDS then holds the positions, one float4
// projected box per triangle and each thread's candidate list.  Without LINES
// (a mesh too large for that LDS) the nodes are visited in x, y, z order.
template <bool LINES>
__global__ __launch_bounds__(VB) void k_visc_grid(VertexDev V, TypeDev T, const int32_t* __restrict__ cells,
                                                  const int64_t* __restrict__ cell_base, LatticeDev L,
                                                  const uint8_t* __restrict__ type, uint8_t cls,
                                                  uint8_t* __restrict__ visc) {
  extern __shared__ __attribute__((aligned(16))) double sp[];  // 3 nv positions [, nt float4, VB x CAND_MAX]
  __shared__ double red[6][VB / 64];
  __shared__ int64_t box[6];
  const int nv = T.nv;
  double* px = sp;
  double* py = sp + nv;
  double* pz = sp + 2 * nv;
  float4* pb = reinterpret_cast<float4*>(sp + ((3 * nv + 1) & ~1));
  const int64_t b = cell_base[cells[blockIdx.x]];
  // a cell removed by k_delete_incomplete has no inner nodes
  if (V.flag && (V.flag[b] & VFLAG_REMOVED)) return;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = threadIdx.x; k < nv; k += VB) {
    const double p[3] = {V.x[b + k], V.x[V.vs + b + k], V.x[2 * V.vs + b + k]};
    px[k] = p[0];
    py[k] = p[1];
    pz[k] = p[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], p[a]);
      hi[a] = fmax(hi[a], p[a]);
    }
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = wave_min(lo[a]);
    hi[a] = wave_max(hi[a]);
    if (lane == 0) {
      red[a][wave] = lo[a];
      red[3 + a][wave] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int dims[3] = {L.nx, L.ny, L.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double l = red[a][0], h = red[3 + a][0];
      for (int w = 1; w < VB / 64; ++w) {
        l = fmin(l, red[a][w]);
        h = fmax(h, red[3 + a][w]);
      }
      // C++ casts truncate toward zero (octree.h:111-121); a cell with
      // non-finite or absurdly spread positions is skipped (empty box)
      int64_t a0 = 0, a1 = -1;
      if (isfinite(l) && isfinite(h) && h - l < 4096.0 && fabs(l) < 1.0e6) {
        a0 = (int64_t)l;
        a1 = (int64_t)h;
      }
      if (!L.periodic[a] && a0 <= a1) {
        a0 = a0 < 0 ? 0 : a0;
        a1 = a1 > dims[a] - 1 ? dims[a] - 1 : a1;
      }
      box[a] = a0;
      box[3 + a] = a1;
    }
  }
  __syncthreads();
  const int64_t ex = box[3] - box[0] + 1, ey = box[4] - box[1] + 1, ez = box[5] - box[2] + 1;
  if (ex <= 0 || ey <= 0 || ez <= 0) return;
  if (LINES) {
    for (int t = threadIdx.x; t < T.nt; t += VB) {
      const int i0 = T.tri[3 * t], i1 = T.tri[3 * t + 1], i2 = T.tri[3 * t + 2];
      const double u0 = px[i0] - pz[i0], u1 = px[i1] - pz[i1], u2 = px[i2] - pz[i2];
      const double v0 = py[i0] - pz[i0], v1 = py[i1] - pz[i1], v2 = py[i2] - pz[i2];
      pb[t] = make_float4(lower_f(fmin(fmin(u0, u1), u2)), upper_f(fmax(fmax(u0, u1), u2)),
                          lower_f(fmin(fmin(v0, v1), v2)), upper_f(fmax(fmax(v0, v1), v2)));
    }
    __syncthreads();
    // one thread per line along the ray direction (fixed u = x - z, v = y - z):
    // the triangles whose projected box holds (u, v) are listed once, then
    // every node of the line, Z in [max(z0, x0 - u, y0 - v), min(z1, x1 - u,
    // y1 - v)], counts its crossings over that list alone
    const int64_t ulo = box[0] - box[5], vlo = box[1] - box[5];
    const int64_t nu = box[3] - box[2] - ulo + 1, nw = box[4] - box[2] - vlo + 1;
    uint16_t* cand = reinterpret_cast<uint16_t*>(pb + T.nt) + threadIdx.x * CAND_MAX;
    for (int64_t l = threadIdx.x; l < nu * nw; l += VB) {
      const int64_t u = ulo + l / nw, v = vlo + l % nw;
      const int64_t zs = max(box[2], max(box[0] - u, box[1] - v));
      const int64_t ze = min(box[5], min(box[3] - u, box[4] - v));
      if (ze < zs) continue;
      const float us = (float)u, vs = (float)v;
      int cnt = 0;
      for (int t = 0; t < T.nt; ++t) {
        const float4 b = pb[t];
        if (us < b.x || us > b.y || vs < b.z || vs > b.w) continue;
        if (cnt < CAND_MAX) cand[cnt] = (uint16_t)t;
        ++cnt;
      }
      for (int64_t Z = zs; Z <= ze; ++Z) {
        if (cnt > CAND_MAX) {  // a crowded line: every triangle, culled per node
          classify_node(u + Z, v + Z, Z, T, px, py, pz, pb, L, type, cls, visc);
          continue;
        }
        const int64_t X = u + Z, Y = v + Z;
        const int64_t n =
            wrap_node(Z, L.nz) + (int64_t)L.nz * (wrap_node(Y, L.ny) + (int64_t)L.ny * wrap_node(X, L.nx));
        if (type != nullptr && type[n] != HC_NODE_FLUID) continue;
        const double sx = (double)X, sy = (double)Y, sz = (double)Z;
        int cross = 0;
        for (int c = 0; c < cnt; ++c) {
          const int t = cand[c];
          const int i0 = T.tri[3 * t], i1 = T.tri[3 * t + 1], i2 = T.tri[3 * t + 2];
          const double ax = px[i0], ay = py[i0], az = pz[i0];
          const double bx = px[i1], by = py[i1], bz = pz[i1];
          const double cx = px[i2], cy = py[i2], cz = pz[i2];
          if (fmin(fmin(ax, bx), cx) > sx || fmin(fmin(ay, by), cy) > sy || fmin(fmin(az, bz), cz) > sz) continue;
          cross += ray_crosses(ax, ay, az, bx, by, bz, cx, cy, cz, sx, sy, sz);
        }
        if (cross & 1) visc[n] = cls;
      }
    }
    return;
  }
  const int64_t total = ex * ey * ez;
  for (int64_t q = threadIdx.x; q < total; q += VB) {
    const int64_t X = box[0] + q / (ey * ez);
    const int64_t rem = q % (ey * ez);
    classify_node(X, box[1] + rem / ez, box[2] + rem % ez, T, px, py, pz, nullptr, L, type, cls, visc);
  }
}

// HemoCell::iterate deletes the incomplete cells (a vertex lost to a wall) before
// the entire-grid pass (core/hemoCell.cpp:347-351, deleteIncompleteCells,
// hemoCellParticleField.cpp:510-553): every vertex of such a cell is removed.
// One workgroup per cell.
__global__ __launch_bounds__(VB) void k_delete_incomplete(VertexDev V, const int64_t* __restrict__ cell_base,
                                                          const int32_t* __restrict__ cell_nv) {
  const int64_t b = cell_base[blockIdx.x];
  const int nv = cell_nv[blockIdx.x];
  bool lrem = false;
  for (int k = threadIdx.x; k < nv; k += VB) lrem |= (V.flag[b + k] & VFLAG_REMOVED) != 0;
  if (__syncthreads_or(lrem))
    for (int k = threadIdx.x; k < nv; k += VB) V.flag[b + k] = VFLAG_REMOVED;
}

template <bool APPLY>
__global__ __launch_bounds__(VB) void k_visc_membrane(VertexDev V, TypeDev T, const int32_t* __restrict__ cells,
                                                      const int64_t* __restrict__ cell_base,
                                                      const double* __restrict__ xold, LatticeDev L,
                                                      const uint8_t* __restrict__ type, uint8_t cls,
                                                      uint32_t* __restrict__ stamp, uint8_t* __restrict__ visc) {
  const int64_t b = cell_base[cells[blockIdx.x]];
  const CellView cv{V.x + b, V.x + V.vs + b, V.x + 2 * V.vs + b};
  for (int k = threadIdx.x; k < T.nv; k += VB) {
    const int64_t i = b + k;
    if (V.flag && (V.flag[i] & VFLAG_REMOVED)) continue;  // removed by wall contact
    // rbcHighOrderModel.cpp:114-121: sum of unit normal x area / area_mean_eq
    // over the vertex's triangles, ascending; zero for the other models
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (T.model == HC_MODEL_RBC) {
      for (int e = T.vtri_off[k]; e < T.vtri_off[k + 1]; ++e) {
        const int4 q = T.vtri4[e];
        const D3 v0 = cv.pos(q.y), v1 = cv.pos(q.z), v2 = cv.pos(q.w);
        D3 n = crossp(v1 - v0, v2 - v0);
        const double nn = normp(n);
        double area = 0.0;
        if (nn != 0.0) {
          area = 0.5 * nn;
          n.x /= nn;
          n.y /= nn;
          n.z /= nn;
        } else {
          n = {0.0, 0.0, 0.0};
        }
        const double s = area / T.area_mean_eq;
        nx += n.x * s;
        ny += n.y * s;
        nz += n.z * s;
      }
    }
    Support sup;
    ibm_support(xold[i], xold[V.vs + i], xold[2 * V.vs + i], L, type, sup);
    const double x0 = V.x[i], x1 = V.x[V.vs + i], x2 = V.x[2 * V.vs + i];
#pragma unroll
    for (int slot = 0; slot < 8; ++slot) {
      if (sup.w[slot] == 0.0) continue;
      const double l0 = (double)(sup.lo[0] + (slot >> 2)) - x0;
      const double l1 = (double)(sup.lo[1] + ((slot >> 1) & 1)) - x1;
      const double l2 = (double)(sup.lo[2] + (slot & 1)) - x2;
      const double len = sqrt(l0 * l0 + l1 * l1 + l2 * l2);
      if (len > T.edge_mean_eq) continue;
      const double d = l0 * nx + l1 * ny + l2 * nz;
      const uint32_t key = ((uint32_t)(i + 1) << 1) | (d < 0.0 ? 1u : 0u);
      const uint32_t node = sup.idx[slot];
      if (!APPLY) {
        atomicMax(&stamp[node], key);
      } else {
        const uint32_t w = stamp[node];
        if ((w >> 1) == (uint32_t)(i + 1)) {
          visc[node] = (w & 1u) ? cls : (uint8_t)0;
          stamp[node] = 0u;
        }
      }
    }
  }
}

__global__ void k_fill_u8(uint8_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0;
}

}  // namespace

hipError_t launch_visc_clear(hipStream_t s, uint8_t* visc, int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)g), dim3(256), 0, s, visc, n);
  return hipGetLastError();
}

hipError_t launch_visc_grid(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                            const int64_t* cell_base, const LatticeDev& L, const uint8_t* type, int cls,
                            uint8_t* visc) {
  if (ncells <= 0) return hipSuccess;
  // 64 KB of LDS per workgroup, less the static arrays
  constexpr size_t kBudget = 65536 - 2048;
  const size_t pos = sizeof(double) * (((size_t)3 * T.nv + 1) & ~(size_t)1);
  const size_t lines = pos + sizeof(float4) * (size_t)T.nt + sizeof(uint16_t) * VB * CAND_MAX;
  if (lines <= kBudget && T.nt < 65536) {
    hipLaunchKernelGGL((k_visc_grid<true>), dim3((unsigned)ncells), dim3(VB), lines, s, V, T, cells, cell_base, L,
                       type, (uint8_t)cls, visc);
  } else {
    if (pos > kBudget) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_visc_grid<false>), dim3((unsigned)ncells), dim3(VB), pos, s, V, T, cells, cell_base, L,
                       type, (uint8_t)cls, visc);
  }
  return hipGetLastError();
}

hipError_t launch_delete_incomplete(hipStream_t s, const VertexDev& V, const int64_t* cell_base, const int32_t* cell_nv,
                                    int ncells) {
  if (ncells <= 0 || !V.flag) return hipSuccess;
  hipLaunchKernelGGL(k_delete_incomplete, dim3((unsigned)ncells), dim3(VB), 0, s, V, cell_base, cell_nv);
  return hipGetLastError();
}

hipError_t launch_visc_membrane(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells,
                                int ncells, const int64_t* cell_base, const double* xold, const LatticeDev& L,
                                const uint8_t* type, int cls, uint32_t* stamp, uint8_t* visc, bool apply) {
  if (ncells <= 0) return hipSuccess;
  if (apply)
    hipLaunchKernelGGL((k_visc_membrane<true>), dim3((unsigned)ncells), dim3(VB), 0, s, V, T, cells, cell_base, xold,
                       L, type, (uint8_t)cls, stamp, visc);
  else
    hipLaunchKernelGGL((k_visc_membrane<false>), dim3((unsigned)ncells), dim3(VB), 0, s, V, T, cells, cell_base,
                       xold, L, type, (uint8_t)cls, stamp, visc);
  return hipGetLastError();
}

}  // namespace hcg
