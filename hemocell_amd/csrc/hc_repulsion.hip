// hc_repulsion.hip -- vertex binning and cell-cell repulsion on MI355X.
//
// Replaces (reference paths relative to the HemoCell tree):
//   HemoCellParticleField::update_pg           core/hemoCellParticleField.cpp:137-168
//   HemoCellParticleField::applyRepulsionForce core/hemoCellParticleField.cpp:677-743
//   populateBoundaryParticles / applyBoundaryRepulsionForce  :865-918
//
// Binning: every vertex gets the key of its nearest node, floor(x + 0.5) per
// axis (wrapped on periodic axes; vertices outside a non-periodic axis get the
// key N and no bin).  A stable radix sort of (key, vertex id) leaves each bin's
// vertices in ascending vertex order, the order in which update_pg appends
// them; a bounds pass records [start, end) of every occupied node and a final
// pass clears exactly those entries again, so no per-step pass over all N
// nodes is needed.  Unlike the reference's fixed 10 slots per node, bins have
// no capacity limit.
//
// Forces: the reference walks every node l with itself and 13 forward
// neighbours m (a half stencil) and, for each ordered pair (i in bin l, j in
// bin m) of vertices of different cells closer than the cut-off, adds
// F = k (r_c / r) r_hat to i and subtracts it from j.  Here each vertex v
// gathers its own terms: it enumerates the nodes l whose half stencil reaches
// its node, in ascending node order, and within each (l, offset) visits the
// pairs in the reference's loop order.  The separation of the reversed pair is
// the exact negation of the forward one (IEEE subtraction, minimum image
// included), so the FP64 sum per vertex is formed from the same terms in the
// same order, without atomics.
#include <rocprim/device/device_radix_sort.hpp>

#include "hc_launch.h"

namespace hcg {
namespace {

// the 14-offset half stencil of applyRepulsionForce, in its visiting order
__constant__ int kOff[14][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},   {1, -1, -1}, {1, -1, 0}, {1, -1, 1},
                                {1, 0, -1}, {1, 0, 0}, {1, 0, 1}, {1, 1, -1}, {1, 1, 0},   {1, 1, 1},  {0, 1, -1}};

// node coordinate along an axis of size n: wrapped when periodic, -1 outside
__device__ __forceinline__ int wrap_axis(int64_t c, int n, int periodic) {
  if (c >= 0 && c < n) return (int)c;
  if (!periodic) return -1;
  int64_t m = c % n;
  if (m < 0) m += n;
  return (int)m;
}

__device__ __forceinline__ int64_t node_index(int x, int y, int z, const LatticeDev& L) {
  return (int64_t)z + (int64_t)L.nz * ((int64_t)y + (int64_t)L.ny * x);
}

// nearest-node key of vertex i (N when it is outside the domain or was removed by
// wall contact: such a vertex is in no bin and repels nothing)
__device__ __forceinline__ uint32_t vertex_key(const VertexDev& V, int64_t i, const LatticeDev& L, int c[3]) {
  const int dims[3] = {L.nx, L.ny, L.nz};
  bool inside = !(V.flag && (V.flag[i] & VFLAG_REMOVED));
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double p = V.x[a * V.vs + i];
    if (!(fabs(p) < 1.0e15)) {
      inside = false;
      c[a] = -1;
      continue;
    }
    c[a] = wrap_axis((int64_t)floor(p + 0.5), dims[a], L.periodic[a]);
    inside &= c[a] >= 0;
  }
  return inside ? (uint32_t)node_index(c[0], c[1], c[2], L) : (uint32_t)L.N;
}

__global__ void k_bin_keys(VertexDev V, LatticeDev L, uint32_t* __restrict__ keys, int32_t* __restrict__ ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V.n) return;
  int c[3];
  keys[i] = vertex_key(V, i, L, c);
  ids[i] = (int32_t)i;
}

// [start, end) of every occupied node in the sorted order
__global__ void k_bin_bounds(const uint32_t* __restrict__ keys, int64_t n, uint32_t N, int2* __restrict__ bins) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t k = keys[p];
  if (k >= N) return;
  if (p == 0 || keys[p - 1] != k) bins[k].x = (int)p;
  if (p == n - 1 || keys[p + 1] != k) bins[k].y = (int)(p + 1);
}

__global__ void k_bin_clear(const uint32_t* __restrict__ keys, int64_t n, uint32_t N, int2* __restrict__ bins) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t k = keys[p];
  if (k < N) bins[k] = make_int2(0, 0);
}

// Boundary particles (populateBoundaryParticles, hemoCellParticleField.cpp:865-887):
// bounce-back nodes with a node that is not bounce-back among their 26
// neighbours; neighbours wrap on periodic axes and end at non-periodic faces.
__global__ void k_boundary_flags(const uint8_t* __restrict__ type, LatticeDev L, uint8_t* __restrict__ flags) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L.N) return;
  uint8_t flag = 0;
  if (type && type[n] == HC_NODE_BOUNCEBACK) {
    const int z = (int)(n % L.nz);
    const int64_t t = n / L.nz;
    const int y = (int)(t % L.ny), x = (int)(t / L.ny);
    for (int dx = -1; dx <= 1 && !flag; ++dx) {
      const int xx = wrap_axis(x + dx, L.nx, L.periodic[0]);
      if (xx < 0) continue;
      for (int dy = -1; dy <= 1 && !flag; ++dy) {
        const int yy = wrap_axis(y + dy, L.ny, L.periodic[1]);
        if (yy < 0) continue;
        for (int dz = -1; dz <= 1 && !flag; ++dz) {
          const int zz = wrap_axis(z + dz, L.nz, L.periodic[2]);
          if (zz < 0) continue;
          if (type[node_index(xx, yy, zz, L)] != HC_NODE_BOUNCEBACK) flag = 1;
        }
      }
    }
  }
  flags[n] = flag;
}

// Wall repulsion (applyBoundaryRepulsionForce, hemoCellParticleField.cpp:889-918):
// every boundary particle b adds k (r_c / r) r_hat, r = x - b, to each vertex of
// the bins in its 3x3x3 neighbourhood closer than r_c.  Gathered per vertex: the
// boundary particles in the 3x3x3 neighbourhood of the vertex's own bin, in
// ascending x, y, z (the reference's order of boundary particles), the separation
// taken to the unwrapped node, i.e. the minimum image.  Overwrites fbr.
__global__ void k_boundary_repulsion(VertexDev V, LatticeDev L, const uint8_t* __restrict__ flags, double k_rep,
                                     double r_c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V.n) return;
  const int64_t vs = V.vs;
  const double p[3] = {V.x[i], V.x[vs + i], V.x[2 * vs + i]};
  double F0 = 0.0, F1 = 0.0, F2 = 0.0;
  const bool removed = V.flag && (V.flag[i] & VFLAG_REMOVED);
  if (!removed && fabs(p[0]) < 1.0e9 && fabs(p[1]) < 1.0e9 && fabs(p[2]) < 1.0e9) {
    const int64_t c0 = (int64_t)floor(p[0] + 0.5), c1 = (int64_t)floor(p[1] + 0.5), c2 = (int64_t)floor(p[2] + 0.5);
    for (int dx = -1; dx <= 1; ++dx) {
      const int xx = wrap_axis(c0 + dx, L.nx, L.periodic[0]);
      if (xx < 0) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int yy = wrap_axis(c1 + dy, L.ny, L.periodic[1]);
        if (yy < 0) continue;
        for (int dz = -1; dz <= 1; ++dz) {
          const int zz = wrap_axis(c2 + dz, L.nz, L.periodic[2]);
          if (zz < 0 || !flags[node_index(xx, yy, zz, L)]) continue;
          const double d0 = p[0] - (double)(c0 + dx), d1 = p[1] - (double)(c1 + dy), d2 = p[2] - (double)(c2 + dz);
          const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
          if (dist < r_c) {
            const double s = k_rep * (1 / (dist / r_c));
            F0 = F0 + s * (d0 / dist);
            F1 = F1 + s * (d1 / dist);
            F2 = F2 + s * (d2 / dist);
          }
        }
      }
    }
  }
  V.fbr[i] = F0;
  V.fbr[vs + i] = F1;
  V.fbr[2 * vs + i] = F2;
}

__global__ void k_vertex_bins(const VertexDev V, LatticeDev L, int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V.n) return;
  int c[3];
  const uint32_t k = vertex_key(V, i, L, c);
  out[i] = k >= (uint32_t)L.N ? -1 : (int64_t)k;
}

struct Pos {
  double x, y, z;
};

// the repulsion term of vertex v from partner w (F on v), zero when excluded
__device__ __forceinline__ void pair_term(const Pos& pv, int cv, int64_t v, int64_t w, const VertexDev& V,
                                          const int32_t* __restrict__ vcell, const LatticeDev& L, double k_rep,
                                          double r_c, double& a0, double& a1, double& a2) {
  if (w == v || vcell[w] == cv) return;
  const int dims[3] = {L.nx, L.ny, L.nz};
  double d[3] = {pv.x - V.x[w], pv.y - V.x[V.vs + w], pv.z - V.x[2 * V.vs + w]};
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (L.periodic[a]) {
      const double n = (double)dims[a];
      d[a] -= n * floor(d[a] / n + 0.5);
    }
  const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (dist < r_c) {
    const double s = k_rep * (1 / (dist / r_c));
    a0 += s * (d[0] / dist);
    a1 += s * (d[1] / dist);
    a2 += s * (d[2] / dist);
  }
}

__global__ __launch_bounds__(256) void k_repulsion(VertexDev V, LatticeDev L, const int32_t* __restrict__ vcell,
                                                   const int2* __restrict__ bins, const int32_t* __restrict__ order,
                                                   double k_rep, double r_c) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V.n) return;
  int c[3];
  const uint32_t key = vertex_key(V, v, L, c);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  if (key < (uint32_t)L.N) {
    const Pos pv{V.x[v], V.x[V.vs + v], V.x[2 * V.vs + v]};
    const int cv = vcell[v];
    const int dims[3] = {L.nx, L.ny, L.nz};
    // nodes l whose half stencil reaches this node: l = node - offset
    int64_t cand[14];
    int nc = 0;
    for (int o = 0; o < 14; ++o) {
      int l3[3];
      bool ok = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        l3[a] = wrap_axis((int64_t)c[a] - kOff[o][a], dims[a], L.periodic[a]);
        ok &= l3[a] >= 0;
      }
      if (!ok) continue;
      const int64_t l = node_index(l3[0], l3[1], l3[2], L);
      bool dup = false;
      for (int q = 0; q < nc; ++q) dup |= cand[q] == l;
      if (dup) continue;
      // insertion into the ascending list
      int q = nc++;
      while (q > 0 && cand[q - 1] > l) {
        cand[q] = cand[q - 1];
        --q;
      }
      cand[q] = l;
    }
    for (int q = 0; q < nc; ++q) {
      const int64_t l = cand[q];
      const int2 bl = bins[l];
      if (bl.y <= bl.x) continue;
      const int lx = (int)(l / ((int64_t)L.ny * L.nz));
      const int ly = (int)((l / L.nz) % L.ny);
      const int lz = (int)(l % L.nz);
      for (int o = 0; o < 14; ++o) {
        const int mx = wrap_axis((int64_t)lx + kOff[o][0], L.nx, L.periodic[0]);
        const int my = wrap_axis((int64_t)ly + kOff[o][1], L.ny, L.periodic[1]);
        const int mz = wrap_axis((int64_t)lz + kOff[o][2], L.nz, L.periodic[2]);
        if (mx < 0 || my < 0 || mz < 0) continue;
        const int64_t m = node_index(mx, my, mz, L);
        if (l == (int64_t)key && m == (int64_t)key) {
          // both ends in this bin: for i in bin, for j in bin
          for (int p = bl.x; p < bl.y; ++p) {
            const int64_t i = order[p];
            if (i == v) {
              for (int r = bl.x; r < bl.y; ++r) pair_term(pv, cv, v, order[r], V, vcell, L, k_rep, r_c, a0, a1, a2);
            } else {
              pair_term(pv, cv, v, i, V, vcell, L, k_rep, r_c, a0, a1, a2);
            }
          }
        } else if (l == (int64_t)key) {
          const int2 bm = bins[m];
          for (int r = bm.x; r < bm.y; ++r) pair_term(pv, cv, v, order[r], V, vcell, L, k_rep, r_c, a0, a1, a2);
        } else if (m == (int64_t)key) {
          for (int p = bl.x; p < bl.y; ++p) pair_term(pv, cv, v, order[p], V, vcell, L, k_rep, r_c, a0, a1, a2);
        }
      }
    }
  }
  V.frep[v] = a0;
  V.frep[V.vs + v] = a1;
  V.frep[2 * V.vs + v] = a2;
}

int key_bits(int64_t N) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) <= N) ++b;
  return b;
}

}  // namespace

size_t repulsion_temp_bytes(int64_t nv, int64_t N) {
  size_t bytes = 0;
  if (nv <= 0) return 0;
  rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const int32_t*)nullptr,
                            (int32_t*)nullptr, (size_t)nv, 0, (unsigned)key_bits(N));
  return bytes;
}

hipError_t launch_repulsion(hipStream_t s, const VertexDev& V, const LatticeDev& L, const RepulsionBufs& B,
                            double k_rep, double r_c) {
  if (V.n <= 0) return hipSuccess;
  if (V.n > INT32_MAX || L.N >= (int64_t)UINT32_MAX) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((V.n + 255) / 256);
  hipLaunchKernelGGL(k_bin_keys, dim3(grid), dim3(256), 0, s, V, L, B.keys_in, B.ids_in);
  size_t bytes = B.temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(B.temp, bytes, B.keys_in, B.keys_out, B.ids_in, B.ids_out, (size_t)V.n, 0,
                                           (unsigned)key_bits(L.N), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bin_bounds, dim3(grid), dim3(256), 0, s, B.keys_out, V.n, (uint32_t)L.N, B.bins);
  hipLaunchKernelGGL(k_repulsion, dim3(grid), dim3(256), 0, s, V, L, B.vcell, B.bins, B.ids_out, k_rep, r_c);
  hipLaunchKernelGGL(k_bin_clear, dim3(grid), dim3(256), 0, s, B.keys_out, V.n, (uint32_t)L.N, B.bins);
  return hipGetLastError();
}

hipError_t launch_boundary_flags(hipStream_t s, const uint8_t* type, const LatticeDev& L, uint8_t* flags) {
  if (L.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_boundary_flags, dim3((unsigned)((L.N + 255) / 256)), dim3(256), 0, s, type, L, flags);
  return hipGetLastError();
}

hipError_t launch_boundary_repulsion(hipStream_t s, const VertexDev& V, const LatticeDev& L, const uint8_t* flags,
                                     double k, double r_c) {
  if (V.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_boundary_repulsion, dim3((unsigned)((V.n + 255) / 256)), dim3(256), 0, s, V, L, flags, k, r_c);
  return hipGetLastError();
}

hipError_t launch_vertex_bins(hipStream_t s, const VertexDev& V, const LatticeDev& L, int64_t* out) {
  if (V.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vertex_bins, dim3((unsigned)((V.n + 255) / 256)), dim3(256), 0, s, V, L, out);
  return hipGetLastError();
}

}  // namespace hcg
