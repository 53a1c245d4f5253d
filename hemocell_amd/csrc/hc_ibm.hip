// hc_ibm.hip -- immersed-boundary coupling on MI355X: phi2 force spreading,
// velocity interpolation and the Euler advance of the membrane vertices.
//
// Replaces (reference paths relative to the HemoCell tree):
//   interpolationCoefficientsPhi2  core/immersedBoundaryMethod.h:37-41,62-138
//   spreadParticleForce            core/hemoCellParticleField.cpp:841-863
//   interpolateFluidVelocity       core/hemoCellParticleField.cpp:819-839
//   advanceParticles               core/hemoCellParticleField.cpp:566-588
//
// One thread per vertex; vertex state is FP64 structure-of-arrays (x, y, z
// planes), so each load of a wavefront is one coalesced 512-byte read.  The
// phi2 weights are recomputed from the position in both the spread and the
// interpolation (the positions do not change in between, so both see the same
// support), never stored.  Support nodes: the 3x3x3 nodes around the nearest
// node, visited x -> y -> z, skipping nodes outside a non-periodic domain,
// zero weights and wall nodes, renormalised to sum 1 -- at most 2 nodes per
// axis have a non-zero weight, so at most 8 survive.
//
// The spread accumulates into the lattice force field with hardware float
// atomics (global_atomic_add_f32/f64).
#include <cmath>

#include "hc_ibm_dev.h"

namespace hcg {
namespace {

constexpr int BLOCK = 256;

// The repulsion force a vertex spreads (force_repulsion of the reference): bit 1
// of use_rep adds the cell-cell term, bit 2 the wall term, cell-cell first.
__device__ __forceinline__ void repulsion_sum(const VertexDev& V, int64_t i, int use_rep, double& r0, double& r1,
                                              double& r2) {
  const int64_t vs = V.vs;
  r0 = 0.0;
  r1 = 0.0;
  r2 = 0.0;
  if (use_rep & 1) {
    r0 = V.frep[i];
    r1 = V.frep[vs + i];
    r2 = V.frep[2 * vs + i];
  }
  if (use_rep & 2) {
    r0 = r0 + V.fbr[i];
    r1 = r1 + V.fbr[vs + i];
    r2 = r2 + V.fbr[2 * vs + i];
  }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_spread(VertexDev V, R* __restrict__ Fibm, const uint8_t* __restrict__ type,
                                                  LatticeDev L, double f_limit, int use_rep) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V.n) return;
  if (V.flag && (V.flag[i] & VFLAG_REMOVED)) return;  // removed by wall contact
  const int64_t vs = V.vs;
  Support sp;
  ibm_support(V.x[i], V.x[vs + i], V.x[2 * vs + i], L, type, sp);
  double f0 = V.f[i], f1 = V.f[vs + i], f2 = V.f[2 * vs + i];
  const double fm = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
  if (fm > f_limit) {
    const double s = f_limit / fm;
    f0 *= s;
    f1 *= s;
    f2 *= s;
    V.f[i] = f0;
    V.f[vs + i] = f1;
    V.f[2 * vs + i] = f2;
  }
  double t0 = f0, t1 = f1, t2 = f2;
  if (use_rep) {
    double r0, r1, r2;
    repulsion_sum(V, i, use_rep, r0, r1, r2);
    t0 = r0 + f0;
    t1 = r1 + f1;
    t2 = r2 + f2;
  }
  const int64_t S = L.stride;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (sp.w[k] == 0.0) continue;
    const int64_t n = sp.idx[k];
    atomic_add(&Fibm[n], (R)spread_term(t0 * sp.w[k], L));
    atomic_add(&Fibm[S + n], (R)spread_term(t1 * sp.w[k], L));
    atomic_add(&Fibm[2 * S + n], (R)spread_term(t2 * sp.w[k], L));
  }
}

// Fluid velocity u = j/rho + F/2 of every node from its post-stream populations
// (phase of the AA pattern as in hc_lbm.hip) and the force field the spread
// filled this step: ExternalForceDynamics::computeVelocity restated.  One
// coalesced pass writing (u0, u1, u2, 0) per node, so the interpolation reads
// each support node with one vector load.
template <typename R, bool PHASE1, bool CHECK>
__device__ __forceinline__ void velocity_node_field(const R* __restrict__ f, const R* __restrict__ Fcur,
                                                    R* __restrict__ uf, const LatticeDev& L, double g0, double g1,
                                                    double g2, const uint32_t n) {
  if (n >= (uint32_t)L.N) return;
  const int64_t S = L.stride;
  const int z = (int)(n % (uint32_t)L.nz);
  const uint32_t t = n / (uint32_t)L.nz;
  const int y = (int)(t % (uint32_t)L.ny);
  const int x = (int)(t / (uint32_t)L.ny);
  const NbrOff o = neighbour_offsets(x, y, z, L);
  R fl[19];
  aa_load<R, PHASE1, CHECK>(f, S, n, o, fl, std::make_integer_sequence<int, 19>{});
  R rhoBar, j0, j1, j2;
  rhobar_j(fl, rhoBar, j0, j1, j2);
  const R invRho = R(1) / (rhoBar + R(1));
  R* q = uf + 3 * (int64_t)n;  // interleaved: one 3-component load per support node
  q[0] = j0 * invRho + (R(g0) + Fcur[n]) / R(2);
  q[1] = j1 * invRho + (R(g1) + Fcur[S + n]) / R(2);
  q[2] = j2 * invRho + (R(g2) + Fcur[2 * S + n]) / R(2);
}

template <typename R, bool PHASE1, bool CHECK>
__global__ __launch_bounds__(BLOCK) void k_velocity_field(const R* __restrict__ f, const R* __restrict__ Fcur,
                                                          R* __restrict__ uf, LatticeDev L, double g0, double g1,
                                                          double g2) {
  velocity_node_field<R, PHASE1, CHECK>(f, Fcur, uf, L, g0, g1, g2, blockIdx.x * BLOCK + threadIdx.x);
}

// several independent lattices in one launch; case c owns the blocks from c.block0
template <typename R, bool PHASE1, bool CHECK>
__global__ __launch_bounds__(BLOCK) void k_velocity_field_multi(const VelCase* __restrict__ cs, int ncase) {
  int lo = 0, hi = ncase - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cs[mid].block0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const VelCase& c = cs[lo];
  velocity_node_field<R, PHASE1, CHECK>((const R*)c.f, (const R*)c.Fcur, (R*)c.uf, c.L, c.g[0], c.g[1], c.g[2],
                                        (blockIdx.x - c.block0) * BLOCK + threadIdx.x);
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_interp_advance(VertexDev V, const R* __restrict__ uf,
                                                          const uint8_t* __restrict__ type, LatticeDev L,
                                                          int do_interp, int do_advance,
                                                          unsigned long long* wall_hits) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V.n) return;
  if (V.flag && (V.flag[i] & VFLAG_REMOVED)) return;  // removed by wall contact: stays put
  const int64_t vs = V.vs;
  double x0 = V.x[i], x1 = V.x[vs + i], x2 = V.x[2 * vs + i];
  double v0, v1, v2;
  if (do_interp) {
    Support sp;
    ibm_support(x0, x1, x2, L, type, sp);
    v0 = 0.0;
    v1 = 0.0;
    v2 = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (sp.w[k] == 0.0) continue;
      const R* q = uf + 3 * (int64_t)sp.idx[k];
      v0 += (double)q[0] * sp.w[k];
      v1 += (double)q[1] * sp.w[k];
      v2 += (double)q[2] * sp.w[k];
    }
    V.v[i] = v0;
    V.v[vs + i] = v1;
    V.v[2 * vs + i] = v2;
  } else {
    v0 = V.v[i];
    v1 = V.v[vs + i];
    v2 = V.v[2 * vs + i];
  }
  if (!do_advance) return;
  x0 += v0;
  x1 += v1;
  x2 += v2;
  V.x[i] = x0;
  V.x[vs + i] = x1;
  V.x[2 * vs + i] = x2;
  if (type && isfinite(x0) && isfinite(x1) && isfinite(x2)) {
    const double p[3] = {x0, x1, x2};
    const int dims[3] = {L.nx, L.ny, L.nz};
    int64_t wc[3];
    bool out = false;
    for (int a = 0; a < 3; ++a) {
      const int64_t c = (int64_t)floor(p[a] + 0.5);
      if (c >= 0 && c < dims[a]) wc[a] = c;
      else if (L.periodic[a]) wc[a] = wrap_mod(c, dims[a]);
      else out = true;
    }
    if (!out && type[wc[2] + (int64_t)L.nz * (wc[1] + (int64_t)L.ny * wc[0])] != HC_NODE_FLUID) {
      // on a wall node: the vertex is removed (hemoCellParticleField.cpp:566-588)
      atomicAdd(wall_hits, 1ull);
      if (V.flag) V.flag[i] = VFLAG_REMOVED;
    }
  }
}

// Cap the vertex force to f_limit in place (hemoCellParticleField.cpp:847-852) and
// return the force to spread (plus the repulsion force when enabled).
__device__ __forceinline__ void spread_force(const VertexDev& V, int64_t i, double f_limit, int use_rep, double& t0,
                                             double& t1, double& t2) {
  const int64_t vs = V.vs;
  if (V.flag && (V.flag[i] & VFLAG_REMOVED)) {  // removed by wall contact: spreads nothing
    t0 = t1 = t2 = 0.0;
    return;
  }
  double f0 = V.f[i], f1 = V.f[vs + i], f2 = V.f[2 * vs + i];
  const double fm = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
  if (fm > f_limit) {
    const double s = f_limit / fm;
    f0 *= s;
    f1 *= s;
    f2 *= s;
    V.f[i] = f0;
    V.f[vs + i] = f1;
    V.f[2 * vs + i] = f2;
  }
  t0 = f0;
  t1 = f1;
  t2 = f2;
  if (use_rep) {
    double r0, r1, r2;
    repulsion_sum(V, i, use_rep, r0, r1, r2);
    t0 = r0 + f0;
    t1 = r1 + f1;
    t2 = r2 + f2;
  }
}

// Force spreading, one workgroup per cell: the cell's contributions are summed
// in an LDS box covering its support (LDS atomics, many vertices share nodes),
// then every non-zero node of the box is added to the force field with one
// global atomic.  A cell whose box does not fit the LDS budget spreads with
// global atomics directly.  The box holds doubles in both precisions (LDS
// double atomic adds are several times faster than float ones here).
template <typename R, int BS>
__global__ __launch_bounds__(BS) void k_spread_cells(VertexDev V, R* __restrict__ Fibm, const uint8_t* __restrict__ type,
                                                     LatticeDev L, const int64_t* __restrict__ cell_base,
                                                     const int32_t* __restrict__ cell_nv, double f_limit, int use_rep,
                                                     int lds_nodes) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* box = (double*)smem;
  __shared__ int bmin[3], bmax[3], bad;
  const int cell = blockIdx.x, tid = threadIdx.x;
  const int64_t base = cell_base[cell];
  const int nv = cell_nv[cell];
  const int64_t vs = V.vs, S = L.stride;
  if (tid == 0) {
    for (int a = 0; a < 3; ++a) {
      bmin[a] = 0x7fffffff;
      bmax[a] = -0x7fffffff;
    }
    bad = 0;
  }
  __syncthreads();
  int lmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, lmax[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
  bool lbad = false;
  for (int k = tid; k < nv; k += BS) {
    const int64_t i = base + k;
    for (int a = 0; a < 3; ++a) {
      const double p = V.x[a * vs + i];
      if (!(fabs(p) < 1.0e9)) {
        lbad = true;
        continue;
      }
      const int c = (int)floor(p);
      lmin[a] = min(lmin[a], c);
      lmax[a] = max(lmax[a], c);
    }
  }
  for (int a = 0; a < 3; ++a) {
    atomicMin(&bmin[a], lmin[a]);
    atomicMax(&bmax[a], lmax[a]);
  }
  if (lbad) atomicOr(&bad, 1);
  __syncthreads();
  // support nodes of a vertex are floor(p) and floor(p) + 1 on each axis
  const int o0 = bmin[0], o1 = bmin[1], o2 = bmin[2];
  const int s0 = bmax[0] - bmin[0] + 2, s1 = bmax[1] - bmin[1] + 2, s2 = bmax[2] - bmin[2] + 2;
  const int64_t total = (int64_t)s0 * s1 * s2;
  const bool direct = bad || total > lds_nodes || s0 <= 0;
  if (!direct)
    for (int64_t e = tid; e < 3 * total; e += BS) box[e] = 0.0;
  __syncthreads();
  // lanes of one wavefront take vertices ~nv/64 apart on the mesh, so they rarely
  // add to the same LDS node at once (consecutive mesh vertices share support nodes)
  const int per = (nv + 63) / 64;
  for (int k = tid; k < 64 * per; k += BS) {
    const int vtx = (k & 63) * per + (k >> 6);
    if (vtx >= nv) continue;
    const int64_t i = base + vtx;
    Support sp;
    ibm_support(V.x[i], V.x[vs + i], V.x[2 * vs + i], L, type, sp);
    double t0, t1, t2;
    spread_force(V, i, f_limit, use_rep, t0, t1, t2);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (sp.w[q] == 0.0) continue;
      if (direct) {
        const int64_t n = sp.idx[q];
        atomic_add(&Fibm[n], (R)spread_term(t0 * sp.w[q], L));
        atomic_add(&Fibm[S + n], (R)spread_term(t1 * sp.w[q], L));
        atomic_add(&Fibm[2 * S + n], (R)spread_term(t2 * sp.w[q], L));
      } else {
        const int e = (sp.lo[2] + (q & 1) - o2) + s2 * ((sp.lo[1] + ((q >> 1) & 1) - o1) + s1 * (sp.lo[0] + (q >> 2) - o0));
        lds_add(&box[e], spread_term(t0 * sp.w[q], L));
        lds_add(&box[total + e], spread_term(t1 * sp.w[q], L));
        lds_add(&box[2 * total + e], spread_term(t2 * sp.w[q], L));
      }
    }
  }
  if (direct) return;
  __syncthreads();
  const int dims[3] = {L.nx, L.ny, L.nz};
  for (int64_t e = tid; e < total; e += BS) {
    const double a = box[e], b = box[total + e], c = box[2 * total + e];
    if (a == 0.0 && b == 0.0 && c == 0.0) continue;
    const int lz = (int)(e % s2);
    const int64_t r = e / s2;
    const int ly = (int)(r % s1), lx = (int)(r / s1);
    const int64_t g[3] = {(int64_t)o0 + lx, (int64_t)o1 + ly, (int64_t)o2 + lz};
    int64_t wc[3];
    for (int q = 0; q < 3; ++q) wc[q] = (g[q] >= 0 && g[q] < dims[q]) ? g[q] : wrap_mod(g[q], dims[q]);
    const int64_t n = wc[2] + (int64_t)L.nz * (wc[1] + (int64_t)L.ny * wc[0]);
    if (a != 0.0) atomic_add(&Fibm[n], (R)a);
    if (b != 0.0) atomic_add(&Fibm[S + n], (R)b);
    if (c != 0.0) atomic_add(&Fibm[2 * S + n], (R)c);
  }
}

__global__ void k_add_extra(VertexDev V, const int64_t* ids, const double* forces, int n) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= n) return;
  const int64_t i = ids[k];
  if (i < 0 || i >= V.n) return;
  V.f[i] += forces[3 * k];
  V.f[V.vs + i] += forces[3 * k + 1];
  V.f[2 * V.vs + i] += forces[3 * k + 2];
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

}  // namespace

template <typename R>
hipError_t launch_spread(hipStream_t s, const VertexDev& V, R* Fibm, const uint8_t* type, const LatticeDev& L,
                         double f_limit, int use_rep) {
  if (V.n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_spread<R>), grid_for(V.n), dim3(BLOCK), 0, s, V, Fibm, type, L, f_limit, use_rep);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_spread_cells(hipStream_t s, const VertexDev& V, R* Fibm, const uint8_t* type, const LatticeDev& L,
                               const int64_t* cell_base, const int32_t* cell_nv, int ncells, double f_limit,
                               int use_rep, int lds_bytes) {
  if (ncells <= 0) return hipSuccess;
  const int lds_nodes = lds_bytes / (3 * (int)sizeof(double));
  hipLaunchKernelGGL((k_spread_cells<R, 256>), dim3((unsigned)ncells), dim3(256), (size_t)lds_bytes, s, V, Fibm, type,
                     L, cell_base, cell_nv, f_limit, use_rep, lds_nodes);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_velocity_field(hipStream_t s, const R* f, const R* Fcur, R* uf, const LatticeDev& L, int phase,
                                 bool general, const double g[3]) {
  dim3 gr = grid_for(L.N);
  if (phase) {
    if (general) hipLaunchKernelGGL((k_velocity_field<R, true, true>), gr, dim3(BLOCK), 0, s, f, Fcur, uf, L, g[0], g[1], g[2]);
    else hipLaunchKernelGGL((k_velocity_field<R, true, false>), gr, dim3(BLOCK), 0, s, f, Fcur, uf, L, g[0], g[1], g[2]);
  } else {
    hipLaunchKernelGGL((k_velocity_field<R, false, false>), gr, dim3(BLOCK), 0, s, f, Fcur, uf, L, g[0], g[1], g[2]);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_velocity_field_multi(hipStream_t s, const VelCase* cases, int ncase, uint32_t blocks, int phase,
                                       bool general) {
  if (ncase <= 0 || blocks == 0) return hipSuccess;
  const dim3 g(blocks), b(BLOCK);
  if (phase) {
    if (general) hipLaunchKernelGGL((k_velocity_field_multi<R, true, true>), g, b, 0, s, cases, ncase);
    else hipLaunchKernelGGL((k_velocity_field_multi<R, true, false>), g, b, 0, s, cases, ncase);
  } else {
    hipLaunchKernelGGL((k_velocity_field_multi<R, false, false>), g, b, 0, s, cases, ncase);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_interp_advance(hipStream_t s, const VertexDev& V, const R* uf, const uint8_t* type,
                                 const LatticeDev& L, int do_interp, int do_advance, unsigned long long* wall_hits) {
  if (V.n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_interp_advance<R>), grid_for(V.n), dim3(BLOCK), 0, s, V, uf, type, L, do_interp, do_advance,
                     wall_hits);
  return hipGetLastError();
}

hipError_t launch_add_extra_force(hipStream_t s, const VertexDev& V, const int64_t* ids, const double* forces, int n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_add_extra, grid_for(n), dim3(BLOCK), 0, s, V, ids, forces, n);
  return hipGetLastError();
}

#define HC_INSTANTIATE(R)                                                                                     \
  template hipError_t launch_spread<R>(hipStream_t, const VertexDev&, R*, const uint8_t*, const LatticeDev&, double,  \
                                       int);                                                                         \
  template hipError_t launch_spread_cells<R>(hipStream_t, const VertexDev&, R*, const uint8_t*, const LatticeDev&,    \
                                             const int64_t*, const int32_t*, int, double, int, int);                  \
  template hipError_t launch_velocity_field<R>(hipStream_t, const R*, const R*, R*, const LatticeDev&, int, bool,     \
                                               const double*);                                                       \
  template hipError_t launch_interp_advance<R>(hipStream_t, const VertexDev&, const R*, const uint8_t*,               \
                                               const LatticeDev&, int, int, unsigned long long*);                     \
  template hipError_t launch_velocity_field_multi<R>(hipStream_t, const VelCase*, int, uint32_t, int, bool);
HC_INSTANTIATE(float)
HC_INSTANTIATE(double)
#undef HC_INSTANTIATE

}  // namespace hcg
