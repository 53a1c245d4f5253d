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

#include "hc_d3q19.h"
#include "hc_launch.h"

namespace hcg {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ int64_t wrap_mod(int64_t c, int n) {
  int64_t m = c % n;
  return m < 0 ? m + n : m;
}

// phi2 support of position p; returns the number of nodes (<= 8).
__device__ int ibm_support(double p0, double p1, double p2, const LatticeDev& L, const uint8_t* type,
                           int64_t (&idx)[8], double (&w)[8]) {
  if (!(isfinite(p0) && isfinite(p1) && isfinite(p2))) return 0;
  const double p[3] = {p0, p1, p2};
  const int dims[3] = {L.nx, L.ny, L.nz};
  int64_t c[3];
  for (int a = 0; a < 3; ++a) c[a] = (int64_t)floor(p[a] + 0.5);
  int cnt = 0;
  double total = 0.0;
  for (int dx = -1; dx < 2; ++dx)
    for (int dy = -1; dy < 2; ++dy)
      for (int dz = -1; dz < 2; ++dz) {
        const int64_t nd[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
        int64_t wc[3];
        bool out = false;
        for (int a = 0; a < 3; ++a) {
          if (nd[a] >= 0 && nd[a] < dims[a]) wc[a] = nd[a];
          else if (L.periodic[a]) wc[a] = wrap_mod(nd[a], dims[a]);
          else out = true;
        }
        if (out) continue;
        const double q0 = fmax(1.0 - fabs(p[0] - (double)nd[0]), 0.0);
        const double q1 = fmax(1.0 - fabs(p[1] - (double)nd[1]), 0.0);
        const double q2 = fmax(1.0 - fabs(p[2] - (double)nd[2]), 0.0);
        const double ww = q0 * q1 * q2;
        if (ww == 0.0) continue;
        const int64_t n = wc[2] + (int64_t)L.nz * (wc[1] + (int64_t)L.ny * wc[0]);
        if (type && type[n] != HC_NODE_FLUID) continue;
        if (cnt < 8) {
          idx[cnt] = n;
          w[cnt] = ww;
          total += ww;
          ++cnt;
        }
      }
  const double coeff = 1.0 / total;
  for (int k = 0; k < cnt; ++k) w[k] *= coeff;
  return cnt;
}

__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_spread(VertexDev V, R* __restrict__ Fibm, const uint8_t* __restrict__ type,
                                                  LatticeDev L, double f_limit, int use_rep) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V.n) return;
  const int64_t vs = V.vs;
  int64_t idx[8];
  double w[8];
  const int cnt = ibm_support(V.x[i], V.x[vs + i], V.x[2 * vs + i], L, type, idx, w);
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
    t0 = V.frep[i] + f0;
    t1 = V.frep[vs + i] + f1;
    t2 = V.frep[2 * vs + i] + f2;
  }
  const int64_t S = L.stride;
  for (int k = 0; k < cnt; ++k) {
    atomic_add(&Fibm[idx[k]], (R)(t0 * w[k]));
    atomic_add(&Fibm[S + idx[k]], (R)(t1 * w[k]));
    atomic_add(&Fibm[2 * S + idx[k]], (R)(t2 * w[k]));
  }
}

// post-stream populations of node n (phase of the AA pattern as in hc_lbm.hip)
template <typename R>
__device__ __forceinline__ void node_velocity(const R* __restrict__ f, const R* __restrict__ Fcur, const LatticeDev& L,
                                              int phase, int64_t n, R g0, R g1, R g2, R (&u)[3]) {
  const int64_t S = L.stride;
  R fl[19];
  if (phase == 0) {
#pragma unroll
    for (int i = 0; i < 19; ++i) fl[i] = f[i * S + n];
  } else {
    const int z = (int)(n % L.nz);
    const int64_t t = n / L.nz;
    const int y = (int)(t % L.ny), x = (int)(t / L.ny);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      const int X = shift_coord(x, -cx(i), L.nx, L.periodic[0]);
      const int Y = shift_coord(y, -cy(i), L.ny, L.periodic[1]);
      const int Z = shift_coord(z, -cz(i), L.nz, L.periodic[2]);
      fl[i] = (X < 0 || Y < 0 || Z < 0) ? R(0)
                                        : f[opp(i) * S + (int64_t)Z + (int64_t)L.nz * ((int64_t)Y + (int64_t)L.ny * X)];
    }
  }
  R rhoBar, j0, j1, j2;
  rhobar_j(fl, rhoBar, j0, j1, j2);
  const R invRho = R(1) / (rhoBar + R(1));
  u[0] = j0 * invRho + (g0 + Fcur[n]) / R(2);
  u[1] = j1 * invRho + (g1 + Fcur[S + n]) / R(2);
  u[2] = j2 * invRho + (g2 + Fcur[2 * S + n]) / R(2);
}

template <typename R, int PHASE>
__global__ __launch_bounds__(BLOCK) void k_interp_advance(VertexDev V, const R* __restrict__ f, const R* __restrict__ Fcur,
                                                          const uint8_t* __restrict__ type, LatticeDev L, double g0,
                                                          double g1, double g2, int do_interp, int do_advance,
                                                          unsigned long long* wall_hits) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V.n) return;
  const int64_t vs = V.vs;
  double x0 = V.x[i], x1 = V.x[vs + i], x2 = V.x[2 * vs + i];
  double v0, v1, v2;
  if (do_interp) {
    int64_t idx[8];
    double w[8];
    const int cnt = ibm_support(x0, x1, x2, L, type, idx, w);
    v0 = 0.0;
    v1 = 0.0;
    v2 = 0.0;
    for (int k = 0; k < cnt; ++k) {
      R u[3];
      node_velocity<R>(f, Fcur, L, PHASE, idx[k], R(g0), R(g1), R(g2), u);
      v0 += (double)u[0] * w[k];
      v1 += (double)u[1] * w[k];
      v2 += (double)u[2] * w[k];
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
    if (!out && type[wc[2] + (int64_t)L.nz * (wc[1] + (int64_t)L.ny * wc[0])] != HC_NODE_FLUID)
      atomicAdd(wall_hits, 1ull);
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
hipError_t launch_interp_advance(hipStream_t s, const VertexDev& V, const R* f, const R* Fcur, const uint8_t* type,
                                 const LatticeDev& L, int phase, const double g[3], int do_interp, int do_advance,
                                 unsigned long long* wall_hits) {
  if (V.n == 0) return hipSuccess;
  if (phase)
    hipLaunchKernelGGL((k_interp_advance<R, 1>), grid_for(V.n), dim3(BLOCK), 0, s, V, f, Fcur, type, L, g[0], g[1], g[2],
                       do_interp, do_advance, wall_hits);
  else
    hipLaunchKernelGGL((k_interp_advance<R, 0>), grid_for(V.n), dim3(BLOCK), 0, s, V, f, Fcur, type, L, g[0], g[1], g[2],
                       do_interp, do_advance, wall_hits);
  return hipGetLastError();
}

hipError_t launch_add_extra_force(hipStream_t s, const VertexDev& V, const int64_t* ids, const double* forces, int n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_add_extra, grid_for(n), dim3(BLOCK), 0, s, V, ids, forces, n);
  return hipGetLastError();
}

template hipError_t launch_spread<float>(hipStream_t, const VertexDev&, float*, const uint8_t*, const LatticeDev&, double, int);
template hipError_t launch_spread<double>(hipStream_t, const VertexDev&, double*, const uint8_t*, const LatticeDev&, double, int);
template hipError_t launch_interp_advance<float>(hipStream_t, const VertexDev&, const float*, const float*, const uint8_t*,
                                                 const LatticeDev&, int, const double*, int, int, unsigned long long*);
template hipError_t launch_interp_advance<double>(hipStream_t, const VertexDev&, const double*, const double*,
                                                  const uint8_t*, const LatticeDev&, int, const double*, int,
                                                  int, unsigned long long*);

}  // namespace hcg
