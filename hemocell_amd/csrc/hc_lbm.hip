// hc_lbm.hip -- D3Q19 Guo-forced BGK collide-and-stream on MI355X.
//
// Replaces Palabos's MultiBlockLattice3D::collideAndStream() with
// GuoExternalForceBGKdynamics (core/hemoCell.cpp:317, :455-469), BounceBack
// (helper/genericFunctions.cpp:144,154) and the local regularized velocity
// boundary (helper/hemocellInit.hh:72-77), plus the per-step force reset
// setExternalVector(..., 0) (core/hemoCell.cpp:369-371).
//
// One in-place kernel per step, AA access pattern over structure-of-arrays
// populations f[i*stride + n] (n = z + nz*(y + ny*x), z fastest so a wavefront
// reads 64 consecutive nodes of one population array: fully coalesced):
//   even step: read f[i][n], collide, write f*_i into f[opp(i)][n];
//   odd step : read f_i(n) = f[opp(i)][n - c_i], collide, write f*_i into f[i][n + c_i].
// Each thread reads and writes the same 19 slots, so no second lattice and no
// race.  After an odd step the array holds natural post-stream populations; after
// an even step the post-stream value of (n, i) is f[opp(i)][n - c_i].  Values
// that would arrive from outside a non-periodic domain are 0 (shifted storage).
//
// The IBM force is double-buffered: the kernel reads the buffer the spread just
// filled and zeroes the other one, which the next step's spread fills -- the
// reference's full-lattice force reset folded into this pass.
//
// The Guo-BGK update is evaluated in an FMA-friendly form that is algebraically
// the Palabos one: with u = j/rho + F/2, mu = rho (1 - omega/2),
//   f_i' = (1-omega) f_i + t_i [A + B cu + C cu^2 + D cF + E cu cF],
//   cu = c_i.u, cF = c_i.F, A = omega (rhoBar - 3/2 rho u^2) - 3 mu u.F,
//   B = 3 omega rho, C = 9/2 omega rho, D = 3 mu, E = 9 mu,
// split into even and odd parts over each opposite pair (i, i+9).
#include <utility>

#include "hc_d3q19.h"
#include "hc_launch.h"

namespace hcg {
namespace {

constexpr int BLOCK = 256;

// c_I . (a0, a1, a2) with the zero components of c_I dropped at compile time.
template <int I, typename R>
__device__ __forceinline__ R cdot(R a0, R a1, R a2) {
  constexpr int X = cx(I), Y = cy(I), Z = cz(I);
  R s = R(0);
  bool first = true;
  if constexpr (X != 0) { s = X > 0 ? a0 : -a0; first = false; }
  if constexpr (Y != 0) { s = first ? (Y > 0 ? a1 : -a1) : (Y > 0 ? s + a1 : s - a1); first = false; }
  if constexpr (Z != 0) { s = first ? (Z > 0 ? a2 : -a2) : (Z > 0 ? s + a2 : s - a2); }
  return s;
}

template <typename R>
struct Coef {
  R A, B, C, D, E, om1, u0, u1, u2, F0, F1, F2;
};

template <int I, typename R>
__device__ __forceinline__ void relax_pair(R (&f)[19], const Coef<R>& k) {
  const R cu = cdot<I>(k.u0, k.u1, k.u2);
  const R cF = cdot<I>(k.F0, k.F1, k.F2);
  const R sym = k.A + cu * (k.C * cu + k.E * cF);
  const R asym = k.B * cu + k.D * cF;
  constexpr R w = wgt<R>(I);
  f[I] = k.om1 * f[I] + w * (sym + asym);
  f[I + 9] = k.om1 * f[I + 9] + w * (sym - asym);
}

template <typename R, int... Is>
__device__ __forceinline__ void relax_all(R (&f)[19], const Coef<R>& k, std::integer_sequence<int, Is...>) {
  (relax_pair<Is + 1>(f, k), ...);
}

template <typename R>
__device__ __forceinline__ void collide_guo(R (&f)[19], R F0, R F1, R F2, R omega) {
  R rhoBar, j0, j1, j2;
  rhobar_j(f, rhoBar, j0, j1, j2);
  const R rho = rhoBar + R(1);
  const R invRho = R(1) / rho;
  Coef<R> k;
  k.u0 = j0 * invRho + F0 * R(0.5);
  k.u1 = j1 * invRho + F1 * R(0.5);
  k.u2 = j2 * invRho + F2 * R(0.5);
  k.F0 = F0;
  k.F1 = F1;
  k.F2 = F2;
  const R mu = rho * (R(1) - omega * R(0.5));
  const R usq = k.u0 * k.u0 + k.u1 * k.u1 + k.u2 * k.u2;
  const R uF = k.u0 * F0 + k.u1 * F1 + k.u2 * F2;
  k.A = omega * (rhoBar - R(1.5) * rho * usq) - R(3) * mu * uF;
  k.B = R(3) * omega * rho;
  k.C = R(4.5) * omega * rho;
  k.D = R(3) * mu;
  k.E = R(9) * mu;
  k.om1 = R(1) - omega;
  f[0] = k.om1 * f[0] + wgt<R>(0) * k.A;
  relax_all(f, k, std::make_integer_sequence<int, 9>{});
}

// Second-order equilibrium in shifted form (used by boundary reconstruction and init).
template <typename R>
__device__ __forceinline__ R feq(int i, R rhoBar, R invRho, R j0, R j1, R j2, R jSqr) {
  R c_j = R(cx(i)) * j0;
  c_j += R(cy(i)) * j1;
  c_j += R(cz(i)) * j2;
  return wgt<R>(i) * (rhoBar + R(3) * c_j + R(3) / R(2) * invRho * (R(3) * c_j * c_j - jSqr));
}

// Local regularized velocity boundary on a domain face: rebuild all populations
// as f_eq(rhoBar, rho u_w) + f_neq(Pi_neq), Pi_neq from the known populations
// (on-wall ones and those arriving from the interior, the latter counted twice:
// bounce-back of the non-equilibrium part).
__device__ __forceinline__ constexpr int c_axis(int i, int a) { return a == 0 ? cx(i) : (a == 1 ? cy(i) : cz(i)); }

template <typename R, int AXIS, int ORIENT>
__device__ __forceinline__ void velocity_boundary(R (&f)[19], const double uw0, const double uw1, const double uw2) {
  // the face is a compile-time (axis, orientation) pair, so every lattice table
  // lookup below folds to a constant
  R rhoOnWall = R(0), rhoNormal = R(0);
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    const int c = c_axis(i, AXIS);
    if (c == 0) rhoOnWall += f[i];
    else if (c == ORIENT) rhoNormal += f[i];
  }
  const R w0 = R(uw0), w1 = R(uw1), w2 = R(uw2);
  const R velNormal = R(ORIENT) * (AXIS == 0 ? w0 : (AXIS == 1 ? w1 : w2));
  const R rhoBar = (R(2) * rhoNormal + rhoOnWall - velNormal) / (R(1) + velNormal);
  const R rho = rhoBar + R(1);
  const R j0 = rho * w0, j1 = rho * w1, j2 = rho * w2;
  const R jSqr = j0 * j0 + j1 * j1 + j2 * j2;
  const R invRho = R(1) / rho;
  R fNeq[19];
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    const int c = c_axis(i, AXIS);
    fNeq[i] = (c == 0 || c == ORIENT) ? f[i] - feq<R>(i, rhoBar, invRho, j0, j1, j2, jSqr) : R(0);
  }
  // Pi_neq (xx, xy, xz, yy, yz, zz): on-wall populations once, incoming ones twice
  R Pi[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    const int c = c_axis(i, AXIS);
    if (c != 0 && c != ORIENT) continue;
    const R m = c == 0 ? fNeq[i] : R(2) * fNeq[i];
    Pi[0] += R(cx(i) * cx(i)) * m;
    Pi[1] += R(cx(i) * cy(i)) * m;
    Pi[2] += R(cx(i) * cz(i)) * m;
    Pi[3] += R(cy(i) * cy(i)) * m;
    Pi[4] += R(cy(i) * cz(i)) * m;
    Pi[5] += R(cz(i) * cz(i)) * m;
  }
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    const R qxx = R(cx(i) * cx(i)) - R(1) / R(3), qyy = R(cy(i) * cy(i)) - R(1) / R(3),
            qzz = R(cz(i) * cz(i)) - R(1) / R(3);
    R fn = qxx * Pi[0] + R(2 * cx(i) * cy(i)) * Pi[1] + R(2 * cx(i) * cz(i)) * Pi[2] + qyy * Pi[3] +
           R(2 * cy(i) * cz(i)) * Pi[4] + qzz * Pi[5];
    fn *= wgt<R>(i) * R(3) * R(3) / R(2);
    f[i] = feq<R>(i, rhoBar, invRho, j0, j1, j2, jSqr) + fn;
  }
}

__device__ __forceinline__ void coords(int64_t n, const LatticeDev& L, int& x, int& y, int& z) {
  z = (int)(n % L.nz);
  int64_t t = n / L.nz;
  y = (int)(t % L.ny);
  x = (int)(t / L.ny);
}

// regularized reconstruction of a velocity node of domain face `face`
template <typename R>
__device__ __forceinline__ void velocity_node(R (&fl)[19], int face, const CollideArgs& a) {
  // one compile-time instance per domain face (no runtime-indexed lattice tables)
  switch (face) {
    case 0: velocity_boundary<R, 0, -1>(fl, a.wall_u[0][0], a.wall_u[0][1], a.wall_u[0][2]); break;
    case 1: velocity_boundary<R, 0, 1>(fl, a.wall_u[1][0], a.wall_u[1][1], a.wall_u[1][2]); break;
    case 2: velocity_boundary<R, 1, -1>(fl, a.wall_u[2][0], a.wall_u[2][1], a.wall_u[2][2]); break;
    case 3: velocity_boundary<R, 1, 1>(fl, a.wall_u[3][0], a.wall_u[3][1], a.wall_u[3][2]); break;
    case 4: velocity_boundary<R, 2, -1>(fl, a.wall_u[4][0], a.wall_u[4][1], a.wall_u[4][2]); break;
    default: velocity_boundary<R, 2, 1>(fl, a.wall_u[5][0], a.wall_u[5][1], a.wall_u[5][2]); break;
  }
}

// GENERAL: node-type mask present and/or some axis non-periodic.  VEL: the
// velocity-boundary nodes are reconstructed in place (each node only touches its
// own 19 slots); without VEL they are skipped and k_velocity_nodes updates them.
// VISC: the relaxation rate of a node comes from its interior-viscosity class.
// relaxation rate of interior-viscosity class vc (1..15): selected with
// constant indices, so the table is read from the kernel arguments without a
// private copy for a run-time index
__device__ __forceinline__ double class_omega(const CollideArgs& a, unsigned vc) {
  double o = a.omega;
#pragma unroll
  for (unsigned k = 1; k < 16; ++k)
    if (vc == k) o = a.omega_cls[k];
  return o;
}

template <typename R, bool ODD, bool GENERAL, bool VEL, bool VISC>
__device__ __forceinline__ void collide_node(R* __restrict__ f, const R* __restrict__ Fcur, R* __restrict__ Fclr,
                                             const uint8_t* __restrict__ type, const LatticeDev& L,
                                             const CollideArgs& a, const uint32_t n) {
  if (n >= (uint32_t)L.N) return;
  const int64_t S = L.stride;
  const uint8_t ty = GENERAL && type ? type[n] : (uint8_t)HC_NODE_FLUID;
  if (GENERAL && !VEL && ty >= HC_NODE_VELOCITY) return;
  const int z = (int)(n % (uint32_t)L.nz);
  const uint32_t t = n / (uint32_t)L.nz;
  const int y = (int)(t % (uint32_t)L.ny);
  const int x = (int)(t / (uint32_t)L.ny);
  const NbrOff o = neighbour_offsets(x, y, z, L);
  R fl[19];
  aa_load<R, ODD, GENERAL>(f, S, n, o, fl, std::make_integer_sequence<int, 19>{});
  if (GENERAL && ty == HC_NODE_BOUNCEBACK) {
#pragma unroll
    for (int i = 1; i <= 9; ++i) {
      R tmp = fl[i];
      fl[i] = fl[i + 9];
      fl[i + 9] = tmp;
    }
  } else {
    if (VEL && ty >= HC_NODE_VELOCITY) velocity_node<R>(fl, ty - HC_NODE_VELOCITY, a);
    const R F0 = R(a.g[0]) + Fcur[n], F1 = R(a.g[1]) + Fcur[S + n], F2 = R(a.g[2]) + Fcur[2 * S + n];
    R om = R(a.omega);
    if (VISC) {
      const unsigned vc = a.visc[n];
      if (vc) om = R(class_omega(a, vc));
    }
    collide_guo<R>(fl, F0, F1, F2, om);
  }
  Fclr[n] = R(0);
  Fclr[S + n] = R(0);
  Fclr[2 * S + n] = R(0);
  aa_store<R, ODD, GENERAL>(f, S, n, o, fl, std::make_integer_sequence<int, 19>{});
}

template <typename R, bool ODD, bool GENERAL, bool VEL = false, bool VISC = false>
__global__ __launch_bounds__(BLOCK) void k_collide_stream(R* __restrict__ f, const R* __restrict__ Fcur,
                                                          R* __restrict__ Fclr, const uint8_t* __restrict__ type,
                                                          LatticeDev L, CollideArgs a) {
  collide_node<R, ODD, GENERAL, VEL, VISC>(f, Fcur, Fclr, type, L, a, blockIdx.x * BLOCK + threadIdx.x);
}

// the case of block b in a batched launch: the last case whose first block is <= b
template <typename C>
__device__ __forceinline__ const C& case_of_block(const C* __restrict__ cs, int ncase, uint32_t b) {
  int lo = 0, hi = ncase - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return cs[lo];
}

// Several independent lattices in one launch (small cases served together):
// each case owns a contiguous range of blocks starting at block0.
template <typename R, bool ODD, bool GENERAL, bool VEL>
__global__ __launch_bounds__(BLOCK) void k_collide_multi(const LatCase* __restrict__ cs, int ncase) {
  const LatCase& c = case_of_block(cs, ncase, blockIdx.x);
  collide_node<R, ODD, GENERAL, VEL, false>((R*)c.f, (const R*)c.Fcur, (R*)c.Fclr, c.type, c.L, c.a,
                                            (blockIdx.x - c.block0) * BLOCK + threadIdx.x);
}

// Velocity-boundary nodes (listed in `nodes`): regularized reconstruction, then
// the Guo-BGK collision, with the same AA read/write as the bulk kernel.
template <typename R, bool ODD>
__global__ __launch_bounds__(64) void k_velocity_nodes(R* __restrict__ f, const R* __restrict__ Fcur,
                                                       R* __restrict__ Fclr, const uint8_t* __restrict__ type,
                                                       const uint32_t* __restrict__ nodes, int count, LatticeDev L,
                                                       CollideArgs a) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= count) return;
  const uint32_t n = nodes[k];
  const int64_t S = L.stride;
  const int z = (int)(n % (uint32_t)L.nz);
  const uint32_t t = n / (uint32_t)L.nz;
  const int y = (int)(t % (uint32_t)L.ny);
  const int x = (int)(t / (uint32_t)L.ny);
  const NbrOff o = neighbour_offsets(x, y, z, L);
  R fl[19];
  aa_load<R, ODD, true>(f, S, n, o, fl, std::make_integer_sequence<int, 19>{});
  velocity_node<R>(fl, type[n] - HC_NODE_VELOCITY, a);
  const R F0 = R(a.g[0]) + Fcur[n], F1 = R(a.g[1]) + Fcur[S + n], F2 = R(a.g[2]) + Fcur[2 * S + n];
  collide_guo<R>(fl, F0, F1, F2, R(a.omega));
  Fclr[n] = R(0);
  Fclr[S + n] = R(0);
  Fclr[2 * S + n] = R(0);
  aa_store<R, ODD, true>(f, S, n, o, fl, std::make_integer_sequence<int, 19>{});
}

template <typename R>
__global__ void k_init_eq(R* f, const uint8_t* type, LatticeDev L, double rho, double u0, double u1, double u2) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (n >= L.N) return;
  const bool bb = type && type[n] == HC_NODE_BOUNCEBACK;
  const R rhoBar = R(rho - 1.0), invRho = R(1.0 / rho);
  const R j0 = R(rho * u0), j1 = R(rho * u1), j2 = R(rho * u2);
  const R jSqr = j0 * j0 + j1 * j1 + j2 * j2;
#pragma unroll
  for (int i = 0; i < 19; ++i) f[i * L.stride + n] = bb ? R(0) : feq<R>(i, rhoBar, invRho, j0, j1, j2, jSqr);
}

// post-stream population (n, i) in the given phase of the AA pattern
template <typename R>
__device__ __forceinline__ R post_stream(const R* f, const LatticeDev& L, int phase, int x, int y, int z, int i,
                                         int64_t n) {
  if (phase == 0) return f[i * L.stride + n];
  int X = shift_coord(x, -cx(i), L.nx, L.periodic[0]);
  int Y = shift_coord(y, -cy(i), L.ny, L.periodic[1]);
  int Z = shift_coord(z, -cz(i), L.nz, L.periodic[2]);
  if (X < 0 || Y < 0 || Z < 0) return R(0);
  return f[opp(i) * L.stride + (int64_t)Z + (int64_t)L.nz * ((int64_t)Y + (int64_t)L.ny * X)];
}

template <typename R>
__global__ void k_natural(const R* f, double* out, LatticeDev L, int phase) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (n >= L.N) return;
  int x, y, z;
  coords(n, L, x, y, z);
  for (int i = 0; i < 19; ++i) out[i * L.N + n] = (double)post_stream(f, L, phase, x, y, z, i, n);
}

template <typename R>
__global__ void k_store(R* f, const double* in, LatticeDev L) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (n >= L.N) return;
  for (int i = 0; i < 19; ++i) f[i * L.stride + n] = (R)in[i * L.N + n];
}

// Density and velocity of node n for output; a velocity-boundary node reports
// its imposed velocity and the density of the mass balance over its known
// populations.
template <typename R>
__device__ void node_moments(const R* f, const R* Fcur, const uint8_t* type, const LatticeDev& L, int phase,
                             const CollideArgs& a, int64_t n, double& rho, double& u0, double& u1, double& u2) {
  int x, y, z;
  coords(n, L, x, y, z);
  R fl[19];
  for (int i = 0; i < 19; ++i) fl[i] = post_stream(f, L, phase, x, y, z, i, n);
  R rhoBar, j0, j1, j2;
  rhobar_j(fl, rhoBar, j0, j1, j2);
  const R invRho = R(1) / (rhoBar + R(1));
  const int64_t S = L.stride;
  const R F0 = Fcur ? Fcur[n] : R(0), F1 = Fcur ? Fcur[S + n] : R(0), F2 = Fcur ? Fcur[2 * S + n] : R(0);
  u0 = (double)(j0 * invRho + (R(a.g[0]) + F0) / R(2));
  u1 = (double)(j1 * invRho + (R(a.g[1]) + F1) / R(2));
  u2 = (double)(j2 * invRho + (R(a.g[2]) + F2) / R(2));
  double rb = (double)rhoBar;
  if (type && type[n] >= HC_NODE_VELOCITY) {
    const int face = type[n] - HC_NODE_VELOCITY, axis = face >> 1, orient = (face & 1) ? 1 : -1;
    double on = 0.0, nrm = 0.0;
    for (int i = 0; i < 19; ++i) {
      const int c = axis == 0 ? cx(i) : (axis == 1 ? cy(i) : cz(i));
      if (c == 0) on += (double)fl[i];
      else if (c == orient) nrm += (double)fl[i];
    }
    const double vn = orient * a.wall_u[face][axis];
    rb = (2.0 * nrm + on - vn) / (1.0 + vn);
    u0 = a.wall_u[face][0];
    u1 = a.wall_u[face][1];
    u2 = a.wall_u[face][2];
  }
  rho = rb + 1.0;
}

template <typename R>
__global__ void k_moments(const R* f, const R* Fcur, const uint8_t* type, LatticeDev L, int phase, CollideArgs a,
                          double* rho, double* u) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (n >= L.N) return;
  double r, u0, u1, u2;
  node_moments<R>(f, Fcur, type, L, phase, a, n, r, u0, u1, u2);
  if (rho) rho[n] = r;
  if (u) {
    u[3 * n + 0] = u0;
    u[3 * n + 1] = u1;
    u[3 * n + 2] = u2;
  }
}

// FluidInfo statistics (helper/fluidInfo.cpp:27-102): minimum, maximum and sum of
// a per-node magnitude over the nodes that are not bounce-back, and their count,
// one partial per block (reduced on the host).  what 0: |u| as k_moments gives
// it; what 1: |F| of the force field Fstat.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_fluid_stats(const R* f, const R* Fcur, const uint8_t* type, LatticeDev L,
                                                       int phase, CollideArgs a, const R* Fstat, int what,
                                                       double* partials) {
  __shared__ double smin[BLOCK], smax[BLOCK], ssum[BLOCK], scnt[BLOCK];
  double mn = INFINITY, mx = -INFINITY, sum = 0.0, cnt = 0.0;
  const int64_t S = L.stride;
  for (int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x; n < L.N; n += (int64_t)gridDim.x * BLOCK) {
    if (type && type[n] == HC_NODE_BOUNCEBACK) continue;
    double m;
    if (what == 0) {
      double r, u0, u1, u2;
      node_moments<R>(f, Fcur, type, L, phase, a, n, r, u0, u1, u2);
      m = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    } else {
      const double F0 = (double)Fstat[n], F1 = (double)Fstat[S + n], F2 = (double)Fstat[2 * S + n];
      m = sqrt(F0 * F0 + F1 * F1 + F2 * F2);
    }
    mn = fmin(mn, m);
    mx = fmax(mx, m);
    sum += m;
    cnt += 1.0;
  }
  const int t = threadIdx.x;
  smin[t] = mn;
  smax[t] = mx;
  ssum[t] = sum;
  scnt[t] = cnt;
  __syncthreads();
  for (int h = BLOCK / 2; h > 0; h >>= 1) {
    if (t < h) {
      smin[t] = fmin(smin[t], smin[t + h]);
      smax[t] = fmax(smax[t], smax[t + h]);
      ssum[t] += ssum[t + h];
      scnt[t] += scnt[t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    double* p = partials + 4 * blockIdx.x;
    p[0] = smin[0];
    p[1] = smax[0];
    p[2] = ssum[0];
    p[3] = scnt[0];
  }
}

template <typename R>
__global__ void k_zero(R* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * BLOCK;
  for (; i < n; i += step) p[i] = R(0);
}

template <typename R>
__global__ void k_to_double(const R* in, double* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * BLOCK;
  for (; i < n; i += step) out[i] = (double)in[i];
}

// Slot that holds the post-stream population (n, i) in the given phase of the
// AA pattern (the one post_stream reads); -1 outside a non-periodic domain.
__device__ __forceinline__ int64_t post_stream_slot(const LatticeDev& L, int phase, int x, int y, int z, int i,
                                                    int64_t n) {
  if (phase == 0) return i * L.stride + n;
  const int X = shift_coord(x, -cx(i), L.nx, L.periodic[0]);
  const int Y = shift_coord(y, -cy(i), L.ny, L.periodic[1]);
  const int Z = shift_coord(z, -cz(i), L.nz, L.periodic[2]);
  if (X < 0 || Y < 0 || Z < 0) return -1;
  return opp(i) * L.stride + (int64_t)Z + (int64_t)L.nz * ((int64_t)Y + (int64_t)L.ny * X);
}

__device__ __forceinline__ int64_t mod_nonneg(int64_t a, int64_t b) { return (a % b + b) % b; }

// Lees-Edwards planes, first half (helper/leesEdwardsBC.h:101-215): for every
// node of the top (z = nz-1, side 0) and bottom (z = 0, side 1) plane, a copy
// of its post-stream populations relaxed once more by BGK, with the node's own
// omega, towards the equilibrium of its rhoBar and the momentum (u_side, 0, 0);
// the five populations that entered across the plane are replaced by the
// linear interpolation, at x + d (top) or x - d (bottom), of the plane's
// post-stream populations, the diagonal pair exchanged.  Wall nodes are
// copied unchanged.  out: 2 nx ny x 19, written back by k_lees_edwards_set.
template <typename R>
__global__ void k_lees_edwards_get(const R* __restrict__ f, const uint8_t* __restrict__ type, LatticeDev L,
                                   int phase, CollideArgs a, double d, double g, double u_top, double u_bottom,
                                   R* __restrict__ out) {
  const int64_t P = (int64_t)L.nx * L.ny;
  const int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k >= 2 * P) return;
  const int side = (int)(k / P);
  const int X = (int)((k % P) / L.ny), Y = (int)((k % P) % L.ny), Z = side == 0 ? L.nz - 1 : 0;
  const int64_t n = (int64_t)Z + (int64_t)L.nz * ((int64_t)Y + (int64_t)L.ny * X);
  R fl[19];
#pragma unroll
  for (int i = 0; i < 19; ++i) fl[i] = post_stream(f, L, phase, X, Y, Z, i, n);
  R* o = out + 19 * k;
  if (type && type[n] != HC_NODE_FLUID) {
#pragma unroll
    for (int i = 0; i < 19; ++i) o[i] = fl[i];
    return;
  }
  R rhoBar, j0, j1, j2;
  rhobar_j(fl, rhoBar, j0, j1, j2);
  const R uw = R(side == 0 ? u_top : u_bottom);
  const R jSqr = uw * uw;
  const R invRho = R(1) / (rhoBar + R(1));
  R om = R(a.omega);
  if (a.visc && a.visc[n]) om = R(class_omega(a, a.visc[n]));
#pragma unroll
  for (int i = 0; i < 19; ++i) o[i] = fl[i] * (R(1) - om) + om * feq<R>(i, rhoBar, invRho, uw, R(0), R(0), jSqr);
  int64_t s1, s2;
  if (side == 0) {
    s1 = mod_nonneg((int64_t)ceil(d + X), L.nx);
    s2 = mod_nonneg((int64_t)floor(d + X), L.nx);
  } else {
    s1 = mod_nonneg((int64_t)floor(-d + X), L.nx);
    s2 = mod_nonneg((int64_t)ceil(-d + X), L.nx);
  }
  const int64_t n1 = (int64_t)Z + (int64_t)L.nz * ((int64_t)Y + (int64_t)L.ny * s1);
  const int64_t n2 = (int64_t)Z + (int64_t)L.nz * ((int64_t)Y + (int64_t)L.ny * s2);
  const R gr = R(g), hr = R(1) - R(g);
  auto lerp = [&](int i) {
    return gr * post_stream(f, L, phase, (int)s1, Y, Z, i, n1) + hr * post_stream(f, L, phase, (int)s2, Y, Z, i, n2);
  };
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

// second half (helper/leesEdwardsBC.h:275-298): the copies back into the planes
template <typename R>
__global__ void k_lees_edwards_set(R* __restrict__ f, LatticeDev L, int phase, const R* __restrict__ out) {
  const int64_t P = (int64_t)L.nx * L.ny;
  const int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k >= 2 * P) return;
  const int side = (int)(k / P);
  const int X = (int)((k % P) / L.ny), Y = (int)((k % P) % L.ny), Z = side == 0 ? L.nz - 1 : 0;
  const int64_t n = (int64_t)Z + (int64_t)L.nz * ((int64_t)Y + (int64_t)L.ny * X);
  for (int i = 0; i < 19; ++i) {
    const int64_t slot = post_stream_slot(L, phase, X, Y, Z, i, n);
    if (slot >= 0) f[slot] = out[19 * k + i];
  }
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }
inline dim3 grid_capped(int64_t n) {
  int64_t b = (n + BLOCK - 1) / BLOCK;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

}  // namespace

template <typename R, bool VISC>
void launch_bulk(hipStream_t s, dim3 g, R* f, const R* Fcur, R* Fclr, const uint8_t* type, const LatticeDev& L,
                 const CollideArgs& a, int odd, bool general, bool vel) {
  if (general && vel) {
    // velocity nodes reconstructed inside the bulk kernel: one launch per step
    if (odd) hipLaunchKernelGGL((k_collide_stream<R, true, true, true, VISC>), g, dim3(BLOCK), 0, s, f, Fcur, Fclr, type, L, a);
    else hipLaunchKernelGGL((k_collide_stream<R, false, true, true, VISC>), g, dim3(BLOCK), 0, s, f, Fcur, Fclr, type, L, a);
  } else if (general) {
    if (odd) hipLaunchKernelGGL((k_collide_stream<R, true, true, false, VISC>), g, dim3(BLOCK), 0, s, f, Fcur, Fclr, type, L, a);
    else hipLaunchKernelGGL((k_collide_stream<R, false, true, false, VISC>), g, dim3(BLOCK), 0, s, f, Fcur, Fclr, type, L, a);
  } else {
    if (odd) hipLaunchKernelGGL((k_collide_stream<R, true, false, false, VISC>), g, dim3(BLOCK), 0, s, f, Fcur, Fclr, type, L, a);
    else hipLaunchKernelGGL((k_collide_stream<R, false, false, false, VISC>), g, dim3(BLOCK), 0, s, f, Fcur, Fclr, type, L, a);
  }
}

template <typename R>
hipError_t launch_collide_stream(hipStream_t s, R* f, const R* Fcur, R* Fclr, const uint8_t* type,
                                 const LatticeDev& L, const CollideArgs& a, int odd, bool general,
                                 const uint32_t* vel_nodes, int n_vel, bool fuse_vel) {
  dim3 g = grid_for(L.N);
  const bool vel = general && n_vel > 0 && fuse_vel;
  if (a.visc) launch_bulk<R, true>(s, g, f, Fcur, Fclr, type, L, a, odd, general, vel);
  else launch_bulk<R, false>(s, g, f, Fcur, Fclr, type, L, a, odd, general, vel);
  if (general && !vel && n_vel > 0) {
    // velocity nodes never hold an interior-viscosity class
    dim3 gv((unsigned)((n_vel + 63) / 64));
    if (odd)
      hipLaunchKernelGGL((k_velocity_nodes<R, true>), gv, dim3(64), 0, s, f, Fcur, Fclr, type, vel_nodes, n_vel, L, a);
    else
      hipLaunchKernelGGL((k_velocity_nodes<R, false>), gv, dim3(64), 0, s, f, Fcur, Fclr, type, vel_nodes, n_vel, L, a);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_init_equilibrium(hipStream_t s, R* f, const uint8_t* type, const LatticeDev& L, double rho,
                                   const double u[3]) {
  hipLaunchKernelGGL((k_init_eq<R>), grid_for(L.N), dim3(BLOCK), 0, s, f, type, L, rho, u[0], u[1], u[2]);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_natural_populations(hipStream_t s, const R* f, double* out, const LatticeDev& L, int phase) {
  hipLaunchKernelGGL((k_natural<R>), grid_for(L.N), dim3(BLOCK), 0, s, f, out, L, phase);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_store_populations(hipStream_t s, R* f, const double* in, const LatticeDev& L) {
  hipLaunchKernelGGL((k_store<R>), grid_for(L.N), dim3(BLOCK), 0, s, f, in, L);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_fluid_moments(hipStream_t s, const R* f, const R* Fcur, const uint8_t* type, const LatticeDev& L,
                                int phase, const CollideArgs& a, double* rho, double* u) {
  hipLaunchKernelGGL((k_moments<R>), grid_for(L.N), dim3(BLOCK), 0, s, f, Fcur, type, L, phase, a, rho, u);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_fluid_stats(hipStream_t s, const R* f, const R* Fcur, const uint8_t* type, const LatticeDev& L,
                              int phase, const CollideArgs& a, const R* Fstat, int what, double* partials,
                              int nblocks) {
  hipLaunchKernelGGL((k_fluid_stats<R>), dim3((unsigned)nblocks), dim3(BLOCK), 0, s, f, Fcur, type, L, phase, a, Fstat,
                     what, partials);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_fill_zero(hipStream_t s, R* p, int64_t n) {
  hipLaunchKernelGGL((k_zero<R>), grid_capped(n), dim3(BLOCK), 0, s, p, n);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_to_double(hipStream_t s, const R* in, double* out, int64_t n) {
  hipLaunchKernelGGL((k_to_double<R>), grid_capped(n), dim3(BLOCK), 0, s, in, out, n);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_collide_multi(hipStream_t s, const LatCase* cases, int ncase, uint32_t blocks, int odd,
                                bool general, bool vel) {
  if (ncase <= 0 || blocks == 0) return hipSuccess;
  const dim3 g(blocks), b(BLOCK);
  if (general && vel) {
    if (odd) hipLaunchKernelGGL((k_collide_multi<R, true, true, true>), g, b, 0, s, cases, ncase);
    else hipLaunchKernelGGL((k_collide_multi<R, false, true, true>), g, b, 0, s, cases, ncase);
  } else if (general) {
    if (odd) hipLaunchKernelGGL((k_collide_multi<R, true, true, false>), g, b, 0, s, cases, ncase);
    else hipLaunchKernelGGL((k_collide_multi<R, false, true, false>), g, b, 0, s, cases, ncase);
  } else {
    if (odd) hipLaunchKernelGGL((k_collide_multi<R, true, false, false>), g, b, 0, s, cases, ncase);
    else hipLaunchKernelGGL((k_collide_multi<R, false, false, false>), g, b, 0, s, cases, ncase);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_lees_edwards(hipStream_t s, R* f, const uint8_t* type, const LatticeDev& L, int phase,
                               const CollideArgs& a, double d, double u_top, double u_bottom, R* scratch) {
  const int64_t n = 2 * (int64_t)L.nx * L.ny;
  const double g = std::fmod(d, 1.0);
  hipLaunchKernelGGL((k_lees_edwards_get<R>), grid_for(n), dim3(BLOCK), 0, s, f, type, L, phase, a, d, g, u_top,
                     u_bottom, scratch);
  hipLaunchKernelGGL((k_lees_edwards_set<R>), grid_for(n), dim3(BLOCK), 0, s, f, L, phase, scratch);
  return hipGetLastError();
}

#define HC_INSTANTIATE(R)                                                                                      \
  template hipError_t launch_collide_stream<R>(hipStream_t, R*, const R*, R*, const uint8_t*, const LatticeDev&, \
                                               const CollideArgs&, int, bool, const uint32_t*, int, bool);    \
  template hipError_t launch_init_equilibrium<R>(hipStream_t, R*, const uint8_t*, const LatticeDev&, double,    \
                                                 const double*);                                              \
  template hipError_t launch_natural_populations<R>(hipStream_t, const R*, double*, const LatticeDev&, int);    \
  template hipError_t launch_store_populations<R>(hipStream_t, R*, const double*, const LatticeDev&);           \
  template hipError_t launch_fluid_moments<R>(hipStream_t, const R*, const R*, const uint8_t*, const LatticeDev&,  \
                                              int, const CollideArgs&, double*, double*);                     \
  template hipError_t launch_fluid_stats<R>(hipStream_t, const R*, const R*, const uint8_t*, const LatticeDev&,    \
                                            int, const CollideArgs&, const R*, int, double*, int);             \
  template hipError_t launch_fill_zero<R>(hipStream_t, R*, int64_t);                                            \
  template hipError_t launch_to_double<R>(hipStream_t, const R*, double*, int64_t);                             \
  template hipError_t launch_lees_edwards<R>(hipStream_t, R*, const uint8_t*, const LatticeDev&, int,             \
                                             const CollideArgs&, double, double, double, R*);                  \
  template hipError_t launch_collide_multi<R>(hipStream_t, const LatCase*, int, uint32_t, int, bool, bool);
HC_INSTANTIATE(float)
HC_INSTANTIATE(double)
#undef HC_INSTANTIATE

}  // namespace hcg
