// hc_d3q19.h -- D3Q19 lattice constants and per-node device helpers.
//
// Velocity set and ordering are Palabos's ForcedD3Q19Descriptor
// (config/constant_defaults.h:58-61): c_0 = 0, c_1..9 as below, c_{i+9} = -c_i
// (the opposite-index rule i <-> i+9 of patch/palabos.patch:493-520).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <utility>

#include "hc_internal.h"

namespace hcg {

// c_1..c_9 (the c_{i+9} are their negatives); written without recursion so that
// calls with a loop index inline and fold
__host__ __device__ __forceinline__ constexpr int cx9(int i) { return (i == 1 || (i >= 4 && i <= 7)) ? -1 : 0; }
__host__ __device__ __forceinline__ constexpr int cy9(int i) {
  return (i == 2 || i == 4 || i == 8 || i == 9) ? -1 : (i == 5 ? 1 : 0);
}
__host__ __device__ __forceinline__ constexpr int cz9(int i) {
  return (i == 3 || i == 6 || i == 8) ? -1 : ((i == 7 || i == 9) ? 1 : 0);
}
__host__ __device__ __forceinline__ constexpr int cx(int i) { return i == 0 ? 0 : (i <= 9 ? cx9(i) : -cx9(i - 9)); }
__host__ __device__ __forceinline__ constexpr int cy(int i) { return i == 0 ? 0 : (i <= 9 ? cy9(i) : -cy9(i - 9)); }
__host__ __device__ __forceinline__ constexpr int cz(int i) { return i == 0 ? 0 : (i <= 9 ? cz9(i) : -cz9(i - 9)); }
__host__ __device__ __forceinline__ constexpr int opp(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }
template <typename R>
__host__ __device__ __forceinline__ constexpr R wgt(int i) {
  return i == 0 ? R(1) / R(3) : (((i - 1) % 9) < 3 ? R(1) / R(18) : R(1) / R(36));
}

static_assert(cx(1) == -1 && cy(1) == 0 && cz(1) == 0, "c1");
static_assert(cx(5) == -1 && cy(5) == 1 && cz(5) == 0, "c5");
static_assert(cx(7) == -1 && cy(7) == 0 && cz(7) == 1, "c7");
static_assert(cx(9) == 0 && cy(9) == -1 && cz(9) == 1, "c9");
static_assert(cx(14) == 1 && cy(14) == -1 && cz(14) == 0, "c14");
static_assert(cx(18) == 0 && cy(18) == 1 && cz(18) == -1, "c18");

// Node coordinate d + delta along an axis of size n: wrapped when periodic,
// -1 when it leaves a non-periodic domain.
__host__ __device__ inline int shift_coord(int d, int delta, int n, int periodic) {
  int s = d + delta;
  if (s >= 0 && s < n) return s;
  if (!periodic) return -1;
  return s < 0 ? s + n : s - n;
}

// Density deviation and momentum from the x = -1/0/+1 partial sums (Palabos
// momentTemplates for D3Q19), in the oracle's summation order.
template <typename R>
__device__ __forceinline__ void rhobar_j(const R (&f)[19], R& rhoBar, R& j0, R& j1, R& j2) {
  R xm = f[1] + f[4] + f[5] + f[6] + f[7];
  R x0 = f[0] + f[2] + f[3] + f[8] + f[9] + f[11] + f[12] + f[17] + f[18];
  R xp = f[10] + f[13] + f[14] + f[15] + f[16];
  R ym = f[2] + f[4] + f[8] + f[9] + f[14];
  R yp = f[5] + f[11] + f[13] + f[17] + f[18];
  R zm = f[3] + f[6] + f[8] + f[16] + f[18];
  R zp = f[7] + f[9] + f[12] + f[15] + f[17];
  rhoBar = xm + x0 + xp;
  j0 = xp - xm;
  j1 = yp - ym;
  j2 = zp - zm;
}

// Index offsets of the six axis neighbours of node (x, y, z); a neighbour
// outside a non-periodic domain gets INVALID.
constexpr int INVALID = 0x40000000;
struct NbrOff {
  int xm, xp, ym, yp, zm, zp;
};
__device__ __forceinline__ NbrOff neighbour_offsets(int x, int y, int z, const LatticeDev& L) {
  const int nyz = L.ny * L.nz;
  NbrOff o;
  o.xm = x > 0 ? -nyz : (L.periodic[0] ? (L.nx - 1) * nyz : INVALID);
  o.xp = x < L.nx - 1 ? nyz : (L.periodic[0] ? -(L.nx - 1) * nyz : INVALID);
  o.ym = y > 0 ? -L.nz : (L.periodic[1] ? (L.ny - 1) * L.nz : INVALID);
  o.yp = y < L.ny - 1 ? L.nz : (L.periodic[1] ? -(L.ny - 1) * L.nz : INVALID);
  o.zm = z > 0 ? -1 : (L.periodic[2] ? L.nz - 1 : INVALID);
  o.zp = z < L.nz - 1 ? 1 : (L.periodic[2] ? -(L.nz - 1) : INVALID);
  return o;
}
// offset of node n + (DX, DY, DZ); sets ok = false when it leaves the domain
template <int DX, int DY, int DZ, bool CHECK>
__device__ __forceinline__ int nbr_offset(const NbrOff& o, bool& ok) {
  int d = 0;
  ok = true;
  if constexpr (DX < 0) { d += o.xm; if (CHECK) ok &= o.xm != INVALID; }
  if constexpr (DX > 0) { d += o.xp; if (CHECK) ok &= o.xp != INVALID; }
  if constexpr (DY < 0) { d += o.ym; if (CHECK) ok &= o.ym != INVALID; }
  if constexpr (DY > 0) { d += o.yp; if (CHECK) ok &= o.yp != INVALID; }
  if constexpr (DZ < 0) { d += o.zm; if (CHECK) ok &= o.zm != INVALID; }
  if constexpr (DZ > 0) { d += o.zp; if (CHECK) ok &= o.zp != INVALID; }
  return d;
}

// AA-pattern read of the 19 post-stream populations of node n
template <typename R, bool ODD, bool CHECK, int... Is>
__device__ __forceinline__ void aa_load(const R* __restrict__ f, int64_t S, uint32_t n, const NbrOff& o, R (&fl)[19],
                                        std::integer_sequence<int, Is...>) {
  auto one = [&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (!ODD) {
      fl[i] = f[i * S + n];
    } else {
      bool ok;
      const int d = nbr_offset<-cx(i), -cy(i), -cz(i), CHECK>(o, ok);
      fl[i] = ok ? f[opp(i) * S + (uint32_t)(n + d)] : R(0);
    }
  };
  (one(std::integral_constant<int, Is>{}), ...);
}

// AA-pattern write of the 19 post-collision populations of node n
template <typename R, bool ODD, bool CHECK, int... Is>
__device__ __forceinline__ void aa_store(R* __restrict__ f, int64_t S, uint32_t n, const NbrOff& o, const R (&fl)[19],
                                         std::integer_sequence<int, Is...>) {
  auto one = [&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (!ODD) {
      f[opp(i) * S + n] = fl[i];
    } else {
      bool ok;
      const int d = nbr_offset<cx(i), cy(i), cz(i), CHECK>(o, ok);
      if (ok) f[i * S + (uint32_t)(n + d)] = fl[i];
      if constexpr (CHECK) {
        bool in;
        (void)nbr_offset<-cx(i), -cy(i), -cz(i), CHECK>(o, in);
        if (!in) f[i * S + n] = R(0);  // nothing arrives from outside the domain
      }
    }
  };
  (one(std::integral_constant<int, Is>{}), ...);
}

}  // namespace hcg
