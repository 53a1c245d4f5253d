// hc_ibm_dev.h -- device helpers of the immersed-boundary coupling, shared by the
// IBM kernels and the fused per-cell step kernel.
#pragma once
#include <cmath>
#include <type_traits>

#include "hc_d3q19.h"
#include "hc_launch.h"

namespace hcg {
namespace {

__device__ __forceinline__ int64_t wrap_mod(int64_t c, int n) {
  int64_t m = c % n;
  return m < 0 ? m + n : m;
}

// phi2 support of position p.  phi2(d) = max(1 - |d|, 0) is non-zero only on
// the nodes floor(p) and floor(p) + 1 of each axis, so the reference's 3x3x3
// candidates around the nearest node reduce to this 2x2x2 block, visited in the
// same x -> y -> z order.  Each of the 8 slots keeps its weight (0 when the
// node lies outside a non-periodic domain, has zero weight or is a wall node);
// adding the zero weights to the total changes nothing, so the normalised
// weights are those of immersedBoundaryMethod.h:62-138.  Fixed slots keep
// everything in registers.  W = double evaluates the weights in FP64 as the
// reference does; W = float (the FP32-shifted form's cell kernels) takes the
// fractional offsets p - floor(p) in FP64 and the weights from them in FP32,
// where 1 - d and d are the two hat values of an axis.
template <typename W>
struct SupportT {
  uint32_t idx[8];
  W w[8];
  int lo[3];  // unwrapped coordinates of the slot-0 node
};
using Support = SupportT<double>;

__device__ __forceinline__ int64_t wrap_far(int64_t c, int n) {
  const double q = floor((double)c / (double)n);
  return c - (int64_t)q * n;
}

template <typename W>
__device__ __forceinline__ void ibm_support(double p0, double p1, double p2, const LatticeDev& L,
                                            const uint8_t* __restrict__ type, SupportT<W>& s) {
  const double p[3] = {p0, p1, p2};
  const int dims[3] = {L.nx, L.ny, L.nz};
  W q[3][2];
  int wi[3][2];
  bool ok[3][2];
  bool finite = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    finite &= fabs(p[a]) < 1.0e15;
    const double fl = finite ? floor(p[a]) : 0.0;
    const int64_t n0 = (int64_t)fl;
    s.lo[a] = (int)n0;
    if (std::is_same<W, float>::value) {
      const W d = (W)(p[a] - fl);  // in [0, 1)
      q[a][0] = W(1) - d;
      q[a][1] = d;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int64_t node = n0 + k;
      if (!std::is_same<W, float>::value) q[a][k] = (W)fmax(1.0 - fabs(p[a] - (double)node), 0.0);
      if (node >= 0 && node < dims[a]) {
        wi[a][k] = (int)node;
        ok[a][k] = true;
      } else if (L.periodic[a]) {
        wi[a][k] = (int)(node == -1 ? dims[a] - 1 : (node == dims[a] ? 0 : wrap_far(node, dims[a])));
        ok[a][k] = true;
      } else {
        wi[a][k] = 0;
        ok[a][k] = false;
      }
    }
  }
  W total = W(0);
#pragma unroll
  for (int ix = 0; ix < 2; ++ix)
#pragma unroll
    for (int iy = 0; iy < 2; ++iy)
#pragma unroll
      for (int iz = 0; iz < 2; ++iz) {
        const int slot = ix * 4 + iy * 2 + iz;
        const uint32_t n = (uint32_t)wi[2][iz] + (uint32_t)L.nz * ((uint32_t)wi[1][iy] + (uint32_t)L.ny * (uint32_t)wi[0][ix]);
        const W ww = q[0][ix] * q[1][iy] * q[2][iz];
        bool keep = finite && ok[0][ix] && ok[1][iy] && ok[2][iz] && ww != W(0);
        if (keep && type) keep = type[n] == HC_NODE_FLUID;
        s.idx[slot] = n;
        s.w[slot] = keep ? ww : W(0);
        total += s.w[slot];
      }
  const W coeff = W(1) / total;
#pragma unroll
  for (int k = 0; k < 8; ++k) s.w[k] = s.w[k] != W(0) ? s.w[k] * coeff : W(0);
}

// One force contribution to the spread: rounded to a multiple of L.spread_q when
// the exact spread is on.  spread_q is a power of two, so both scalings are
// exact and rint is the only rounding; the sums of such multiples stay exact
// (hence independent of the order of the adds) while they are below 2^53
// spread_q in magnitude.
__device__ __forceinline__ double spread_term(double v, const LatticeDev& L) {
  return L.spread_q > 0.0 ? rint(v * L.spread_inv_q) * L.spread_q : v;
}

__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

__device__ __forceinline__ void lds_add(float* p, float v) { atomicAdd(p, v); }
__device__ __forceinline__ void lds_add(double* p, double v) { atomicAdd(p, v); }

}  // namespace
}  // namespace hcg
