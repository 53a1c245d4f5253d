// hc_cell_body.h -- the vertex half of the IB-LBM step, fused per cell.
//
// One workgroup per cell runs, in the order of HemoCell::iterate()
// (core/hemoCell.cpp:299-376):
//   1. interpolateFluidVelocity + advanceParticles of step t
//      (core/hemoCellParticleField.cpp:566-588, :819-839): the cell's vertices
//      read the fluid velocity of their phi2 support, move, and land in LDS;
//   2. applyConstitutiveModel of step t (:633-675 and the model's
//      ParticleMechanics) on the positions in LDS;
//   3. spreadParticleForce of step t+1 (:841-863): the forces just computed (plus
//      any extra force, capped to f_limit in place) are spread from the same
//      positions into the force buffer of the next step, through an LDS box
//      covering the cell's support and one global atomic per non-zero node.
//      The box holds doubles in both precisions: LDS double atomic adds run
//      several times faster than float ones on this chip.
// Step t+1's spread needs exactly x_t and F(x_t), which steps 1-2 produce, so
// doing it here saves reloading the cell.  Positions stay in LDS for the whole
// kernel; per-vertex bending forces and velocities too when they fit.
//
// Instantiated by hc_cell_f64.hip (FP64, compiled without contracting
// multiply-adds so the FP64 form follows the oracle's operation order) and
// hc_cell_f32.hip (the FP32-shifted form, which may fuse them).
#pragma once
#include <algorithm>
#include <type_traits>

#include "hc_ibm_dev.h"
#include "hc_lds_plan.h"
#include "hc_mech_dev.h"

namespace hcg {
namespace {

enum { CELL_INTERP = 1, CELL_MECH = 2, CELL_SPREAD = 4, CELL_REPULSION = 8 };
// x-planes of a cell's support box up to which the spread buckets its vertices
constexpr int kMaxPlanes = 64;
// 64-triangle chunks of a mesh whose volume is summed in an order independent of
// the block size (meshes up to 8192 triangles)
constexpr int kVolChunks = 128;

// Scalar type of the positions in LDS and of the membrane mechanics: the FP32-
// shifted form's staged kernels keep float offsets from the cell's first vertex
// (the mechanics only takes differences of positions; the volume stays FP64);
// everything else keeps absolute double positions.
template <typename R, bool STAGED, bool GPOS>
using MechT = typename std::conditional<std::is_same<R, float>::value && STAGED && !GPOS, float, double>::type;

// STAGED: positions, velocities and bending forces in LDS.  Otherwise positions
// in LDS and the rest in global memory, or (GPOS) the positions too: then the
// workgroup needs LDS only for the spread box and fits beside the many small
// workgroups of another cell type running at the same time.
template <typename R, int MODEL, int BS, bool STAGED, bool GPOS>
__device__ __forceinline__ void cell_step_body(const VertexDev& V, const TypeDev& T, const int64_t base,
                                               const R* __restrict__ uf, const uint8_t* __restrict__ type,
                                               const LatticeDev& L, R* __restrict__ Fnext,
                                               const double* __restrict__ extra, double* __restrict__ scratch,
                                               double f_limit, int flags, int box_nodes,
                                               unsigned long long* wall_hits) {
  using P = MechT<R, STAGED, GPOS>;
  constexpr bool REL = !std::is_same<P, double>::value;  // LDS holds offsets from the cell origin
  // scalar of the support weights, the interpolated velocity and the spread's
  // contributions: FP32 in the FP32-shifted form (whose velocity field and
  // force field are FP32), FP64 otherwise
  using SW = typename std::conditional<std::is_same<R, float>::value, float, double>::type;
  extern __shared__ __align__(16) double sm[];
  __shared__ double red[kVolChunks];
  __shared__ double vforce_sh;
  __shared__ int bmin[3], bmax[3], bad, farpk;
  __shared__ int pcount[kMaxPlanes + 1], pstart[kMaxPlanes + 1];
  const int tid = threadIdx.x;
  const int nv = T.nv;
  const int64_t vs = V.vs;
  const bool visc = (MODEL != HC_MODEL_RBC) || (T.eta_m != 0.0);
  // the spread's support box bounds, first used after the barrier ending phase 1
  if (tid == 0) {
    for (int a = 0; a < 3; ++a) {
      bmin[a] = 0x7fffffff;
      bmax[a] = -0x7fffffff;
    }
    bad = 0;
    farpk = 0;
  }
  // origin of the offsets: the first vertex before this step's advance, read by
  // every thread before any thread moves it (its advanced position is stored
  // only after the barrier that ends phase 1)
  double org0 = 0.0, org1 = 0.0, org2 = 0.0;
  if (REL) {
    org0 = V.x[base];
    org1 = V.x[vs + base];
    org2 = V.x[2 * vs + base];
  }
  __shared__ double held_sh[3];  // REL: vertex 0's advanced position until then
  bool held = false;
  // LDS: positions, then (STAGED) velocities, then a region shared by the bending
  // forces (phase 2) and the spread box (phase 3)
  P* const sp_ = reinterpret_cast<P*>(sm);
  P* px = GPOS ? reinterpret_cast<P*>(V.x + base) : sp_;
  P* py = GPOS ? reinterpret_cast<P*>(V.x + vs + base) : sp_ + nv;
  P* pz = GPOS ? reinterpret_cast<P*>(V.x + 2 * vs + base) : sp_ + 2 * nv;
  P *vx, *vy, *vz, *bx, *by, *bz;
  // the region (bending forces, then the box) after the fixed part (hc_lds_plan.h)
  double* const region =
      reinterpret_cast<double*>(reinterpret_cast<char*>(sm) + cell_fixed_bytes(nv, visc, sizeof(P), STAGED, GPOS));
  if (STAGED) {
    vx = sp_ + 3 * nv;
    vy = sp_ + 4 * nv;
    vz = sp_ + 5 * nv;
    bx = reinterpret_cast<P*>(region);
    by = bx + nv;
    bz = bx + 2 * nv;
  } else {
    vx = reinterpret_cast<P*>(V.v + base);
    vy = reinterpret_cast<P*>(V.v + vs + base);
    vz = reinterpret_cast<P*>(V.v + 2 * vs + base);
    bx = reinterpret_cast<P*>(scratch + base);
    by = reinterpret_cast<P*>(scratch + vs + base);
    bz = reinterpret_cast<P*>(scratch + 2 * vs + base);
  }
  // Absolute position of vertex k for the spread: LDS (double), or with float
  // offsets in LDS (REL) a packed copy that phase 3 writes over the offsets once
  // the mechanics is done: per axis the node floor(x) relative to the origin's
  // node (8 bits) and x - floor(x) in 24-bit fixed point, which rebuilds floor(x)
  // exactly and the fraction to 2^-24 (a float's resolution near 1) without
  // rereading the vertex planes; the planes when a cell spans more than 127 nodes.
  bool use_pk = false;
  const uint32_t* const pk = reinterpret_cast<const uint32_t*>(sp_);
  const double fo0 = floor(org0), fo1 = floor(org1), fo2 = floor(org2);
  auto unpack = [&](double fo, uint32_t w) -> double {
    return fo + (double)((int)(w >> 24) - 128) + (double)(w & 0xFFFFFFu) * (1.0 / 16777216.0);
  };
  auto pos_x = [&](int k) -> double {
    return REL ? (use_pk ? unpack(fo0, pk[k]) : V.x[base + k]) : (double)px[k];
  };
  auto pos_y = [&](int k) -> double {
    return REL ? (use_pk ? unpack(fo1, pk[nv + k]) : V.x[vs + base + k]) : (double)py[k];
  };
  auto pos_z = [&](int k) -> double {
    return REL ? (use_pk ? unpack(fo2, pk[2 * nv + k]) : V.x[2 * vs + base + k]) : (double)pz[k];
  };
  double* box = region;
  // after the box: the spread's vertex list, bucketed by x-plane
  uint16_t* vlist = reinterpret_cast<uint16_t*>(region + 3 * (int64_t)box_nodes);

  // ---- 1. interpolation and advance ------------------------------------------
  // A thread takes up to U of its vertices at once, so their position loads and
  // support-node gathers are in flight together.
  // A vertex removed by wall contact (flag) stays where it is and spreads
  // nothing; its cell is incomplete from then on (hemoCellParticleField.cpp:
  // 566-588 removes it; applyConstitutiveModel, :633-675, skips cells with a
  // vertex missing and zeroes the rest's forces).
  constexpr int U = 3;
  bool lrem = false;
  for (int k0 = tid; k0 < nv; k0 += U * BS) {
    double xa[U][3], va[U][3];
    bool gone[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * BS;
      if (k >= nv) continue;
      const int64_t i = base + k;
      xa[u][0] = V.x[i];
      xa[u][1] = V.x[vs + i];
      xa[u][2] = V.x[2 * vs + i];
      gone[u] = V.flag && (V.flag[i] & VFLAG_REMOVED);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * BS;
      if (k >= nv) continue;
      const int64_t i = base + k;
      if (gone[u]) {
        lrem = true;
        if (!GPOS) {
          px[k] = (P)(xa[u][0] - org0);
          py[k] = (P)(xa[u][1] - org1);
          pz[k] = (P)(xa[u][2] - org2);
        }
        if (STAGED && visc) vx[k] = vy[k] = vz[k] = P(0);
        continue;
      }
      if (flags & CELL_INTERP) {
        SupportT<SW> sp;
        ibm_support(xa[u][0], xa[u][1], xa[u][2], L, type, sp);
        SW v0 = SW(0), v1 = SW(0), v2 = SW(0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (sp.w[q] == SW(0)) continue;
          const R* un = uf + 3 * (int64_t)sp.idx[q];
          v0 += (SW)un[0] * sp.w[q];
          v1 += (SW)un[1] * sp.w[q];
          v2 += (SW)un[2] * sp.w[q];
        }
        va[u][0] = (double)v0;
        va[u][1] = (double)v1;
        va[u][2] = (double)v2;
      } else {
        va[u][0] = V.v[i];
        va[u][1] = V.v[vs + i];
        va[u][2] = V.v[2 * vs + i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * BS;
      if (k >= nv || gone[u]) continue;
      const int64_t i = base + k;
      const double v0 = va[u][0], v1 = va[u][1], v2 = va[u][2];
      if (flags & CELL_INTERP) {
        V.v[i] = v0;
        V.v[vs + i] = v1;
        V.v[2 * vs + i] = v2;
      }
      const double x0 = xa[u][0] + v0, x1 = xa[u][1] + v1, x2 = xa[u][2] + v2;
      if (REL && k == 0) {
        held = true;
        held_sh[0] = x0;
        held_sh[1] = x1;
        held_sh[2] = x2;
      } else {
        V.x[i] = x0;
        V.x[vs + i] = x1;
        V.x[2 * vs + i] = x2;
      }
      if (!GPOS) {
        px[k] = (P)(x0 - org0);
        py[k] = (P)(x1 - org1);
        pz[k] = (P)(x2 - org2);
      }
      if (STAGED && visc) {
        vx[k] = (P)v0;
        vy[k] = (P)v1;
        vz[k] = (P)v2;
      }
      if (type && isfinite(x0) && isfinite(x1) && isfinite(x2)) {
        const double p[3] = {x0, x1, x2};
        const int dims[3] = {L.nx, L.ny, L.nz};
        int64_t wc[3];
        bool out = false;
        for (int a = 0; a < 3; ++a) {
          const int64_t cc = (int64_t)floor(p[a] + 0.5);
          if (cc >= 0 && cc < dims[a]) wc[a] = cc;
          else if (L.periodic[a]) wc[a] = wrap_mod(cc, dims[a]);
          else out = true;
        }
        if (!out && type[wc[2] + (int64_t)L.nz * (wc[1] + (int64_t)L.ny * wc[0])] != HC_NODE_FLUID) {
          atomicAdd(wall_hits, 1ull);
          if (V.flag) V.flag[i] = VFLAG_REMOVED;
          lrem = true;
        }
      }
    }
  }
  const bool incomplete = __syncthreads_or(lrem) != 0;
  if (held) {
    V.x[base] = held_sh[0];
    V.x[vs + base] = held_sh[1];
    V.x[2 * vs + base] = held_sh[2];
  }
  const CellViewT<P> c{px, py, pz};

  // The force a vertex spreads: its force plus any extra force, capped to f_limit
  // (hemoCellParticleField.cpp:846-852).  Applied where the mechanics stores the
  // force when it ran on this cell, else by the spread.
  auto cap_force = [&](int64_t i, double& f0, double& f1, double& f2) -> bool {
    bool changed = false;
    if (extra) {
      f0 += extra[i];
      f1 += extra[vs + i];
      f2 += extra[2 * vs + i];
      changed = true;
    }
    const double fm = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
    if (fm > f_limit) {
      const double sc = f_limit / fm;
      f0 *= sc;
      f1 *= sc;
      f2 *= sc;
      changed = true;
    }
    return changed;
  };
  const bool capped = (flags & CELL_MECH) && !incomplete;

  // ---- 2. mechanics ------------------------------------------------------------
  if ((flags & CELL_MECH) && incomplete) {
    for (int k = tid; k < nv; k += BS) {
      const int64_t i = base + k;
      V.f[i] = 0.0;
      V.f[vs + i] = 0.0;
      V.f[2 * vs + i] = 0.0;
    }
    __syncthreads();
  } else if (flags & CELL_MECH) {
    // the volume: triangle t is lane t % 64 of chunk t / 64; a wavefront sums a
    // chunk by a butterfly over its lanes and thread 0 adds the chunks in order,
    // so the sum is the same at every block size (batched, mixed and alone
    // launches agree bit for bit); larger meshes sum per thread, then per
    // wavefront
    const int nchunk = (T.nt + 63) >> 6;
    int nred = nchunk;
    if (nchunk <= kVolChunks) {
      const V3T<P> rm = c.pos(0);
      const D3 r = {(double)rm.x, (double)rm.y, (double)rm.z};
      for (int ch = tid >> 6; ch < nchunk; ch += BS / 64) {
        double part = volume_term(c, T, (ch << 6) | (tid & 63), r);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
        if ((tid & 63) == 0) red[ch] = part;
      }
    } else {
      double part = volume_partial<BS>(c, T, tid);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
      if ((tid & 63) == 0) red[tid >> 6] = part;
      nred = BS / 64;
    }
    __syncthreads();
    if (tid == 0) {
      double vsum = red[0];
      for (int w = 1; w < nred; ++w) vsum += red[w];
      const double volume = vsum * (1.0 / 6.0);
      const double vf = (volume - T.volume_eq) / T.volume_eq;
      vforce_sh = -T.k_volume * vf / fabs(MaxCellVolumetricChange - vf * vf);
    }
    // FP32 RBC: the ring-ordered tables, else the compact per-slot ones (both
    // built beside the padded tables)
    constexpr bool kCompact = MODEL == HC_MODEL_RBC && std::is_same<P, float>::value;
    const bool compact = kCompact && T.cslot != nullptr;
    if (MODEL != HC_MODEL_PLT) {
      for (int k = tid; k < nv; k += BS) {
        V3T<P> b;
        if constexpr (kCompact) b = compact ? bending_force_c(c, T, k) : bending_force(c, T, k);
        else b = bending_force(c, T, k);
        bx[k] = b.x;
        by[k] = b.y;
        bz[k] = b.z;
      }
    }
    __syncthreads();
    const P vforce = (P)vforce_sh;
    for (int k = tid; k < nv; k += BS) {
      V3T<P> F;
      if constexpr (kCompact)
        F = T.rpar ? vertex_force_rbc_r(c, T, k, vforce, bx, by, bz, vx, vy, vz, visc)
            : compact ? vertex_force_rbc_c(c, T, k, vforce, bx, by, bz, vx, vy, vz, visc)
                      : vertex_force<MODEL, TERM_ALL, P>(c, T, k, vforce, bx, by, bz, vx, vy, vz, visc);
      else F = vertex_force<MODEL, TERM_ALL, P>(c, T, k, vforce, bx, by, bz, vx, vy, vz, visc);
      const int64_t i = base + k;
      double f0 = (double)F.x, f1 = (double)F.y, f2 = (double)F.z;
      if (flags & CELL_SPREAD) cap_force(i, f0, f1, f2);
      V.f[i] = f0;
      V.f[vs + i] = f1;
      V.f[2 * vs + i] = f2;
    }
    __syncthreads();  // forces visible to the whole workgroup; the region is free again
  }
  if (!(flags & CELL_SPREAD)) return;

  // ---- 3. spread of the next step ---------------------------------------------
  // the force each vertex spreads: its force plus any extra force, capped to
  // f_limit and written back in place when that changed it
  // (hemoCellParticleField.cpp:846-852; done by the mechanics above when it ran
  // on this cell), plus the repulsion force when enabled.  The packed positions
  // below overwrite the offsets: the mechanics ended with a barrier.
  {
    int lmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, lmax[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
    bool lbad = false, lfar = false;
    uint32_t* const pkw = reinterpret_cast<uint32_t*>(sp_);
    const double fo[3] = {fo0, fo1, fo2};
    for (int k = tid; k < nv; k += BS) {
      const double p[3] = {pos_x(k), pos_y(k), pos_z(k)};  // use_pk is still false here
      for (int a = 0; a < 3; ++a) {
        if (!(fabs(p[a]) < 1.0e9)) {
          lbad = true;
          continue;
        }
        const double fl = floor(p[a]);
        const int cc = (int)fl;
        lmin[a] = min(lmin[a], cc);
        lmax[a] = max(lmax[a], cc);
        if (REL) {
          const double rel = fl - fo[a];
          if (!(rel >= -128.0 && rel <= 127.0)) lfar = true;
          const uint32_t f24 = min((uint32_t)((p[a] - fl) * 16777216.0), 0xFFFFFFu);
          pkw[a * nv + k] = ((uint32_t)((int)rel + 128) << 24) | f24;
        }
      }
      const int64_t i = base + k;
      if (capped || (incomplete && (V.flag[i] & VFLAG_REMOVED))) continue;
      double f0 = V.f[i], f1 = V.f[vs + i], f2 = V.f[2 * vs + i];
      if (cap_force(i, f0, f1, f2)) {
        V.f[i] = f0;
        V.f[vs + i] = f1;
        V.f[2 * vs + i] = f2;
      }
    }
    for (int a = 0; a < 3; ++a) {
      atomicMin(&bmin[a], lmin[a]);
      atomicMax(&bmax[a], lmax[a]);
    }
    if (lbad) atomicOr(&bad, 1);
    if (lfar) atomicOr(&farpk, 1);
    // the spread box starts at zero (each slab's flush clears what it used), and
    // so do the plane counts of the bucketing
    for (int e = tid; e < 3 * box_nodes; e += BS) box[e] = 0.0;
    for (int b = tid; b <= kMaxPlanes; b += BS) pcount[b] = 0;
  }
  __syncthreads();
  use_pk = REL && !bad && !farpk;
  const int o0 = bmin[0], o1 = bmin[1], o2 = bmin[2];
  const int s0 = bmax[0] - bmin[0] + 2, s1 = bmax[1] - bmin[1] + 2, s2 = bmax[2] - bmin[2] + 2;
  const int64_t plane = (int64_t)s1 * s2;
  // the box is spread in slabs of whole x-planes when it does not fit at once;
  // only a cell whose single x-plane does not fit uses global atomics per vertex
  const bool direct = bad || s0 <= 0 || plane > box_nodes;
  const int64_t S = L.stride;
  auto spread_force = [&](int64_t i, double& f0, double& f1, double& f2) {
    if (incomplete && (V.flag[i] & VFLAG_REMOVED)) {
      f0 = f1 = f2 = 0.0;
      return;
    }
    f0 = V.f[i];
    f1 = V.f[vs + i];
    f2 = V.f[2 * vs + i];
    if (flags & CELL_REPULSION) {
      f0 = V.frep[i] + f0;
      f1 = V.frep[vs + i] + f1;
      f2 = V.frep[2 * vs + i] + f2;
    }
  };
  if (direct) {
    for (int k = tid; k < nv; k += BS) {
      const int64_t i = base + k;
      double f0, f1, f2;
      spread_force(i, f0, f1, f2);
      const SW g0 = (SW)f0, g1 = (SW)f1, g2 = (SW)f2;
      SupportT<SW> sp;
      ibm_support(pos_x(k), pos_y(k), pos_z(k), L, type, sp);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (sp.w[q] == SW(0)) continue;
        const int64_t n = sp.idx[q];
        atomic_add(&Fnext[n], (R)spread_term((double)(g0 * sp.w[q]), L));
        atomic_add(&Fnext[S + n], (R)spread_term((double)(g1 * sp.w[q]), L));
        atomic_add(&Fnext[2 * S + n], (R)spread_term((double)(g2 * sp.w[q]), L));
      }
    }
    return;
  }
  const int sx = (int)std::min<int64_t>(s0, box_nodes / plane);  // x-planes per slab
  // With several slabs, the vertices are bucketed by the x-plane of their
  // support's low corner (a vertex adds to that plane and the next), so a slab
  // visits only the vertices that reach it.
  const bool bucketed = sx < s0 && s0 <= kMaxPlanes && nv <= 65535;
  if (bucketed) {
    for (int k = tid; k < nv; k += BS) atomicAdd(&pcount[(int)floor(pos_x(k)) - o0], 1);
    __syncthreads();
    if (tid < 64) {
      // exclusive prefix sum of the s0 + 1 <= 65 plane counts over wavefront 0;
      // the counts become the insertion cursors
      const int c = tid <= s0 ? pcount[tid] : 0;
      int inc = c;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (tid >= o) inc += y;
      }
      if (tid <= s0) {
        pstart[tid] = inc - c;
        pcount[tid] = inc - c;
      }
      if (tid == 63 && s0 == 64) {
        pstart[64] = inc;
        pcount[64] = inc;
      }
    }
    __syncthreads();
    for (int k = tid; k < nv; k += BS) vlist[atomicAdd(&pcount[(int)floor(pos_x(k)) - o0], 1)] = (uint16_t)k;
    __syncthreads();
  }
  const int dims[3] = {L.nx, L.ny, L.nz};
  for (int xs = 0; xs < s0; xs += sx) {
    const int nxp = std::min(sx, s0 - xs);
    // the slab fits the box (plane <= box_nodes), so 32-bit indices suffice
    const int total = nxp * (int)plane;
    // the vertices to visit: a range of the bucketed list, or all of them
    const int first = bucketed ? pstart[max(xs - 1, 0)] : 0;
    const int count = bucketed ? pstart[min(xs + nxp, s0)] - first : nv;
    const int per = (count + 63) / 64;
    for (int k = tid; k < 64 * per; k += BS) {
      // lanes of one wavefront take vertices ~count/64 apart (fewer LDS same-address adds)
      const int j = (k & 63) * per + (k >> 6);
      if (j >= count) continue;
      const int vtx = bucketed ? (int)vlist[first + j] : j;
      double f0, f1, f2;
      spread_force(base + vtx, f0, f1, f2);
      const SW g0 = (SW)f0, g1 = (SW)f1, g2 = (SW)f2;
      SupportT<SW> sp;
      ibm_support(pos_x(vtx), pos_y(vtx), pos_z(vtx), L, type, sp);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int lx = sp.lo[0] + (q >> 2) - o0 - xs;
        if (sp.w[q] == SW(0) || lx < 0 || lx >= nxp) continue;
        const int e = (sp.lo[2] + (q & 1) - o2) + s2 * ((sp.lo[1] + ((q >> 1) & 1) - o1) + s1 * lx);
        lds_add(&box[e], spread_term((double)(g0 * sp.w[q]), L));
        lds_add(&box[total + e], spread_term((double)(g1 * sp.w[q]), L));
        lds_add(&box[2 * total + e], spread_term((double)(g2 * sp.w[q]), L));
      }
    }
    __syncthreads();
    const unsigned us1 = (unsigned)s1, us2 = (unsigned)s2;
    for (int e = tid; e < total; e += BS) {
      const double a = box[e], b = box[total + e], cz = box[2 * total + e];
      if (a == 0.0 && b == 0.0 && cz == 0.0) continue;
      box[e] = 0.0;  // zero for the next slab, whose box is no larger
      box[total + e] = 0.0;
      box[2 * total + e] = 0.0;
      const unsigned r = (unsigned)e / us2;
      const int lz = (int)((unsigned)e - r * us2);
      const unsigned r2 = r / us1;
      const int ly = (int)(r - r2 * us1), lx = (int)r2;
      const int64_t g[3] = {(int64_t)o0 + xs + lx, (int64_t)o1 + ly, (int64_t)o2 + lz};
      int64_t wc[3];
      for (int q = 0; q < 3; ++q) wc[q] = (g[q] >= 0 && g[q] < dims[q]) ? g[q] : wrap_mod(g[q], dims[q]);
      const int64_t n = wc[2] + (int64_t)L.nz * (wc[1] + (int64_t)L.ny * wc[0]);
      if (a != 0.0) atomic_add(&Fnext[n], (R)a);
      if (b != 0.0) atomic_add(&Fnext[S + n], (R)b);
      if (cz != 0.0) atomic_add(&Fnext[2 * S + n], (R)cz);
    }
    __syncthreads();  // the box is zero again for the next slab
  }
}

template <typename R, int MODEL, int BS, bool STAGED, bool GPOS>
__global__ __launch_bounds__(BS) void k_cell_step(VertexDev V, TypeDev T, const int32_t* __restrict__ cells,
                                                  const int64_t* __restrict__ cell_base, const R* __restrict__ uf,
                                                  const uint8_t* __restrict__ type, LatticeDev L, R* __restrict__ Fnext,
                                                  const double* __restrict__ extra, double* __restrict__ scratch,
                                                  double f_limit, int flags, int box_nodes,
                                                  unsigned long long* wall_hits) {
  cell_step_body<R, MODEL, BS, STAGED, GPOS>(V, T, cell_base[cells[blockIdx.x]], uf, type, L, Fnext, extra, scratch,
                                             f_limit, flags, box_nodes, wall_hits);
}

// cells of several independent cases in one launch, one workgroup per cell
template <typename R, int MODEL, int BS, bool STAGED, bool GPOS>
__global__ __launch_bounds__(BS) void k_cell_step_multi(const CellCase* __restrict__ cs, int box_nodes) {
  const CellCase& c = cs[blockIdx.x];
  cell_step_body<R, MODEL, BS, STAGED, GPOS>(c.V, c.T, c.base, (const R*)c.uf, c.type, c.L, (R*)c.Fnext, c.extra,
                                             c.scratch, c.f_limit, c.flags, box_nodes, c.wall_hits);
}

// Cells of every type in one persistent launch: each workgroup takes the next
// entry of the queue (cell, type index; the types with the most vertices first)
// until the queue is empty.  RBC and platelet cells run the staged body, WBCs
// the one with positions in global memory; all share the launch's dynamic LDS.
template <typename R>
__global__ __launch_bounds__(kMixedBlock) void k_cell_step_mixed(
    VertexDev V, const TypeDev* __restrict__ types, MixedArgs A, const int2* __restrict__ queue, int nqueue,
    unsigned int* __restrict__ head, const int64_t* __restrict__ cell_base, const R* __restrict__ uf,
    const uint8_t* __restrict__ type, LatticeDev L, R* __restrict__ Fnext, const double* __restrict__ extra,
    double* __restrict__ scratch, double f_limit, unsigned long long* wall_hits) {
  __shared__ int slot;
  for (;;) {
    if (threadIdx.x == 0) slot = (int)atomicAdd(head, 1u);
    __syncthreads();
    const int q = slot;
    __syncthreads();  // every thread has read the slot before thread 0 takes the next one
    if (q >= nqueue) break;
    const int2 e = queue[q];
    const TypeDev& T = types[e.y];
    const int64_t base = cell_base[e.x];
    if (T.model == HC_MODEL_RBC)
      cell_step_body<R, HC_MODEL_RBC, kMixedBlock, true, false>(V, T, base, uf, type, L, Fnext, extra, scratch,
                                                                 f_limit, A.flags[e.y], A.box_nodes[e.y], wall_hits);
    else if (T.model == HC_MODEL_PLT)
      cell_step_body<R, HC_MODEL_PLT, kMixedBlock, true, false>(V, T, base, uf, type, L, Fnext, extra, scratch,
                                                                 f_limit, A.flags[e.y], A.box_nodes[e.y], wall_hits);
    else
      cell_step_body<R, HC_MODEL_WBC, kMixedBlock, false, true>(V, T, base, uf, type, L, Fnext, extra, scratch,
                                                                 f_limit, A.flags[e.y], A.box_nodes[e.y], wall_hits);
    __syncthreads();  // the next cell reuses the LDS
  }
}

}  // namespace

// Static LDS of one kernel instantiation (the reduction array grows with the block).
template <typename K>
static size_t static_lds(K kernel) {
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kernel)) != hipSuccess) return 32768;
  return a.sharedSizeBytes;
}

// Lay out the dynamic LDS of one cell type for block size B and launch: staged
// (positions, velocities and a region shared by bending forces and the box) when
// it fits beside the kernel's static LDS, else positions plus as much box as
// fits; with gpos only the box.  The spread's vertex list (2 bytes per vertex)
// follows the region.
template <typename R, int M, int B>
static hipError_t launch_cell(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                              const int64_t* cell_base, const R* uf, const uint8_t* type, const LatticeDev& L, R* Fnext,
                              const double* extra, double* scratch, double f_limit, int flags, int box_bytes,
                              int lds_device, unsigned long long* wall_hits, bool gpos,
                              const CellCase* multi = nullptr) {
  static const size_t st_staged = std::max(static_lds(k_cell_step<R, M, B, true, false>),
                                           static_lds(k_cell_step_multi<R, M, B, true, false>));
  static const size_t st_plain = std::max(static_lds(k_cell_step<R, M, B, false, false>),
                                          static_lds(k_cell_step_multi<R, M, B, false, false>));
  static const size_t st_gpos = static_lds(k_cell_step<R, M, B, false, true>);
  const bool visc = (T.model != HC_MODEL_RBC) || (T.eta_m != 0.0);
  // the staged kernel's position scalar: float offsets in the FP32-shifted form
  const CellLdsPlan plan = cell_lds_plan(T.model, T.nv, visc, sizeof(MechT<R, true, false>), (size_t)box_bytes,
                                         (size_t)lds_device, st_staged, st_plain, st_gpos, gpos);
  if (!plan.ok) return hipErrorInvalidConfiguration;
  const bool staged = plan.staged;
  const int box_nodes = plan.box_nodes;
  const size_t dyn = plan.dyn;
  if (multi) {
    if (gpos) return hipErrorInvalidValue;
    if (staged)
      hipLaunchKernelGGL((k_cell_step_multi<R, M, B, true, false>), dim3((unsigned)ncells), dim3(B), dyn, s, multi,
                         box_nodes);
    else
      hipLaunchKernelGGL((k_cell_step_multi<R, M, B, false, false>), dim3((unsigned)ncells), dim3(B), dyn, s, multi,
                         box_nodes);
    return hipGetLastError();
  }
  if (staged)
    hipLaunchKernelGGL((k_cell_step<R, M, B, true, false>), dim3((unsigned)ncells), dim3(B), dyn, s, V, T, cells,
                       cell_base, uf, type, L, Fnext, extra, scratch, f_limit, flags, box_nodes, wall_hits);
  else if (gpos)
    hipLaunchKernelGGL((k_cell_step<R, M, B, false, true>), dim3((unsigned)ncells), dim3(B), dyn, s, V, T, cells,
                       cell_base, uf, type, L, Fnext, extra, scratch, f_limit, flags, box_nodes, wall_hits);
  else
    hipLaunchKernelGGL((k_cell_step<R, M, B, false, false>), dim3((unsigned)ncells), dim3(B), dyn, s, V, T, cells,
                       cell_base, uf, type, L, Fnext, extra, scratch, f_limit, flags, box_nodes, wall_hits);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_cell_step(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                            const int64_t* cell_base, const R* uf, const uint8_t* type, const LatticeDev& L,
                            R* Fnext, const double* extra, double* scratch, double f_limit, int flags, int box_bytes,
                            int lds_device, unsigned long long* wall_hits, int block, bool gpos) {
  if (ncells <= 0) return hipSuccess;
  // threads per cell for the RBC and WBC meshes: 256, 512 or 1024; a block whose
  // static LDS leaves too little room for the cell falls back to a smaller one
  hipError_t e = hipErrorInvalidConfiguration;
#define HC_CELL(M, B)                                                                                                \
  if (e == hipErrorInvalidConfiguration)                                                                             \
    e = launch_cell<R, M, B>(s, V, T, cells, ncells, cell_base, uf, type, L, Fnext, extra, scratch, f_limit, flags, \
                             box_bytes, lds_device, wall_hits, gpos);
  switch (T.model) {
    case HC_MODEL_RBC:
      if (block >= 1024) HC_CELL(HC_MODEL_RBC, 1024)
      if (block >= 512) HC_CELL(HC_MODEL_RBC, 512)
      HC_CELL(HC_MODEL_RBC, 256)
      break;
    case HC_MODEL_PLT:
      HC_CELL(HC_MODEL_PLT, 64)
      break;
    case HC_MODEL_WBC:
      if (block >= 1024) HC_CELL(HC_MODEL_WBC, 1024)
      if (block >= 512) HC_CELL(HC_MODEL_WBC, 512)
      HC_CELL(HC_MODEL_WBC, 256)
      break;
    default:
      return hipErrorInvalidValue;
  }
#undef HC_CELL
  // a mesh whose positions do not fit the workgroup's LDS at any block (a
  // 2562-vertex WBC on a 64 KB device) keeps them in global memory
  if (e == hipErrorInvalidConfiguration && !gpos)
    return launch_cell_step<R>(s, V, T, cells, ncells, cell_base, uf, type, L, Fnext, extra, scratch, f_limit, flags,
                               box_bytes, lds_device, wall_hits, 256, true);
  return e;
}

// The persistent mixed launch: dynamic LDS for the largest need of the types
// present (staged RBC and platelet cells, WBCs with positions in global memory),
// each type's box filling what its layout leaves; `grid` workgroups.
template <typename R>
hipError_t launch_cell_step_mixed(hipStream_t s, const VertexDev& V, const TypeDev* d_types, const TypeDev* h_types,
                                  int ntypes, const int* flags, const int2* queue, int nqueue, unsigned int* head,
                                  int grid, const int64_t* cell_base, const R* uf, const uint8_t* type,
                                  const LatticeDev& L, R* Fnext, const double* extra, double* scratch, double f_limit,
                                  int box_bytes, int lds_device, unsigned long long* wall_hits) {
  if (nqueue <= 0) return hipSuccess;
  if (ntypes > kMaxMixedTypes || grid <= 0) return hipErrorInvalidValue;
  static const size_t st = static_lds(k_cell_step_mixed<R>);
  int model[kMaxMixedTypes], nv[kMaxMixedTypes];
  bool visc[kMaxMixedTypes];
  for (int t = 0; t < ntypes; ++t) {
    model[t] = h_types[t].model;
    nv[t] = h_types[t].nv;
    visc[t] = (h_types[t].model != HC_MODEL_RBC) || (h_types[t].eta_m != 0.0);
  }
  MixedArgs A{};
  const size_t dyn = mixed_lds_plan(ntypes, model, nv, visc, sizeof(MechT<R, true, false>), (size_t)box_bytes,
                                    (size_t)lds_device, st, A.box_nodes);
  if (dyn == 0) return hipErrorInvalidConfiguration;
  for (int t = 0; t < ntypes; ++t) A.flags[t] = flags[t];
  hipError_t e = launch_fill_zero<float>(s, reinterpret_cast<float*>(head), 1);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_cell_step_mixed<R>), dim3((unsigned)std::min(grid, nqueue)), dim3(kMixedBlock), dyn, s, V,
                     d_types, A, queue, nqueue, head, cell_base, uf, type, L, Fnext, extra, scratch, f_limit,
                     wall_hits);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_cell_step_multi(hipStream_t s, const TypeDev& T, const CellCase* cases, int ncells, int box_bytes,
                                  int lds_device, int block) {
  if (ncells <= 0) return hipSuccess;
  const VertexDev V{};
  const LatticeDev L{};
  hipError_t e = hipErrorInvalidConfiguration;
#define HC_CELL(M, B)                                                                                             \
  if (e == hipErrorInvalidConfiguration)                                                                          \
    e = launch_cell<R, M, B>(s, V, T, nullptr, ncells, nullptr, nullptr, nullptr, L, nullptr, nullptr, nullptr, 0.0, \
                             0, box_bytes, lds_device, nullptr, false, cases);
  switch (T.model) {
    case HC_MODEL_RBC:
      if (block >= 1024) HC_CELL(HC_MODEL_RBC, 1024)
      if (block >= 512) HC_CELL(HC_MODEL_RBC, 512)
      HC_CELL(HC_MODEL_RBC, 256)
      return e;
    case HC_MODEL_PLT:
      // fewer platelets than compute units (the caller passes block 1024 then):
      // four waves per platelet shorten its box clear and flush
      if (block >= 256) HC_CELL(HC_MODEL_PLT, 256)
      HC_CELL(HC_MODEL_PLT, 64)
      return e;
    case HC_MODEL_WBC:
      if (block >= 1024) HC_CELL(HC_MODEL_WBC, 1024)
      if (block >= 512) HC_CELL(HC_MODEL_WBC, 512)
      HC_CELL(HC_MODEL_WBC, 256)
      return e;
    default:
      return hipErrorInvalidValue;
  }
#undef HC_CELL
}

}  // namespace hcg
