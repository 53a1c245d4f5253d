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
// so every per-vertex sum is formed in the same order without atomics (the area
// and volume sums of a vertex are formed separately and then added).  The cell
// volume is a workgroup tree reduction (FP64).  Per-triangle and per-edge
// terms are recomputed by each incident vertex (cheap, no storage); the
// per-vertex bending force, needed by the whole 1-ring, is computed once into a
// scratch array and read back after a barrier.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "hc_mech_dev.h"

namespace hcg {
namespace {

// STAGED: the cell's positions, bending forces and (with membrane viscosity)
// velocities live in LDS for the whole kernel (9 x 8 bytes per vertex at most).
// TERMS: the force terms computed (all for the step; one at a time, into `out`
// planes of stride V.vs, for output).
template <int MODEL, int BS, bool STAGED, int TERMS = TERM_ALL>
__global__ __launch_bounds__(BS) void k_mechanics(VertexDev V, TypeDev T, const int32_t* __restrict__ cells,
                                                  const int64_t* __restrict__ cell_base, double* __restrict__ bend,
                                                  double* __restrict__ out) {
  extern __shared__ double sm[];
  __shared__ double red[BS];
  __shared__ double vforce_sh;
  const int cell = cells[blockIdx.x];
  const int64_t base = cell_base[cell];
  const int tid = threadIdx.x;
  const int64_t vs = V.vs;
  const int nv = T.nv;
  const bool visc = (MODEL != HC_MODEL_RBC) || (T.eta_m != 0.0);
  // a cell that lost a vertex to a wall gets no mechanics and zero force
  // (hemoCellParticleField.cpp:633-675 leaves incomplete cells out)
  bool lrem = false;
  if (V.flag)
    for (int k = tid; k < nv; k += BS) lrem |= (V.flag[base + k] & VFLAG_REMOVED) != 0;
  if (__syncthreads_or(lrem)) {
    double* dst = out ? out : V.f;
    for (int k = tid; k < nv; k += BS) {
      dst[base + k] = 0.0;
      dst[vs + base + k] = 0.0;
      dst[2 * vs + base + k] = 0.0;
    }
    return;
  }
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
  const D3 r = c.pos(0);  // corners relative to vertex 0 (closed mesh: same volume, no cancellation)
  for (int t = tid; t < T.nt; t += BS) {
    const D3 v0 = c.pos(T.tri[3 * t]) - r, v1 = c.pos(T.tri[3 * t + 1]) - r, v2 = c.pos(T.tri[3 * t + 2]) - r;
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
    const D3 F = vertex_force<MODEL, TERMS>(c, T, k, vforce, bx, by, bz, vx, vy, vz, visc);
    const double F0 = F.x, F1 = F.y, F2 = F.z;
    const int64_t q = base + k;
    double* dst = out ? out : V.f;
    dst[q] = F0;
    dst[vs + q] = F1;
    dst[2 * vs + q] = F2;
  }
}

// Per-cell statistics of CellInformationFunctionals (helper/cellInfo.cpp:170-266):
// volume (sum over triangles of the signed triple product / 6), area, centroid
// and mean velocity (vertex means), bounding box and the largest vertex-vertex
// distance.  out[15*c + ...] = vol, area, centroid[3], velocity[3], bbox[6]
// (xmin, xmax, ymin, ymax, zmin, zmax), stretch.
template <int BS>
__global__ __launch_bounds__(BS) void k_cell_stats(VertexDev V, TypeDev T, const int32_t* __restrict__ cells,
                                                   const int64_t* __restrict__ cell_base, double* __restrict__ out) {
  constexpr int NQ = 15;
  __shared__ double red[NQ][BS];
  const int cell = cells[blockIdx.x];
  const int64_t base = cell_base[cell];
  const int64_t vs = V.vs;
  const CellView c{V.x + base, V.x + vs + base, V.x + 2 * vs + base};
  const int tid = threadIdx.x;
  double q[NQ];
  for (int k = 0; k < NQ; ++k) q[k] = 0.0;
  q[8] = q[10] = q[12] = INFINITY;
  q[9] = q[11] = q[13] = -INFINITY;
  const D3 r = c.pos(0);  // volume corners relative to vertex 0, as the mechanics takes them
  for (int t = tid; t < T.nt; t += BS) {
    const D3 v0 = c.pos(T.tri[3 * t]), v1 = c.pos(T.tri[3 * t + 1]), v2 = c.pos(T.tri[3 * t + 2]);
    const D3 w0 = v0 - r, w1 = v1 - r, w2 = v2 - r;
    const double v210 = w2.x * w1.y * w0.z;
    const double v120 = w1.x * w2.y * w0.z;
    const double v201 = w2.x * w0.y * w1.z;
    const double v021 = w0.x * w2.y * w1.z;
    const double v102 = w1.x * w0.y * w2.z;
    const double v012 = w0.x * w1.y * w2.z;
    q[0] += (1.0 / 6.0) * (-v210 + v120 + v201 - v021 - v102 + v012);
    q[1] += 0.5 * normp(crossp(v1 - v0, v2 - v0));
  }
  for (int i = tid; i < T.nv; i += BS) {
    const D3 p = c.pos(i);
    q[2] += p.x;
    q[3] += p.y;
    q[4] += p.z;
    q[5] += V.v[base + i];
    q[6] += V.v[vs + base + i];
    q[7] += V.v[2 * vs + base + i];
    q[8] = fmin(q[8], p.x);
    q[9] = fmax(q[9], p.x);
    q[10] = fmin(q[10], p.y);
    q[11] = fmax(q[11], p.y);
    q[12] = fmin(q[12], p.z);
    q[13] = fmax(q[13], p.z);
    for (int j = i + 1; j < T.nv; ++j) {
      const D3 r = c.pos(j);
      const double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
      q[14] = fmax(q[14], sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  for (int k = 0; k < NQ; ++k) red[k][tid] = q[k];
  __syncthreads();
  for (int h = BS / 2; h > 0; h >>= 1) {
    if (tid < h) {
      for (int k = 0; k < 8; ++k) red[k][tid] += red[k][tid + h];
      for (int k = 8; k < 14; k += 2) {
        red[k][tid] = fmin(red[k][tid], red[k][tid + h]);
        red[k + 1][tid] = fmax(red[k + 1][tid], red[k + 1][tid + h]);
      }
      red[14][tid] = fmax(red[14][tid], red[14][tid + h]);
    }
    __syncthreads();
  }
  if (tid < NQ) {
    double r = red[tid][0];
    if (tid >= 2 && tid < 8) r /= (double)T.nv;
    out[(int64_t)NQ * cell + tid] = r;
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
                            const int64_t* cell_base, double* bend, int lds_limit, int term, double* out) {
  if (ncells <= 0) return hipSuccess;
  dim3 g((unsigned)ncells);
  const bool visc = (T.model != HC_MODEL_RBC) || (T.eta_m != 0.0);
  const size_t lds = sizeof(double) * (size_t)T.nv * (visc ? 9 : 6);
  const bool staged = lds <= (size_t)std::min(lds_limit, 65536);
  const size_t dyn = staged ? lds : 0;
#define HC_MECH_T(M, B, TM)                                                                                      \
  if (staged) hipLaunchKernelGGL((k_mechanics<M, B, true, TM>), g, dim3(B), dyn, s, V, T, cells, cell_base, bend,  \
                                 out);                                                                            \
  else hipLaunchKernelGGL((k_mechanics<M, B, false, TM>), g, dim3(B), 0, s, V, T, cells, cell_base, bend, out);
#define HC_MECH(M, B)                                                      \
  switch (term) {                                                          \
    case TERM_AREA: HC_MECH_T(M, B, TERM_AREA) break;                      \
    case TERM_VOLUME: HC_MECH_T(M, B, TERM_VOLUME) break;                  \
    case TERM_BENDING: HC_MECH_T(M, B, TERM_BENDING) break;                \
    case TERM_LINK: HC_MECH_T(M, B, TERM_LINK) break;                      \
    case TERM_VISC: HC_MECH_T(M, B, TERM_VISC) break;                      \
    case TERM_INNER: HC_MECH_T(M, B, TERM_INNER) break;                    \
    default: HC_MECH_T(M, B, TERM_ALL) break;                              \
  }
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
#undef HC_MECH_T
  return hipGetLastError();
}

hipError_t launch_cell_stats(hipStream_t s, const VertexDev& V, const TypeDev& T, const int32_t* cells, int ncells,
                             const int64_t* cell_base, double* out) {
  if (ncells <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_cell_stats<128>), dim3((unsigned)ncells), dim3(128), 0, s, V, T, cells, cell_base, out);
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


// ---- statistics staging without allocation -----------------------------------
// hc_get_cell_stats_noalloc stages its per-cell statistics (15 doubles per cell,
// the layout k_cell_stats writes) in this file-scope array instead of a buffer
// allocated per call: HC_CELL_STATS_NOALLOC_MAX cells, 31.5 MB on each device
// the library runs on, held until the process ends.  In host code the array's
// name is not a device address: its address on each device is looked up once
// (noalloc_scratch_setup, with that device current) and only that pointer is
// used.  Every call on one device shares the array, so calls that overlap on a
// device race; each call overwrites the slots of all its cells before reading.
static __device__ __align__(256) double scratch[HC_CELL_STATS_QUANTITIES * (size_t)HC_CELL_STATS_NOALLOC_MAX];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<double*>(p);
  return hipSuccess;
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace hcg
