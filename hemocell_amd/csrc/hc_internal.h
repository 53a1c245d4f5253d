// hc_internal.h -- shared declarations of libhcgpu (host side and kernels).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/hemocell_gpu.h"

namespace hcg {

// Fixed-width per-vertex copies of the gather lists (entries past the end have a
// negative first field): a vertex then issues all its table loads at once.
constexpr int kPadW = 7;  // incident triangles / edge entries
constexpr int kPadB = 8;  // bending contributors (the vertex and its 1-ring)

// Rest-geometry tables of one cell type, in the layout the mechanics kernels read.
// The first block mirrors CommonCellConstants (mechanics/commonCellConstants.cpp:70-409);
// the *_off/_list pairs are per-vertex gather lists that visit every force
// contribution in the order the reference's scatter loops add it to sv.force.
struct TypeHost {
  int model = HC_MODEL_RBC;
  int nv = 0, nt = 0, ne = 0, ni = 0;
  std::vector<int32_t> tri;         // nt*3
  std::vector<double> rest;         // nv*3
  std::vector<int32_t> edge;        // ne*2 (a<b, triangle order)
  std::vector<int32_t> edge_tris;   // ne*2, adjacent triangles ascending
  std::vector<int32_t> edge_outer;  // ne*2, vertex opposite the edge in each
  std::vector<int32_t> ring;        // nv*6, ordered 1-ring
  std::vector<int32_t> nring;       // nv
  std::vector<int32_t> inner;       // ni*2
  std::vector<double> edge_len_eq, edge_angle_eq, tri_area_eq, patch_dist_eq, inner_len_eq;
  double volume_eq = 0, area_mean_eq = 0, edge_mean_eq = 0, angle_mean_eq = 0;
  // gather lists (CSR): entries per vertex in accumulation order
  std::vector<int32_t> vtri_off, vtri;      // t*4 + corner slot
  std::vector<int32_t> vbend_off, vbend;    // contributing vertex i (own term when i == k)
  std::vector<int32_t> vedge_off, vedge;    // e*4 + kind (0 first end, 1 second end, 2 opposite vertex)
  std::vector<int32_t> vinner_off, vinner;  // ie*2 + end (0 first, 1 second)
  std::vector<int32_t> vtri4;   // per vtri entry: triangle, its three corners
  std::vector<int32_t> vedge4;  // per vedge entry: e*4 + kind, the edge's two ends, 0
  bool padded = false;          // every list fits kPadW / kPadB
  std::vector<int32_t> vtri4p, vedge4p, vbendp;  // nv*kPadW int4, nv*kPadW int4, nv*kPadB int
  // per padded slot: the rest area of its triangle, the rest length of its edge and
  // the ring size of its bending contributor (addresses known from the vertex alone)
  std::vector<double> vtri_aeq, vedge_l0, vbend_n;
  // Compact copies for the FP32 RBC kernels (padded lists, nv < 65535): per vertex
  // and slot j < kPadW one uint4 {corner0 | corner1 << 16, corner2 | edge partner
  // << 16, rest area (float bits), rest length (float bits)} of its j-th triangle
  // and j-th edge (0xFFFF: none); per bending slot j < kPadB the contributor i |
  // its ring size << 16 (all ones: none); per vertex one uint4 {ring0 | ring1 <<
  // 16, ring2 | ring3 << 16, ring4 | ring5 << 16, rest patch distance (float)}.
  // Edges are stored from the vertex's side: the force of an edge on its second
  // end, -F(b - a), equals F(a - b) up to the rounding of fused multiply-adds.
  std::vector<uint32_t> cslot, cbend, cring;
  // Ring-ordered tables for the FP32 RBC staged kernel, built when every vertex's
  // triangles are (v, ring[j], ring[j+1]) up to rotation (the ring is walked along
  // the triangles' winding, so this holds for a closed, consistently wound mesh),
  // its edges go to ring[j] and its bending contributors are v and its ring: per
  // vertex 12 floats (rest areas of triangles j = 0..5, then rest lengths of edges
  // j) and one word: ring size | ring size of ring[j] << (4 + 3j).
  std::vector<float> rpar;
  std::vector<uint32_t> rnb;
  hc_model_params p{};
  int timescale = 1;
};

bool build_type_tables(TypeHost& t, std::string& err);
bool build_mesh(int shape, double radius_lu, int min_triangles, double aspect_ratio,
                std::vector<double>& verts, std::vector<int32_t>& tris, std::string& err);

// Device view of one cell type (passed by value to the mechanics kernels).
struct TypeDev {
  int model, nv, nt, ne, ni;
  const int32_t *tri, *edge, *edge_tris, *edge_outer, *ring, *nring, *inner;
  const double *edge_len_eq, *edge_angle_eq, *tri_area_eq, *patch_dist_eq, *inner_len_eq;
  const int32_t *vtri_off, *vtri, *vbend_off, *vbend, *vedge_off, *vedge, *vinner_off, *vinner;
  const int4 *vtri4, *vedge4;
  const int4 *vtri4p, *vedge4p;  // fixed-width copies (null when a list does not fit)
  const int32_t* vbendp;
  const double *vtri_aeq, *vedge_l0, *vbend_n;
  const uint4 *cslot, *cring;  // compact FP32 tables (null when not built)
  const uint32_t* cbend;
  const float4* rpar;  // ring-ordered FP32 tables (null when not built)
  const uint32_t* rnb;
  double volume_eq, area_mean_eq, edge_mean_eq;
  double k_link, k_bend, k_volume, k_area, eta_m, k_inner_rigid, k_cyto, core_radius, radius;
};

// Lattice geometry passed to every lattice / IBM kernel.
struct LatticeDev {
  int nx, ny, nz;
  int periodic[3];
  int64_t N;       // nodes
  int64_t stride;  // population / force component stride (>= N, padded)
  // Exact spread (FP64 only): every force contribution is rounded to a multiple
  // of the power of two spread_q before it is added, so the sums are exact and
  // do not depend on the order of the atomic adds; 0 = plain adds.
  double spread_q, spread_inv_q;
};

}  // namespace hcg
