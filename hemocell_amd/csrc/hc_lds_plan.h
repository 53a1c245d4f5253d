// hc_lds_plan.h -- the dynamic LDS layout of the per-cell kernels as plain
// arithmetic (no HIP types), shared by the kernels, their launchers and a CPU
// test that checks every layout against the workgroup limit.
//
// Dynamic LDS of one cell, in order:
//   fixed   positions (3 nv scalars) and, staged with a viscous membrane, the
//           vertex velocities (3 nv more), rounded up to 16 bytes; in the plain
//           layout 3 nv doubles; nothing when the positions stay in global memory;
//   region  the bending forces of the mechanics (staged, 3 nv scalars) and then
//           the spread box (3 doubles per node) over the same bytes;
//   lists   the spread's bucketed vertex list (2 bytes per vertex), which the
//           kernel places right after the box's last node.
#pragma once
#include <algorithm>
#include <cstddef>

namespace hcg {

enum { LDS_MODEL_RBC = 0, LDS_MODEL_PLT = 1, LDS_MODEL_WBC = 2 };  // the hc_model values
constexpr size_t kLdsMargin = 256;  // bytes kept free beside static and dynamic LDS

// bytes before the region
constexpr size_t cell_fixed_bytes(int nv, bool visc, size_t psize, bool staged, bool gpos) {
  return gpos ? 0 : (staged ? (((visc ? 6 : 3) * (size_t)nv * psize + 15) & ~(size_t)15) : 3 * (size_t)nv * sizeof(double));
}

constexpr size_t cell_list_bytes(int nv) { return (2 * (size_t)nv + 15) & ~(size_t)15; }

struct CellLdsPlan {
  bool ok;        // fits beside the kernel's static LDS within the workgroup limit
  bool staged;    // the layout the launch uses (staged may fall back to plain)
  size_t fixed, region, lists, dyn;
  int box_nodes;  // spread box nodes: 3 doubles each, within the region
};

// One cell type's layout for a launch of the per-type kernel.  st_* are the
// static LDS bytes of the staged, plain and global-position instantiations.
// Staged when positions, velocities, bending forces and the box fit; otherwise
// positions plus as much box as fits; with gpos only the box.
inline CellLdsPlan cell_lds_plan(int model, int nv, bool visc, size_t psize, size_t box_bytes, size_t lds_device,
                                 size_t st_staged, size_t st_plain, size_t st_gpos, bool gpos) {
  CellLdsPlan p{};
  const size_t bend = model != LDS_MODEL_PLT ? 3 * (size_t)nv * psize : 0;
  p.lists = cell_list_bytes(nv);
  p.staged = !gpos;
  p.fixed = cell_fixed_bytes(nv, visc, psize, p.staged, gpos);
  p.region = p.staged ? std::max(bend, box_bytes) : box_bytes;
  if (p.staged && p.fixed + p.region + p.lists + st_staged + kLdsMargin > lds_device) {
    p.staged = false;
    p.fixed = cell_fixed_bytes(nv, visc, sizeof(double), false, false);
    const size_t used = p.fixed + p.lists + st_plain + kLdsMargin;
    p.region = std::min(box_bytes, lds_device > used ? lds_device - used : 0);
  }
  const size_t st = p.staged ? st_staged : (gpos ? st_gpos : st_plain);
  p.dyn = p.fixed + p.region + p.lists;
  p.ok = p.dyn + st + kLdsMargin <= lds_device;
  p.box_nodes = (int)(p.region / (3 * sizeof(double)));
  return p;
}

// The persistent mixed launch: one dynamic size for every type (staged RBC and
// platelets, WBCs with positions in global memory), each type's box filling
// what its own fixed part and list leave.  Returns the dynamic bytes, 0 when
// they do not fit beside the kernel's static LDS `st`.
inline size_t mixed_lds_plan(int ntypes, const int* model, const int* nv, const bool* visc, size_t psize,
                             size_t box_bytes, size_t lds_device, size_t st, int* box_nodes) {
  size_t dyn = 0;
  for (int t = 0; t < ntypes; ++t) {
    const bool gpos = model[t] == LDS_MODEL_WBC;
    const size_t bend = (!gpos && model[t] != LDS_MODEL_PLT) ? 3 * (size_t)nv[t] * psize : 0;
    const size_t fixed = cell_fixed_bytes(nv[t], visc[t], psize, !gpos, gpos);
    dyn = std::max(dyn, fixed + std::max(bend, box_bytes) + cell_list_bytes(nv[t]));
  }
  dyn = (dyn + 15) & ~(size_t)15;
  if (dyn + st + kLdsMargin > lds_device) return 0;
  for (int t = 0; t < ntypes; ++t) {
    const bool gpos = model[t] == LDS_MODEL_WBC;
    const size_t fixed = cell_fixed_bytes(nv[t], visc[t], psize, !gpos, gpos);
    box_nodes[t] = (int)((dyn - fixed - cell_list_bytes(nv[t])) / (3 * sizeof(double)));
  }
  return dyn;
}

}  // namespace hcg
