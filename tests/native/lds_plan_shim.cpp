// C entry points over hc_lds_plan.h for the CPU test of the per-cell kernels'
// LDS layouts (tests/test_lds_plan.py builds this file with g++).
#include <cstdint>

#include "../../hemocell_amd/csrc/hc_lds_plan.h"

extern "C" {

// out: ok, staged, fixed, region, lists, dyn, box_nodes
void plan_cell(int model, int nv, int visc, int64_t psize, int64_t box_bytes, int64_t lds_device, int64_t st_staged,
               int64_t st_plain, int64_t st_gpos, int gpos, int64_t* out) {
  const hcg::CellLdsPlan p = hcg::cell_lds_plan(model, nv, visc != 0, (size_t)psize, (size_t)box_bytes,
                                                (size_t)lds_device, (size_t)st_staged, (size_t)st_plain,
                                                (size_t)st_gpos, gpos != 0);
  out[0] = p.ok;
  out[1] = p.staged;
  out[2] = (int64_t)p.fixed;
  out[3] = (int64_t)p.region;
  out[4] = (int64_t)p.lists;
  out[5] = (int64_t)p.dyn;
  out[6] = p.box_nodes;
}

int64_t plan_mixed(int ntypes, const int* model, const int* nv, const int* visc, int64_t psize, int64_t box_bytes,
                   int64_t lds_device, int64_t st, int* box_nodes) {
  bool v[16];
  for (int t = 0; t < ntypes && t < 16; ++t) v[t] = visc[t] != 0;
  return (int64_t)hcg::mixed_lds_plan(ntypes, model, nv, v, (size_t)psize, (size_t)box_bytes, (size_t)lds_device,
                                      (size_t)st, box_nodes);
}

int64_t fixed_bytes(int nv, int visc, int64_t psize, int staged, int gpos) {
  return (int64_t)hcg::cell_fixed_bytes(nv, visc != 0, (size_t)psize, staged != 0, gpos != 0);
}
}
