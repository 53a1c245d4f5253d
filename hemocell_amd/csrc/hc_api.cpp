// hc_api.cpp -- the C ABI of libhcgpu (include/hemocell_gpu.h): device state,
// the time step in the order of HemoCell::iterate() (core/hemoCell.cpp:299-376),
// state access and timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <map>
#include <tuple>
#include <sstream>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "hc_launch.h"

using namespace hcg;

namespace {

struct TypeBufs {
  TypeHost host;
  std::vector<void*> allocs;
  TypeDev dev{};
  int32_t* d_cells = nullptr;  // cells of this type
  int ncells = 0;
};

struct Prof {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  double total_ms[HC_PROF_NSTAGES] = {0};
  int64_t count[HC_PROF_NSTAGES] = {0};
};

}  // namespace

struct hc_ctx {
  int device = 0;
  bool noalloc_ok = false;  // the statistics staging array of this device was found
  hipStream_t stream = nullptr;
  // the cell kernels of different types run side by side: two more streams that
  // fork from and join back into the context stream within each step
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork_ev = nullptr, join_ev[2] = {nullptr, nullptr};
  std::string err;
  const char* err_fixed = nullptr;  // a static message set without allocating (takes precedence)
  // lattice
  bool lattice_ready = false;
  LatticeDev L{};
  int precision = HC_FP64;
  double tau = 1.0;
  void* f = nullptr;
  void* Fibm[2] = {nullptr, nullptr};
  void* ufield = nullptr;  // fluid velocity per node, filled before the interpolation
  uint8_t* type = nullptr;
  uint32_t* vel_nodes = nullptr;  // nodes with a velocity boundary
  int n_vel = 0;
  bool general = false;
  bool has_walls = false;  // a node of the mask is not bulk fluid
  int cur = 0;
  bool other_clean = true;
  int phase = 0;  // AA layout: 0 natural, 1 after an even step
  double g[3] = {0, 0, 0};
  double wall_u[6][3] = {{0}};
  // cells
  std::vector<std::unique_ptr<TypeBufs>> types;
  std::vector<int32_t> cell_type, cell_id;
  std::vector<double> cell_ctr;  // per cell: vertex mean when added (3 doubles), for the launch order
  bool cell_order = true;        // launch cells in spatial order (HC_CELL_ORDER=0: in the order added)
  std::vector<int64_t> cell_base;
  int64_t nv = 0, vcap = 0;
  double *x = nullptr, *v = nullptr, *fv = nullptr, *frep = nullptr, *bend = nullptr, *fbr = nullptr;
  uint8_t* vflag = nullptr;  // per vertex: VFLAG_REMOVED
  int64_t* d_cell_base = nullptr;
  int32_t* d_cell_nv = nullptr;
  bool cells_dirty = true;
  int lds_limit = 65536;        // dynamic LDS per workgroup available on the device
  int lds_device = 65536;       // LDS per workgroup of the device (static + dynamic)
  bool spread_per_cell = true;  // LDS-staged spread (false: one thread per vertex, global atomics)
  int spread_lds = 0;           // LDS box budget in bytes (0: default per precision)
  bool side_gpos = true;        // large-cell side types read positions from global memory
  int cell_block = 256;         // threads per RBC cell in the fused vertex kernel
  int wbc_block = 1024;         // threads per WBC cell (falls back when the LDS does not fit)
  bool cell_block_env = false;  // the RBC block was set explicitly
  int n_cu = 256;               // compute units of the device
  bool fuse_vel = true;         // velocity-wall nodes inside the collide-stream kernel
  // every cell type in one persistent launch (instead of one kernel per type side
  // by side): the queue of (cell, type) entries, most vertices first, the types'
  // device descriptors, the queue head, and workgroups per compute unit
  bool mixed = false;
  int mixed_per_cu = 8;
  int2* d_queue = nullptr;
  TypeDev* d_types = nullptr;
  unsigned int* d_head = nullptr;
  int nqueue = 0;
  // extra (stretch) forces
  int64_t* d_extra_ids = nullptr;
  double* d_extra_f = nullptr;
  int n_extra = 0;
  double* d_extra_dense = nullptr;  // the same forces per vertex (3 planes), for the fused step
  std::vector<int64_t> h_extra_ids;
  std::vector<double> h_extra_f;
  // fused step: the vertex kernels of step t also spread the forces of step t+1
  bool fuse = true;
  bool spread_ready = false;  // Fibm[cur] already holds this step's spread
  // steady fused steps are replayed from a captured graph of two steps (the layout
  // parity and the force buffers alternate, so two steps bring every argument back)
  bool use_graphs = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> graph_sig;  // launch arguments the graph was captured with
  // controls
  int64_t iter = 0;
  int pv_timescale = 1;
  double f_limit = 1e300;
  bool f_limit_set = false;
  bool exact_spread = false;  // FP64: spread contributions rounded to L.spread_q
  bool repulsion = false;
  double rep_k = 0, rep_cutoff = 0;
  int rep_timescale = 1;
  // wall repulsion from boundary particles (flags per node, built on first use)
  bool brep = false;
  double brep_k = 0, brep_cutoff = 0;
  int brep_timescale = 1;
  uint8_t* d_bflag = nullptr;
  // Lees-Edwards planes z = 0 and z = nz-1: plane velocities, displacement per
  // step and the planes' scratch populations
  bool le = false;
  double le_top_u = 0, le_bottom_u = 0, le_disp = 0;
  void* d_le = nullptr;
  // batched steps led by this context: two descriptor buffers (device and
  // pinned host staging) used in turn, each with the event of the last batch
  // that read it, so the host prepares one batch while the previous one runs
  void* d_batch[2] = {nullptr, nullptr};
  void* h_batch[2] = {nullptr, nullptr};
  size_t batch_cap[2] = {0, 0};
  hipEvent_t batch_slot_ev[2] = {nullptr, nullptr};
  int batch_slot = 0;
  hipEvent_t batch_ev = nullptr;  // this context's queued work / the batch's end
  // interior viscosity: tau inside the membranes of each type (0 = off) and the
  // two refresh timescales; node classes, last-writer stamps of the membrane
  // pass (zero between passes) and the positions before the advance
  std::vector<double> iv_tau;
  int iv_timescale = 1, iv_grid_timescale = 1;
  uint8_t* d_visc = nullptr;
  uint32_t* d_stamp = nullptr;
  double* d_xold = nullptr;
  int64_t xold_cap = 0;
  // binning and repulsion buffers, built on first use and whenever the vertex count changed
  RepulsionBufs rb{};
  int64_t rb_nv = -1;
  int32_t* d_vcell = nullptr;
  unsigned long long* d_wall_hits = nullptr;
  unsigned int* d_flag = nullptr;
  Prof prof;
};

namespace {

hc_status fail(hc_ctx* c, hc_status s, const std::string& m) {
  if (c) {
    c->err = m;
    c->err_fixed = nullptr;
  }
  return s;
}
// an error with a static message: sets nothing that allocates
hc_status fail_fixed(hc_ctx* c, hc_status s, const char* m) {
  if (c) c->err_fixed = m;
  return s;
}
hc_status hipfail(hc_ctx* c, hipError_t e, const char* where) {
  return fail(c, HC_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}
#define HC_TRY(expr, where)                  \
  do {                                       \
    hipError_t e_ = (expr);                  \
    if (e_ != hipSuccess) return hipfail(ctx, e_, where); \
  } while (0)

size_t rsize(const hc_ctx* c) { return c->precision == HC_FP32_SHIFTED ? sizeof(float) : sizeof(double); }

hc_status dalloc(hc_ctx* ctx, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, HC_OOM, std::string("device allocation of ") + what + " (" + std::to_string(bytes) + " bytes) failed");
  }
  return HC_OK;
}

VertexDev vdev(hc_ctx* c) { return VertexDev{c->x, c->v, c->fv, c->frep, c->fbr, c->nv, c->vcap, c->vflag}; }

void prof_begin(hc_ctx* c, int stage, hipEvent_t* start) {
  *start = nullptr;
  if (!c->prof.on) return;
  if (c->prof.used + 2 > c->prof.pool.size()) {
    for (int k = 0; k < 256; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      c->prof.pool.push_back(e);
    }
  }
  *start = c->prof.pool[c->prof.used++];
  hipEventRecord(*start, c->stream);
  (void)stage;
}
void prof_end(hc_ctx* c, int stage, hipEvent_t start) {
  if (!c->prof.on || !start) return;
  hipEvent_t e = c->prof.pool[c->prof.used++];
  hipEventRecord(e, c->stream);
  c->prof.recs.push_back({stage, start, e});
}
// fold recorded event pairs into the totals (synchronises the stream)
void prof_flush(hc_ctx* c) {
  if (c->prof.recs.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& r : c->prof.recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->prof.total_ms[r.stage] += ms;
      c->prof.count[r.stage] += 1;
    }
  }
  c->prof.recs.clear();
  c->prof.used = 0;
}

hc_status need_lattice(hc_ctx* ctx) {
  if (!ctx) return HC_INVALID_ARG;
  if (!ctx->lattice_ready) return fail(ctx, HC_STATE, "lattice not initialised (call hc_lattice_init first)");
  return HC_OK;
}

// grow the vertex arrays to hold `need` vertices (keeps the contents)
hc_status grow_vertices(hc_ctx* ctx, int64_t need) {
  if (need <= ctx->vcap) return HC_OK;
  int64_t cap = std::max<int64_t>(need, ctx->vcap * 2);
  cap = (cap + 63) / 64 * 64;
  double* arrs[6] = {ctx->x, ctx->v, ctx->fv, ctx->frep, ctx->bend, ctx->fbr};
  double* nw[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int a = 0; a < 6; ++a) {
    hc_status s = dalloc(ctx, (void**)&nw[a], sizeof(double) * 3 * cap, "vertex state");
    if (s != HC_OK) {
      for (int b = 0; b < a; ++b) hipFree(nw[b]);
      return s;
    }
    HC_TRY(hipMemsetAsync(nw[a], 0, sizeof(double) * 3 * cap, ctx->stream), "hipMemsetAsync");
    if (arrs[a] && ctx->nv > 0)
      for (int c = 0; c < 3; ++c)
        HC_TRY(hipMemcpyAsync(nw[a] + c * cap, arrs[a] + c * ctx->vcap, sizeof(double) * ctx->nv,
                              hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpyAsync");
  }
  uint8_t* nflag = nullptr;
  hc_status fs = dalloc(ctx, (void**)&nflag, (size_t)cap, "vertex flags");
  if (fs != HC_OK) {
    for (int b = 0; b < 6; ++b) hipFree(nw[b]);
    return fs;
  }
  HC_TRY(hipMemsetAsync(nflag, 0, (size_t)cap, ctx->stream), "hipMemsetAsync");
  if (ctx->vflag && ctx->nv > 0)
    HC_TRY(hipMemcpyAsync(nflag, ctx->vflag, (size_t)ctx->nv, hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpyAsync");
  HC_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  for (int a = 0; a < 6; ++a)
    if (arrs[a]) hipFree(arrs[a]);
  if (ctx->vflag) hipFree(ctx->vflag);
  ctx->vflag = nflag;
  ctx->x = nw[0];
  ctx->v = nw[1];
  ctx->fv = nw[2];
  ctx->frep = nw[3];
  ctx->bend = nw[4];
  ctx->fbr = nw[5];
  ctx->vcap = cap;
  return HC_OK;
}

// upload per-type cell lists and cell bases after cells were added
hc_status sync_cells(hc_ctx* ctx) {
  if (!ctx->cells_dirty) return HC_OK;
  if (ctx->d_cell_base) hipFree(ctx->d_cell_base);
  ctx->d_cell_base = nullptr;
  size_t nc = ctx->cell_base.size();
  hc_status s = dalloc(ctx, (void**)&ctx->d_cell_base, sizeof(int64_t) * std::max<size_t>(nc, 1), "cell bases");
  if (s != HC_OK) return s;
  if (nc) HC_TRY(hipMemcpy(ctx->d_cell_base, ctx->cell_base.data(), sizeof(int64_t) * nc, hipMemcpyHostToDevice), "hipMemcpy");
  if (ctx->d_cell_nv) hipFree(ctx->d_cell_nv);
  ctx->d_cell_nv = nullptr;
  s = dalloc(ctx, (void**)&ctx->d_cell_nv, sizeof(int32_t) * std::max<size_t>(nc, 1), "cell sizes");
  if (s != HC_OK) return s;
  {
    std::vector<int32_t> cnv(nc);
    for (size_t c = 0; c < nc; ++c) cnv[c] = ctx->types[ctx->cell_type[c]]->host.nv;
    if (nc) HC_TRY(hipMemcpy(ctx->d_cell_nv, cnv.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice), "hipMemcpy");
  }
  // Launch order: the cells of a type by the 8-node tile of their centre, tiles in
  // the lattice's x, y, z order (z fastest, as the node index runs).  Workgroups
  // resident at the same time then cover a compact region: the support nodes
  // they gather velocities from and spread forces into share cache lines.
  std::vector<int64_t> key(nc, 0);
  if (ctx->cell_order && ctx->lattice_ready) {
    const int dims[3] = {ctx->L.nx, ctx->L.ny, ctx->L.nz};
    int64_t tiles[3];
    for (int a = 0; a < 3; ++a) tiles[a] = (dims[a] + 7) / 8;
    for (size_t c = 0; c < nc; ++c) {
      int64_t q[3];
      for (int a = 0; a < 3; ++a) {
        const double v = ctx->cell_ctr[3 * c + a];
        int64_t n = std::isfinite(v) ? (int64_t)std::floor(v) : 0;
        n = ((n % dims[a]) + dims[a]) % dims[a];
        q[a] = n / 8;
      }
      key[c] = (q[0] * tiles[1] + q[1]) * tiles[2] + q[2];
    }
  }
  for (size_t t = 0; t < ctx->types.size(); ++t) {
    TypeBufs& tb = *ctx->types[t];
    std::vector<int32_t> list;
    for (size_t c = 0; c < nc; ++c)
      if (ctx->cell_type[c] == (int)t) list.push_back((int32_t)c);
    std::stable_sort(list.begin(), list.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    if (tb.d_cells) hipFree(tb.d_cells);
    tb.d_cells = nullptr;
    s = dalloc(ctx, (void**)&tb.d_cells, sizeof(int32_t) * std::max<size_t>(list.size(), 1), "cell list");
    if (s != HC_OK) return s;
    if (!list.empty())
      HC_TRY(hipMemcpy(tb.d_cells, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice), "hipMemcpy");
    tb.ncells = (int)list.size();
  }
  // the persistent launch's queue: types by vertex count, descending
  if (ctx->d_queue) hipFree(ctx->d_queue);
  if (ctx->d_types) hipFree(ctx->d_types);
  ctx->d_queue = nullptr;
  ctx->d_types = nullptr;
  ctx->nqueue = 0;
  if (nc && ctx->types.size() <= (size_t)kMaxMixedTypes) {
    std::vector<int> order(ctx->types.size());
    for (size_t t = 0; t < order.size(); ++t) order[t] = (int)t;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return ctx->types[a]->host.nv > ctx->types[b]->host.nv; });
    std::vector<int2> q;
    q.reserve(nc);
    for (int t : order) {
      const size_t first = q.size();
      for (size_t c = 0; c < nc; ++c)
        if (ctx->cell_type[c] == t) q.push_back(make_int2((int)c, t));
      std::stable_sort(q.begin() + first, q.end(), [&](int2 a, int2 b) { return key[a.x] < key[b.x]; });
    }
    std::vector<TypeDev> td(ctx->types.size());
    for (size_t t = 0; t < td.size(); ++t) td[t] = ctx->types[t]->dev;
    if ((s = dalloc(ctx, (void**)&ctx->d_queue, sizeof(int2) * q.size(), "cell queue")) != HC_OK) return s;
    if ((s = dalloc(ctx, (void**)&ctx->d_types, sizeof(TypeDev) * td.size(), "cell type descriptors")) != HC_OK)
      return s;
    HC_TRY(hipMemcpy(ctx->d_queue, q.data(), sizeof(int2) * q.size(), hipMemcpyHostToDevice), "hipMemcpy");
    HC_TRY(hipMemcpy(ctx->d_types, td.data(), sizeof(TypeDev) * td.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (!ctx->d_head && (s = dalloc(ctx, (void**)&ctx->d_head, 256, "cell queue head")) != HC_OK) return s;
    ctx->nqueue = (int)q.size();
  }
  ctx->cells_dirty = false;
  return HC_OK;
}

// tuning: HC_MECH_TABLES=csr keeps the mechanics on the variable-length lists
bool padded_off(const hc_ctx*) {
  const char* e = std::getenv("HC_MECH_TABLES");
  return e && std::string(e) == "csr";
}
// tuning: HC_MECH_TABLES=compact keeps the FP32 RBC mechanics on the per-slot tables
static bool ring_tables_off() {
  const char* e = std::getenv("HC_MECH_TABLES");
  return e && std::string(e) == "compact";
}

template <typename T>
hc_status upload(hc_ctx* ctx, TypeBufs& tb, const std::vector<T>& h, const T** out) {
  void* p = nullptr;
  hc_status s = dalloc(ctx, &p, sizeof(T) * std::max<size_t>(h.size(), 1), "cell type tables");
  if (s != HC_OK) return s;
  tb.allocs.push_back(p);
  if (!h.empty()) HC_TRY(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), "hipMemcpy");
  *out = (const T*)p;
  return HC_OK;
}

// Drop a spread done ahead by the fused step (the host changed what it used).
hc_status invalidate_spread(hc_ctx* ctx) {
  if (!ctx->spread_ready) return HC_OK;
  hipError_t e = ctx->precision == HC_FP32_SHIFTED
                     ? launch_fill_zero<float>(ctx->stream, (float*)ctx->Fibm[ctx->cur], 3 * ctx->L.stride)
                     : launch_fill_zero<double>(ctx->stream, (double*)ctx->Fibm[ctx->cur], 3 * ctx->L.stride);
  if (e != hipSuccess) return hipfail(ctx, e, "invalidate spread");
  ctx->spread_ready = false;
  return HC_OK;
}

// Per-vertex planes of the extra forces, read by the fused step.
hc_status build_dense_extra(hc_ctx* ctx) {
  if (ctx->d_extra_dense) hipFree(ctx->d_extra_dense);
  ctx->d_extra_dense = nullptr;
  if (ctx->h_extra_ids.empty()) return HC_OK;
  hc_status s = dalloc(ctx, (void**)&ctx->d_extra_dense, sizeof(double) * 3 * ctx->vcap, "extra forces");
  if (s != HC_OK) return s;
  std::vector<double> h(3 * ctx->vcap, 0.0);
  for (size_t k = 0; k < ctx->h_extra_ids.size(); ++k)
    for (int c = 0; c < 3; ++c) h[c * ctx->vcap + ctx->h_extra_ids[k]] += ctx->h_extra_f[3 * k + c];
  HC_TRY(hipMemcpy(ctx->d_extra_dense, h.data(), sizeof(double) * 3 * ctx->vcap, hipMemcpyHostToDevice), "hipMemcpy");
  return HC_OK;
}

void free_repulsion(hc_ctx* ctx) {
  void* ps[] = {ctx->rb.keys_in, ctx->rb.keys_out, ctx->rb.ids_in, ctx->rb.ids_out, ctx->rb.temp, ctx->d_vcell};
  for (void* p : ps)
    if (p) hipFree(p);
  RepulsionBufs z{};
  z.bins = ctx->rb.bins;
  ctx->rb = z;
  ctx->d_vcell = nullptr;
  ctx->rb_nv = -1;
}

// Sort buffers sized for the current vertices, the cell index of every vertex,
// and the per-node bin bounds (zeroed once; the repulsion launch leaves them zero).
hc_status ensure_repulsion(hc_ctx* ctx) {
  hc_status s;
  if (!ctx->rb.bins) {
    if ((s = dalloc(ctx, (void**)&ctx->rb.bins, sizeof(int2) * ctx->L.N, "repulsion bins")) != HC_OK) return s;
    HC_TRY(hipMemset(ctx->rb.bins, 0, sizeof(int2) * ctx->L.N), "hipMemset");
  }
  if (ctx->rb_nv == ctx->nv) return HC_OK;
  HC_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  free_repulsion(ctx);
  const int64_t nv = std::max<int64_t>(ctx->nv, 1);
  void** bufs[4] = {(void**)&ctx->rb.keys_in, (void**)&ctx->rb.keys_out, (void**)&ctx->rb.ids_in, (void**)&ctx->rb.ids_out};
  for (void** b : bufs)
    if ((s = dalloc(ctx, b, 4 * nv, "repulsion sort buffers")) != HC_OK) return s;
  ctx->rb.temp_bytes = repulsion_temp_bytes(ctx->nv, ctx->L.N);
  if ((s = dalloc(ctx, &ctx->rb.temp, std::max<size_t>(ctx->rb.temp_bytes, 16), "radix-sort scratch")) != HC_OK) return s;
  std::vector<int32_t> vcell((size_t)nv, 0);
  for (size_t c = 0; c < ctx->cell_base.size(); ++c) {
    const int n = ctx->types[ctx->cell_type[c]]->host.nv;
    for (int k = 0; k < n; ++k) vcell[(size_t)ctx->cell_base[c] + k] = (int32_t)c;
  }
  if ((s = dalloc(ctx, (void**)&ctx->d_vcell, sizeof(int32_t) * nv, "vertex cells")) != HC_OK) return s;
  HC_TRY(hipMemcpy(ctx->d_vcell, vcell.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice), "hipMemcpy");
  ctx->rb.vcell = ctx->d_vcell;
  ctx->rb_nv = ctx->nv;
  return HC_OK;
}

hc_status do_repulsion(hc_ctx* ctx) {
  hc_status s = ensure_repulsion(ctx);
  if (s != HC_OK) return s;
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_REPULSION, &ev);
  HC_TRY(launch_repulsion(ctx->stream, vdev(ctx), ctx->L, ctx->rb, ctx->rep_k, ctx->rep_cutoff), "repulsion");
  prof_end(ctx, HC_PROF_REPULSION, ev);
  return HC_OK;
}

hc_status do_boundary_repulsion(hc_ctx* ctx) {
  hc_status s;
  if (!ctx->d_bflag) {
    if ((s = dalloc(ctx, (void**)&ctx->d_bflag, std::max<int64_t>(ctx->L.N, 1), "boundary particles")) != HC_OK)
      return s;
    HC_TRY(launch_boundary_flags(ctx->stream, ctx->type, ctx->L, ctx->d_bflag), "boundary particles");
  }
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_REPULSION, &ev);
  HC_TRY(launch_boundary_repulsion(ctx->stream, vdev(ctx), ctx->L, ctx->d_bflag, ctx->brep_k, ctx->brep_cutoff),
         "boundary repulsion");
  prof_end(ctx, HC_PROF_REPULSION, ev);
  return HC_OK;
}

// the repulsion terms the spread adds: 1 cell-cell, 2 wall
int rep_mask(const hc_ctx* c) { return (c->repulsion ? 1 : 0) | (c->brep ? 2 : 0); }

double iv_tau_of(const hc_ctx* c, size_t t) { return t < c->iv_tau.size() ? c->iv_tau[t] : 0.0; }
bool iv_active(const hc_ctx* c) {
  for (double t : c->iv_tau)
    if (t > 0.0) return true;
  return false;
}

// node classes and stamps (zeroed once), positions-before-advance buffer
hc_status ensure_visc(hc_ctx* ctx) {
  hc_status s;
  const int64_t N = std::max<int64_t>(ctx->L.N, 1);
  if (!ctx->d_visc) {
    if ((s = dalloc(ctx, (void**)&ctx->d_visc, N, "interior viscosity classes")) != HC_OK) return s;
    HC_TRY(launch_visc_clear(ctx->stream, ctx->d_visc, ctx->L.N), "interior viscosity classes");
  }
  if (!ctx->d_stamp) {
    if ((s = dalloc(ctx, (void**)&ctx->d_stamp, sizeof(uint32_t) * N, "interior viscosity stamps")) != HC_OK) return s;
    HC_TRY(hipMemsetAsync(ctx->d_stamp, 0, sizeof(uint32_t) * N, ctx->stream), "hipMemsetAsync");
  }
  if (ctx->xold_cap != ctx->vcap) {
    HC_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (ctx->d_xold) hipFree(ctx->d_xold);
    ctx->d_xold = nullptr;
    ctx->xold_cap = 0;
    if ((s = dalloc(ctx, (void**)&ctx->d_xold, sizeof(double) * 3 * std::max<int64_t>(ctx->vcap, 1),
                    "positions before the advance")) != HC_OK)
      return s;
    ctx->xold_cap = ctx->vcap;
  }
  return HC_OK;
}

// the collide's per-class relaxation rates (classes of disabled types relax
// with the background rate)
void visc_args(const hc_ctx* ctx, CollideArgs& a) {
  if (!ctx->d_visc || !iv_active(ctx)) return;
  a.visc = ctx->d_visc;
  for (int k = 0; k < 16; ++k) a.omega_cls[k] = a.omega;
  for (size_t t = 0; t < ctx->iv_tau.size() && t < 15; ++t)
    if (ctx->iv_tau[t] > 0.0) a.omega_cls[t + 1] = 1.0 / ctx->iv_tau[t];
}

// entire-grid pass: every class back to 0, then the inner nodes of each cell
hc_status do_visc_grid(hc_ctx* ctx) {
  hc_status s = ensure_visc(ctx);
  if (s != HC_OK) return s;
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_MECH, &ev);
  HC_TRY(launch_visc_clear(ctx->stream, ctx->d_visc, ctx->L.N), "interior viscosity classes");
  HC_TRY(launch_delete_incomplete(ctx->stream, vdev(ctx), ctx->d_cell_base, ctx->d_cell_nv,
                                  (int)ctx->cell_base.size()),
         "delete incomplete cells");
  for (size_t t = 0; t < ctx->types.size(); ++t) {
    TypeBufs& tb = *ctx->types[t];
    if (!(iv_tau_of(ctx, t) > 0.0) || tb.ncells == 0) continue;
    HC_TRY(launch_visc_grid(ctx->stream, vdev(ctx), tb.dev, tb.d_cells, tb.ncells, ctx->d_cell_base, ctx->L,
                            ctx->type, (int)t + 1, ctx->d_visc),
           "interior viscosity (entire grid)");
  }
  prof_end(ctx, HC_PROF_MECH, ev);
  return HC_OK;
}

// membrane pass with the support of the positions xold (3 planes, stride vcap)
hc_status do_visc_membrane(hc_ctx* ctx, const double* xold) {
  hc_status s = ensure_visc(ctx);
  if (s != HC_OK) return s;
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_MECH, &ev);
  for (int apply = 0; apply < 2; ++apply)
    for (size_t t = 0; t < ctx->types.size(); ++t) {
      TypeBufs& tb = *ctx->types[t];
      if (!(iv_tau_of(ctx, t) > 0.0) || tb.ncells == 0) continue;
      HC_TRY(launch_visc_membrane(ctx->stream, vdev(ctx), tb.dev, tb.d_cells, tb.ncells, ctx->d_cell_base, xold,
                                  ctx->L, ctx->type, (int)t + 1, ctx->d_stamp, ctx->d_visc, apply != 0),
             "interior viscosity (membrane)");
    }
  prof_end(ctx, HC_PROF_MECH, ev);
  return HC_OK;
}

// every type in one persistent launch on the context stream; false when the
// launch does not apply (no queue, or the layout does not fit the LDS)
template <typename R>
bool mixed_cell_step(hc_ctx* ctx, int do_interp, hc_status& st) {
  st = HC_OK;
  if (!ctx->mixed || !ctx->d_queue || ctx->nqueue == 0) return false;
  std::vector<TypeDev> ht(ctx->types.size());
  std::vector<int> flags(ctx->types.size());
  for (size_t t = 0; t < ht.size(); ++t) {
    const TypeBufs& tb = *ctx->types[t];
    ht[t] = tb.dev;
    int f = CELL_FLAG_SPREAD;
    if (do_interp) f |= CELL_FLAG_INTERP;
    if ((ctx->iter % tb.host.timescale) == 0) f |= CELL_FLAG_MECH;
    if (ctx->repulsion) f |= CELL_FLAG_REPULSION;
    flags[t] = f;
  }
  const int box = ctx->spread_lds > 0 ? ctx->spread_lds : 12288;
  const hipError_t e = launch_cell_step_mixed<R>(
      ctx->stream, vdev(ctx), ctx->d_types, ht.data(), (int)ht.size(), flags.data(), ctx->d_queue, ctx->nqueue,
      ctx->d_head, ctx->mixed_per_cu * ctx->n_cu, ctx->d_cell_base, (const R*)ctx->ufield, ctx->type, ctx->L,
      (R*)ctx->Fibm[1 - ctx->cur], ctx->d_extra_dense, ctx->bend, ctx->f_limit, box, ctx->lds_device,
      ctx->d_wall_hits);
  if (e == hipErrorInvalidConfiguration) return false;
  if (e != hipSuccess) st = hipfail(ctx, e, "cell_step (all types)");
  return true;
}

template <typename R>
hc_status do_cell_step(hc_ctx* ctx, int do_interp) {
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_MECH, &ev);
  {
    hc_status st;
    if (mixed_cell_step<R>(ctx, do_interp, st)) {
      prof_end(ctx, HC_PROF_MECH, ev);
      return st;
    }
  }
  // the type with the most vertices runs on the context stream, the others beside it
  std::vector<TypeBufs*> act;
  for (auto& tp : ctx->types)
    if (tp->ncells) act.push_back(tp.get());
  std::stable_sort(act.begin(), act.end(), [](const TypeBufs* a, const TypeBufs* b) {
    return (int64_t)a->ncells * a->host.nv > (int64_t)b->ncells * b->host.nv;
  });
  const int nside = std::min<int>((int)act.size() - 1, ctx->side[0] ? 2 : 0);
  if (nside > 0) {
    HC_TRY(hipEventRecord(ctx->fork_ev, ctx->stream), "hipEventRecord");
    for (int j = 0; j < nside; ++j) HC_TRY(hipStreamWaitEvent(ctx->side[j], ctx->fork_ev, 0), "hipStreamWaitEvent");
  }
  // The side kernels are launched first.  A side type with large cells (the
  // 2562-vertex WBCs) keeps its positions in global memory and runs 256 threads
  // per cell: a workgroup staging them in LDS needs a whole compute unit and
  // would wait until the main type's stream of small workgroups drains.
  for (size_t kk = 0; kk < act.size(); ++kk) {
    const size_t k = (kk + 1) % act.size();
    TypeBufs& tb = *act[k];
    hipStream_t st = (k == 0 || nside == 0) ? ctx->stream : ctx->side[(k - 1) % nside];
    int flags = CELL_FLAG_SPREAD;
    if (do_interp) flags |= CELL_FLAG_INTERP;
    if ((ctx->iter % tb.host.timescale) == 0) flags |= CELL_FLAG_MECH;
    if (ctx->repulsion) flags |= CELL_FLAG_REPULSION;
    // spread box in LDS; larger supports are spread in slabs of x-planes
    const int box = ctx->spread_lds > 0 ? ctx->spread_lds : 12288;
    // few cells: their latency, not the occupancy, sets the time: the widest block
    int block = tb.host.model == HC_MODEL_WBC ? ctx->wbc_block : ctx->cell_block;
    if (tb.ncells < ctx->n_cu && !(ctx->cell_block_env && tb.host.model == HC_MODEL_RBC)) block = 1024;
    const bool gpos = k != 0 && nside > 0 && ctx->side_gpos && 3 * (size_t)tb.host.nv * sizeof(double) > 32768;
    if (gpos) block = 256;
    HC_TRY(launch_cell_step<R>(st, vdev(ctx), tb.dev, tb.d_cells, tb.ncells, ctx->d_cell_base, (const R*)ctx->ufield,
                               ctx->type, ctx->L, (R*)ctx->Fibm[1 - ctx->cur], ctx->d_extra_dense, ctx->bend,
                               ctx->f_limit, flags, box, ctx->lds_device, ctx->d_wall_hits, block, gpos),
           "cell_step");
  }
  for (int j = 0; j < nside; ++j) {
    HC_TRY(hipEventRecord(ctx->join_ev[j], ctx->side[j]), "hipEventRecord");
    HC_TRY(hipStreamWaitEvent(ctx->stream, ctx->join_ev[j], 0), "hipStreamWaitEvent");
  }
  prof_end(ctx, HC_PROF_MECH, ev);
  return HC_OK;
}

template <typename R>
hc_status do_spread(hc_ctx* ctx) {
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_SPREAD, &ev);
  if (ctx->n_extra > 0) HC_TRY(launch_add_extra_force(ctx->stream, vdev(ctx), ctx->d_extra_ids, ctx->d_extra_f, ctx->n_extra), "extra force");
  if (ctx->spread_per_cell) {
    // LDS budget (a box of doubles): room for the support box of an aligned RBC
    const int want = ctx->spread_lds > 0 ? ctx->spread_lds : 65536;
    HC_TRY(launch_spread_cells<R>(ctx->stream, vdev(ctx), (R*)ctx->Fibm[ctx->cur], ctx->type, ctx->L, ctx->d_cell_base,
                                  ctx->d_cell_nv, (int)ctx->cell_base.size(), ctx->f_limit, rep_mask(ctx),
                                  std::min(want, ctx->lds_limit)),
           "spread");
  } else {
    HC_TRY(launch_spread<R>(ctx->stream, vdev(ctx), (R*)ctx->Fibm[ctx->cur], ctx->type, ctx->L, ctx->f_limit,
                            rep_mask(ctx)), "spread");
  }
  prof_end(ctx, HC_PROF_SPREAD, ev);
  return HC_OK;
}

template <typename R>
hc_status do_collide(hc_ctx* ctx) {
  CollideArgs a{};
  a.omega = 1.0 / ctx->tau;
  for (int c = 0; c < 3; ++c) a.g[c] = ctx->g[c];
  std::memcpy(a.wall_u, ctx->wall_u, sizeof(a.wall_u));
  visc_args(ctx, a);
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_COLLIDE, &ev);
  HC_TRY(launch_collide_stream<R>(ctx->stream, (R*)ctx->f, (const R*)ctx->Fibm[ctx->cur], (R*)ctx->Fibm[1 - ctx->cur],
                                  ctx->type, ctx->L, a, ctx->phase, ctx->general, ctx->vel_nodes, ctx->n_vel,
                                  ctx->fuse_vel),
         "collide_stream");
  prof_end(ctx, HC_PROF_COLLIDE, ev);
  ctx->phase ^= 1;
  ctx->other_clean = true;
  return HC_OK;
}

// Lees-Edwards planes after the collide-stream, with the displacement of this
// iteration, fmod(displacement per step x iteration, nx) (leesEdwardsBC.h:360-363)
template <typename R>
hc_status do_lees_edwards(hc_ctx* ctx) {
  hc_status s;
  if (!ctx->d_le &&
      (s = dalloc(ctx, &ctx->d_le, sizeof(double) * 19 * 2 * (size_t)ctx->L.nx * ctx->L.ny, "Lees-Edwards planes")) !=
          HC_OK)
    return s;
  CollideArgs a{};
  a.omega = 1.0 / ctx->tau;
  visc_args(ctx, a);
  const double d = std::fmod(ctx->le_disp * (double)ctx->iter, (double)ctx->L.nx);
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_COLLIDE, &ev);
  HC_TRY(launch_lees_edwards<R>(ctx->stream, (R*)ctx->f, ctx->type, ctx->L, ctx->phase, a, d, ctx->le_top_u,
                                ctx->le_bottom_u, (R*)ctx->d_le),
         "Lees-Edwards");
  prof_end(ctx, HC_PROF_COLLIDE, ev);
  return HC_OK;
}

template <typename R>
hc_status do_interp_advance(hc_ctx* ctx, int do_interp, int do_advance) {
  if (ctx->nv == 0) return HC_OK;
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_INTERP, &ev);
  if (do_interp)
    HC_TRY(launch_velocity_field<R>(ctx->stream, (const R*)ctx->f, (const R*)ctx->Fibm[ctx->cur], (R*)ctx->ufield,
                                    ctx->L, ctx->phase, ctx->general, ctx->g),
           "velocity_field");
  HC_TRY(launch_interp_advance<R>(ctx->stream, vdev(ctx), (const R*)ctx->ufield, ctx->type, ctx->L, do_interp,
                                  do_advance, ctx->d_wall_hits),
         "interp_advance");
  prof_end(ctx, HC_PROF_INTERP, ev);
  return HC_OK;
}

hc_status do_mechanics(hc_ctx* ctx, bool forced) {
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_MECH, &ev);
  for (auto& tp : ctx->types) {
    TypeBufs& tb = *tp;
    if (tb.ncells == 0) continue;
    if (!forced && (ctx->iter % tb.host.timescale) != 0) continue;
    HC_TRY(launch_mechanics(ctx->stream, vdev(ctx), tb.dev, tb.d_cells, tb.ncells, ctx->d_cell_base, ctx->bend,
                            ctx->lds_limit),
           "mechanics");
  }
  prof_end(ctx, HC_PROF_MECH, ev);
  return HC_OK;
}

template <typename R>
hc_status do_end(hc_ctx* ctx) {
  if (!ctx->other_clean) {
    HC_TRY(launch_fill_zero<R>(ctx->stream, (R*)ctx->Fibm[1 - ctx->cur], 3 * ctx->L.stride), "zero force");
  }
  // the buffer just used is never read again before the collide that zeroes it
  ctx->cur ^= 1;
  ctx->other_clean = false;
  ctx->iter++;
  return HC_OK;
}

template <typename R>
hc_status one_step(hc_ctx* ctx) {
  hc_status s;
  hipEvent_t ev;
  prof_begin(ctx, HC_PROF_STEP, &ev);
  const bool iv = iv_active(ctx) && ctx->nv > 0;
  const bool fused = ctx->fuse && !ctx->repulsion && !ctx->brep && ctx->nv > 0;
  // interior viscosity (hemoCell.cpp:347-357): after the constitutive model; the
  // membrane pass takes the support of the positions before the advance
  const bool memb = iv && ctx->iter % ctx->iv_timescale == 0;
  const bool grid = iv && ctx->iter % ctx->iv_grid_timescale == 0;
  if (ctx->repulsion && ctx->nv > 0 && ctx->iter % ctx->rep_timescale == 0 && (s = do_repulsion(ctx)) != HC_OK)
    return s;
  if (ctx->brep && ctx->nv > 0 && ctx->iter % ctx->brep_timescale == 0 && (s = do_boundary_repulsion(ctx)) != HC_OK)
    return s;
  if (ctx->nv > 0 && !ctx->spread_ready && (s = do_spread<R>(ctx)) != HC_OK) return s;
  if ((s = do_collide<R>(ctx)) != HC_OK) return s;
  if (ctx->le && (s = do_lees_edwards<R>(ctx)) != HC_OK) return s;
  const int do_interp = ctx->iter % ctx->pv_timescale == 0;
  if (iv && (s = ensure_visc(ctx)) != HC_OK) return s;
  if (memb)
    HC_TRY(hipMemcpyAsync(ctx->d_xold, ctx->x, sizeof(double) * 3 * ctx->vcap, hipMemcpyDeviceToDevice, ctx->stream),
           "hipMemcpyAsync");
  if (fused) {
    if (do_interp) {
      hipEvent_t e2;
      prof_begin(ctx, HC_PROF_INTERP, &e2);
      HC_TRY(launch_velocity_field<R>(ctx->stream, (const R*)ctx->f, (const R*)ctx->Fibm[ctx->cur], (R*)ctx->ufield,
                                      ctx->L, ctx->phase, ctx->general, ctx->g),
             "velocity_field");
      prof_end(ctx, HC_PROF_INTERP, e2);
    }
    if ((s = do_cell_step<R>(ctx, do_interp)) != HC_OK) return s;
  } else {
    if ((s = do_interp_advance<R>(ctx, do_interp, 1)) != HC_OK) return s;
    if ((s = do_mechanics(ctx, false)) != HC_OK) return s;
  }
  if (grid && (s = do_visc_grid(ctx)) != HC_OK) return s;
  if (memb && (s = do_visc_membrane(ctx, ctx->d_xold)) != HC_OK) return s;
  if ((s = do_end<R>(ctx)) != HC_OK) return s;
  ctx->spread_ready = fused;
  prof_end(ctx, HC_PROF_STEP, ev);
  if (ctx->prof.on && ctx->prof.used > 4096) prof_flush(ctx);
  return HC_OK;
}

template <typename V>
void put(std::vector<unsigned char>& sig, const V& x) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&x);
  sig.insert(sig.end(), p, p + sizeof(V));
}

// Everything a fused step passes to its kernels: equal signatures mean equal launches.
std::vector<unsigned char> launch_signature(const hc_ctx* c) {
  std::vector<unsigned char> v;
  put(v, c->phase);
  put(v, c->cur);
  put(v, c->other_clean);
  put(v, c->spread_ready);
  put(v, c->precision);
  put(v, c->tau);
  put(v, c->g);
  put(v, c->wall_u);
  put(v, c->f_limit);
  put(v, c->L);
  put(v, c->general);
  put(v, c->f);
  put(v, c->Fibm);
  put(v, c->ufield);
  put(v, c->type);
  put(v, c->vel_nodes);
  put(v, c->n_vel);
  put(v, c->x);
  put(v, c->v);
  put(v, c->fv);
  put(v, c->frep);
  put(v, c->vflag);
  put(v, c->bend);
  put(v, c->nv);
  put(v, c->vcap);
  put(v, c->d_cell_base);
  put(v, c->d_extra_dense);
  put(v, c->d_wall_hits);
  put(v, c->lds_limit);
  put(v, c->lds_device);
  put(v, c->spread_lds);
  put(v, c->cell_block);
  put(v, c->wbc_block);
  put(v, c->n_cu);
  put(v, c->fuse_vel);
  put(v, c->side_gpos);
  put(v, c->mixed);
  put(v, c->mixed_per_cu);
  put(v, c->d_queue);
  put(v, c->d_types);
  put(v, c->d_head);
  put(v, c->nqueue);
  put(v, c->pv_timescale);
  for (const auto& tp : c->types) {
    put(v, tp->dev);
    put(v, tp->d_cells);
    put(v, tp->ncells);
    put(v, tp->host.timescale);
  }
  return v;
}

// Two steps repeat exactly when the step is fused and every timescale is 1.
bool graph_eligible(const hc_ctx* c) {
  if (!c->use_graphs || c->prof.on || !c->fuse || c->repulsion || c->brep || iv_active(c) || c->le || c->nv == 0 ||
      !c->spread_ready)
    return false;
  if (c->pv_timescale != 1) return false;
  for (const auto& tp : c->types)
    if (tp->ncells && tp->host.timescale != 1) return false;
  return true;
}

void drop_graph(hc_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_sig.clear();
}

// Capture two steps into a graph.  The host-side state the capture advanced is
// put back: the captured work has not run yet.
template <typename R>
hc_status capture_pair(hc_ctx* ctx, std::vector<unsigned char> sig) {
  drop_graph(ctx);
  const int64_t iter0 = ctx->iter;
  const int phase0 = ctx->phase, cur0 = ctx->cur;
  const bool clean0 = ctx->other_clean, ready0 = ctx->spread_ready;
  HC_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  hc_status s = one_step<R>(ctx);
  if (s == HC_OK) s = one_step<R>(ctx);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  const bool periodic = ctx->phase == phase0 && ctx->cur == cur0 && ctx->other_clean == clean0 &&
                        ctx->spread_ready == ready0;
  ctx->iter = iter0;
  ctx->phase = phase0;
  ctx->cur = cur0;
  ctx->other_clean = clean0;
  ctx->spread_ready = ready0;
  if (s != HC_OK || e != hipSuccess || !periodic) {
    if (g) hipGraphDestroy(g);
    if (s != HC_OK) return s;
    if (e != hipSuccess) return hipfail(ctx, e, "hipStreamEndCapture");
    ctx->use_graphs = false;  // not a two-step cycle: stay eager
    return HC_OK;
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    hipGraphDestroy(g);
    return hipfail(ctx, ie, "hipGraphInstantiate");
  }
  ctx->graph = g;
  ctx->gexec = x;
  ctx->graph_sig = std::move(sig);
  return HC_OK;
}

template <typename R>
hc_status step_impl(hc_ctx* ctx, int nsteps) {
  hc_status s;
  int k = 0;
  while (k < nsteps) {
    if (nsteps - k >= 2 && graph_eligible(ctx)) {
      std::vector<unsigned char> sig = launch_signature(ctx);
      if (!ctx->gexec || sig != ctx->graph_sig) {
        if ((s = capture_pair<R>(ctx, std::move(sig))) != HC_OK) return s;
        if (!ctx->gexec) continue;
      }
      HC_TRY(hipGraphLaunch(ctx->gexec, ctx->stream), "hipGraphLaunch");
      ctx->iter += 2;
      k += 2;
      continue;
    }
    if ((s = one_step<R>(ctx)) != HC_OK) return s;
    ++k;
  }
  return HC_OK;
}

// ---- batched steps of several contexts ---------------------------------------
// A context can join a batch when its step is the plain fused one: no
// repulsion, interior viscosity or Lees-Edwards planes, every timescale 1,
// velocity nodes (if any) inside the collide kernel, no profiling.
bool batchable(const hc_ctx* c) {
  if (!c->fuse || c->repulsion || c->brep || iv_active(c) || c->le || c->nv == 0 || c->prof.on) return false;
  if (c->pv_timescale != 1) return false;
  if (c->general && c->n_vel > 0 && !c->fuse_vel) return false;
  for (const auto& tp : c->types)
    if (tp->ncells && tp->host.timescale != 1) return false;
  return true;
}

template <typename T>
void append_bytes(std::vector<unsigned char>& buf, const std::vector<T>& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(v.data());
  buf.insert(buf.end(), p, p + v.size() * sizeof(T));
}

// Steps the contexts together: every step launches the collide-stream of all
// cases of one AA parity and boundary kind at once, then their velocity
// fields, then the cells of one mesh at once (each cell spreading into its own
// case's next force buffer).  The launches go to the first context's stream,
// which waits for every context's queued work; every context's stream then
// waits for the batch.  Each case computes exactly what its own fused step
// computes.
template <typename R>
hc_status batch_impl(hc_ctx* const* cs, int n, int nsteps) {
  hc_ctx* lead = cs[0];
  hc_ctx* ctx = lead;  // for HC_TRY
  for (int i = 0; i < n; ++i) {
    hc_ctx* c = cs[i];
    if (!c->batch_ev) HC_TRY(hipEventCreateWithFlags(&c->batch_ev, hipEventDisableTiming), "hipEventCreate");
    if (!c->spread_ready) {
      hc_status s = do_spread<R>(c);
      if (s != HC_OK) return s;
      c->spread_ready = true;
    }
  }
  // descriptors for the two parities of the step counter
  struct Group {
    int kind;  // 0 collide, 1 velocity field, 2 cells
    int a, b, c;
    size_t offset;
    int count;
    uint32_t blocks;
    TypeDev T;
  };
  std::vector<Group> groups[2];
  std::vector<unsigned char> bytes;
  for (int p = 0; p < 2; ++p) {
    // collide: key (odd, general, vel)
    std::map<std::tuple<int, int, int>, std::vector<LatCase>> lat;
    std::map<std::tuple<int, int>, std::vector<VelCase>> vel;
    std::map<std::tuple<int, int, int>, std::vector<CellCase>> cel;
    std::map<std::tuple<int, int, int>, TypeDev> cel_t;
    for (int i = 0; i < n; ++i) {
      hc_ctx* c = cs[i];
      const int phase = c->phase ^ p, cur = c->cur ^ p;
      LatCase lc{};
      lc.f = c->f;
      lc.Fcur = c->Fibm[cur];
      lc.Fclr = c->Fibm[1 - cur];
      lc.type = c->type;
      lc.L = c->L;
      lc.a.omega = 1.0 / c->tau;
      for (int k = 0; k < 3; ++k) lc.a.g[k] = c->g[k];
      std::memcpy(lc.a.wall_u, c->wall_u, sizeof(lc.a.wall_u));
      const int vflag = c->general && c->n_vel > 0 && c->fuse_vel;
      lat[std::make_tuple(phase, (int)c->general, vflag)].push_back(lc);
      VelCase vc{};
      vc.f = c->f;
      vc.Fcur = c->Fibm[cur];
      vc.uf = c->ufield;
      vc.L = c->L;
      for (int k = 0; k < 3; ++k) vc.g[k] = c->g[k];
      vel[std::make_tuple(phase ^ 1, (int)c->general)].push_back(vc);
      for (size_t k = 0; k < c->cell_base.size(); ++k) {
        const TypeBufs& tb = *c->types[(size_t)c->cell_type[k]];
        CellCase cc{};
        cc.V = vdev(c);
        cc.T = tb.dev;
        cc.base = c->cell_base[k];
        cc.uf = c->ufield;
        cc.type = c->type;
        cc.L = c->L;
        cc.Fnext = c->Fibm[1 - cur];
        cc.extra = c->d_extra_dense;
        cc.scratch = c->bend;
        cc.f_limit = c->f_limit;
        cc.flags = CELL_FLAG_SPREAD | CELL_FLAG_INTERP | CELL_FLAG_MECH;
        cc.wall_hits = c->d_wall_hits;
        const int visc = (tb.dev.model != HC_MODEL_RBC) || (tb.dev.eta_m != 0.0);
        const auto key = std::make_tuple(tb.dev.model, tb.dev.nv, visc);
        cel[key].push_back(cc);
        cel_t[key] = tb.dev;
      }
    }
    auto align = [&]() {
      while (bytes.size() % 64) bytes.push_back(0);
    };
    for (auto& kv : lat) {
      uint32_t blocks = 0;
      for (auto& lc : kv.second) {
        lc.block0 = blocks;
        blocks += (uint32_t)((lc.L.N + 255) / 256);
      }
      align();
      Group g{0, std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), bytes.size(),
              (int)kv.second.size(), blocks, TypeDev{}};
      append_bytes(bytes, kv.second);
      groups[p].push_back(g);
    }
    for (auto& kv : vel) {
      uint32_t blocks = 0;
      for (auto& vc : kv.second) {
        vc.block0 = blocks;
        blocks += (uint32_t)((vc.L.N + 255) / 256);
      }
      align();
      Group g{1, std::get<0>(kv.first), std::get<1>(kv.first), 0, bytes.size(), (int)kv.second.size(), blocks,
              TypeDev{}};
      append_bytes(bytes, kv.second);
      groups[p].push_back(g);
    }
    for (auto& kv : cel) {
      align();
      Group g{2, 0, 0, 0, bytes.size(), (int)kv.second.size(), (uint32_t)kv.second.size(), cel_t[kv.first]};
      append_bytes(bytes, kv.second);
      groups[p].push_back(g);
    }
  }
  // this slot's buffers may still be read by the batch before the previous one
  const int slot = lead->batch_slot;
  lead->batch_slot ^= 1;
  if (!lead->batch_slot_ev[slot])
    HC_TRY(hipEventCreateWithFlags(&lead->batch_slot_ev[slot], hipEventDisableTiming), "hipEventCreate");
  else
    HC_TRY(hipEventSynchronize(lead->batch_slot_ev[slot]), "hipEventSynchronize");
  if (bytes.size() > lead->batch_cap[slot]) {
    if (lead->d_batch[slot]) hipFree(lead->d_batch[slot]);
    if (lead->h_batch[slot]) hipHostFree(lead->h_batch[slot]);
    lead->d_batch[slot] = lead->h_batch[slot] = nullptr;
    lead->batch_cap[slot] = 0;
    const size_t cap = bytes.size() * 2;
    hc_status s = dalloc(lead, &lead->d_batch[slot], cap, "batch descriptors");
    if (s != HC_OK) return s;
    HC_TRY(hipHostMalloc(&lead->h_batch[slot], cap, hipHostMallocDefault), "hipHostMalloc");
    lead->batch_cap[slot] = cap;
  }
  std::memcpy(lead->h_batch[slot], bytes.data(), bytes.size());
  for (int i = 1; i < n; ++i) {
    HC_TRY(hipEventRecord(cs[i]->batch_ev, cs[i]->stream), "hipEventRecord");
    HC_TRY(hipStreamWaitEvent(lead->stream, cs[i]->batch_ev, 0), "hipStreamWaitEvent");
  }
  HC_TRY(hipMemcpyAsync(lead->d_batch[slot], lead->h_batch[slot], bytes.size(), hipMemcpyHostToDevice, lead->stream),
         "hipMemcpyAsync");
  const unsigned char* base = (const unsigned char*)lead->d_batch[slot];
  const int box = lead->spread_lds > 0 ? lead->spread_lds : 12288;
  for (int k = 0; k < nsteps; ++k) {
    for (const Group& g : groups[k & 1]) {
      if (g.kind == 0) {
        HC_TRY(launch_collide_multi<R>(lead->stream, (const LatCase*)(base + g.offset), g.count, g.blocks, g.a,
                                       g.b != 0, g.c != 0),
               "batched collide_stream");
      } else if (g.kind == 1) {
        HC_TRY(launch_velocity_field_multi<R>(lead->stream, (const VelCase*)(base + g.offset), g.count, g.blocks,
                                              g.a, g.b != 0),
               "batched velocity_field");
      }
    }
    // the cell groups (one per mesh) run side by side: the first on the lead
    // stream, the others on its two side streams, joined before the next step
    int ncell = 0;
    for (const Group& g : groups[k & 1]) ncell += g.kind == 2;
    const int nside = std::min(ncell - 1, lead->side[0] && lead->side[1] ? 2 : 0);
    if (nside > 0) {
      HC_TRY(hipEventRecord(lead->fork_ev, lead->stream), "hipEventRecord");
      for (int j = 0; j < nside; ++j) HC_TRY(hipStreamWaitEvent(lead->side[j], lead->fork_ev, 0), "hipStreamWaitEvent");
    }
    int gi = 0;
    for (const Group& g : groups[k & 1])
      if (g.kind == 2) {
        hipStream_t st = (gi == 0 || nside == 0) ? lead->stream : lead->side[(gi - 1) % nside];
        ++gi;
        // the lead context's launch shapes, chosen as do_cell_step chooses them:
        // fewer cells than compute units take the widest block (platelets: 256
        // threads instead of 64), unless HC_CELL_BLOCK fixes the RBC block
        int block = g.T.model == HC_MODEL_WBC ? lead->wbc_block
                                              : (g.T.model == HC_MODEL_PLT ? 64 : lead->cell_block);
        if (g.count < lead->n_cu && !(lead->cell_block_env && g.T.model == HC_MODEL_RBC)) block = 1024;
        HC_TRY(launch_cell_step_multi<R>(st, g.T, (const CellCase*)(base + g.offset), g.count, box, lead->lds_device,
                                         block),
               "batched cell_step");
      }
    for (int j = 0; j < nside; ++j) {
      HC_TRY(hipEventRecord(lead->join_ev[j], lead->side[j]), "hipEventRecord");
      HC_TRY(hipStreamWaitEvent(lead->stream, lead->join_ev[j], 0), "hipStreamWaitEvent");
    }
  }
  HC_TRY(hipEventRecord(lead->batch_ev, lead->stream), "hipEventRecord");
  HC_TRY(hipEventRecord(lead->batch_slot_ev[slot], lead->stream), "hipEventRecord");
  for (int i = 0; i < n; ++i) {
    hc_ctx* c = cs[i];
    if (i > 0) HC_TRY(hipStreamWaitEvent(c->stream, lead->batch_ev, 0), "hipStreamWaitEvent");
    c->phase ^= nsteps & 1;
    c->cur ^= nsteps & 1;
    c->iter += nsteps;
    c->other_clean = false;
    c->spread_ready = true;
  }
  return HC_OK;
}

// Autotuner: times real steps of this context (HIP events on its stream, steps
// launched eagerly) for each value of one parameter at a time -- the spread
// box, the threads per RBC cell, the side types' layout, the threads per WBC
// cell -- and keeps the fastest.  The steps advance the simulation as any other.
template <typename R>
hc_status autotune_impl(hc_ctx* ctx, int steps, std::string& report) {
  const bool graphs = ctx->use_graphs;
  ctx->use_graphs = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HC_TRY(hipEventCreate(&e0), "hipEventCreate");
  HC_TRY(hipEventCreate(&e1), "hipEventCreate");
  hc_status s = HC_OK;
  auto trial = [&](double& ms) -> hc_status {
    hc_status t;
    if ((t = one_step<R>(ctx)) != HC_OK) return t;  // first launch of a new shape: not timed
    HC_TRY(hipEventRecord(e0, ctx->stream), "hipEventRecord");
    for (int k = 0; k < steps; ++k)
      if ((t = one_step<R>(ctx)) != HC_OK) return t;
    HC_TRY(hipEventRecord(e1, ctx->stream), "hipEventRecord");
    HC_TRY(hipEventSynchronize(e1), "hipEventSynchronize");
    float f = 0.0f;
    HC_TRY(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime");
    ms = (double)f / steps;
    return HC_OK;
  };
  bool has_rbc = false, has_wbc = false, big_side = false;
  int ntypes = 0;
  for (auto& tp : ctx->types) {
    if (!tp->ncells) continue;
    ++ntypes;
    has_rbc |= tp->host.model == HC_MODEL_RBC;
    has_wbc |= tp->host.model == HC_MODEL_WBC;
    big_side |= 3 * (size_t)tp->host.nv * sizeof(double) > 32768;
  }
  struct Param {
    const char* name;
    std::vector<int> values;
    std::function<void(int)> set;
    std::function<bool()> on;
  };
  // the per-type launches side by side first, then the one persistent launch of
  // all types against them, and its workgroups per compute unit when it wins
  const bool queue = ctx->d_queue && ntypes > 1;
  std::vector<Param> params = {
      {"spread_box_bytes", {8192, 12288, 16384, 20480, 24576}, [&](int v) { ctx->spread_lds = v; },
       [] { return true; }},
      {"rbc_threads", {256, 512, 1024}, [&](int v) { ctx->cell_block = v; ctx->cell_block_env = true; },
       [&] { return has_rbc; }},
      {"side_positions_global", {1, 0}, [&](int v) { ctx->side_gpos = v != 0; },
       [&] { return ntypes > 1 && big_side; }},
      // a side WBC with positions in global memory always runs 256 threads
      {"wbc_threads", {256, 512, 1024}, [&](int v) { ctx->wbc_block = v; },
       [&] { return has_wbc && !(ctx->side_gpos && ntypes > 1 && big_side); }},
      {"persistent_launch", {0, 1}, [&](int v) { ctx->mixed = v != 0; }, [&] { return queue; }},
      {"persistent_workgroups_per_cu", {4, 8, 16}, [&](int v) { ctx->mixed_per_cu = v; },
       [&] { return ctx->mixed; }},
  };
  std::ostringstream out;
  out << "{";
  bool first_param = true;
  for (auto& p : params) {
    if (!p.on() || s != HC_OK) continue;
    int best = p.values[0];
    double best_ms = 1e300;
    out << (first_param ? "" : ", ") << "\"" << p.name << "\": {";
    first_param = false;
    for (size_t k = 0; k < p.values.size(); ++k) {
      p.set(p.values[k]);
      double ms = 0;
      if ((s = trial(ms)) != HC_OK) break;
      out << (k ? ", " : "") << "\"" << p.values[k] << "\": " << ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = p.values[k];
      }
    }
    p.set(best);
    out << ", \"chosen\": " << best << "}";
  }
  out << "}";
  report = out.str();
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  ctx->use_graphs = graphs;
  drop_graph(ctx);
  return s;
}

bool fp32(const hc_ctx* c) { return c->precision == HC_FP32_SHIFTED; }

hc_status ready_for_step(hc_ctx* ctx) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  return sync_cells(ctx);
}

}  // namespace

template <typename R>
static hc_status fluid_stats(hc_ctx* ctx, int what, double* out) {
  const int nb = (int)std::min<int64_t>(4096, (ctx->L.N + 255) / 256);
  double* part = nullptr;
  R* Fs = nullptr;
  hc_status s = dalloc(ctx, (void**)&part, sizeof(double) * 4 * nb, "statistics partials");
  if (s != HC_OK) return s;
  CollideArgs a{};
  a.omega = 1.0 / ctx->tau;
  for (int c = 0; c < 3; ++c) a.g[c] = ctx->g[c];
  std::memcpy(a.wall_u, ctx->wall_u, sizeof(a.wall_u));
  const R* Fc = ctx->spread_ready ? nullptr : (const R*)ctx->Fibm[ctx->cur];  // reset field after a fused step
  hipError_t e = hipSuccess;
  if (what == 1) {
    // calculateForceStatistics spreads the vertex forces into a reset field first
    // (helper/fluidInfo.cpp:117-120): here into a field of its own
    if ((s = dalloc(ctx, (void**)&Fs, sizeof(R) * 3 * ctx->L.stride, "statistics force field")) != HC_OK) {
      hipFree(part);
      return s;
    }
    e = launch_fill_zero<R>(ctx->stream, Fs, 3 * ctx->L.stride);
    if (e == hipSuccess && ctx->nv > 0 && sync_cells(ctx) == HC_OK)
      e = launch_spread_cells<R>(ctx->stream, vdev(ctx), Fs, ctx->type, ctx->L, ctx->d_cell_base, ctx->d_cell_nv,
                                 (int)ctx->cell_base.size(), ctx->f_limit, rep_mask(ctx),
                                 std::min(65536, ctx->lds_limit));
  }
  if (e == hipSuccess)
    e = launch_fluid_stats<R>(ctx->stream, (const R*)ctx->f, Fc, ctx->type, ctx->L, ctx->phase, a, Fs, what, part, nb);
  std::vector<double> h(4 * (size_t)nb);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(part);
  if (Fs) hipFree(Fs);
  if (e != hipSuccess) return hipfail(ctx, e, "fluid statistics");
  double mn = INFINITY, mx = -INFINITY, sum = 0.0, cnt = 0.0;
  for (int b = 0; b < nb; ++b) {
    mn = std::min(mn, h[4 * b]);
    mx = std::max(mx, h[4 * b + 1]);
    sum += h[4 * b + 2];
    cnt += h[4 * b + 3];
  }
  out[0] = cnt > 0 ? mn : 0.0;
  out[1] = cnt > 0 ? mx : 0.0;
  out[2] = cnt > 0 ? sum / cnt : 0.0;
  out[3] = cnt;
  return HC_OK;
}

extern "C" {

hc_status hc_create(int device, hc_ctx** out) {
  if (!out) return HC_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return HC_DEVICE;
  if (device < 0 || device >= n) return HC_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return HC_DEVICE;
  hc_ctx* ctx = new hc_ctx();
  ctx->device = device;
  // the statistics staging array of hc_get_cell_stats_noalloc on this device
  // (looked up once per device; a failure disables only that entry point)
  ctx->noalloc_ok = noalloc_scratch_setup(device) == hipSuccess;
  if (hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return HC_DEVICE;
  }
  bool side_ok = hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) == hipSuccess;
  for (int j = 0; j < 2; ++j)
    side_ok = side_ok && hipStreamCreate(&ctx->side[j]) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->join_ev[j], hipEventDisableTiming) == hipSuccess;
  if (!side_ok) {  // run every type on the context stream
    for (int j = 0; j < 2; ++j)
      if (ctx->side[j]) hipStreamDestroy(ctx->side[j]);
    ctx->side[0] = ctx->side[1] = nullptr;
  }
  if (hipMalloc((void**)&ctx->d_wall_hits, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc((void**)&ctx->d_flag, sizeof(unsigned int)) != hipSuccess ||
      hipMemset(ctx->d_wall_hits, 0, sizeof(unsigned long long)) != hipSuccess) {
    if (ctx->d_wall_hits) (void)hipFree(ctx->d_wall_hits);
    if (ctx->d_flag) (void)hipFree(ctx->d_flag);
    for (int j = 0; j < 2; ++j) {
      if (ctx->side[j]) (void)hipStreamDestroy(ctx->side[j]);
      if (ctx->join_ev[j]) (void)hipEventDestroy(ctx->join_ev[j]);
    }
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HC_OOM;
  }
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
  {
    ctx->lds_limit = lds - 256;  // the kernel's own static LDS
    ctx->lds_device = lds;
  }
  const char* sp = std::getenv("HC_SPREAD");
  if (sp && std::string(sp) == "vertex") ctx->spread_per_cell = false;
  const char* fu = std::getenv("HC_FUSE");  // tuning: 0 runs the unfused vertex kernels
  if (fu && std::string(fu) == "0") ctx->fuse = false;
  const char* cb = std::getenv("HC_CELL_BLOCK");  // tuning: 256, 512 or 1024 threads per RBC cell
  if (cb) {
    ctx->cell_block = std::atoi(cb);
    ctx->cell_block_env = true;
  }
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    ctx->n_cu = ncu;
  const char* fv = std::getenv("HC_FUSE_VEL");  // 0: velocity-wall nodes in their own kernel
  if (fv && std::string(fv) == "0") ctx->fuse_vel = false;
  const char* co = std::getenv("HC_CELL_ORDER");  // tuning: 0 launches cells in the order added
  if (co && std::string(co) == "0") ctx->cell_order = false;
  const char* wb = std::getenv("HC_WBC_BLOCK");  // tuning: threads per WBC cell
  if (wb) ctx->wbc_block = std::atoi(wb);
  const char* gr = std::getenv("HC_GRAPH");  // 0: launch every step eagerly
  if (gr && std::string(gr) == "0") ctx->use_graphs = false;
  const char* sl = std::getenv("HC_SPREAD_LDS");  // tuning: LDS bytes of the per-cell box
  if (sl) ctx->spread_lds = std::atoi(sl);
  const char* sg = std::getenv("HC_SIDE_GPOS");  // tuning: 0 stages every type's positions in LDS
  if (sg) ctx->side_gpos = std::atoi(sg) != 0;
  const char* mx = std::getenv("HC_MIXED");  // tuning: 1 runs every cell type in one persistent launch
  if (mx) ctx->mixed = std::atoi(mx) != 0;
  const char* mp = std::getenv("HC_MIXED_PER_CU");  // tuning: its workgroups per compute unit
  if (mp && std::atoi(mp) > 0) ctx->mixed_per_cu = std::atoi(mp);
  *out = ctx;
  return HC_OK;
}

void hc_destroy(hc_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  void* ps[] = {ctx->d_queue, ctx->d_types, ctx->d_head,
                ctx->f, ctx->Fibm[0], ctx->Fibm[1], ctx->ufield, ctx->d_cell_nv, ctx->d_extra_dense, ctx->type, ctx->vel_nodes, ctx->x, ctx->v, ctx->fv, ctx->frep, ctx->bend,
                ctx->fbr, ctx->vflag, ctx->d_bflag, ctx->d_cell_base, ctx->d_extra_ids, ctx->d_extra_f, ctx->d_wall_hits, ctx->d_flag,
                ctx->d_visc, ctx->d_stamp, ctx->d_xold, ctx->d_le, ctx->d_batch[0], ctx->d_batch[1]};
  for (void* p : ps)
    if (p) hipFree(p);
  for (auto& tp : ctx->types) {
    for (void* p : tp->allocs) hipFree(p);
    if (tp->d_cells) hipFree(tp->d_cells);
  }
  drop_graph(ctx);
  free_repulsion(ctx);
  if (ctx->rb.bins) hipFree(ctx->rb.bins);
  for (hipEvent_t e : ctx->prof.pool) hipEventDestroy(e);
  for (int j = 0; j < 2; ++j) {
    if (ctx->side[j]) hipStreamDestroy(ctx->side[j]);
    if (ctx->join_ev[j]) hipEventDestroy(ctx->join_ev[j]);
  }
  if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
  if (ctx->batch_ev) hipEventDestroy(ctx->batch_ev);
  for (int k = 0; k < 2; ++k) {
    if (ctx->batch_slot_ev[k]) hipEventDestroy(ctx->batch_slot_ev[k]);
    if (ctx->h_batch[k]) hipHostFree(ctx->h_batch[k]);
  }
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* hc_last_error(const hc_ctx* ctx) {
  if (!ctx) return "null context";
  return ctx->err_fixed ? ctx->err_fixed : ctx->err.c_str();
}

hc_status hc_lattice_init(hc_ctx* ctx, int nx, int ny, int nz, const uint8_t* node_type, double tau,
                          const int periodic[3], int precision) {
  if (!ctx) return HC_INVALID_ARG;
  if (ctx->lattice_ready) return fail(ctx, HC_STATE, "lattice already initialised");
  if (nx < 1 || ny < 1 || nz < 1) return fail(ctx, HC_INVALID_ARG, "lattice sizes must be >= 1");
  if (!(tau > 0.5)) return fail(ctx, HC_INVALID_ARG, "tau must be > 0.5");
  if (precision != HC_FP64 && precision != HC_FP32_SHIFTED) return fail(ctx, HC_INVALID_ARG, "unknown precision");
  const int64_t N = (int64_t)nx * ny * nz;
  if (N >= ((int64_t)1 << 31)) return fail(ctx, HC_INVALID_ARG, "lattice too large (at most 2^31-1 nodes)");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, HC_DEVICE, "hipSetDevice failed");
  ctx->precision = precision;
  ctx->tau = tau;
  ctx->L.nx = nx;
  ctx->L.ny = ny;
  ctx->L.nz = nz;
  bool all_periodic = true;
  for (int a = 0; a < 3; ++a) {
    ctx->L.periodic[a] = periodic ? (periodic[a] != 0) : 1;
    all_periodic &= ctx->L.periodic[a] != 0;
  }
  ctx->L.N = N;
  ctx->L.stride = (N + 255) / 256 * 256;
  ctx->L.spread_q = ctx->L.spread_inv_q = 0.0;  // plain adds until hc_set_exact_spread
  bool has_walls = false;
  std::vector<uint32_t> vel;
  if (node_type) {
    for (int64_t n = 0; n < N; ++n) {
      uint8_t t = node_type[n];
      if (t > HC_NODE_VELOCITY + 5) return fail(ctx, HC_INVALID_ARG, "unknown node type in mask");
      if (t != HC_NODE_FLUID) has_walls = true;
      if (t >= HC_NODE_VELOCITY) {
        // a velocity node must sit on the domain face it names
        const int face = t - HC_NODE_VELOCITY, axis = face >> 1;
        const int64_t z = n % nz, y = (n / nz) % ny, x = n / ((int64_t)ny * nz);
        const int64_t c = axis == 0 ? x : (axis == 1 ? y : z);
        const int64_t last = (axis == 0 ? nx : (axis == 1 ? ny : nz)) - 1;
        if (c != ((face & 1) ? last : 0))
          return fail(ctx, HC_INVALID_ARG, "velocity-boundary node not on its domain face");
        vel.push_back((uint32_t)n);
      }
    }
  }
  // grouped by face, so that the lanes of a wavefront take the same face branch
  std::stable_sort(vel.begin(), vel.end(), [&](uint32_t a, uint32_t b) { return node_type[a] < node_type[b]; });
  ctx->has_walls = has_walls;
  ctx->general = has_walls || !all_periodic;
  const size_t rs = rsize(ctx);
  hc_status s;
  if ((s = dalloc(ctx, &ctx->f, rs * 19 * ctx->L.stride, "populations")) != HC_OK) return s;
  for (int b = 0; b < 2; ++b)
    if ((s = dalloc(ctx, &ctx->Fibm[b], rs * 3 * ctx->L.stride, "force field")) != HC_OK) return s;
  if ((s = dalloc(ctx, &ctx->ufield, rs * 3 * ctx->L.stride, "velocity field")) != HC_OK) return s;
  HC_TRY(hipMemsetAsync(ctx->f, 0, rs * 19 * ctx->L.stride, ctx->stream), "hipMemsetAsync");
  for (int b = 0; b < 2; ++b) HC_TRY(hipMemsetAsync(ctx->Fibm[b], 0, rs * 3 * ctx->L.stride, ctx->stream), "hipMemsetAsync");
  if (has_walls) {
    if ((s = dalloc(ctx, (void**)&ctx->type, N, "node types")) != HC_OK) return s;
    HC_TRY(hipMemcpy(ctx->type, node_type, N, hipMemcpyHostToDevice), "hipMemcpy");
  }
  if (!vel.empty()) {
    if ((s = dalloc(ctx, (void**)&ctx->vel_nodes, sizeof(uint32_t) * vel.size(), "velocity nodes")) != HC_OK) return s;
    HC_TRY(hipMemcpy(ctx->vel_nodes, vel.data(), sizeof(uint32_t) * vel.size(), hipMemcpyHostToDevice), "hipMemcpy");
    ctx->n_vel = (int)vel.size();
  }
  HC_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  ctx->lattice_ready = true;
  ctx->phase = 0;
  ctx->cur = 0;
  ctx->other_clean = true;
  return HC_OK;
}

hc_status hc_set_wall_velocity(hc_ctx* ctx, int face, const double u[3]) {
  if (!ctx || !u) return HC_INVALID_ARG;
  if (face < 0 || face > 5) return fail(ctx, HC_INVALID_ARG, "face must be 0..5");
  for (int c = 0; c < 3; ++c) ctx->wall_u[face][c] = u[c];
  return HC_OK;
}

hc_status hc_set_body_force(hc_ctx* ctx, const double g[3]) {
  if (!ctx || !g) return HC_INVALID_ARG;
  for (int c = 0; c < 3; ++c) ctx->g[c] = g[c];
  return HC_OK;
}

hc_status hc_init_equilibrium(hc_ctx* ctx, double rho, const double u[3]) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!(rho > 0.0)) return fail(ctx, HC_INVALID_ARG, "rho must be > 0");
  const double zero[3] = {0, 0, 0};
  const double* uu = u ? u : zero;
  if (fp32(ctx)) HC_TRY(launch_init_equilibrium<float>(ctx->stream, (float*)ctx->f, ctx->type, ctx->L, rho, uu), "init");
  else HC_TRY(launch_init_equilibrium<double>(ctx->stream, (double*)ctx->f, ctx->type, ctx->L, rho, uu), "init");
  ctx->phase = 0;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  return HC_OK;
}

hc_status hc_build_mesh(int shape, double radius_lu, int min_triangles, double aspect_ratio, int* nv, int* nt,
                        double* vertices, int32_t* triangles) {
  std::vector<double> v;
  std::vector<int32_t> t;
  std::string err;
  if (!build_mesh(shape, radius_lu, min_triangles, aspect_ratio, v, t, err)) return HC_INVALID_ARG;
  if (nv) *nv = (int)(v.size() / 3);
  if (nt) *nt = (int)(t.size() / 3);
  if (vertices) std::copy(v.begin(), v.end(), vertices);
  if (triangles) std::copy(t.begin(), t.end(), triangles);
  return HC_OK;
}

hc_status hc_add_cell_type(hc_ctx* ctx, int model, int nv, int nt, const int32_t* triangles, const double* rest,
                           int n_inner, const int32_t* inner, const hc_model_params* params, int timescale,
                           int* type_id) {
  if (!ctx) return HC_INVALID_ARG;
  if (model < HC_MODEL_RBC || model > HC_MODEL_WBC) return fail(ctx, HC_INVALID_ARG, "unknown model");
  if (nv < 4 || nt < 4 || !triangles || !rest || !params) return fail(ctx, HC_INVALID_ARG, "bad mesh arguments");
  if (n_inner < 0 || (n_inner > 0 && !inner)) return fail(ctx, HC_INVALID_ARG, "bad inner edges");
  auto tb = std::make_unique<TypeBufs>();
  TypeHost& T = tb->host;
  T.model = model;
  T.nv = nv;
  T.nt = nt;
  T.tri.assign(triangles, triangles + 3 * nt);
  T.rest.assign(rest, rest + 3 * nv);
  if (n_inner > 0) T.inner.assign(inner, inner + 2 * n_inner);
  T.p = *params;
  T.timescale = timescale > 0 ? timescale : 1;
  std::string err;
  if (!build_type_tables(T, err)) return fail(ctx, HC_INVALID_ARG, "cell type: " + err);
  TypeDev& D = tb->dev;
  hc_status s;
#define UP(field, dst) \
  if ((s = upload(ctx, *tb, T.field, &D.dst)) != HC_OK) return s;
  UP(tri, tri) UP(edge, edge) UP(edge_tris, edge_tris) UP(edge_outer, edge_outer) UP(ring, ring) UP(nring, nring)
  UP(inner, inner) UP(edge_len_eq, edge_len_eq) UP(edge_angle_eq, edge_angle_eq) UP(tri_area_eq, tri_area_eq)
  UP(patch_dist_eq, patch_dist_eq) UP(inner_len_eq, inner_len_eq) UP(vtri_off, vtri_off) UP(vtri, vtri)
  UP(vbend_off, vbend_off) UP(vbend, vbend) UP(vedge_off, vedge_off) UP(vedge, vedge) UP(vinner_off, vinner_off)
  UP(vinner, vinner)
#undef UP
  {
    const int32_t* p4 = nullptr;
    if ((s = upload(ctx, *tb, T.vtri4, &p4)) != HC_OK) return s;
    D.vtri4 = (const int4*)p4;
    if ((s = upload(ctx, *tb, T.vedge4, &p4)) != HC_OK) return s;
    D.vedge4 = (const int4*)p4;
    D.vtri4p = D.vedge4p = nullptr;
    D.vbendp = nullptr;
    D.vtri_aeq = D.vedge_l0 = D.vbend_n = nullptr;
    D.cslot = D.cring = nullptr;
    D.cbend = nullptr;
    D.rpar = nullptr;
    D.rnb = nullptr;
    if (!T.rpar.empty() && !padded_off(ctx) && !ring_tables_off()) {
      const float* pf = nullptr;
      if ((s = upload(ctx, *tb, T.rpar, &pf)) != HC_OK) return s;
      D.rpar = (const float4*)pf;
      if ((s = upload(ctx, *tb, T.rnb, &D.rnb)) != HC_OK) return s;
    }
    if (!T.cslot.empty() && !padded_off(ctx)) {
      const uint32_t* pc = nullptr;
      if ((s = upload(ctx, *tb, T.cslot, &pc)) != HC_OK) return s;
      D.cslot = (const uint4*)pc;
      if ((s = upload(ctx, *tb, T.cring, &pc)) != HC_OK) return s;
      D.cring = (const uint4*)pc;
      if ((s = upload(ctx, *tb, T.cbend, &D.cbend)) != HC_OK) return s;
    }
    if (T.padded && !padded_off(ctx)) {
      if ((s = upload(ctx, *tb, T.vtri4p, &p4)) != HC_OK) return s;
      D.vtri4p = (const int4*)p4;
      if ((s = upload(ctx, *tb, T.vedge4p, &p4)) != HC_OK) return s;
      D.vedge4p = (const int4*)p4;
      if ((s = upload(ctx, *tb, T.vbendp, &D.vbendp)) != HC_OK) return s;
      if ((s = upload(ctx, *tb, T.vtri_aeq, &D.vtri_aeq)) != HC_OK) return s;
      if ((s = upload(ctx, *tb, T.vedge_l0, &D.vedge_l0)) != HC_OK) return s;
      if ((s = upload(ctx, *tb, T.vbend_n, &D.vbend_n)) != HC_OK) return s;
    }
  }
  D.model = model;
  D.nv = T.nv;
  D.nt = T.nt;
  D.ne = T.ne;
  D.ni = T.ni;
  D.volume_eq = T.volume_eq;
  D.area_mean_eq = T.area_mean_eq;
  D.edge_mean_eq = T.edge_mean_eq;
  D.k_link = params->k_link;
  D.k_bend = params->k_bend;
  D.k_volume = params->k_volume;
  D.k_area = params->k_area;
  D.eta_m = params->eta_m;
  D.k_inner_rigid = params->k_inner_rigid;
  D.k_cyto = params->k_cytoskeleton;
  D.core_radius = params->core_radius;
  D.radius = params->radius;
  ctx->types.push_back(std::move(tb));
  if (type_id) *type_id = (int)ctx->types.size() - 1;
  return HC_OK;
}

hc_status hc_add_cells(hc_ctx* ctx, int type_id, int ncells, const double* xyz, const int32_t* cell_ids) {
  if (!ctx) return HC_INVALID_ARG;
  if (type_id < 0 || type_id >= (int)ctx->types.size()) return fail(ctx, HC_INVALID_ARG, "unknown cell type");
  if (ncells < 0 || (ncells > 0 && !xyz)) return fail(ctx, HC_INVALID_ARG, "bad cell arguments");
  if (ncells == 0) return HC_OK;
  const int nvt = ctx->types[type_id]->host.nv;
  const int64_t add = (int64_t)ncells * nvt;
  hc_status s;
  if (ctx->lattice_ready && (s = invalidate_spread(ctx)) != HC_OK) return s;
  const int64_t old_cap = ctx->vcap;
  if ((s = grow_vertices(ctx, ctx->nv + add)) != HC_OK) return s;
  if (ctx->vcap != old_cap && (s = build_dense_extra(ctx)) != HC_OK) return s;
  std::vector<double> plane(add);
  for (int c = 0; c < 3; ++c) {
    for (int64_t k = 0; k < add; ++k) plane[k] = xyz[3 * k + c];
    HC_TRY(hipMemcpy(ctx->x + c * ctx->vcap + ctx->nv, plane.data(), sizeof(double) * add, hipMemcpyHostToDevice),
           "hipMemcpy");
  }
  for (int c = 0; c < ncells; ++c) {
    double ctr[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < nvt; ++k)
      for (int a = 0; a < 3; ++a) ctr[a] += xyz[3 * ((int64_t)c * nvt + k) + a];
    for (int a = 0; a < 3; ++a) ctx->cell_ctr.push_back(ctr[a] / nvt);
    ctx->cell_base.push_back(ctx->nv + (int64_t)c * nvt);
    ctx->cell_type.push_back(type_id);
    ctx->cell_id.push_back(cell_ids ? cell_ids[c] : (int32_t)ctx->cell_id.size());
  }
  ctx->nv += add;
  ctx->cells_dirty = true;
  return HC_OK;
}

hc_status hc_set_velocity_timescale(hc_ctx* ctx, int ts) {
  if (!ctx || ts < 1) return fail(ctx, HC_INVALID_ARG, "timescale must be >= 1");
  ctx->pv_timescale = ts;
  return HC_OK;
}
hc_status hc_set_material_timescale(hc_ctx* ctx, int type_id, int ts) {
  if (!ctx || ts < 1) return fail(ctx, HC_INVALID_ARG, "timescale must be >= 1");
  if (type_id < 0 || type_id >= (int)ctx->types.size()) return fail(ctx, HC_INVALID_ARG, "unknown cell type");
  ctx->types[type_id]->host.timescale = ts;
  return HC_OK;
}
// The exact spread's quantum: 2^(e-49) for a force limit in [2^e, 2^(e+1)).  A
// node's force sum stays exact up to 2^53 quanta = 2^(e+4) >= 8 force limits,
// and a vertex force at the limit keeps 49 bits.
static void set_spread_quantum(hc_ctx* ctx) {
  if (ctx->exact_spread) {
    const int e = std::ilogb(ctx->f_limit);
    ctx->L.spread_q = std::ldexp(1.0, e - 49);
    ctx->L.spread_inv_q = std::ldexp(1.0, 49 - e);
  } else {
    ctx->L.spread_q = 0.0;
    ctx->L.spread_inv_q = 0.0;
  }
}

hc_status hc_set_force_limit(hc_ctx* ctx, double f_limit) {
  if (!ctx || !(f_limit > 0) || !std::isfinite(f_limit))
    return fail(ctx, HC_INVALID_ARG, "force limit must be finite and > 0");
  if (ctx->lattice_ready) {
    hc_status s = invalidate_spread(ctx);
    if (s != HC_OK) return s;
  }
  ctx->f_limit = f_limit;
  ctx->f_limit_set = true;
  set_spread_quantum(ctx);
  return HC_OK;
}
hc_status hc_set_exact_spread(hc_ctx* ctx, int on) {
  if (!ctx) return HC_INVALID_ARG;
  if (!ctx->lattice_ready) return fail(ctx, HC_STATE, "lattice not initialised");
  if (on && ctx->precision != HC_FP64) return fail(ctx, HC_INVALID_ARG, "the exact spread needs the FP64 form");
  if (on && !ctx->f_limit_set) return fail(ctx, HC_STATE, "the exact spread needs a force limit (hc_set_force_limit)");
  hc_status s = invalidate_spread(ctx);
  if (s != HC_OK) return s;
  ctx->exact_spread = on != 0;
  set_spread_quantum(ctx);
  return HC_OK;
}
hc_status hc_set_repulsion(hc_ctx* ctx, double k_rep, double cutoff_lu, int timescale) {
  if (!ctx) return HC_INVALID_ARG;
  if (!std::isfinite(k_rep) || !(cutoff_lu > 0.0) || !std::isfinite(cutoff_lu) || timescale < 1)
    return fail(ctx, HC_INVALID_ARG, "repulsion needs a finite constant, a cut-off > 0 and a timescale >= 1");
  if (ctx->lattice_ready) {
    hc_status s = invalidate_spread(ctx);
    if (s != HC_OK) return s;
  }
  ctx->repulsion = true;
  ctx->rep_k = k_rep;
  ctx->rep_cutoff = cutoff_lu;
  ctx->rep_timescale = timescale;
  return HC_OK;
}

hc_status hc_set_boundary_repulsion(hc_ctx* ctx, double k_rep, double cutoff_lu, int timescale) {
  if (!ctx) return HC_INVALID_ARG;
  if (!std::isfinite(k_rep) || !(cutoff_lu > 0.0) || !std::isfinite(cutoff_lu) || timescale < 1)
    return fail(ctx, HC_INVALID_ARG, "boundary repulsion needs a finite constant, a cut-off > 0 and a timescale >= 1");
  if (ctx->lattice_ready) {
    hc_status s = invalidate_spread(ctx);
    if (s != HC_OK) return s;
  }
  ctx->brep = true;
  ctx->brep_k = k_rep;
  ctx->brep_cutoff = cutoff_lu;
  ctx->brep_timescale = timescale;
  return HC_OK;
}

hc_status hc_set_periodicity(hc_ctx* ctx, const int periodic[3]) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!periodic) return fail(ctx, HC_INVALID_ARG, "null periodicity");
  if (ctx->le && (!periodic[0] || !periodic[2]))
    return fail(ctx, HC_INVALID_ARG, "Lees-Edwards planes need a lattice periodic in x and z");
  HC_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  bool all = true;
  for (int a = 0; a < 3; ++a) {
    ctx->L.periodic[a] = periodic[a] ? 1 : 0;
    all = all && periodic[a];
  }
  ctx->general = ctx->has_walls || !all;
  if (ctx->d_bflag) hipFree(ctx->d_bflag);  // boundary particles are flagged again on first use
  ctx->d_bflag = nullptr;
  drop_graph(ctx);
  return HC_OK;
}

hc_status hc_set_lees_edwards(hc_ctx* ctx, double shear_rate, double displacement_per_step) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!std::isfinite(shear_rate) || !std::isfinite(displacement_per_step))
    return fail(ctx, HC_INVALID_ARG, "Lees-Edwards needs a finite shear rate and displacement");
  if (!ctx->L.periodic[0] || !ctx->L.periodic[2] || ctx->L.nz < 2)
    return fail(ctx, HC_INVALID_ARG, "Lees-Edwards planes need a lattice periodic in x and z (nz >= 2)");
  const double vhalf = (ctx->L.nz - 1) * shear_rate * 0.5;  // leesEdwardsBC.h:322-324
  ctx->le = true;
  ctx->le_top_u = -vhalf;
  ctx->le_bottom_u = vhalf;
  ctx->le_disp = displacement_per_step;
  return HC_OK;
}

hc_status hc_stage_lees_edwards(hc_ctx* ctx) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!ctx->le) return fail(ctx, HC_STATE, "Lees-Edwards planes not set (hc_set_lees_edwards)");
  return fp32(ctx) ? do_lees_edwards<float>(ctx) : do_lees_edwards<double>(ctx);
}

hc_status hc_set_interior_viscosity(hc_ctx* ctx, int type_id, double tau_inner) {
  if (!ctx) return HC_INVALID_ARG;
  if (type_id < 0 || type_id >= (int)ctx->types.size()) return fail(ctx, HC_INVALID_ARG, "bad type id");
  if (type_id >= 15) return fail(ctx, HC_INVALID_ARG, "interior viscosity is available for type ids 0..14");
  if (!std::isfinite(tau_inner) || (tau_inner > 0.0 && !(tau_inner > 0.5)))
    return fail(ctx, HC_INVALID_ARG, "tau_inner must be > 0.5 (or <= 0 to turn interior viscosity off)");
  if (ctx->lattice_ready) {
    hc_status s = invalidate_spread(ctx);
    if (s != HC_OK) return s;
  }
  if (ctx->iv_tau.size() < ctx->types.size()) ctx->iv_tau.resize(ctx->types.size(), 0.0);
  ctx->iv_tau[type_id] = tau_inner > 0.0 ? tau_inner : 0.0;
  return HC_OK;
}

hc_status hc_set_interior_viscosity_timescales(hc_ctx* ctx, int membrane, int entire_grid) {
  if (!ctx) return HC_INVALID_ARG;
  if (membrane < 1 || entire_grid < 1) return fail(ctx, HC_INVALID_ARG, "timescales must be >= 1");
  ctx->iv_timescale = membrane;
  ctx->iv_grid_timescale = entire_grid;
  return HC_OK;
}

hc_status hc_get_interior_viscosity(hc_ctx* ctx, double* tau) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!tau) return fail(ctx, HC_INVALID_ARG, "null output");
  std::vector<uint8_t> cls((size_t)ctx->L.N, 0);
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  if (ctx->d_visc) HC_TRY(hipMemcpy(cls.data(), ctx->d_visc, (size_t)ctx->L.N, hipMemcpyDeviceToHost), "hipMemcpy");
  for (int64_t n = 0; n < ctx->L.N; ++n) tau[n] = cls[n] ? iv_tau_of(ctx, cls[n] - 1) : 0.0;
  return HC_OK;
}

hc_status hc_set_interior_viscosity_field(hc_ctx* ctx, const double* tau) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!tau) return fail(ctx, HC_INVALID_ARG, "null input");
  std::vector<uint8_t> cls((size_t)ctx->L.N, 0);
  for (int64_t n = 0; n < ctx->L.N; ++n) {
    if (tau[n] == 0.0) continue;
    size_t t = 0;
    while (t < ctx->iv_tau.size() && !(ctx->iv_tau[t] > 0.0 && ctx->iv_tau[t] == tau[n])) ++t;
    if (t == ctx->iv_tau.size())
      return fail(ctx, HC_INVALID_ARG, "a node's tau is neither 0 nor the tau_inner of a type with interior viscosity");
    cls[n] = (uint8_t)(t + 1);
  }
  if ((s = ensure_visc(ctx)) != HC_OK) return s;
  HC_TRY(hipMemcpyAsync(ctx->d_visc, cls.data(), (size_t)ctx->L.N, hipMemcpyHostToDevice, ctx->stream),
         "hipMemcpyAsync");
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  return HC_OK;
}

hc_status hc_stage_interior_viscosity(hc_ctx* ctx, int pass) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if (pass != 0 && pass != 1) return fail(ctx, HC_INVALID_ARG, "pass must be 0 (entire grid) or 1 (membrane)");
  if (!iv_active(ctx)) return fail(ctx, HC_STATE, "interior viscosity not enabled (hc_set_interior_viscosity)");
  if (ctx->nv == 0) return HC_OK;
  if ((s = ensure_visc(ctx)) != HC_OK) return s;
  if (pass == 0) return do_visc_grid(ctx);
  HC_TRY(hipMemcpyAsync(ctx->d_xold, ctx->x, sizeof(double) * 3 * ctx->vcap, hipMemcpyDeviceToDevice, ctx->stream),
         "hipMemcpyAsync");
  return do_visc_membrane(ctx, ctx->d_xold);
}

hc_status hc_set_extra_vertex_force(hc_ctx* ctx, int n, const int64_t* ids, const double* forces) {
  if (!ctx || n < 0 || (n > 0 && (!ids || !forces))) return fail(ctx, HC_INVALID_ARG, "bad extra-force arguments");
  if (ctx->lattice_ready) {
    hc_status s0 = invalidate_spread(ctx);
    if (s0 != HC_OK) return s0;
  }
  if (ctx->d_extra_ids) hipFree(ctx->d_extra_ids);
  if (ctx->d_extra_f) hipFree(ctx->d_extra_f);
  ctx->d_extra_ids = nullptr;
  ctx->d_extra_f = nullptr;
  ctx->n_extra = 0;
  ctx->h_extra_ids.clear();
  ctx->h_extra_f.clear();
  if (n == 0) return build_dense_extra(ctx);
  for (int k = 0; k < n; ++k)
    if (ids[k] < 0 || ids[k] >= ctx->nv) return fail(ctx, HC_INVALID_ARG, "extra-force vertex id out of range");
  hc_status s;
  if ((s = dalloc(ctx, (void**)&ctx->d_extra_ids, sizeof(int64_t) * n, "extra ids")) != HC_OK) return s;
  if ((s = dalloc(ctx, (void**)&ctx->d_extra_f, sizeof(double) * 3 * n, "extra forces")) != HC_OK) return s;
  HC_TRY(hipMemcpy(ctx->d_extra_ids, ids, sizeof(int64_t) * n, hipMemcpyHostToDevice), "hipMemcpy");
  HC_TRY(hipMemcpy(ctx->d_extra_f, forces, sizeof(double) * 3 * n, hipMemcpyHostToDevice), "hipMemcpy");
  ctx->n_extra = n;
  ctx->h_extra_ids.assign(ids, ids + n);
  ctx->h_extra_f.assign(forces, forces + 3 * n);
  return build_dense_extra(ctx);
}

hc_status hc_step(hc_ctx* ctx, int nsteps) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if (nsteps < 0) return fail(ctx, HC_INVALID_ARG, "nsteps must be >= 0");
  // HemoCell::sanityCheck (hemoCell.cpp:614-620)
  if (iv_active(ctx) &&
      (ctx->iv_grid_timescale % ctx->pv_timescale != 0 || ctx->iv_timescale % ctx->pv_timescale != 0))
    return fail(ctx, HC_INVALID_ARG,
                "interior viscosity timescales must be multiples of the particle velocity timescale");
  return fp32(ctx) ? step_impl<float>(ctx, nsteps) : step_impl<double>(ctx, nsteps);
}

hc_status hc_step_batch(hc_ctx* const* ctxs, int n, int nsteps) {
  if (!ctxs || n < 0 || nsteps < 0) return HC_INVALID_ARG;
  if (n == 0 || nsteps == 0) return HC_OK;
  std::vector<hc_ctx*> batch;
  for (int i = 0; i < n; ++i) {
    hc_ctx* c = ctxs[i];
    if (!c) return HC_INVALID_ARG;
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == c) return fail(c, HC_INVALID_ARG, "a context appears twice in the batch");
    hc_status s = ready_for_step(c);
    if (s != HC_OK) return s;
    if (c->device != ctxs[0]->device || fp32(c) != fp32(ctxs[0]))
      return fail(c, HC_INVALID_ARG, "batched contexts need one device and one precision");
  }
  for (int i = 0; i < n; ++i) {
    hc_ctx* c = ctxs[i];
    if (batchable(c)) {
      batch.push_back(c);
    } else {
      // its own steps on its own stream
      hc_status s = hc_step(c, nsteps);
      if (s != HC_OK) return s;
    }
  }
  if (batch.empty()) return HC_OK;
  return fp32(batch[0]) ? batch_impl<float>(batch.data(), (int)batch.size(), nsteps)
                        : batch_impl<double>(batch.data(), (int)batch.size(), nsteps);
}

hc_status hc_autotune(hc_ctx* ctx, int steps_per_trial, char* report, int report_len) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if (steps_per_trial < 1) return fail(ctx, HC_INVALID_ARG, "steps_per_trial must be >= 1");
  std::string rep = "{}";
  if (ctx->nv > 0 && ctx->fuse && !ctx->repulsion && !ctx->brep && !iv_active(ctx))
    s = fp32(ctx) ? autotune_impl<float>(ctx, steps_per_trial, rep) : autotune_impl<double>(ctx, steps_per_trial, rep);
  if (report && report_len > 0) {
    const size_t n = std::min<size_t>(rep.size(), (size_t)report_len - 1);
    std::memcpy(report, rep.data(), n);
    report[n] = '\0';
  }
  return s;
}

int hc_busy(hc_ctx* ctx) {
  if (!ctx) return -1;
  const hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipSuccess) return 0;
  if (e == hipErrorNotReady) return 1;
  ctx->err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
  return -1;
}

hc_status hc_synchronize(hc_ctx* ctx) {
  if (!ctx) return HC_INVALID_ARG;
  HC_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return HC_OK;
}

int64_t hc_iteration(const hc_ctx* ctx) { return ctx ? ctx->iter : -1; }

hc_status hc_set_iteration(hc_ctx* ctx, int64_t iter) {
  if (!ctx) return HC_INVALID_ARG;
  if (iter < 0) return fail(ctx, HC_INVALID_ARG, "iteration must be >= 0");
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  ctx->iter = iter;
  return HC_OK;
}

hc_status hc_stage_spread(hc_ctx* ctx) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  if (ctx->nv == 0) return HC_OK;
  return fp32(ctx) ? do_spread<float>(ctx) : do_spread<double>(ctx);
}
hc_status hc_stage_collide_stream(hc_ctx* ctx) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  // the lattice's data processors (the Lees-Edwards planes) run with it
  if (fp32(ctx)) {
    if ((s = do_collide<float>(ctx)) != HC_OK) return s;
    return ctx->le ? do_lees_edwards<float>(ctx) : HC_OK;
  }
  if ((s = do_collide<double>(ctx)) != HC_OK) return s;
  return ctx->le ? do_lees_edwards<double>(ctx) : HC_OK;
}
hc_status hc_stage_interpolate(hc_ctx* ctx) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  return fp32(ctx) ? do_interp_advance<float>(ctx, 1, 0) : do_interp_advance<double>(ctx, 1, 0);
}
hc_status hc_stage_advance(hc_ctx* ctx) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  return fp32(ctx) ? do_interp_advance<float>(ctx, 0, 1) : do_interp_advance<double>(ctx, 0, 1);
}
hc_status hc_stage_mechanics(hc_ctx* ctx, int forced) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  return do_mechanics(ctx, forced != 0);
}
hc_status hc_stage_repulsion(hc_ctx* ctx) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if (!ctx->repulsion) return fail(ctx, HC_STATE, "repulsion not enabled (hc_set_repulsion)");
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  return ctx->nv > 0 ? do_repulsion(ctx) : HC_OK;
}
hc_status hc_stage_boundary_repulsion(hc_ctx* ctx) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if (!ctx->brep) return fail(ctx, HC_STATE, "boundary repulsion not enabled (hc_set_boundary_repulsion)");
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  return ctx->nv > 0 ? do_boundary_repulsion(ctx) : HC_OK;
}
hc_status hc_stage_end(hc_ctx* ctx) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if ((s = invalidate_spread(ctx)) != HC_OK) return s;
  return fp32(ctx) ? do_end<float>(ctx) : do_end<double>(ctx);
}

int64_t hc_num_nodes(const hc_ctx* ctx) { return ctx && ctx->lattice_ready ? ctx->L.N : 0; }
int64_t hc_num_vertices(const hc_ctx* ctx) { return ctx ? ctx->nv : 0; }
int hc_num_cells(const hc_ctx* ctx) { return ctx ? (int)ctx->cell_base.size() : 0; }

hc_status hc_get_populations(hc_ctx* ctx, double* f) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!f) return fail(ctx, HC_INVALID_ARG, "null output");
  double* d = nullptr;
  if ((s = dalloc(ctx, (void**)&d, sizeof(double) * 19 * ctx->L.N, "population readback")) != HC_OK) return s;
  hipError_t e = fp32(ctx) ? launch_natural_populations<float>(ctx->stream, (float*)ctx->f, d, ctx->L, ctx->phase)
                           : launch_natural_populations<double>(ctx->stream, (double*)ctx->f, d, ctx->L, ctx->phase);
  if (e == hipSuccess) e = hipMemcpyAsync(f, d, sizeof(double) * 19 * ctx->L.N, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hipfail(ctx, e, "get_populations");
  return HC_OK;
}

hc_status hc_set_populations(hc_ctx* ctx, const double* f) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!f) return fail(ctx, HC_INVALID_ARG, "null input");
  double* d = nullptr;
  if ((s = dalloc(ctx, (void**)&d, sizeof(double) * 19 * ctx->L.N, "population upload")) != HC_OK) return s;
  hipError_t e = hipMemcpyAsync(d, f, sizeof(double) * 19 * ctx->L.N, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = fp32(ctx) ? launch_store_populations<float>(ctx->stream, (float*)ctx->f, d, ctx->L)
                  : launch_store_populations<double>(ctx->stream, (double*)ctx->f, d, ctx->L);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hipfail(ctx, e, "set_populations");
  ctx->phase = 0;
  return HC_OK;
}

hc_status hc_get_force_field(hc_ctx* ctx, double* F) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!F) return fail(ctx, HC_INVALID_ARG, "null output");
  const int64_t S = ctx->L.stride, N = ctx->L.N;
  double* d = nullptr;
  if ((s = dalloc(ctx, (void**)&d, sizeof(double) * 3 * S, "force readback")) != HC_OK) return s;
  // after a fused step the buffer already holds the next step's spread; the
  // reference's field at this point is the reset one
  hipError_t e = ctx->spread_ready ? hipMemsetAsync(d, 0, sizeof(double) * 3 * S, ctx->stream)
                 : fp32(ctx) ? launch_to_double<float>(ctx->stream, (const float*)ctx->Fibm[ctx->cur], d, 3 * S)
                             : launch_to_double<double>(ctx->stream, (const double*)ctx->Fibm[ctx->cur], d, 3 * S);
  for (int c = 0; c < 3 && e == hipSuccess; ++c)
    e = hipMemcpyAsync(F + c * N, d + c * S, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hipfail(ctx, e, "get_force_field");
  return HC_OK;
}

hc_status hc_get_fluid(hc_ctx* ctx, double* rho, double* u) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  const int64_t N = ctx->L.N;
  double *dr = nullptr, *du = nullptr;
  if (rho && (s = dalloc(ctx, (void**)&dr, sizeof(double) * N, "rho readback")) != HC_OK) return s;
  if (u && (s = dalloc(ctx, (void**)&du, sizeof(double) * 3 * N, "u readback")) != HC_OK) {
    if (dr) hipFree(dr);
    return s;
  }
  CollideArgs a{};
  a.omega = 1.0 / ctx->tau;
  for (int c = 0; c < 3; ++c) a.g[c] = ctx->g[c];
  std::memcpy(a.wall_u, ctx->wall_u, sizeof(a.wall_u));
  const void* Fc = ctx->spread_ready ? nullptr : ctx->Fibm[ctx->cur];  // reset field after a fused step
  hipError_t e = fp32(ctx) ? launch_fluid_moments<float>(ctx->stream, (const float*)ctx->f, (const float*)Fc,
                                                         ctx->type, ctx->L, ctx->phase, a, dr, du)
                           : launch_fluid_moments<double>(ctx->stream, (const double*)ctx->f, (const double*)Fc,
                                                          ctx->type, ctx->L, ctx->phase, a, dr, du);
  if (e == hipSuccess && rho) e = hipMemcpyAsync(rho, dr, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && u) e = hipMemcpyAsync(u, du, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (dr) hipFree(dr);
  if (du) hipFree(du);
  if (e != hipSuccess) return hipfail(ctx, e, "get_fluid");
  return HC_OK;
}

static hc_status get_planes(hc_ctx* ctx, const double* d, double* out);

hc_status hc_get_force_terms(hc_ctx* ctx, double* area, double* volume, double* bending, double* link,
                             double* visc, double* inner_link) {
  if (!ctx) return HC_INVALID_ARG;
  hc_status s = sync_cells(ctx);
  if (s != HC_OK) return s;
  if (ctx->nv == 0) return HC_OK;
  double* outs[6] = {area, volume, bending, link, visc, inner_link};
  const int terms[6] = {1, 2, 4, 8, 16, 32};  // TERM_AREA ... TERM_INNER of the mechanics kernels
  double* d = nullptr;
  if ((s = dalloc(ctx, (void**)&d, sizeof(double) * 3 * ctx->vcap, "force terms")) != HC_OK) return s;
  for (int t = 0; t < 6 && s == HC_OK; ++t) {
    if (!outs[t]) continue;
    hipError_t e = launch_fill_zero<double>(ctx->stream, d, 3 * ctx->vcap);
    for (auto& tp : ctx->types) {
      if (e != hipSuccess) break;
      TypeBufs& tb = *tp;
      if (tb.ncells == 0) continue;
      e = launch_mechanics(ctx->stream, vdev(ctx), tb.dev, tb.d_cells, tb.ncells, ctx->d_cell_base, ctx->bend,
                           ctx->lds_limit, terms[t], d);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) s = hipfail(ctx, e, "force terms");
    else s = get_planes(ctx, d, outs[t]);
  }
  hipFree(d);
  return s;
}

hc_status hc_get_fluid_stats(hc_ctx* ctx, int what, double* out) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  if (!out || (what != 0 && what != 1)) return fail(ctx, HC_INVALID_ARG, "fluid statistics: what is 0 or 1");
  return fp32(ctx) ? fluid_stats<float>(ctx, what, out) : fluid_stats<double>(ctx, what, out);
}

static hc_status get_planes(hc_ctx* ctx, const double* d, double* out) {
  const int64_t nv = ctx->nv;
  std::vector<double> plane(nv);
  for (int c = 0; c < 3; ++c) {
    HC_TRY(hipMemcpy(plane.data(), d + c * ctx->vcap, sizeof(double) * nv, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int64_t k = 0; k < nv; ++k) out[3 * k + c] = plane[k];
  }
  return HC_OK;
}
static hc_status set_planes(hc_ctx* ctx, double* d, const double* in) {
  const int64_t nv = ctx->nv;
  std::vector<double> plane(nv);
  for (int c = 0; c < 3; ++c) {
    for (int64_t k = 0; k < nv; ++k) plane[k] = in[3 * k + c];
    HC_TRY(hipMemcpy(d + c * ctx->vcap, plane.data(), sizeof(double) * nv, hipMemcpyHostToDevice), "hipMemcpy");
  }
  return HC_OK;
}

hc_status hc_get_vertices(hc_ctx* ctx, double* x, double* v, double* f) {
  if (!ctx) return HC_INVALID_ARG;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  hc_status s;
  if (x && (s = get_planes(ctx, ctx->x, x)) != HC_OK) return s;
  if (v && (s = get_planes(ctx, ctx->v, v)) != HC_OK) return s;
  if (f && (s = get_planes(ctx, ctx->fv, f)) != HC_OK) return s;
  return HC_OK;
}
hc_status hc_set_vertices(hc_ctx* ctx, const double* x, const double* v, const double* f) {
  if (!ctx) return HC_INVALID_ARG;
  hc_status s;
  if (ctx->lattice_ready && (s = invalidate_spread(ctx)) != HC_OK) return s;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  if (x && (s = set_planes(ctx, ctx->x, x)) != HC_OK) return s;
  if (v && (s = set_planes(ctx, ctx->v, v)) != HC_OK) return s;
  if (f && (s = set_planes(ctx, ctx->fv, f)) != HC_OK) return s;
  return HC_OK;
}
hc_status hc_get_removed_vertices(hc_ctx* ctx, uint8_t* removed) {
  if (!ctx || (!removed && ctx->nv > 0)) return HC_INVALID_ARG;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  if (ctx->nv > 0)
    HC_TRY(hipMemcpy(removed, ctx->vflag, (size_t)ctx->nv, hipMemcpyDeviceToHost), "hipMemcpy");
  for (int64_t k = 0; k < ctx->nv; ++k) removed[k] = (removed[k] & VFLAG_REMOVED) ? 1 : 0;
  return HC_OK;
}
hc_status hc_set_removed_vertices(hc_ctx* ctx, const uint8_t* removed) {
  if (!ctx || (!removed && ctx->nv > 0)) return HC_INVALID_ARG;
  hc_status s;
  if (ctx->lattice_ready && (s = invalidate_spread(ctx)) != HC_OK) return s;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  std::vector<uint8_t> h((size_t)ctx->nv);
  for (int64_t k = 0; k < ctx->nv; ++k) h[k] = removed[k] ? VFLAG_REMOVED : 0;
  if (ctx->nv > 0) HC_TRY(hipMemcpy(ctx->vflag, h.data(), (size_t)ctx->nv, hipMemcpyHostToDevice), "hipMemcpy");
  return HC_OK;
}
hc_status hc_get_incomplete_cells(hc_ctx* ctx, uint8_t* incomplete) {
  if (!ctx || (!incomplete && !ctx->cell_base.empty())) return HC_INVALID_ARG;
  std::vector<uint8_t> rm((size_t)ctx->nv);
  hc_status s = hc_get_removed_vertices(ctx, rm.data());
  if (s != HC_OK) return s;
  const size_t nc = ctx->cell_base.size();
  for (size_t c = 0; c < nc; ++c) {
    const int64_t b = ctx->cell_base[c], e = c + 1 < nc ? ctx->cell_base[c + 1] : ctx->nv;
    uint8_t any = 0;
    for (int64_t k = b; k < e; ++k) any |= rm[k];
    incomplete[c] = any;
  }
  return HC_OK;
}
hc_status hc_get_repulsion_force(hc_ctx* ctx, double* frep) {
  if (!ctx || !frep) return HC_INVALID_ARG;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  return get_planes(ctx, ctx->frep, frep);
}
hc_status hc_get_boundary_repulsion_force(hc_ctx* ctx, double* fbr) {
  if (!ctx || !fbr) return HC_INVALID_ARG;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  return get_planes(ctx, ctx->fbr, fbr);
}
hc_status hc_set_repulsion_force(hc_ctx* ctx, const double* frep) {
  if (!ctx) return HC_INVALID_ARG;
  if (!frep) return fail(ctx, HC_INVALID_ARG, "null input");
  hc_status s;
  if (ctx->lattice_ready && (s = invalidate_spread(ctx)) != HC_OK) return s;
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  return set_planes(ctx, ctx->frep, frep);
}

hc_status hc_get_cells(hc_ctx* ctx, int32_t* type_ids, int32_t* cell_ids) {
  if (!ctx) return HC_INVALID_ARG;
  const size_t nc = ctx->cell_base.size();
  for (size_t c = 0; c < nc; ++c) {
    if (type_ids) type_ids[c] = ctx->cell_type[c];
    if (cell_ids) cell_ids[c] = ctx->cell_id[c];
  }
  return HC_OK;
}

hc_status hc_get_cell_stats(hc_ctx* ctx, double* volume, double* area, double* centroid, double* velocity,
                            double* bbox, double* stretch) {
  if (!ctx) return HC_INVALID_ARG;
  hc_status s = sync_cells(ctx);
  if (s != HC_OK) return s;
  const int nc = (int)ctx->cell_base.size();
  if (nc == 0) return HC_OK;
  constexpr int NQ = 15;
  double* d = nullptr;
  if ((s = dalloc(ctx, (void**)&d, sizeof(double) * NQ * nc, "cell statistics")) != HC_OK) return s;
  hipError_t e = hipSuccess;
  for (auto& tp : ctx->types)
    if (e == hipSuccess && tp->ncells)
      e = launch_cell_stats(ctx->stream, vdev(ctx), tp->dev, tp->d_cells, tp->ncells, ctx->d_cell_base, d);
  std::vector<double> h((size_t)NQ * nc);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(double) * NQ * nc, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hipfail(ctx, e, "cell statistics");
  for (int c = 0; c < nc; ++c) {
    const double* q = &h[(size_t)NQ * c];
    if (volume) volume[c] = q[0];
    if (area) area[c] = q[1];
    for (int k = 0; k < 3; ++k) {
      if (centroid) centroid[3 * c + k] = q[2 + k];
      if (velocity) velocity[3 * c + k] = q[5 + k];
    }
    if (bbox)
      for (int k = 0; k < 6; ++k) bbox[6 * c + k] = q[8 + k];
    if (stretch) stretch[c] = q[14];
  }
  return HC_OK;
}

hc_status hc_get_cell_stats_noalloc(hc_ctx* ctx, double* stats, int64_t max_cells) {
  // no allocation on any path: the staging array is static, errors carry
  // static messages
  if (!ctx) return HC_INVALID_ARG;
  if (!stats || max_cells < 0) return fail_fixed(ctx, HC_INVALID_ARG, "hc_get_cell_stats_noalloc: null output");
  if (ctx->cells_dirty)
    return fail_fixed(ctx, HC_STATE,
                      "hc_get_cell_stats_noalloc: cells were added since the last step or statistics call and their "
                      "device tables are not built (building them allocates): call hc_step or hc_get_cell_stats "
                      "first");
  const int64_t nc = (int64_t)ctx->cell_base.size();
  if (nc > (int64_t)HC_CELL_STATS_NOALLOC_MAX)
    return fail_fixed(ctx, HC_INVALID_ARG,
                      "hc_get_cell_stats_noalloc: more cells than HC_CELL_STATS_NOALLOC_MAX (262144); use "
                      "hc_get_cell_stats");
  if (nc > max_cells)
    return fail_fixed(ctx, HC_INVALID_ARG, "hc_get_cell_stats_noalloc: the output holds fewer cells than the context");
  if (nc == 0) return HC_OK;
  double* d = ctx->noalloc_ok ? noalloc_scratch(ctx->device) : nullptr;
  if (!d) return fail_fixed(ctx, HC_DEVICE, "hc_get_cell_stats_noalloc: no staging array on this device");
  hipError_t e = hipSuccess;
  for (auto& tp : ctx->types)
    if (e == hipSuccess && tp->ncells)
      e = launch_cell_stats(ctx->stream, vdev(ctx), tp->dev, tp->d_cells, tp->ncells, ctx->d_cell_base, d);
  if (e == hipSuccess)
    e = hipMemcpyAsync(stats, d, sizeof(double) * HC_CELL_STATS_QUANTITIES * (size_t)nc, hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail_fixed(ctx, HC_DEVICE, "hc_get_cell_stats_noalloc: a HIP call failed");
  return HC_OK;
}

hc_status hc_get_cell_volume_area(hc_ctx* ctx, double* volume, double* area) {
  return hc_get_cell_stats(ctx, volume, area, nullptr, nullptr, nullptr, nullptr);
}

hc_status hc_get_vertex_bins(hc_ctx* ctx, int64_t* bins) {
  hc_status s = ready_for_step(ctx);
  if (s != HC_OK) return s;
  if (!bins) return fail(ctx, HC_INVALID_ARG, "null output");
  if (ctx->nv == 0) return HC_OK;
  int64_t* d = nullptr;
  if ((s = dalloc(ctx, (void**)&d, sizeof(int64_t) * ctx->nv, "vertex bins")) != HC_OK) return s;
  hipError_t e = launch_vertex_bins(ctx->stream, vdev(ctx), ctx->L, d);
  if (e == hipSuccess) e = hipMemcpyAsync(bins, d, sizeof(int64_t) * ctx->nv, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hipfail(ctx, e, "vertex bins");
  return HC_OK;
}

int64_t hc_wall_hits(hc_ctx* ctx) {
  if (!ctx) return -1;
  unsigned long long h = 0;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
  if (hipMemcpy(&h, ctx->d_wall_hits, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return (int64_t)h;
}

hc_status hc_check_finite(hc_ctx* ctx) {
  hc_status s = need_lattice(ctx);
  if (s != HC_OK) return s;
  HC_TRY(hipMemsetAsync(ctx->d_flag, 0, sizeof(unsigned int), ctx->stream), "memset");
  if (fp32(ctx)) HC_TRY(launch_check_finite_f(ctx->stream, (const float*)ctx->f, 19 * ctx->L.stride, ctx->d_flag), "check");
  else HC_TRY(launch_check_finite(ctx->stream, (const double*)ctx->f, 19 * ctx->L.stride, ctx->d_flag), "check");
  if (ctx->nv) HC_TRY(launch_check_finite(ctx->stream, ctx->x, 3 * ctx->vcap, ctx->d_flag), "check");
  unsigned int bad = 0;
  HC_TRY(hipMemcpyAsync(&bad, ctx->d_flag, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream), "copy");
  HC_TRY(hipStreamSynchronize(ctx->stream), "sync");
  return bad ? fail(ctx, HC_NUMERIC, "non-finite population or vertex coordinate") : HC_OK;
}

hc_status hc_profile_enable(hc_ctx* ctx, int enable) {
  if (!ctx) return HC_INVALID_ARG;
  prof_flush(ctx);
  ctx->prof.on = enable != 0;
  for (int k = 0; k < HC_PROF_NSTAGES; ++k) {
    ctx->prof.total_ms[k] = 0;
    ctx->prof.count[k] = 0;
  }
  return HC_OK;
}

hc_status hc_profile_read(hc_ctx* ctx, int stage, double* total_ms, int64_t* count) {
  if (!ctx || stage < 0 || stage >= HC_PROF_NSTAGES) return HC_INVALID_ARG;
  prof_flush(ctx);
  if (total_ms) *total_ms = ctx->prof.total_ms[stage];
  if (count) *count = ctx->prof.count[stage];
  return HC_OK;
}

}  // extern "C"
