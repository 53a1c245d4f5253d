// hc_server.cpp -- a native request loop for many small single-cell cases.
//
// No counterpart in the reference (one HemoCell per process, one case per
// run); it serves the stream of independent shear and stretch cases that the
// reference's examples/oneCellShear and tests/validation/stretch_cell runs
// would each be.  Built only on the public C ABI of this library:
//   * a case is one context (its own buffers and stream): a cubic box with
//     velocity walls on the z faces (shear, periodic in x and y, walls at
//     +-u along x; helper/hemocellInit.hh:58-92) or on every face at rest
//     (stretch, with +-F/n on the n lowest and highest vertices in x;
//     helper/hemoCellStretch.cpp:30-110), one cell centred in it;
//   * hc_server_poll retires the cases whose stream has drained (keeping their
//     final positions and cell statistics), admits queued cases in arrival
//     order while fewer than max_active run and their footprints fit the HBM
//     budget, and advances every running case by min(chunk, its steps left)
//     with hc_step_batch (one launch per kernel and case group per step; one
//     batch per precision and step count).  The launches are asynchronous.
//   * a case fails alone: a set-up, batch, stream or read-back error marks it
//     HC_CASE_FAILED and releases its context; the other cases go on and poll
//     returns the first error of the pass.
#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hemocell_gpu.h"

struct hc_server {
  int device = 0;
  int max_active = 32;
  double budget = 0;
  struct Case {
    int64_t id = 0;
    hc_case_spec spec{};
    std::vector<double> verts;
    std::vector<int32_t> tris, inner;
    hc_ctx* ctx = nullptr;
    int64_t left = 0;
    int status = HC_CASE_QUEUED;
    double bytes = 0;
    std::vector<double> x;   // final positions
    double stats[3] = {0, 0, 0};  // volume, area, largest vertex-vertex distance
    int inject = HC_SERVER_INJECT_NONE;  // host-side failure forced by hc_server_inject_failure
  };
  std::map<int64_t, std::unique_ptr<Case>> cases;
  std::deque<Case*> queued;
  std::vector<Case*> active;
  int64_t next_id = 0;
  std::string err;
};

namespace {

hc_status serr(hc_server* s, hc_status st, const std::string& m) {
  if (s) s->err = m;
  return st;
}

double footprint(const hc_case_spec& c) {
  const double e = (double)c.edge, r = c.precision == HC_FP32_SHIFTED ? 4.0 : 8.0;
  return e * e * e * ((19 + 6 + 3) * r + 1) + (double)c.nv * 7 * 3 * 8;
}

// the context of one case, set up and ready to step
hc_status build(hc_server* s, hc_server::Case& c) {
  const hc_case_spec& sp = c.spec;
  const int E = sp.edge;
  hc_ctx* ctx = nullptr;
  if (hc_create(s->device, &ctx) != HC_OK) return serr(s, HC_DEVICE, "hc_create failed");
  c.ctx = ctx;
  std::vector<uint8_t> mask((size_t)E * E * E, HC_NODE_FLUID);
  auto at = [&](int x, int y, int z) -> uint8_t& { return mask[(size_t)z + (size_t)E * ((size_t)y + (size_t)E * x)]; };
  int periodic[3] = {0, 0, 0};
  if (sp.kind == HC_CASE_SHEAR) {
    periodic[0] = periodic[1] = 1;
    for (int x = 0; x < E; ++x)
      for (int y = 0; y < E; ++y) {
        at(x, y, 0) = (uint8_t)(HC_NODE_VELOCITY + 4);
        at(x, y, E - 1) = (uint8_t)(HC_NODE_VELOCITY + 5);
      }
  } else {
    // every face at rest; an edge or corner node takes the first of its faces
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side)
        for (int i = 0; i < E; ++i)
          for (int j = 0; j < E; ++j) {
            const int k = side ? E - 1 : 0;
            int p[3];
            p[a] = k;
            p[(a + 1) % 3] = i;
            p[(a + 2) % 3] = j;
            uint8_t& m = at(p[0], p[1], p[2]);
            if (m == HC_NODE_FLUID) m = (uint8_t)(HC_NODE_VELOCITY + 2 * a + side);
          }
  }
  hc_status st;
  if ((st = hc_lattice_init(ctx, E, E, E, mask.data(), sp.tau, periodic, sp.precision)) != HC_OK) return st;
  if (sp.kind == HC_CASE_SHEAR) {
    const double up[3] = {sp.wall_u, 0.0, 0.0}, dn[3] = {-sp.wall_u, 0.0, 0.0};
    if ((st = hc_set_wall_velocity(ctx, 4, up)) != HC_OK || (st = hc_set_wall_velocity(ctx, 5, dn)) != HC_OK) return st;
  }
  const double zero[3] = {0, 0, 0};
  if ((st = hc_init_equilibrium(ctx, 1.0, zero)) != HC_OK) return st;
  if ((st = hc_set_force_limit(ctx, sp.f_limit)) != HC_OK) return st;
  int tid = 0;
  if ((st = hc_add_cell_type(ctx, sp.model, sp.nv, sp.nt, c.tris.data(), c.verts.data(), sp.n_inner,
                             c.inner.empty() ? nullptr : c.inner.data(), &sp.params, 1, &tid)) != HC_OK)
    return st;
  // the cell centred in the box (bounding-box centre at (E-1)/2)
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = c.verts[(size_t)a];
  for (int k = 1; k < sp.nv; ++k)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], c.verts[(size_t)3 * k + a]);
      hi[a] = std::max(hi[a], c.verts[(size_t)3 * k + a]);
    }
  std::vector<double> x((size_t)3 * sp.nv);
  for (int k = 0; k < sp.nv; ++k)
    for (int a = 0; a < 3; ++a)
      x[(size_t)3 * k + a] = c.verts[(size_t)3 * k + a] - 0.5 * (lo[a] + hi[a]) + (E - 1) / 2.0;
  if ((st = hc_add_cells(ctx, tid, 1, x.data(), nullptr)) != HC_OK) return st;
  if (sp.kind == HC_CASE_STRETCH && sp.stretch_n > 0) {
    const int n = std::min(sp.stretch_n, sp.nv / 2);
    std::vector<int> order((size_t)sp.nv);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return x[(size_t)3 * a] < x[(size_t)3 * b]; });
    std::vector<int64_t> ids((size_t)2 * n);
    std::vector<double> f((size_t)6 * n, 0.0);
    for (int k = 0; k < n; ++k) {
      ids[(size_t)k] = order[(size_t)k];
      ids[(size_t)(n + k)] = order[(size_t)(sp.nv - 1 - k)];
      f[(size_t)3 * k] = -sp.stretch_force / n;
      f[(size_t)3 * (n + k)] = sp.stretch_force / n;
    }
    if ((st = hc_set_extra_vertex_force(ctx, 2 * n, ids.data(), f.data())) != HC_OK) return st;
  }
  return HC_OK;
}

hc_status retire(hc_server* s, hc_server::Case& c) {
  hc_status st;
  if (c.inject == HC_SERVER_INJECT_RETIRE) return serr(s, HC_STATE, "injected retire failure");
  c.x.assign((size_t)3 * c.spec.nv, 0.0);
  std::vector<double> v(c.x.size()), f(c.x.size());
  if ((st = hc_get_vertices(c.ctx, c.x.data(), v.data(), f.data())) != HC_OK) return serr(s, st, hc_last_error(c.ctx));
  if ((st = hc_get_cell_stats(c.ctx, &c.stats[0], &c.stats[1], nullptr, nullptr, nullptr, &c.stats[2])) != HC_OK)
    return serr(s, st, hc_last_error(c.ctx));
  hc_destroy(c.ctx);
  c.ctx = nullptr;
  c.status = HC_CASE_DONE;
  return HC_OK;
}

}  // namespace

extern "C" {

hc_status hc_server_create(int device, int max_active, double hbm_budget_bytes, hc_server** out) {
  if (!out || max_active < 1 || !(hbm_budget_bytes > 0)) return HC_INVALID_ARG;
  hc_server* s = new hc_server();
  s->device = device;
  s->max_active = max_active;
  s->budget = hbm_budget_bytes;
  *out = s;
  return HC_OK;
}

void hc_server_destroy(hc_server* s) {
  if (!s) return;
  for (auto& kv : s->cases)
    if (kv.second->ctx) hc_destroy(kv.second->ctx);
  delete s;
}

const char* hc_server_last_error(const hc_server* s) { return s ? s->err.c_str() : "null server"; }

hc_status hc_case_submit(hc_server* s, const hc_case_spec* spec, int64_t* case_id) {
  if (!s || !spec || !case_id) return HC_INVALID_ARG;
  const hc_case_spec& c = *spec;
  if (c.edge < 4 || c.steps < 0 || (c.kind != HC_CASE_SHEAR && c.kind != HC_CASE_STRETCH) || c.nv < 4 ||
      (c.precision != HC_FP64 && c.precision != HC_FP32_SHIFTED) ||
      c.nt < 4 || !c.vertices || !c.triangles || (c.n_inner > 0 && !c.inner_edges) || !(c.tau > 0.5))
    return serr(s, HC_INVALID_ARG, "bad case specification");
  auto p = std::make_unique<hc_server::Case>();
  p->id = s->next_id++;
  p->spec = c;
  p->verts.assign(c.vertices, c.vertices + (size_t)3 * c.nv);
  p->tris.assign(c.triangles, c.triangles + (size_t)3 * c.nt);
  if (c.n_inner > 0) p->inner.assign(c.inner_edges, c.inner_edges + (size_t)2 * c.n_inner);
  p->spec.vertices = nullptr;  // the copies above are used from here on
  p->spec.triangles = nullptr;
  p->spec.inner_edges = nullptr;
  p->left = c.steps;
  p->bytes = footprint(c);
  s->queued.push_back(p.get());
  *case_id = p->id;
  s->cases[p->id] = std::move(p);
  return HC_OK;
}

hc_status hc_server_poll(hc_server* s, int chunk, int* n_running) {
  if (!s || chunk < 1) return HC_INVALID_ARG;
  hc_status st, first = HC_OK;
  // a case that cannot go on: its context is released, the others keep going,
  // and the first such error of this poll is reported
  auto fail_case = [&](hc_server::Case* c, hc_status why, const std::string& m) {
    if (c->ctx) hc_destroy(c->ctx);
    c->ctx = nullptr;
    c->status = HC_CASE_FAILED;
    if (first == HC_OK) {
      first = why;
      s->err = "case " + std::to_string(c->id) + ": " + m;
    }
  };
  // retire the cases that have taken all their steps and whose stream drained;
  // a stream that reports an error fails its case
  std::vector<hc_server::Case*> still;
  for (auto* c : s->active) {
    const int busy = c->inject == HC_SERVER_INJECT_STREAM_ERROR ? -1 : hc_busy(c->ctx);
    if (busy < 0) {
      const std::string m = c->inject == HC_SERVER_INJECT_STREAM_ERROR ? "injected stream error" : hc_last_error(c->ctx);
      fail_case(c, HC_DEVICE, m);
    } else if (c->left == 0 && busy == 0) {
      if ((st = retire(s, *c)) != HC_OK) fail_case(c, st, s->err);
    } else {
      still.push_back(c);
    }
  }
  s->active.swap(still);
  // admit in arrival order while slots and the HBM budget allow (one case always fits)
  double used = 0;
  for (auto* c : s->active) used += c->bytes;
  while (!s->queued.empty() && (int)s->active.size() < s->max_active &&
         (s->active.empty() || used + s->queued.front()->bytes <= s->budget)) {
    hc_server::Case* c = s->queued.front();
    s->queued.pop_front();
    if ((st = build(s, *c)) != HC_OK) {
      fail_case(c, st, c->ctx ? hc_last_error(c->ctx) : s->err);
      continue;
    }
    c->status = HC_CASE_RUNNING;
    used += c->bytes;
    s->active.push_back(c);
  }
  // every case with steps left takes min(chunk, its steps left): one batch per
  // precision and step count (a case shorter than the chunk ends in its own
  // batch instead of shortening everybody's)
  std::map<std::pair<int, int64_t>, std::vector<hc_server::Case*>> groups;
  for (auto* c : s->active)
    if (c->left > 0) groups[{(int)c->spec.precision, std::min<int64_t>(chunk, c->left)}].push_back(c);
  for (auto& kv : groups) {
    const int64_t n = kv.first.second;
    std::vector<hc_ctx*> ctxs;
    for (auto* c : kv.second) ctxs.push_back(c->ctx);
    if ((st = hc_step_batch(ctxs.data(), (int)ctxs.size(), (int)n)) == HC_OK) {
      for (auto* c : kv.second) c->left -= n;
      continue;
    }
    // a batch that failed leaves its cases in an unknown state: they fail, the
    // other batches keep their steps
    std::string m;
    for (hc_ctx* x : ctxs)
      if (*hc_last_error(x)) m = hc_last_error(x);
    for (auto* c : kv.second) fail_case(c, st, m);
  }
  s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                 [](hc_server::Case* c) { return c->status == HC_CASE_FAILED; }),
                  s->active.end());
  if (n_running) *n_running = (int)(s->active.size() + s->queued.size());
  return first;
}

hc_status hc_server_inject_failure(hc_server* s, int64_t case_id, int what) {
  if (!s || (what != HC_SERVER_INJECT_NONE && what != HC_SERVER_INJECT_RETIRE &&
             what != HC_SERVER_INJECT_STREAM_ERROR))
    return HC_INVALID_ARG;
  auto it = s->cases.find(case_id);
  if (it == s->cases.end()) return serr(s, HC_INVALID_ARG, "unknown case");
  it->second->inject = what;
  return HC_OK;
}

int hc_case_status(const hc_server* s, int64_t case_id) {
  if (!s) return -1;
  auto it = s->cases.find(case_id);
  return it == s->cases.end() ? -1 : it->second->status;
}

hc_status hc_case_result(hc_server* s, int64_t case_id, double* x, double* stats) {
  if (!s) return HC_INVALID_ARG;
  auto it = s->cases.find(case_id);
  if (it == s->cases.end()) return serr(s, HC_INVALID_ARG, "unknown case");
  hc_server::Case& c = *it->second;
  if (c.status != HC_CASE_DONE) return serr(s, HC_STATE, "the case has not finished");
  if (x) std::memcpy(x, c.x.data(), sizeof(double) * c.x.size());
  if (stats) std::memcpy(stats, c.stats, sizeof(c.stats));
  return HC_OK;
}

}  // extern "C"
