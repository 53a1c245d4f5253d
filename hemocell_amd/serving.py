"""Serving many small single-cell cases on one GPU.

Each case is one cell (RBC, PLT or WBC) in a cubic box, either sheared between
two velocity walls (the Couette set-up of helper/hemocellInit.hh:58-92) or
pulled by a stretching force on 7 vertices per end inside zero-velocity walls
(helper/hemoCellStretch.cpp:30-110, as in tests/validation/stretch_cell).
Cases arrive as a Poisson stream; every case gets its own context (its own
HIP stream and device buffers), so cases never share state and a case's
result does not depend on what else runs.  One host thread admits arrivals
while slots and the HBM budget allow (continuous batching: a case joins at the
next chunk boundary), steps all active cases together in chunks
(hc_step_batch: one launch per kernel and case group per step, the launches
asynchronous, the next chunk prepared while the previous one runs) and
retires the cases whose stream has drained; the latency of a case is its
completion time minus its arrival time, set-up included.
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np

import hemocell_amd as hc
from hemocell_amd import suspension


@dataclass
class Case:
    ident: int
    kind: str            # "shear" or "stretch"
    cell: str            # "RBC", "PLT" or "WBC"
    edge: int            # lattice nodes per side
    steps: int
    shear_rate: float = 0.0  # 1/s (shear)
    force_pN: float = 0.0    # total pulling force per end (stretch)
    arrival: float = 0.0     # s after the start of the stream
    # filled in by the server
    started: float = math.nan
    finished: float = math.nan
    result: dict = field(default_factory=dict)
    error: str = ""          # why the library failed the case (serve_native)

    @property
    def latency(self):
        return self.finished - self.arrival


def make_cases(n, rate, seed=7, edge=(32, 96), steps=(2000, 20000)):
    """A seeded stream of n cases arriving at `rate` cases/s (Poisson): shear rate
    log-uniform in 10..1000 1/s or stretch force uniform in 0..200 pN, cell kind,
    cubic edge and step count uniform in their ranges."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        kind = "shear" if rng.uniform() < 0.5 else "stretch"
        cell = ("RBC", "PLT", "WBC")[int(rng.integers(0, 3))]
        c = Case(ident=i, kind=kind, cell=cell, edge=int(rng.integers(edge[0], edge[1] + 1)),
                 steps=int(rng.integers(steps[0], steps[1] + 1)), arrival=t)
        if kind == "shear":
            c.shear_rate = float(10.0 ** rng.uniform(1.0, 3.0))
        else:
            c.force_pN = float(rng.uniform(0.0, 200.0))
        out.append(c)
    return out


_TYPES = {}


def _cell_types(units):
    key = (units.dx, units.dt)
    if key not in _TYPES:
        _TYPES[key] = suspension.cell_types(units, wbc_triangles=600)
    return _TYPES[key]


def build(case, units, precision=hc.HC_FP32_SHIFTED, device=0, exact_spread=False):
    """The context of one case, set up and ready to step (exact_spread: the
    order-independent spread of the FP64 form, for bit-reproducible runs)."""
    E = case.edge
    types = _cell_types(units)
    t = types[case.cell]
    sim = hc.Simulation(device)
    mask = np.zeros((E, E, E), dtype=np.uint8)
    if case.kind == "shear":
        # velocity walls z = 0 and z = E-1, periodic in x and y
        mask[:, :, 0] = hc.NODE_VELOCITY + 4
        mask[:, :, E - 1] = hc.NODE_VELOCITY + 5
        sim.lattice_init(E, E, E, units.tau, periodic=(1, 1, 0), node_type=mask.reshape(-1), precision=precision)
        v_half = (E - 1) * case.shear_rate * units.dt * 0.5
        sim.set_wall_velocity(4, (v_half, 0.0, 0.0))
        sim.set_wall_velocity(5, (-v_half, 0.0, 0.0))
    else:
        # zero-velocity walls on every face (the stretch validation's box)
        for a in range(3):
            for side, face in ((0, 2 * a), (E - 1, 2 * a + 1)):
                sl = [slice(None)] * 3
                sl[a] = side
                face_mask = mask[tuple(sl)]
                face_mask[face_mask == 0] = hc.NODE_VELOCITY + face
        sim.lattice_init(E, E, E, units.tau, periodic=(0, 0, 0), node_type=mask.reshape(-1), precision=precision)
    sim.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    sim.set_force_limit(units.f_limit)
    if exact_spread:
        sim.set_exact_spread(True)
    tid = sim.add_cell_type(t.model, t.verts, t.tris, t.params, t.inner)
    x = t.verts - 0.5 * (t.verts.min(axis=0) + t.verts.max(axis=0)) + (E - 1) / 2.0
    sim.add_cells(tid, x[None])
    if case.kind == "stretch":
        order = np.argsort(x[:, 0], kind="stable")
        n = 7
        f = case.force_pN * 1e-12 / units.df / n
        ids = np.concatenate([order[:n], order[::-1][:n]]).astype(np.int64)
        forces = np.zeros((2 * n, 3))
        forces[:n, 0] = -f
        forces[n:, 0] = f
        sim.set_extra_vertex_force(ids, forces)
    return sim


def summarise(sim):
    st = sim.cell_stats()
    return dict(volume=float(st["volume"][0]), stretch=float(st["stretch"][0]),
                bbox=[float(b) for b in st["bbox"][0]], centroid=[float(c) for c in st["centroid"][0]])


def run_alone(case, units, precision=hc.HC_FP32_SHIFTED, exact_spread=False):
    sim = build(case, units, precision, exact_spread=exact_spread)
    sim.step(case.steps)
    sim.synchronize()
    r = summarise(sim)
    x = sim.vertices()[0]
    sim.close()
    return r, x


def footprint(case, precision=hc.HC_FP32_SHIFTED):
    """Device bytes of one case: populations, two force buffers, the velocity
    field and the node types per node, plus the cell's vertex arrays."""
    s = 4 if precision == hc.HC_FP32_SHIFTED else 8
    nv = {"RBC": 642, "PLT": 66, "WBC": 642}[case.cell]
    return case.edge ** 3 * ((19 + 6 + 3) * s + 1) + nv * 7 * 3 * 8


def serve(cases, units, max_active=32, chunk=250, precision=hc.HC_FP32_SHIFTED, keep_vertices=False,
          progress=None, hbm_budget=64 << 30, batched=True, exact_spread=False):
    """Run the stream in real time; returns the cases with latencies and results.
    Arrivals are admitted while fewer than max_active cases run and their
    footprints fit in hbm_budget bytes."""
    pending = sorted(cases, key=lambda c: c.arrival)
    active = []  # (case, sim, steps_left)
    done = []
    t0 = time.perf_counter()
    k = 0
    while k < len(pending) or active:
        now = time.perf_counter() - t0
        used = sum(footprint(a[0], precision) for a in active)
        while (k < len(pending) and pending[k].arrival <= now and len(active) < max_active
               and (not active or used + footprint(pending[k], precision) <= hbm_budget)):
            c = pending[k]
            k += 1
            c.started = time.perf_counter() - t0
            used += footprint(c, precision)
            active.append([c, build(c, units, precision, exact_spread=exact_spread), c.steps])
        idle = True
        stepping = [a for a in active if a[2] > 0]
        if batched and stepping:
            # every active case advances by min(chunk, its steps left): one batch
            # per step count, so a case about to finish does not shorten the others
            by_n = {}
            for a in stepping:
                by_n.setdefault(min(chunk, a[2]), []).append(a)
            for n, group in by_n.items():
                hc.step_batch([a[1] for a in group], n)
                for a in group:
                    a[2] -= n
            idle = False
        for a in ([] if batched else active):
            c, sim, left = a
            if left > 0 and not sim.busy():
                n = min(chunk, left)
                sim.step(n)
                a[2] = left - n
                idle = False
        still = []
        for a in active:
            c, sim, left = a
            if left == 0 and not sim.busy():
                c.finished = time.perf_counter() - t0
                c.result = summarise(sim)
                if keep_vertices:
                    c.result["x"] = sim.vertices()[0]
                sim.close()
                done.append(c)
                if progress:
                    progress(c, len(done))
            else:
                still.append(a)
        active = still
        if idle and k < len(pending) and not active:
            # nothing in flight: wait for the next arrival
            time.sleep(max(0.0, min(0.01, pending[k].arrival - (time.perf_counter() - t0))))
        elif idle:
            time.sleep(0.0002)
    return done, time.perf_counter() - t0


def spec(case, units):
    """Keyword arguments of Server.submit for one case (the set-up of build())."""
    t = _cell_types(units)[case.cell]
    kw = dict(kind=hc.CASE_SHEAR if case.kind == "shear" else hc.CASE_STRETCH, edge=case.edge, steps=case.steps,
              model=t.model, vertices=t.verts, triangles=t.tris, params=t.params, inner_edges=t.inner,
              tau=units.tau, f_limit=units.f_limit)
    if case.kind == "shear":
        kw["wall_u"] = (case.edge - 1) * case.shear_rate * units.dt * 0.5
    else:
        kw["stretch_force"] = case.force_pN * 1e-12 / units.df
        kw["stretch_n"] = 7
    return kw


def serve_native(cases, units, max_active=32, chunk=250, precision=hc.HC_FP32_SHIFTED, keep_vertices=False,
                 hbm_budget=64 << 30, spec_of=None):
    """The same stream through the library's own request loop (hc_server_*):
    arrivals are submitted as they come, every poll retires, admits and steps.
    A case the library fails (set-up, batch, stream or read-back error) ends
    with c.error set and no result; the other cases go on.  Returns the
    finished cases (failed ones included) and the wall time.  spec_of(case,
    units) gives a case's submission arguments (default: spec)."""
    spec_of = spec_of or spec
    server = hc.Server(0, max_active, hbm_budget)
    pending = sorted(cases, key=lambda c: c.arrival)
    ids = {}
    done = []
    last_error = ""
    t0 = time.perf_counter()
    k = 0
    try:
        while k < len(pending) or ids:
            now = time.perf_counter() - t0
            while k < len(pending) and pending[k].arrival <= now:
                c = pending[k]
                k += 1
                c.started = now
                ids[c.ident] = (c, server.submit(precision=precision, **spec_of(c, units)))
            try:
                server.poll(chunk)
            except hc.HcError as e:
                # the failed cases read CASE_FAILED below; the others keep their steps
                last_error = str(e)
            for ident, (c, cid) in list(ids.items()):
                st = server.status(cid)
                if st == hc.CASE_DONE:
                    c.finished = time.perf_counter() - t0
                    x, cs = server.result(cid)
                    c.result = dict(volume=float(cs["volume"]), stretch=float(cs["stretch"]))
                    if keep_vertices:
                        c.result["x"] = x
                elif st == hc.CASE_FAILED:
                    c.finished = time.perf_counter() - t0
                    c.result = None
                    c.error = last_error
                else:
                    continue
                done.append(c)
                del ids[ident]
            if not ids and k < len(pending):
                time.sleep(max(0.0, min(0.01, pending[k].arrival - (time.perf_counter() - t0))))
    finally:
        server.close()
    return done, time.perf_counter() - t0


def latency_stats(done, wall):
    """p50/p99/mean latency and throughput of the cases that finished; failed
    cases (serve_native, result None) are counted apart."""
    ok = [c for c in done if c.result is not None]
    lat = np.array([c.latency for c in ok])
    return dict(cases=len(ok), failed=len(done) - len(ok), p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()), cases_per_s=len(ok) / wall,
                steps_per_s=float(sum(c.steps for c in ok) / wall),
                lups_per_s=float(sum(c.steps * c.edge ** 3 for c in ok) / wall))


def main(argv=None):
    """Serve a seeded Poisson stream and print its latency statistics as one JSON line:
    python -m hemocell_amd.serving --cases 24 --rate 50 [--native] [--per-case]"""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(prog="python -m hemocell_amd.serving")
    ap.add_argument("--cases", type=int, default=24)
    ap.add_argument("--rate", type=float, default=50.0, help="Poisson arrival rate, cases/s")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--max-active", type=int, default=32)
    ap.add_argument("--chunk", type=int, default=250)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--native", action="store_true", help="the library's request loop (hc_server_*)")
    ap.add_argument("--per-case", action="store_true", help="step each case on its own (no hc_step_batch)")
    a = ap.parse_args(argv)
    prec = hc.HC_FP32_SHIFTED if a.precision == "fp32" else hc.HC_FP64
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = make_cases(a.cases, a.rate, seed=a.seed)
    print("[serving] %d cases at %.1f/s, %s" % (a.cases, a.rate, "native" if a.native else "python loop"),
          file=sys.stderr, flush=True)
    if a.native:
        done, wall = serve_native(stream, units, a.max_active, a.chunk, prec)
    else:
        done, wall = serve(stream, units, a.max_active, a.chunk, prec, batched=not a.per_case)
    st = latency_stats(done, wall)
    st.update(arrival_rate=a.rate, max_active=a.max_active, chunk=a.chunk, precision=a.precision,
              loop="native" if a.native else ("per-case" if a.per_case else "batched"))
    print(json.dumps(st))


if __name__ == "__main__":
    main()
