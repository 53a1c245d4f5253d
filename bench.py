"""Benchmark of the HemoCell IB-LBM time step on one MI355X.

Workload (the headline "Ht 30%" configuration): a 256x256x512 periodic D3Q19 box,
body force along z, 30% hematocrit of 642-vertex RBCs, 66-vertex platelets
(1 per 10 RBC) and 2562-vertex WBCs (1 per 500 RBC) at seeded uniform random
positions and orientations, no two cells overlapping (the committed packing
hemocell_amd/data/suspension_256x256x512_ht30.npz, written by
`python -m hemocell_amd.packing`); dx = 0.5 um, dt = 1e-7 s (tau = 1.82),
reference default stiffness.

One step = HemoCell::iterate(): spread, collide-stream, interpolate + advance,
mechanics, force reset.  The timed region runs K steps back to back with inputs
resident in HBM, through the production path (steady steps replayed from the
captured HIP graph); the device stream is synchronised on both sides.  `value`
is MLUPS (fluid nodes x steps / s) of the FP32-shifted production form; the
FP64 form is measured in the same run.  A second pass of K steps, launched
eagerly with HIP events recorded on the library's stream around every kernel
launch, gives the per-stage times and the collide-stream kernel's mean duration
(the roofline figure).

python bench.py [--steps K] [--warmup W] [--gpus 1] [--precision both|fp32|fp64] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS & % HBM roofline, FP32-shifted vs FP64 at Ht 30%; p50/p99 case latency"
# algorithmic bytes per fluid-node update of the collide-stream kernel: 19 population
# reads + 19 writes + 3 IBM-force reads (the force reset counted with the read)
BYTES_PER_LUP = {"fp32": 41 * 4, "fp64": 41 * 8}
# with the 3 writes that clear the other force buffer, which the kernel also does
BYTES_PER_LUP_WITH_CLEAR = {"fp32": 44 * 4, "fp64": 44 * 8}
PACKED = os.path.join(ROOT, "hemocell_amd", "data", "suspension_256x256x512_ht30.npz")
CHANNEL_TILE = os.path.join(ROOT, "hemocell_amd", "data", "channel_tile_128x512x128_ht35.npz")
TUBE = os.path.join(ROOT, "hemocell_amd", "data", "tube_256x64x64_ht20.npz")


PROFILE = os.path.join(ROOT, "profiles", "step_suspension_256x256x512_ht30.md")


def profile_traffic(label):
    """Mean HBM bytes per launch of the collide-stream kernel (even and odd steps) from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, or None."""
    rtype = "float" if label == "fp32" else "double"
    tot = []
    try:
        for line in open(PROFILE):
            if line.startswith("| `k_collide_stream<%s," % rtype) and line.count("|") == 5:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                tot.append(float(cols[3]) * 1e6)
    except OSError:
        return None
    return sum(tot) / len(tot) if tot else None


# the two one-GPU workloads: the Ht 30% periodic suspension (the headline) and the
# largest domain, a 512x512x1024 channel between bounce-back planes y = 0 and
# y = 511 at Ht 35% (RBC and PLT 1:10)
WORKLOADS = {
    "suspension": dict(shape="256,256,512", hematocrit=0.30, wbc_per_rbc=1.0 / 500, walls=False,
                       text="periodic %dx%dx%d body-force suspension, Ht 30%%: RBC 642v, PLT 66v 1:10, "
                            "WBC 2562v 1:500, %s"),
    "channel": dict(shape="512,512,1024", hematocrit=0.35, wbc_per_rbc=0.0, walls=True,
                    text="%dx%dx%d channel, bounce-back walls y = 0 and y = ny-1, periodic x and z, body force "
                         "along z, Ht 35%%: RBC 642v, PLT 66v 1:10, %s (a 128x512x128 tile, 4 x 8)"),
}
PLACEMENT = {
    "random": "seeded uniform random positions and orientations, non-overlapping (force-biased packing)",
    "aligned": "seeded non-overlapping slot lattice, RBC disc axes along z (aligned variant)",
}


def place_cells(susp, types, shape, w, placement, seed=20261202):
    """{name: [n, nv, 3]} for the workload: the committed random packing of the
    headline box, a random packing made now for other shapes, or the aligned slot
    lattice."""
    from hemocell_amd import packing
    if placement == "aligned":
        return susp.pack(shape, types, hematocrit=w["hematocrit"], wbc_per_rbc=w["wbc_per_rbc"], seed=seed,
                         wall_axis=1 if w["walls"] else None)
    if not w["walls"] and shape == (256, 256, 512) and w["hematocrit"] == 0.30 and os.path.exists(PACKED):
        pk, _ = packing.load(PACKED)
    elif w["walls"] and shape == (512, 512, 1024) and os.path.exists(CHANNEL_TILE):
        # the channel tile (128 x 512 x 128, walls y = 0 and 511, periodic x and z) repeated 4 x 8
        pk = packing.tile(packing.load(CHANNEL_TILE)[0], (128, 512, 128), (4, 1, 8))
    else:
        fluid = float(np.prod(shape)) * ((shape[1] - 2) / shape[1] if w["walls"] else 1.0)
        counts = susp.cell_counts(types, fluid, w["hematocrit"], 0.1, w["wbc_per_rbc"])
        print("[bench] packing %s at random positions and orientations (seed %d)" % (counts, seed),
              file=sys.stderr, flush=True)
        pk = packing.pack(shape, types, counts, seed=seed, walls=1 if w["walls"] else None,
                          log=lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
    return packing.instances(types, pk)


def build_case(hc, susp, shape, precision, seed=20261202, workload="suspension", placement="random"):
    w = WORKLOADS[workload]
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    types = susp.cell_types(units)
    cells = place_cells(susp, types, shape, w, placement, seed)
    sim = hc.Simulation(0)
    if w["walls"]:
        mask = np.zeros(shape, dtype=np.uint8)
        mask[:, 0, :] = mask[:, shape[1] - 1, :] = hc.NODE_BOUNCEBACK
        sim.lattice_init(*shape, units.tau, periodic=(1, 0, 1), node_type=mask.reshape(-1), precision=precision)
        del mask
    else:
        sim.lattice_init(*shape, units.tau, periodic=(1, 1, 1), precision=precision)
    sim.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    sim.set_body_force((0.0, 0.0, 1e-6))
    sim.set_force_limit(units.f_limit)
    for name in ("RBC", "PLT", "WBC"):
        if name not in cells:
            continue
        t = types[name]
        tid = sim.add_cell_type(t.model, t.verts, t.tris, t.params, t.inner)
        sim.add_cells(tid, cells[name])
    return sim, units, types, cells


def run_gpu(hc, susp, shape, precision, steps, warmup, workload="suspension", autotune=True, placement="random"):
    label = "fp32" if precision == hc.HC_FP32_SHIFTED else "fp64"
    print("[bench] building %s %s case %s" % (label, workload, shape), file=sys.stderr, flush=True)
    t0 = time.time()
    sim, units, types, cells = build_case(hc, susp, shape, precision, workload=workload, placement=placement)
    nv = sim.num_vertices
    x0 = sim.vertices()[0]
    print("[bench] %d cells, %d vertices, Ht %.3f, setup %.1fs" % (
        sum(len(c) for c in cells.values()), nv, susp.hematocrit(types, cells, shape), time.time() - t0),
        file=sys.stderr, flush=True)
    sim.step(warmup)
    sim.synchronize()
    tuning = sim.autotune(2) if autotune else None
    if tuning:
        print("[bench] autotune: %s" % json.dumps({k: v["chosen"] for k, v in tuning.items()}), file=sys.stderr,
              flush=True)
    sim.step(2)  # the step graph is captured here, outside the timed region
    sim.synchronize()
    done_before = sim.iteration
    t1 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    dt = time.perf_counter() - t1
    x1 = sim.vertices()[0]
    # the same number of steps again, launched eagerly with HIP events around every
    # kernel launch: per-stage times (fused path: "interp" is the velocity field and
    # "mechanics" the per-cell kernel doing interpolation, mechanics and next spread)
    sim.profile_enable(True)
    sim.synchronize()
    tp = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    dtp = time.perf_counter() - tp
    stages = {}
    for name, k in (("spread", hc.PROF_SPREAD), ("collide_stream", hc.PROF_COLLIDE),
                    ("velocity_field_or_interp", hc.PROF_INTERP), ("cell_step_or_mechanics", hc.PROF_MECH)):
        ms, cnt = sim.profile_read(k)
        stages[name] = ms / steps
        if k == hc.PROF_COLLIDE:
            collide_ms = ms / max(cnt, 1)
    sim.profile_enable(False)
    sim.check_finite()
    wall_hits = sim.wall_hits()
    N = int(np.prod(shape))
    res = dict(label=label, ms_per_step=1e3 * dt / steps, mlups=N * steps / dt / 1e6,
               ms_per_step_eager=1e3 * dtp / steps, vertex_updates_per_s=nv * steps / dt, stage_ms=stages,
               vertices=nv, nodes=N, cells={k: int(len(v)) for k, v in cells.items()})
    res["collide_ms"] = collide_ms
    res["support_nodes"] = support_nodes(x1, shape)
    res["tuning"] = {k: v["chosen"] for k, v in tuning.items()} if tuning else None
    # every trial's ms per step (eager steps, one parameter varied at a time)
    res["tuning_ms"] = ({k: {c: round(t, 4) for c, t in v.items() if c != "chosen"} for k, v in tuning.items()}
                        if tuning else None)
    res["steps_before"] = done_before
    res["wall_hits"] = int(wall_hits)
    res["x0"], res["x1"] = x0, x1
    res["collide_gbs"] = BYTES_PER_LUP[label] * N / (collide_ms * 1e-3) / 1e9
    res["collide_gbs_with_clear"] = BYTES_PER_LUP_WITH_CLEAR[label] * N / (collide_ms * 1e-3) / 1e9
    print("[bench] %s: %.3f ms/step (graph), %.3f eager, %.1f MLUPS, collide %.3f ms = %.0f GB/s" % (
        label, res["ms_per_step"], res["ms_per_step_eager"], res["mlups"], collide_ms, res["collide_gbs"]),
        file=sys.stderr, flush=True)
    sim.close()
    return res


def copy_ceiling(mib=2048, reps=10):
    """GB/s of a plain device-to-device copy (read + write) of mib MiB, HIP events."""
    import torch
    a = torch.ones(mib * 2 ** 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * a.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def serve_stream(n, rate, label="serving"):
    """p50/p99 latency of a seeded Poisson stream of single-cell cases (shear 10-1000 1/s
    or stretch 0-200 pN on one RBC, PLT or WBC; cubic edge 32-96; 2000-20000 steps;
    FP32-shifted), each case in its own context and stream, the active cases stepped
    together in chunks (hc_step_batch)."""
    import hemocell_amd as hc
    from hemocell_amd import serving
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = serving.make_cases(n, rate, seed=7)
    print("[bench] %s: %d cases at %.1f/s" % (label, n, rate), file=sys.stderr, flush=True)
    done, wall = serving.serve(stream, units, max_active=32, chunk=250,
                               progress=lambda c, k: print("[bench] case %d done (%d/%d) %.2fs" % (
                                   c.ident, k, n, c.latency), file=sys.stderr, flush=True) if k % 8 == 0 else None)
    st = serving.latency_stats(done, wall)
    st.update(arrival_rate=rate, max_active=32, chunk=250, batched=True,
              workload="Poisson seed 7, shear or stretch, one RBC/PLT/WBC, edge 32-96, 2000-20000 steps, "
                       "FP32-shifted")
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()}


def support_nodes(x, shape):
    """Distinct lattice nodes in the phi2 supports (floor(x) + {0,1} per axis, periodic
    images) of the vertex positions x (n, 3)."""
    lo = np.floor(x).astype(np.int64)
    dims = np.asarray(shape, dtype=np.int64)
    keys = []
    for d in ((0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)):
        n = np.mod(lo + np.asarray(d), dims)
        keys.append((n[:, 0] * dims[1] + n[:, 1]) * dims[2] + n[:, 2])
    return int(np.unique(np.concatenate(keys)).size)


def whole_step(r):
    """Algorithmic HBM bytes of one step over the step time: collide-stream (164/328 B per
    node), velocity field (22 reads + 3 writes per node) and the vertex state of the cell
    kernel (x read, x, v and f written: 96 B per vertex).  `with_support` adds the IBM
    support traffic counted once per distinct support node touched in the step: the
    interpolation's 3 velocity reads and the spread's 3 force read-modify-writes."""
    rs = 4 if r["label"] == "fp32" else 8
    b = (BYTES_PER_LUP[r["label"]] + 25 * rs) * r["nodes"] + 96 * r["vertices"]
    gbs = b / (r["ms_per_step"] * 1e-3) / 1e9
    bs = b + 9 * rs * r["support_nodes"]
    gbs_s = bs / (r["ms_per_step"] * 1e-3) / 1e9
    return {"algorithmic_bytes": b, "achieved": round(gbs, 1), "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "cell_kernel_ms": round(r["stage_ms"]["cell_step_or_mechanics"], 4),
            "with_support": {"support_nodes": r["support_nodes"], "algorithmic_bytes": bs, "achieved": round(gbs_s, 1),
                             "frac": round(gbs_s / HBM_PEAK_GBS, 4)}}


def _oracle_rate(o, nodes, warmup, reps, steps):
    """Oracle MLUPS: `warmup` untimed steps, then `reps` timed repetitions of `steps`."""
    o.step(warmup)
    rates = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.step(steps)
        rates.append(nodes * steps / (time.perf_counter() - t0) / 1e6)
    return rates


def cpu_baseline(hc, susp):
    """The FP64 CPU oracle (a restatement of the reference step, not the reference
    binary, which cannot be built here: Palabos is not available; OpenMP over nodes
    and cells on the host's cores) on the single-cell shear box (50^3, one RBC), the
    Ht 20% tube (256x64x64) and a 128x128x256 crop of the Ht 30% suspension (the
    cells of the headline packing that lie wholly inside it): 10 warm-up steps, then
    3 timed repetitions each.  `value` is the crop's median MLUPS."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import OracleSim, library
    from hemocell_amd import packing
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    types = susp.cell_types(units)
    out = {}

    def add(o, cells):
        o.init_equilibrium(1.0, (0, 0, 0))
        o.set_force_limit(units.f_limit)
        for name in ("RBC", "PLT", "WBC"):
            if name in cells and len(cells[name]):
                t = types[name]
                o.add_cells(o.add_cell_type(t.model, t.verts, t.tris, t.params, t.inner), cells[name])

    # C1: 50^3 Couette between velocity walls z = 0, 49 at 100 1/s, one RBC
    n = 50
    mask = np.zeros((n, n, n), dtype=np.uint8)
    mask[:, :, 0] = hc.NODE_VELOCITY + 4
    mask[:, :, n - 1] = hc.NODE_VELOCITY + 5
    o = OracleSim(n, n, n, units.tau, periodic=(1, 1, 0), node_type=mask.reshape(-1))
    uw = (n - 1) * 100.0 * units.dt * 0.5
    o.set_wall_velocity(4, (uw, 0, 0))
    o.set_wall_velocity(5, (-uw, 0, 0))
    v = types["RBC"].verts
    add(o, {"RBC": (v - v.mean(axis=0) + (n - 1) / 2.0)[None]})
    out["single_cell_shear_50"] = _oracle_rate(o, n ** 3, 10, 3, 20)
    del o
    # C2: the Ht 20% tube
    L, D = 256, 64
    fluid = susp.tube_fluid(L, D)
    o = OracleSim(L, D, D, units.tau, periodic=(1, 0, 0),
                  node_type=np.where(fluid, 0, hc.NODE_BOUNCEBACK).astype(np.uint8).reshape(-1))
    R = D / 2.0 - 0.5
    o.set_body_force((4.0 * units.nu_lbm * (1e-3 * units.dt / units.dx) / (R * R), 0.0, 0.0))
    add(o, packing.instances(types, packing.load(TUBE)[0]))
    out["tube_ht20_256x64x64"] = _oracle_rate(o, int(fluid.sum()), 10, 3, 5)
    del o
    # 128x128x256 crop of the headline packing
    shape = (128, 128, 256)
    full = packing.instances(types, packing.load(PACKED)[0])
    crop = {}
    for name, x in full.items():
        lo, hi = x.min(axis=1), x.max(axis=1)
        inside = np.all((lo >= 1.0) & (hi <= np.array(shape) - 2.0), axis=1)
        crop[name] = x[inside]
    o = OracleSim(*shape, units.tau)
    o.set_body_force((0.0, 0.0, 1e-6))
    add(o, crop)
    rates = _oracle_rate(o, int(np.prod(shape)), 10, 3, 5)
    ht = susp.hematocrit(types, crop, shape)
    del o
    threads = int(library().ho_num_threads())
    return dict(value=float(np.median(rates)), unit="MLUPS", cores=threads, kind="port",
                sample="the FP64 oracle (restatement, not the reference binary) on a 128x128x256 crop of the Ht 30%% "
                       "suspension (%d RBC, %d PLT, %d WBC wholly inside, Ht %.3f), 10 warm-up steps then 3 x 5 "
                       "timed steps; median MLUPS" % (len(crop["RBC"]), len(crop["PLT"]), len(crop.get("WBC", [])),
                                                      ht),
                repetitions_mlups={"crop_128x128x256": [round(r, 3) for r in rates],
                                   **{k: [round(r, 3) for r in v] for k, v in out.items()}},
                host={"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
                      "omp_threads": threads})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="both", choices=["both", "fp32", "fp64"])
    ap.add_argument("--workload", default="suspension", choices=sorted(WORKLOADS),
                    help="suspension: the Ht 30%% periodic box (the headline line); channel: the largest domain")
    ap.add_argument("--placement", default="random", choices=sorted(PLACEMENT),
                    help="random: random positions and orientations (the workload as specified); aligned: the "
                         "slot-lattice variant with RBC disc axes along z")
    ap.add_argument("--shape", default=None, help="nx,ny,nz (default: the workload's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", type=int, default=200, metavar="N",
                    help="cases of the steady serving stream (0 skips the latency measurement)")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="Poisson arrival rate of the steady stream, cases/s (default: 0.9 x the sustained rate "
                         "measured on a saturated burst first)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a device-to-device copy of MIB MiB (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import hemocell_amd as hc
    from hemocell_amd import suspension as susp
    wl = WORKLOADS[args.workload]
    shape = tuple(int(s) for s in (args.shape or wl["shape"]).split(","))
    if args.copy_calibration:
        import torch
        a = torch.ones(args.copy_calibration * 2 ** 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print("[bench] calibration copy: %d MiB read + %d MiB written" % (args.copy_calibration, args.copy_calibration),
              file=sys.stderr, flush=True)
        del a, b
    # the copy ceiling first: torch's runtime initialises the device before the library does
    ceiling = copy_ceiling()
    print("[bench] device-to-device copy: %.0f GB/s" % ceiling, file=sys.stderr, flush=True)
    res = {}
    if args.precision in ("both", "fp32"):
        res["fp32"] = run_gpu(hc, susp, shape, hc.HC_FP32_SHIFTED, args.steps, args.warmup, args.workload,
                              not args.no_autotune, args.placement)
    if args.precision in ("both", "fp64"):
        res["fp64"] = run_gpu(hc, susp, shape, hc.HC_FP64, args.steps, args.warmup, args.workload,
                              not args.no_autotune, args.placement)
    head = res.get("fp32") or res["fp64"]
    traffic = profile_traffic(head["label"])
    if traffic is not None and (shape != (256, 256, 512) or args.workload != "suspension"):
        traffic = None  # the committed summary is for the 256x256x512 suspension only
    line = {
        "metric": METRIC,
        "value": round(head["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head["label"],
        "data": "synthetic",
        "config": {"workload": wl["text"] % (shape + (PLACEMENT[args.placement],)),
                   "placement": args.placement, "cells": head["cells"], "vertices": head["vertices"],
                   "precision": head["label"], "wall_hits": head["wall_hits"], "timed_path": "hip graph replay"},
        "roofline": {"bound": "hbm", "achieved": round(head["collide_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["collide_gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_collide_stream", "mean_ms": round(head["collide_ms"], 4),
                     "bytes_per_lup": BYTES_PER_LUP[head["label"]],
                     "algorithmic_bytes": BYTES_PER_LUP[head["label"]] * head["nodes"],
                     "achieved_with_force_clear": round(head["collide_gbs_with_clear"], 1),
                     "bytes_per_lup_with_force_clear": BYTES_PER_LUP_WITH_CLEAR[head["label"]],
                     "copy_ceiling_gbs": round(ceiling, 1),
                     "frac_of_copy_ceiling": round(head["collide_gbs_with_clear"] / ceiling, 4)},
        "whole_step": whole_step(head),
        "vertex_updates_per_s": round(head["vertex_updates_per_s"], 1),
        "stage_ms": {k: round(v, 4) for k, v in head["stage_ms"].items()},
        "ms_per_step_eager_profiled": round(head["ms_per_step_eager"], 4),
        "tuning": head["tuning"],
        "tuning_ms": head["tuning_ms"],
    }
    if "fp64" in res and "fp32" in res:
        f64 = res["fp64"]
        # FP32-shifted against FP64 after the same steps from the same suspension
        d32 = res["fp32"]["x1"] - res["fp32"]["x0"]
        d64 = f64["x1"] - f64["x0"]
        line["fp32_accuracy"] = {
            "vertex_position_rel_l2": float(np.linalg.norm(res["fp32"]["x1"] - f64["x1"]) / np.linalg.norm(f64["x1"])),
            "vertex_displacement_rel_l2": float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64)),
            "steps": res["fp32"]["steps_before"] + args.steps, "reference": "the FP64 form of this build"}
        line["fp64"] = {"value": round(f64["mlups"], 2), "ms_per_step": round(f64["ms_per_step"], 4),
                        "collide_gbs": round(f64["collide_gbs"], 1),
                        "collide_frac": round(f64["collide_gbs"] / HBM_PEAK_GBS, 4),
                        "stage_ms": {k: round(v, 4) for k, v in f64["stage_ms"].items()},
                        "tuning": f64["tuning"], "tuning_ms": f64["tuning_ms"]}
        line["fp32_over_fp64"] = round(res["fp32"]["mlups"] / f64["mlups"], 3)
    # the serving stream and the CPU baseline belong to the headline workload
    if args.serve > 0 and args.workload == "suspension":
        # a burst arriving faster than it can be served gives the sustained rate;
        # the steady stream then arrives at 90% of it
        sat = serve_stream(24, 50.0, "serving (saturated burst)")
        line["serving_saturated"] = sat
        rate = args.serve_rate or round(0.9 * sat["cases_per_s"], 3)
        line["serving"] = serve_stream(args.serve, rate)
    if not args.no_cpu and args.workload == "suspension":
        line["cpu_baseline"] = cpu_baseline(hc, susp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
