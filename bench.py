"""Benchmark of the HemoCell IB-LBM time step on one MI355X.

Workload (the headline "Ht 30%" configuration): a 256x256x512 periodic D3Q19 box,
body force along z, 30% hematocrit of 642-vertex RBCs, 66-vertex platelets
(1 per 10 RBC) and 2562-vertex WBCs (1 per 500 RBC), seeded non-overlapping
placement; dx = 0.5 um, dt = 1e-7 s (tau = 1.82), reference default stiffness.

One step = HemoCell::iterate(): spread, collide-stream, interpolate + advance,
mechanics, force reset.  The timed region runs K steps back to back with inputs
resident in HBM; the device stream is synchronised on both sides.  `value` is
MLUPS (fluid nodes x steps / s) of the FP32-shifted production form; the FP64
form is measured in the same run.  The collide-stream kernel's mean duration is
taken from HIP events recorded on the library's stream around every launch in
the timed region.

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
# algorithmic bytes per fluid-node update of the collide-stream kernel:
# 19 population reads + 19 writes + 3 IBM-force reads + 3 writes zeroing the other force buffer
BYTES_PER_LUP = {"fp32": 44 * 4, "fp64": 44 * 8}


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
                            "WBC 2562v 1:500, seeded non-overlapping"),
    "channel": dict(shape="512,512,1024", hematocrit=0.35, wbc_per_rbc=0.0, walls=True,
                    text="%dx%dx%d channel, bounce-back walls y = 0 and y = ny-1, periodic x and z, body force "
                         "along z, Ht 35%%: RBC 642v, PLT 66v 1:10, seeded non-overlapping"),
}


def build_case(hc, susp, shape, precision, seed=20261202, workload="suspension"):
    w = WORKLOADS[workload]
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    types = susp.cell_types(units)
    cells = susp.pack(shape, types, hematocrit=w["hematocrit"], wbc_per_rbc=w["wbc_per_rbc"], seed=seed,
                      wall_axis=1 if w["walls"] else None)
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


def run_gpu(hc, susp, shape, precision, steps, warmup, workload="suspension", autotune=True):
    label = "fp32" if precision == hc.HC_FP32_SHIFTED else "fp64"
    print("[bench] building %s %s case %s" % (label, workload, shape), file=sys.stderr, flush=True)
    t0 = time.time()
    sim, units, types, cells = build_case(hc, susp, shape, precision, workload=workload)
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
    done_before = sim.iteration
    sim.profile_enable(True)
    sim.synchronize()
    t1 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    dt = time.perf_counter() - t1
    # per-step mean of each stage (fused path: "interp" is the velocity field and
    # "mechanics" the per-cell kernel doing interpolation, mechanics and next spread)
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
    x1 = sim.vertices()[0]
    N = int(np.prod(shape))
    res = dict(label=label, ms_per_step=1e3 * dt / steps, mlups=N * steps / dt / 1e6,
               vertex_updates_per_s=nv * steps / dt, stage_ms=stages, vertices=nv, nodes=N,
               cells={k: int(len(v)) for k, v in cells.items()})
    res["collide_ms"] = collide_ms
    res["tuning"] = {k: v["chosen"] for k, v in tuning.items()} if tuning else None
    res["steps_before"] = done_before
    res["wall_hits"] = int(wall_hits)
    res["x0"], res["x1"] = x0, x1
    res["collide_gbs"] = BYTES_PER_LUP[label] * N / (stages["collide_stream"] * 1e-3) / 1e9
    print("[bench] %s: %.3f ms/step, %.1f MLUPS, collide %.3f ms = %.0f GB/s" % (
        label, res["ms_per_step"], res["mlups"], res["collide_ms"], res["collide_gbs"]), file=sys.stderr, flush=True)
    sim.close()
    return res


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


def whole_step(r):
    """Algorithmic HBM bytes of one step over the step time: collide-stream (176/352 B per
    node), velocity field (22 reads + 4 writes per node) and the vertex state of the cell
    kernel (x read, x, v and f written: 96 B per vertex); the spread's atomic updates are
    not counted."""
    rs = 4 if r["label"] == "fp32" else 8
    b = (BYTES_PER_LUP[r["label"]] + 26 * rs) * r["nodes"] + 96 * r["vertices"]
    gbs = b / (r["ms_per_step"] * 1e-3) / 1e9
    return {"algorithmic_bytes": b, "achieved": round(gbs, 1), "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "cell_kernel_ms": round(r["stage_ms"]["cell_step_or_mechanics"], 4)}


def cpu_baseline(hc, susp, seconds=12.0):
    """FP64 CPU oracle (a restatement of the reference step, OpenMP over nodes and
    cells) on a 64x64x128 crop of the workload at the same hematocrit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import OracleSim, library
    shape = (64, 64, 128)
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    types = susp.cell_types(units)
    cells = susp.pack(shape, types, seed=7)
    o = OracleSim(*shape, units.tau)
    o.init_equilibrium(1.0, (0, 0, 0))
    o.set_body_force((0, 0, 1e-6))
    o.set_force_limit(units.f_limit)
    for name in ("RBC", "PLT", "WBC"):
        if name in cells:
            t = types[name]
            tid = o.add_cell_type(t.model, t.verts, t.tris, t.params, t.inner)
            o.add_cells(tid, cells[name])
    o.step(1)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds and n < 200:
        o.step(1)
        n += 1
    el = time.perf_counter() - t0
    threads = library().ho_num_threads()
    N = int(np.prod(shape))
    return dict(value=N * n / el / 1e6, unit="MLUPS", cores=int(threads), kind="port",
                sample="64x64x128 periodic crop at Ht %.2f (%d RBC, %d PLT, %d WBC), %d steps FP64 after 1 warm-up" % (
                    susp.hematocrit(types, cells, shape), len(cells["RBC"]), len(cells.get("PLT", [])),
                    len(cells.get("WBC", [])), n))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="both", choices=["both", "fp32", "fp64"])
    ap.add_argument("--workload", default="suspension", choices=sorted(WORKLOADS),
                    help="suspension: the Ht 30%% periodic box (the headline line); channel: the largest domain")
    ap.add_argument("--shape", default=None, help="nx,ny,nz (default: the workload's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", type=int, default=32, metavar="N",
                    help="cases of the serving stream (0 skips the latency measurement)")
    ap.add_argument("--serve-rate", type=float, default=1.5, help="Poisson arrival rate, cases/s")
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
    res = {}
    if args.precision in ("both", "fp32"):
        res["fp32"] = run_gpu(hc, susp, shape, hc.HC_FP32_SHIFTED, args.steps, args.warmup, args.workload,
                              not args.no_autotune)
    if args.precision in ("both", "fp64"):
        res["fp64"] = run_gpu(hc, susp, shape, hc.HC_FP64, args.steps, args.warmup, args.workload,
                              not args.no_autotune)
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
        "config": {"workload": wl["text"] % shape,
                   "cells": head["cells"], "vertices": head["vertices"], "precision": head["label"],
                   "wall_hits": head["wall_hits"]},
        "roofline": {"bound": "hbm", "achieved": round(head["collide_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["collide_gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_collide_stream", "mean_ms": round(head["collide_ms"], 4),
                     "bytes_per_lup": BYTES_PER_LUP[head["label"]],
                     "algorithmic_bytes": BYTES_PER_LUP[head["label"]] * head["nodes"]},
        "whole_step": whole_step(head),
        "vertex_updates_per_s": round(head["vertex_updates_per_s"], 1),
        "stage_ms": {k: round(v, 4) for k, v in head["stage_ms"].items()},
        "tuning": head["tuning"],
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
                        "stage_ms": {k: round(v, 4) for k, v in f64["stage_ms"].items()}}
        line["fp32_over_fp64"] = round(res["fp32"]["mlups"] / f64["mlups"], 3)
    # the serving stream and the CPU baseline belong to the headline workload
    if args.serve > 0 and args.workload == "suspension":
        line["serving"] = serve_stream(args.serve, args.serve_rate)
        # the same kind of stream arriving faster than it can be served: its
        # cases/s is the sustained rate, with the latencies at that rate
        line["serving_saturated"] = serve_stream(max(8, (3 * args.serve) // 4), 50.0, "serving (saturated)")
    if not args.no_cpu and args.workload == "suspension":
        line["cpu_baseline"] = cpu_baseline(hc, susp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
