"""Checkpoint and resume on the GPU.

* Through the C ABI: the state a step boundary carries (populations, vertex
  position, velocity, force and repulsion force, cells, iteration) read from one
  context and written into a fresh one continues the run as the uninterrupted
  context does, with the material, particle-velocity and repulsion timescales
  above one so that the restored iteration and the kept repulsion force matter
  (HemoCellFields::save / load, core/hemoCellFields.cpp:240-319).
* Through hemo::HemoCell: saveCheckPoint() writes checkpoint.xml, lattice.dat and
  particleField.dat, renames the previous set to *.old, and a run started from
  checkpoint.xml (loadCheckPoint(), the reference examples' restart path) ends
  where the uninterrupted run ends.

* Against the oracle: with the material timescale 2 (every type), the particle
  velocity timescale 2 and the repulsion timescale 3, a GPU run checkpointed at
  iteration 8 and resumed in a fresh context through the C ABI ends, after 17
  steps, where the uninterrupted FP64 oracle ends: positions, displacements and
  populations within 1e-10 relative L2 (the resumed timescale phases line up
  with the oracle's only if the restored iteration is right).

The resumed run spreads its first step's forces with the stand-alone spread
kernel where the uninterrupted one used the fused kernel's spread, and float
atomics commit in no fixed order, so the runs agree to round-off, not bit for
bit: FP64 positions within 1e-12 and populations within 1e-11 (relative L2),
FP32 positions within 1e-6 and populations within 1e-6, and in every case no
further apart than 100 times what two uninterrupted runs differ by (floored at
the FP64 or FP32 unit round-off).
"""
import json
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import cases
import hemocell_amd as hc
from oracle import OracleSim

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hemocell_amd", "bin")
SHAPE = (40, 30, 36)


def _types(u):
    return [cases.mesh(kind, u, hc.build_mesh) for kind in ("RBC", "PLT", "WBC")]


def _configure(s, u, types, variant):
    """Everything a driver sets up again before it loads a checkpoint."""
    s.set_body_force([0, 0, 2e-6])
    s.set_force_limit(u.f_limit)
    tids = [s.add_cell_type(model, v, t, p, inner) for model, v, t, p, inner in types]
    if variant == "timescales":
        for tid in tids:
            s.set_material_timescale(tid, 2)
        s.set_velocity_timescale(2)
        s.set_repulsion(2e-3, 1.2, 3)
    return tids


def _fresh(u, types, variant, precision):
    s = hc.Simulation(0)
    s.lattice_init(*SHAPE, u.tau, precision=precision)
    return s, _configure(s, u, types, variant)


def _load_cells(s, tids, types):
    rng = np.random.default_rng(8)
    for tid, (kind, centers), (model, v, t, p, inner) in zip(
            tids, (("RBC", [(10, 10, 10), (28, 18, 22)]), ("PLT", [(20, 25, 5), (33, 5, 30)]),
                   ("WBC", [(25, 12, 28)])), types):
        xs = np.stack([cases.perturb(cases.place(v, c, cases.rotation(rng)), rng, 0.03) for c in centers])
        s.add_cells(tid, xs)


def _save(s):
    x, v, f = s.vertices()
    tid, cid = s.cells()
    return dict(pop=s.populations(), x=x, v=v, f=f, frep=s.repulsion_force(), tid=tid, cid=cid, it=s.iteration)


def _restore(s, tids, types, st):
    nvs = {tid: len(types[k][1]) for k, tid in enumerate(tids)}
    off, c = 0, 0
    while c < len(st["tid"]):
        e = c
        while e < len(st["tid"]) and st["tid"][e] == st["tid"][c]:
            e += 1
        nv = nvs[int(st["tid"][c])]
        xs = st["x"][off:off + (e - c) * nv].reshape(e - c, nv, 3)
        s.add_cells(int(st["tid"][c]), xs, st["cid"][c:e])
        off += (e - c) * nv
        c = e
    s.set_vertices(st["x"], st["v"], st["f"])
    s.set_repulsion_force(st["frep"])
    s.set_populations(st["pop"])
    s.set_iteration(st["it"])


@pytest.mark.parametrize("variant,precision", [("fused", hc.HC_FP64), ("timescales", hc.HC_FP64),
                                               ("fused", hc.HC_FP32_SHIFTED)])
def test_state_round_trip_through_the_c_abi_resumes_the_run(variant, precision):
    print("[checkpoint] c-abi", variant, precision, flush=True)
    u = cases.units()
    types = _types(u)
    runs = []
    for _ in range(3):
        s, tids = _fresh(u, types, variant, precision)
        _load_cells(s, tids, types)
        runs.append(s)
    whole, whole2, first = runs
    whole.step(17)
    whole2.step(17)
    first.step(8)
    st = _save(first)
    assert st["it"] == 8
    if variant == "timescales":
        assert np.abs(st["frep"]).max() > 0  # computed at iteration 6, still acting at 8
    resumed, tids = _fresh(u, types, variant, precision)
    _restore(resumed, tids, types, st)
    assert resumed.iteration == 8
    t2, c2 = resumed.cells()
    assert np.array_equal(t2, st["tid"]) and np.array_equal(c2, st["cid"])
    resumed.step(9)
    assert resumed.iteration == whole.iteration == 17
    (xa, va, _), (x2, v2, _), (xb, vb, _) = whole.vertices(), whole2.vertices(), resumed.vertices()
    fa, f2, fb = whole.populations(), whole2.populations(), resumed.populations()
    fp64 = precision == hc.HC_FP64
    floor = 1e-15 if fp64 else 1e-7
    assert cases.rel_l2(xa, xb) < (1e-12 if fp64 else 1e-6)
    assert cases.rel_l2(fa, fb) < (1e-11 if fp64 else 1e-6)
    assert cases.rel_l2(xa, xb) <= 100 * max(cases.rel_l2(xa, x2), floor)
    assert cases.rel_l2(fa, fb) <= 100 * max(cases.rel_l2(fa, f2), floor)
    assert cases.rel_l2(va, vb) <= 100 * max(cases.rel_l2(va, v2), floor)


def _oracle_timescales(u, types):
    o = OracleSim(*SHAPE, u.tau)
    o.set_body_force([0, 0, 2e-6])
    o.set_force_limit(u.f_limit)
    tids = [o.add_cell_type(model, v, t, p, inner, timescale=2) for model, v, t, p, inner in types]
    o.set_velocity_timescale(2)
    o.set_repulsion(2e-3, 1.2, 3)
    _load_cells(o, tids, types)
    return o


def test_resumed_run_with_timescales_matches_the_oracle():
    print("[checkpoint] resumed vs oracle", flush=True)
    u = cases.units()
    types = _types(u)
    o = _oracle_timescales(u, types)
    first, tids = _fresh(u, types, "timescales", hc.HC_FP64)
    _load_cells(first, tids, types)
    x0 = first.vertices()[0]
    first.step(8)
    st = _save(first)
    first.close()
    resumed, tids = _fresh(u, types, "timescales", hc.HC_FP64)
    _restore(resumed, tids, types, st)
    resumed.step(9)
    o.step(17)
    xg, xo = resumed.vertices()[0], o.vertices()[0]
    e = dict(x=cases.rel_l2(xg, xo), disp=cases.rel_l2(xg - x0, xo - x0),
             pop=cases.rel_l2(resumed.populations(), o.populations()))
    print("[checkpoint] resumed GPU vs uninterrupted oracle after 17 steps:", e)
    assert e["x"] < 1e-10 and e["disp"] < 1e-10 and e["pop"] < 1e-10


def test_iteration_and_repulsion_force_setters_validate():
    print("[checkpoint] setters", flush=True)
    s = hc.Simulation(0)
    s.lattice_init(8, 8, 8, 1.0)
    with pytest.raises(hc.HcError):
        s.set_iteration(-1)
    s.set_iteration(12)
    assert s.iteration == 12
    assert s.cells()[0].shape == (0,)


def _read_out(path):
    with open(path, "rb") as f:
        it, nc, nv = np.frombuffer(f.read(24), dtype=np.int64)
        ids = np.frombuffer(f.read(4 * nc), dtype=np.int32)
        x = np.frombuffer(f.read(24 * nv), dtype=np.float64).reshape(nv, 3)
        v = np.frombuffer(f.read(24 * nv), dtype=np.float64).reshape(nv, 3)
        fv = np.frombuffer(f.read(24 * nv), dtype=np.float64).reshape(nv, 3)
        pop = np.frombuffer(f.read(), dtype=np.float64)
    return dict(it=int(it), ids=ids, x=x, v=v, f=fv, pop=pop)


def _run(args, cwd, env):
    print("[checkpoint] host", " ".join(args), flush=True)
    r = subprocess.run(args, cwd=cwd, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_host_layer_checkpoint_files_and_restart(tmp_path):
    nx, ny, nz = 48, 40, 40
    (tmp_path / "RBC.xml").write_text(cases.material_xml(cases.RBC_MAT, "RBC"))
    (tmp_path / "PLT.xml").write_text(cases.material_xml(cases.PLT_MAT, "PLT", inner_edges=21))
    (tmp_path / "RBC.pos").write_text("2\n6 7 7 90 0 0\n17 12 12 0 30 0\n")
    (tmp_path / "PLT.pos").write_text("1\n12 4 10 10 20 30\n")
    env = dict(os.environ, HEMOCELL_TIMESCALES="2,2,3", HEMOCELL_REPULSION="2e-3,0.6")
    drv = os.path.join(BIN, "checkpoint_run")
    common = [str(nx), str(ny), str(nz), "2e-3"]
    specs = ["RBC:RBC", "PLT:PLT"]
    # two uninterrupted runs in their own folders (./tmp, then ./tmp_0 ...)
    whole, whole2 = tmp_path / "whole", tmp_path / "whole2"
    for d in (whole, whole2):
        d.mkdir()
        for name in ("RBC.xml", "PLT.xml", "RBC.pos", "PLT.pos"):
            (d / name).write_text((tmp_path / name).read_text())
        (d / "config.xml").write_text(cases.config_xml())
        r = _run([drv, "config.xml"] + common + ["14", "0", "out.bin"] + specs, cwd=d, env=env)
        (d / "stats.json").write_text(json.dumps(r))
    # a run that checkpoints at 4 and 8 and stops at 8
    split = tmp_path / "split"
    split.mkdir()
    for name in ("RBC.xml", "PLT.xml", "RBC.pos", "PLT.pos"):
        (split / name).write_text((tmp_path / name).read_text())
    (split / "config.xml").write_text(cases.config_xml())
    res = _run([drv, "config.xml"] + common + ["8", "4", "first.bin"] + specs, cwd=split, env=env)
    assert res["iteration"] == 8 and res["start"] == 0
    chk = split / "tmp" / "checkpoint"
    for name in ("checkpoint.xml", "lattice.dat", "particleField.dat"):
        assert (chk / name).is_file() and (chk / (name + ".old")).is_file(), name
    root = ET.parse(chk / "checkpoint.xml").getroot()
    assert root.tag == "Checkpoint"
    assert root.find("General/Iteration").text.strip() == "8"
    assert root.find("General/OutDirectory").text.strip() == "./tmp/"
    assert root.find("hemocell/domain/dx") is not None
    old = ET.parse(chk / "checkpoint.xml.old").getroot()
    assert old.find("General/Iteration").text.strip() == "4"
    # the restart reads its configuration from the checkpoint and writes on in ./tmp
    res = _run([drv, "tmp/checkpoint/checkpoint.xml"] + common + ["14", "0", "resumed.bin"] + specs, cwd=split,
               env=env)
    assert res["start"] == 8 and res["iteration"] == 14 and res["out"] == "./tmp/"
    assert res["fluid_nodes"] == nx * ny * nz  # velocity walls are not bounce-back: every node counts
    assert not (split / "tmp_0").exists()
    whole_res = json.loads((whole / "stats.json").read_text())
    assert abs(res["u_max"] - whole_res["u_max"]) <= 1e-10 * whole_res["u_max"]
    assert abs(res["u_mean"] - whole_res["u_mean"]) <= 1e-10 * whole_res["u_mean"]
    a, a2, b = _read_out(whole / "out.bin"), _read_out(whole2 / "out.bin"), _read_out(split / "resumed.bin")
    assert a["it"] == b["it"] == 14
    assert np.array_equal(a["ids"], b["ids"]) and len(a["ids"]) == 3
    assert cases.rel_l2(a["x"], b["x"]) < 1e-12
    assert cases.rel_l2(a["pop"], b["pop"]) < 1e-11
    assert cases.rel_l2(a["x"], b["x"]) <= 100 * max(cases.rel_l2(a["x"], a2["x"]), 1e-15)
    assert cases.rel_l2(a["v"], b["v"]) <= 100 * max(cases.rel_l2(a["v"], a2["v"]), 1e-15)


def test_host_layer_checkpoint_with_interior_viscosity(tmp_path):
    """Interior viscosity through hemo::HemoCell: the RBC material enables it
    (<enableInteriorViscosity>, <viscosityRatio> 5), setInteriorViscosityTimeScaleSeperation
    sets the two passes, saveCheckPoint writes internalViscosity.dat and a restart
    ends where the uninterrupted run ends, with the same inner nodes."""
    nx, ny, nz = 48, 40, 40
    rbc = dict(cases.RBC_MAT, viscosityRatio=5.0, enableInteriorViscosity=1)
    files = {"RBC.xml": cases.material_xml(rbc, "RBC"),
             "PLT.xml": cases.material_xml(cases.PLT_MAT, "PLT", inner_edges=21),
             "RBC.pos": "2\n6 7 7 90 0 0\n17 12 12 0 30 0\n", "PLT.pos": "1\n12 4 10 10 20 30\n",
             "config.xml": cases.config_xml()}
    env = dict(os.environ, HEMOCELL_TIMESCALES="1,1,1", HEMOCELL_INTERIOR_VISCOSITY="2,4")
    drv = os.path.join(BIN, "checkpoint_run")
    common = [str(nx), str(ny), str(nz), "2e-3"]
    specs = ["RBC:RBC", "PLT:PLT"]
    runs = {}
    for name, args in (("whole", ["config.xml"] + common + ["14", "0", "out.bin"]),
                       ("split", ["config.xml"] + common + ["8", "4", "first.bin"])):
        d = tmp_path / name
        d.mkdir()
        for k, v in files.items():
            (d / k).write_text(v)
        runs[name] = _run([drv] + args + specs, cwd=d, env=env)
    chk = tmp_path / "split" / "tmp" / "checkpoint"
    assert (chk / "internalViscosity.dat").is_file() and (chk / "internalViscosity.dat.old").is_file()
    assert runs["split"]["inner_nodes"] > 500
    res = _run([drv, "tmp/checkpoint/checkpoint.xml"] + common + ["14", "0", "resumed.bin"] + specs,
               cwd=tmp_path / "split", env=env)
    assert res["start"] == 8 and res["iteration"] == 14
    assert res["inner_nodes"] == runs["whole"]["inner_nodes"] > 500
    a, b = _read_out(tmp_path / "whole" / "out.bin"), _read_out(tmp_path / "split" / "resumed.bin")
    assert cases.rel_l2(a["x"], b["x"]) < 1e-12
    assert cases.rel_l2(a["pop"], b["pop"]) < 1e-11
    # the inner nodes relax with the larger tau: the run differs from one without it
    plain = tmp_path / "plain"
    plain.mkdir()
    for k, v in files.items():
        (plain / k).write_text(v.replace("<enableInteriorViscosity> 1 </enableInteriorViscosity>", ""))
    rp = _run([drv, "config.xml"] + common + ["14", "0", "out.bin"] + specs, cwd=plain, env=env)
    assert rp["inner_nodes"] == 0
    c = _read_out(plain / "out.bin")
    assert cases.rel_l2(a["pop"], c["pop"]) > 1e3 * max(cases.rel_l2(a["pop"], b["pop"]), 1e-15)
