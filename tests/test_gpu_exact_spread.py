"""The exact spread (hc_set_exact_spread, FP64 form): every contribution to the
force field is rounded to a multiple of a power-of-two quantum derived from the
force limit before the LDS and global atomic adds, so the sums are exact and
the result no longer depends on the order in which the adds land.  The
reference gets the same property from its serial particle-order spread
(core/hemoCellParticleField.cpp:841-863).

* Two runs of the 64x64x128 Ht 30% crop (242 RBC, 24 platelets, one 2562-vertex
  WBC; 100 steps) agree bit for bit: vertex positions, velocities and forces,
  populations; so does a run with an 8 KB instead of a 24 KB spread box (other
  slabs, other commit order of the box nodes).
* The exact-spread run stays within 1e-10 relative L2 of the oracle (positions,
  displacements, vertex and fluid velocities), as the plain-add run does.
* A run checkpointed at iteration 8 and resumed in a fresh context, whose first
  spread then runs in the stand-alone spread kernel, ends bit for bit where the
  uninterrupted run ends (material and particle-velocity timescales 2,
  repulsion timescale 3).
* The setter refuses the FP32 form and a context without a force limit.
"""
import os

import numpy as np
import pytest

import cases
import hemocell_amd as hc
from hemocell_amd import packing
from hemocell_amd import suspension as susp
from oracle import OracleSim

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CROP = os.path.join(HERE, "golden", "c3_crop_64x64x128.npz")
G = (0.0, 0.0, 1e-6)


def _crop(sims):
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    types = susp.cell_types(units)
    pk, _ = packing.load(CROP)
    cells = packing.instances(types, pk)
    for s in sims:
        s.init_equilibrium(1.0, (0.0, 0.0, 0.0))
        s.set_body_force(G)
        s.set_force_limit(units.f_limit)
        for name in ("RBC", "PLT", "WBC"):
            t = types[name]
            tid = s.add_cell_type(t.model, t.verts, t.tris, t.params, t.inner)
            s.add_cells(tid, cells[name])
    return units


def _gpu_crop(exact, monkeypatch, box=None):
    monkeypatch.setenv("HC_CELL_BLOCK", "256")
    monkeypatch.setenv("HC_MIXED", "0")
    if box:
        monkeypatch.setenv("HC_SPREAD_LDS", str(box))
    else:
        monkeypatch.delenv("HC_SPREAD_LDS", raising=False)
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    g = hc.Simulation(0)
    g.lattice_init(64, 64, 128, units.tau, periodic=(1, 1, 1), precision=hc.HC_FP64)
    _crop([g])
    if exact:
        g.set_exact_spread(True)
    return g


def _state(s):
    x, v, f = s.vertices()
    return dict(x=x, v=v, f=f, pop=s.populations())


def _run(exact, monkeypatch, box=None, steps=100):
    print("[exact spread] crop run exact=%s box=%s" % (exact, box), flush=True)
    g = _gpu_crop(exact, monkeypatch, box)
    x0 = g.vertices()[0]
    for k in range(steps // 20):
        g.step(20)
    out = _state(g)
    out["x0"] = x0
    g.close()
    return out


def test_exact_spread_runs_repeat_bit_for_bit(monkeypatch):
    a = _run(True, monkeypatch, box=24576)
    b = _run(True, monkeypatch, box=24576)
    c = _run(True, monkeypatch, box=8192)
    for k in ("x", "v", "f", "pop"):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k
    p = _run(False, monkeypatch, box=24576)
    q = _run(False, monkeypatch, box=8192)
    print("[exact spread] plain adds, 24 KB vs 8 KB box: positions %.2e, populations %.2e; bit-equal: %s" %
          (cases.rel_l2(p["x"], q["x"]), cases.rel_l2(p["pop"], q["pop"]),
           all(np.array_equal(p[k], q[k]) for k in ("x", "pop"))))
    # the rounding moves the result by far less than the oracle tolerance
    assert cases.rel_l2(a["x"] - a["x0"], p["x"] - p["x0"]) < 1e-10
    assert cases.rel_l2(a["pop"], p["pop"]) < 1e-12


def test_exact_spread_matches_the_oracle(monkeypatch):
    g = _gpu_crop(True, monkeypatch)
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    o = OracleSim(64, 64, 128, units.tau)
    _crop([o])
    x0 = g.vertices()[0]
    for k in range(5):
        print("[exact spread] vs oracle steps %d-%d" % (20 * k, 20 * k + 20), flush=True)
        g.step(20)
        o.step(20)
    xg, vg, _ = g.vertices()
    xo, vo, _ = o.vertices()
    e = dict(x=cases.rel_l2(xg, xo), disp=cases.rel_l2(xg - x0, xo - x0), v=cases.rel_l2(vg, vo),
             u=cases.rel_l2(g.fluid()[1], o.fluid()[1]))
    print("[exact spread] relative L2 vs oracle after 100 steps:", e)
    assert e["x"] < 1e-10 and e["disp"] < 1e-10 and e["v"] < 1e-10 and e["u"] < 1e-10


SHAPE = (40, 30, 36)


def _small(u, types):
    s = hc.Simulation(0)
    s.lattice_init(*SHAPE, u.tau, precision=hc.HC_FP64)
    s.set_body_force([0, 0, 2e-6])
    s.set_force_limit(u.f_limit)
    tids = [s.add_cell_type(model, v, t, p, inner) for model, v, t, p, inner in types]
    for tid in tids:
        s.set_material_timescale(tid, 2)
    s.set_velocity_timescale(2)
    s.set_repulsion(2e-3, 1.2, 3)
    s.set_exact_spread(True)
    return s, tids


def _cells(s, tids, types):
    rng = np.random.default_rng(8)
    for tid, centers, (model, v, t, p, inner) in zip(
            tids, ([(10, 10, 10), (28, 18, 22)], [(20, 25, 5), (33, 5, 30)], [(25, 12, 28)]), types):
        xs = np.stack([cases.perturb(cases.place(v, c, cases.rotation(rng)), rng, 0.03) for c in centers])
        s.add_cells(tid, xs)


def test_resumed_run_is_bit_identical_with_the_exact_spread():
    print("[exact spread] checkpoint resume", flush=True)
    u = cases.units()
    types = [cases.mesh(kind, u, hc.build_mesh) for kind in ("RBC", "PLT", "WBC")]
    whole, tids = _small(u, types)
    _cells(whole, tids, types)
    first, tids = _small(u, types)
    _cells(first, tids, types)
    whole.step(17)
    first.step(8)
    x, v, f = first.vertices()
    tid, cid = first.cells()
    pop, frep, it = first.populations(), first.repulsion_force(), first.iteration
    first.close()
    resumed, tids2 = _small(u, types)
    nvs = {t: len(types[k][1]) for k, t in enumerate(tids2)}
    off, c = 0, 0
    while c < len(tid):
        e = c
        while e < len(tid) and tid[e] == tid[c]:
            e += 1
        nv = nvs[int(tid[c])]
        resumed.add_cells(int(tid[c]), x[off:off + (e - c) * nv].reshape(e - c, nv, 3), cid[c:e])
        off += (e - c) * nv
        c = e
    resumed.set_vertices(x, v, f)
    resumed.set_repulsion_force(frep)
    resumed.set_populations(pop)
    resumed.set_iteration(it)
    resumed.step(9)
    a, b = _state(whole), _state(resumed)
    for k in ("x", "v", "f", "pop"):
        assert np.array_equal(a[k], b[k]), (k, cases.rel_l2(a[k], b[k]))


def test_exact_spread_setter_validates():
    print("[exact spread] setter", flush=True)
    u = cases.units()
    s = hc.Simulation(0)
    s.lattice_init(8, 8, 8, u.tau, precision=hc.HC_FP64)
    with pytest.raises(hc.HcError):
        s.set_exact_spread(True)  # no force limit yet
    s.set_force_limit(u.f_limit)
    s.set_exact_spread(True)
    s.set_exact_spread(False)
    s.close()
    s = hc.Simulation(0)
    s.lattice_init(8, 8, 8, u.tau, precision=hc.HC_FP32_SHIFTED)
    s.set_force_limit(u.f_limit)
    with pytest.raises(hc.HcError):
        s.set_exact_spread(True)
    s.close()
