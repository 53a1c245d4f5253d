"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the FP64 oracle) against the oracle and the
product's host-side mesh builder on the CPU: a change of the oracle or of the
mesh numbering shows up here before any GPU comparison."""
import os

import numpy as np
import pytest

import cases
import hemocell_amd as hc
import mesh_oracle
from oracle import OracleSim, collide_one

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


def test_collide_one_fixture():
    d = load("collide_one.npz")
    for k in range(3):
        assert np.array_equal(collide_one(d["f"][k], d["F"][k], float(d["omega"])), d["out"][k])


def test_periodic8_fixture():
    d = load("periodic8.npz")
    o = OracleSim(8, 8, 8, float(d["tau"]))
    o.set_populations(d["f0"])
    o.set_body_force(d["g"])
    o.step(int(d["steps"]))
    assert cases.rel_l2(o.populations(), d["f"]) < 1e-14


def test_rbc_mesh_and_tables_fixture():
    d = load("rbc_tables.npz")
    u = cases.units()
    v, t = hc.build_mesh(1, cases.RBC_MAT["radius"] / u.dx, cases.RBC_MAT["minNumTriangles"])
    assert np.array_equal(v, d["vertices"]) and np.array_equal(t, d["triangles"])
    model, v2, t2, p, inner = cases.mesh("RBC", u, mesh_oracle.build_mesh)
    o = OracleSim(40, 40, 40, u.tau)
    tab = o.type_tables(o.add_cell_type(model, v2, t2, p, inner))
    for k in ("edge", "edge_tris", "edge_outer", "ring", "nring"):
        assert np.array_equal(tab[k], d[k]), k
    for k in ("edge_len", "edge_angle", "tri_area", "patch_dist", "volume_eq", "area_mean_eq", "edge_mean_eq"):
        assert np.allclose(tab[k], d[k], rtol=1e-14, atol=0), k
    assert abs(float(d["volume_eq"]) * (u.dx * 1e6) ** 3 - 81.116) < 5e-3  # RBC_template.xml:16


def test_rbc_forces_fixture():
    d = load("rbc_forces.npz")
    u = cases.units()
    model, v, t, p, inner = cases.mesh("RBC", u, mesh_oracle.build_mesh)
    o = OracleSim(40, 40, 40, u.tau)
    o.add_cells(o.add_cell_type(model, v, t, p, inner), d["x"][None])
    o.stage("mechanics", 1)
    assert cases.rel_l2(o.vertices()[2], d["f"]) < 1e-14


def test_ibm_weights_fixture():
    d = load("ibm_weights.npz")
    nx, ny, nz = [int(s) for s in d["shape"]]
    o = OracleSim(nx, ny, nz, 1.0, periodic=(1, 0, 1), node_type=d["mask"])
    for k, p in enumerate(d["pos"]):
        i, w = o.ibm_kernel(p)
        assert np.array_equal(i, d["idx"][k, :len(i)]) and np.array_equal(w, d["w"][k, :len(w)])
        assert abs(w.sum() - 1.0) < 1e-14
        walls = d["mask"][i] != 0
        assert not walls.any()


def test_shear32_fixture():
    d = load("rbc_shear32.npz")
    u = cases.units()
    o = OracleSim(32, 32, 32, float(d["tau"]), periodic=(1, 1, 0), node_type=d["mask"])
    uw = float(d["wall_u"])
    o.set_wall_velocity(4, (uw, 0, 0))
    o.set_wall_velocity(5, (-uw, 0, 0))
    o.init_equilibrium(1.0, (0, 0, 0))
    o.set_force_limit(u.f_limit)
    model, v, t, p, inner = cases.mesh("RBC", u, mesh_oracle.build_mesh)
    o.add_cells(o.add_cell_type(model, v, t, p, inner), d["x0"][None])
    o.step(int(d["steps"]))
    x, vel, _ = o.vertices()
    assert cases.rel_l2(x, d["x"]) < 1e-13
    assert cases.rel_l2(vel, d["v"]) < 1e-11
    # the cell moved and turned under the shear
    assert np.abs(d["x"] - d["x0"]).max() > 1e-3
