"""The HIP path against the committed golden fixtures (tests/golden/*.npz, FP64
oracle outputs), without running the oracle.  Tolerances: relative L2, FP64."""
import os

import numpy as np
import pytest

import cases
import hemocell_amd as hc

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


def test_collide_one_golden():
    # a 1x1x1 periodic lattice streams every population back to its node, so one
    # step is exactly one collision
    d = load("collide_one.npz")
    for k in range(3):
        g = hc.Simulation(0)
        g.lattice_init(1, 1, 1, 1.0 / float(d["omega"]))
        g.set_populations(d["f"][k].reshape(19, 1))
        g.set_body_force(d["F"][k])
        g.step(1)
        out = g.populations()[:, 0]
        assert np.abs(out - d["out"][k]).max() <= 1e-15 + 1e-13 * np.abs(d["out"][k]).max()


def test_periodic8_golden():
    d = load("periodic8.npz")
    g = hc.Simulation(0)
    g.lattice_init(8, 8, 8, float(d["tau"]))
    g.set_populations(d["f0"])
    g.set_body_force(d["g"])
    g.step(int(d["steps"]))
    assert cases.rel_l2(g.populations(), d["f"]) < 1e-12


def test_rbc_forces_golden():
    d = load("rbc_forces.npz")
    u = cases.units()
    model, v, t, p, inner = cases.mesh("RBC", u, hc.build_mesh)
    g = hc.Simulation(0)
    g.lattice_init(40, 40, 40, u.tau)
    g.add_cells(g.add_cell_type(model, v, t, p, inner), d["x"][None])
    g.stage("mechanics", 1)
    assert cases.rel_l2(g.vertices()[2], d["f"]) < 1e-11


def test_ibm_weights_golden():
    # a unit force on one vertex spreads exactly its phi2 weights onto its support
    d = load("ibm_weights.npz")
    nx, ny, nz = [int(s) for s in d["shape"]]
    u = cases.units()
    model, v, t, p, inner = cases.mesh("PLT", u, hc.build_mesh)
    for k, pos in enumerate(d["pos"]):
        g = hc.Simulation(0)
        g.lattice_init(nx, ny, nz, 1.0, periodic=(1, 0, 1), node_type=d["mask"])
        tid = g.add_cell_type(model, v, t, p, inner)
        x = np.tile(pos, (len(v), 1))
        g.add_cells(tid, x[None])
        f = np.zeros((len(v), 3))
        f[0] = (1e-3, 0, 0)
        g.set_vertices(None, None, f)
        g.stage("spread")
        F = g.force_field()[0]
        n = np.count_nonzero(d["idx"][k] >= 0)
        want = np.zeros(nx * ny * nz)
        want[d["idx"][k, :n]] = 1e-3 * d["w"][k, :n]
        assert np.abs(F - want).max() < 1e-17, k


def test_rbc_shear32_golden():
    d = load("rbc_shear32.npz")
    u = cases.units()
    g = hc.Simulation(0)
    g.lattice_init(32, 32, 32, float(d["tau"]), periodic=(1, 1, 0), node_type=d["mask"])
    uw = float(d["wall_u"])
    g.set_wall_velocity(4, (uw, 0, 0))
    g.set_wall_velocity(5, (-uw, 0, 0))
    g.init_equilibrium(1.0, (0, 0, 0))
    g.set_force_limit(u.f_limit)
    model, v, t, p, inner = cases.mesh("RBC", u, hc.build_mesh)
    g.add_cells(g.add_cell_type(model, v, t, p, inner), d["x0"][None])
    g.step(int(d["steps"]))
    x, vel, _ = g.vertices()
    assert cases.rel_l2(x, d["x"]) < 1e-10
    assert cases.rel_l2(x - d["x0"], d["x"] - d["x0"]) < 1e-10
    assert cases.rel_l2(vel, d["v"]) < 1e-9
    _, uf = g.fluid()
    assert cases.rel_l2(uf.reshape(32, 32, 32, 3)[:, 16, :, :], d["u_slice_y16"]) < 1e-10
