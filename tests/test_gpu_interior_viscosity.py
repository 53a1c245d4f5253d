"""Interior viscosity on the GPU against the FP64 oracle.

The node classes are integer data and must match the oracle exactly: the
entire-grid ray-casting pass (core/hemoCellParticleField.cpp:776-808), the
membrane pass (:745-774) and the classes after whole steps with both passes on
their timescales.  Positions after whole FP64 steps agree within 1e-10
(relative L2); a lattice whose nodes are all inner steps as a plain lattice
with tau = tau_inner up to round-off (1e-13 FP64, 1e-6 FP32).
"""
import numpy as np
import pytest

import cases
import hemocell_amd as hc
from oracle import OracleSim

pytestmark = pytest.mark.gpu

SHAPE = (40, 30, 36)
TAU_RBC, TAU_PLT = 2.5, 3.0


def both(g, o, name, *args):
    getattr(g, name)(*args)
    getattr(o, name)(*args)


def _pair(shape=SHAPE, periodic=(1, 1, 1), mask=None, precision=hc.HC_FP64, seed=8, straddle=True):
    u = cases.units()
    g = hc.Simulation(0)
    g.lattice_init(*shape, u.tau, periodic=periodic, node_type=mask, precision=precision)
    o = OracleSim(*shape, u.tau, periodic=periodic, node_type=mask)
    rng = np.random.default_rng(seed)
    both(g, o, "set_body_force", [0, 0, 2e-6])
    both(g, o, "set_force_limit", u.f_limit)
    # one RBC straddles the periodic x and z faces when the axes are periodic
    rbc = [(10.3, 10.7, 10.2), (0.6 if straddle else 28.4, 18.2, 35.4 if straddle else 22.1)]
    tids = []
    for kind, centers in (("RBC", rbc), ("PLT", [(20.2, 25.1, 5.3), (33.4, 5.6, 30.2)]), ("WBC", [(25.3, 12.4, 24.8)])):
        model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
        xs = np.stack([cases.perturb(cases.place(v, c, cases.rotation(rng)), rng, 0.03) for c in centers])
        tid = g.add_cell_type(model, v, t, p, inner)
        assert o.add_cell_type(model, v, t, p, inner) == tid
        both(g, o, "add_cells", tid, xs)
        tids.append(tid)
    return g, o, tids


def test_entire_grid_pass_matches_oracle():
    print("[iv] entire grid", flush=True)
    g, o, (rbc, plt, wbc) = _pair()
    both(g, o, "set_interior_viscosity", rbc, TAU_RBC)
    both(g, o, "set_interior_viscosity", plt, TAU_PLT)
    g.stage("interior_viscosity", 0)
    o.stage_interior_viscosity(0)
    tg, to = g.interior_viscosity(), o.interior_viscosity()
    assert (to == TAU_RBC).sum() > 500 and (to == TAU_PLT).sum() > 0
    assert np.array_equal(tg, to)


def test_entire_grid_pass_with_walls_matches_oracle():
    print("[iv] entire grid, walls", flush=True)
    shape = (40, 30, 36)
    mask = cases.channel_mask(*shape, axis=1)
    g, o, (rbc, plt, wbc) = _pair(shape, periodic=(1, 0, 1), mask=mask, straddle=False)
    both(g, o, "set_interior_viscosity", rbc, TAU_RBC)
    both(g, o, "set_interior_viscosity", wbc, TAU_PLT)
    g.stage("interior_viscosity", 0)
    o.stage_interior_viscosity(0)
    assert np.array_equal(g.interior_viscosity(), o.interior_viscosity())


def test_entire_grid_pass_large_cell_matches_oracle():
    """A 2562-vertex WBC (its positions fill the LDS: nodes in x, y, z order)
    beside an RBC stretched flat in x and z (6660 lines along the ray)."""
    print("[iv] entire grid, large cells", flush=True)
    u = cases.units()
    shape = (72, 44, 72)
    g = hc.Simulation(0)
    g.lattice_init(*shape, u.tau)
    o = OracleSim(*shape, u.tau)
    rng = np.random.default_rng(21)
    tids = []
    # the RBC lies flat in y, stretched in x and z
    for kind, center, rot, stretch in (("WBC2562", (12.3, 10.6, 11.8), cases.rotation(rng), (1.0, 1.0, 1.0)),
                                       ("RBC", (40.2, 31.7, 40.3), np.eye(3), (3.5, 1.0, 3.5))):
        model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
        v0 = v - v.mean(axis=0)
        flat = v0[:, np.argsort(np.ptp(v0, axis=0))[::-1][[0, 2, 1]]]  # thin axis to y
        x = cases.place(flat if kind == "RBC" else v0, (0, 0, 0), rot) * np.array(stretch) + np.array(center)
        tid = g.add_cell_type(model, v, t, p, inner)
        assert o.add_cell_type(model, v, t, p, inner) == tid
        both(g, o, "add_cells", tid, x[None])
        both(g, o, "set_interior_viscosity", tid, TAU_RBC + tid)
        tids.append(tid)
    g.stage("interior_viscosity", 0)
    o.stage_interior_viscosity(0)
    to = o.interior_viscosity()
    assert (to == TAU_RBC).sum() > 2000 and (to == TAU_RBC + 1).sum() > 500
    assert np.array_equal(g.interior_viscosity(), to)


def test_membrane_pass_matches_oracle():
    print("[iv] membrane", flush=True)
    g, o, (rbc, plt, wbc) = _pair(seed=9)
    both(g, o, "set_interior_viscosity", rbc, TAU_RBC)
    both(g, o, "set_interior_viscosity", plt, TAU_PLT)
    # start from the ray-cast classes, then let every vertex re-classify its support
    g.stage("interior_viscosity", 0)
    o.stage_interior_viscosity(0)
    before = o.interior_viscosity()
    g.stage("interior_viscosity", 1)
    o.stage_interior_viscosity(1)
    after = o.interior_viscosity()
    assert np.count_nonzero(after != before) > 0
    assert np.array_equal(g.interior_viscosity(), after)


@pytest.mark.parametrize("timescales", [(1, 3), (2, 4)])
def test_steps_with_interior_viscosity_fp64(timescales):
    print("[iv] steps", timescales, flush=True)
    g, o, (rbc, plt, wbc) = _pair(seed=10)
    both(g, o, "set_interior_viscosity", rbc, TAU_RBC)
    both(g, o, "set_interior_viscosity", wbc, TAU_PLT)
    both(g, o, "set_interior_viscosity_timescales", *timescales)
    for _ in range(3):
        both(g, o, "step", 5)
        assert np.array_equal(g.interior_viscosity(), o.interior_viscosity())
    assert cases.rel_l2(g.vertices()[0], o.vertices()[0]) < 1e-10
    assert cases.rel_l2(g.populations(), o.populations()) < 1e-10
    assert (o.interior_viscosity() != 0).sum() > 500


def test_fp32_steps_keep_the_fp64_classes():
    print("[iv] fp32 steps", flush=True)
    runs = []
    for precision in (hc.HC_FP64, hc.HC_FP32_SHIFTED):
        g, o, (rbc, plt, wbc) = _pair(seed=11, precision=precision)
        g.set_interior_viscosity(rbc, TAU_RBC)
        g.set_interior_viscosity_timescales(2, 6)
        g.step(12)
        runs.append(g.interior_viscosity())
    a, b = runs
    assert (a != 0).sum() > 500
    # FP32 moves the vertices by round-off only: a handful of membrane nodes may flip
    assert np.count_nonzero(a != b) <= 0.01 * (a != 0).sum()


@pytest.mark.parametrize("precision,tol", [(hc.HC_FP64, 1e-13), (hc.HC_FP32_SHIFTED, 1e-6)])
def test_all_inner_lattice_steps_as_uniform_tau(precision, tol):
    """Every node inner with tau_inner against a plain lattice with tau = tau_inner:
    equal up to round-off (the two collide kernels may contract different
    multiply-adds), and far from the lattice with the background tau."""
    print("[iv] uniform", precision, flush=True)
    u = cases.units()
    shape = (12, 10, 11)
    rng = np.random.default_rng(3)
    f0 = cases.random_populations(int(np.prod(shape)), rng)
    model, v, t, p, inner = cases.mesh("PLT", u, hc.build_mesh)
    runs = []
    for tau, iv in ((u.tau, True), (1.37, False), (u.tau, False)):
        g = hc.Simulation(0)
        g.lattice_init(*shape, tau, precision=precision)
        tid = g.add_cell_type(model, v, t, p, inner)
        g.set_body_force([1e-5, -2e-5, 3e-5])
        g.set_populations(f0)
        if iv:
            g.set_interior_viscosity(tid, 1.37)
            g.set_interior_viscosity_field(np.full(shape, 1.37))
        g.step(4)
        runs.append(g.populations())
    inner_run, uniform, background = runs
    assert cases.rel_l2(inner_run, uniform) < tol
    assert cases.rel_l2(background, uniform) > 100 * max(cases.rel_l2(inner_run, uniform), tol)


def test_interior_viscosity_arguments_and_field_round_trip():
    print("[iv] arguments", flush=True)
    g, o, (rbc, plt, wbc) = _pair(seed=12)
    with pytest.raises(hc.HcError):
        g.set_interior_viscosity(rbc, 0.4)
    with pytest.raises(hc.HcError):
        g.set_interior_viscosity(7, 2.0)
    with pytest.raises(hc.HcError):
        g.set_interior_viscosity_timescales(0, 5)
    with pytest.raises(hc.HcError):
        g.stage("interior_viscosity", 0)  # not enabled
    g.set_interior_viscosity(rbc, TAU_RBC)
    g.stage("interior_viscosity", 0)
    t = g.interior_viscosity()
    g2, _, (r2, _, _) = _pair(seed=12)
    g2.set_interior_viscosity(r2, TAU_RBC)
    g2.set_interior_viscosity_field(t)
    assert np.array_equal(g2.interior_viscosity(), t)
    bad = t.copy()
    bad[0, 0, 0] = 1.7
    with pytest.raises(hc.HcError):
        g2.set_interior_viscosity_field(bad)
    # the timescales must be multiples of the particle-velocity timescale
    g2.set_velocity_timescale(2)
    g2.set_interior_viscosity_timescales(3, 4)
    with pytest.raises(hc.HcError):
        g2.step(1)
