"""GPU parity: libhcgpu.so (through the C ABI) against the FP64 CPU oracle.

Tolerances (relative L2 unless stated):
  FP64 lattice, stage or step          <= 1e-12 (summation-order round-off only)
  FP64 full step, 100 steps            <= 1e-10 (velocity field and vertex positions)
  FP32-shifted, fluid velocity         <= 1e-5 of the velocity scale
  FP32-shifted, vertex positions       <= 1e-4 of the displacement
Integer data (node masks, support node sets, cell/vertex counts) is bit-exact.
"""
import numpy as np
import pytest

import cases
import hemocell_amd as hc
from oracle import OracleSim

pytestmark = pytest.mark.gpu


def pair(nx, ny, nz, tau, periodic=(1, 1, 1), mask=None, precision=hc.HC_FP64):
    g = hc.Simulation(0)
    g.lattice_init(nx, ny, nz, tau, periodic=periodic, node_type=mask, precision=precision)
    o = OracleSim(nx, ny, nz, tau, periodic=periodic, node_type=mask)
    return g, o


def both(g, o, name, *args):
    getattr(g, name)(*args)
    getattr(o, name)(*args)


# ------------------------------------------------------------------ lattice --
@pytest.mark.parametrize("shape", [(8, 8, 8), (5, 7, 9)])
@pytest.mark.parametrize("nsteps", [1, 2, 5])
def test_collide_stream_periodic_fp64(shape, nsteps):
    rng = np.random.default_rng(1)
    g, o = pair(*shape, tau=0.8)
    f0 = cases.random_populations(np.prod(shape), rng)
    both(g, o, "set_populations", f0)
    both(g, o, "set_body_force", [1e-5, -2e-5, 3e-5])
    for _ in range(nsteps):
        both(g, o, "step", 1)
    assert cases.rel_l2(g.populations(), o.populations()) < 1e-12
    rg, ug = g.fluid()
    ro, uo = o.fluid()
    assert cases.rel_l2(rg, ro) < 1e-14
    assert cases.rel_l2(ug, uo) < 1e-12


@pytest.mark.parametrize("nsteps", [1, 4])
def test_collide_stream_bounceback_channel_fp64(nsteps):
    rng = np.random.default_rng(2)
    nx, ny, nz = 6, 9, 7
    mask = cases.channel_mask(nx, ny, nz, axis=1)
    g, o = pair(nx, ny, nz, tau=1.1, periodic=(1, 0, 1), mask=mask)
    f0 = cases.random_populations(nx * ny * nz, rng)
    both(g, o, "set_populations", f0)
    both(g, o, "set_body_force", [2e-5, 0, 1e-5])
    both(g, o, "step", nsteps)
    fg, fo = g.populations(), o.populations()
    fluid = mask == 0
    assert cases.rel_l2(fg[:, fluid], fo[:, fluid]) < 1e-12
    assert cases.rel_l2(fg, fo) < 1e-12


@pytest.mark.parametrize("nsteps", [1, 2, 7])
def test_collide_stream_velocity_walls_fp64(nsteps):
    rng = np.random.default_rng(3)
    nx, ny, nz = 6, 5, 8
    mask = cases.couette_mask(nx, ny, nz)
    g, o = pair(nx, ny, nz, tau=1.0, periodic=(1, 1, 0), mask=mask)
    both(g, o, "set_wall_velocity", 4, [1e-3, 0, 0])
    both(g, o, "set_wall_velocity", 5, [-1e-3, 2e-4, 0])
    f0 = cases.random_populations(nx * ny * nz, rng)
    both(g, o, "set_populations", f0)
    both(g, o, "step", nsteps)
    assert cases.rel_l2(g.populations(), o.populations()) < 1e-12


def test_collide_stream_tube_fp32():
    rng = np.random.default_rng(4)
    nx, ny, nz = 12, 10, 10
    mask = cases.tube_mask(nx, ny, nz, axis=0)
    g, o = pair(nx, ny, nz, tau=0.9, periodic=(1, 0, 0), mask=mask, precision=hc.HC_FP32_SHIFTED)
    f0 = cases.random_populations(nx * ny * nz, rng, amp=1e-4)
    both(g, o, "set_populations", f0)
    both(g, o, "set_body_force", [1e-6, 0, 0])
    both(g, o, "step", 6)
    _, ug = g.fluid()
    _, uo = o.fluid()
    assert cases.rel_l2(ug, uo) < 1e-5


def test_init_equilibrium_matches():
    nx, ny, nz = 4, 5, 6
    mask = cases.channel_mask(nx, ny, nz, axis=2)
    g, o = pair(nx, ny, nz, tau=1.0, periodic=(1, 1, 0), mask=mask)
    both(g, o, "init_equilibrium", 1.01, [1e-3, -2e-3, 5e-4])
    assert cases.rel_l2(g.populations(), o.populations()) < 1e-15


# ---------------------------------------------------------------- IBM stages --
def _one_cell(g, o, u, kind, center, rng, amp=0.05):
    model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
    tg = g.add_cell_type(model, v, t, p, inner)
    to = o.add_cell_type(model, v, t, p, inner)
    x = cases.perturb(cases.place(v, center, cases.rotation(rng)), rng, amp)
    g.add_cells(tg, x[None])
    o.add_cells(to, x[None])
    return x


def test_spread_and_interpolate_stage_fp64():
    rng = np.random.default_rng(5)
    u = cases.units()
    nx, ny, nz = 24, 26, 28
    mask = cases.channel_mask(nx, ny, nz, axis=2)
    g, o = pair(nx, ny, nz, tau=u.tau, periodic=(1, 1, 0), mask=mask)
    _one_cell(g, o, u, "RBC", (12.3, 13.1, 13.7), rng)
    both(g, o, "set_populations", cases.random_populations(nx * ny * nz, rng, 1e-4))
    nv = g.num_vertices
    forces = 1e-3 * rng.normal(size=(nv, 3))
    forces[::7] *= 50.0  # some exceed f_limit and get capped in place
    both(g, o, "set_vertices", None, None, forces)
    both(g, o, "set_force_limit", u.f_limit)
    g.stage("spread")
    o.stage("spread")
    assert cases.rel_l2(g.force_field(), o.force_field()) < 1e-13
    _, _, fg = g.vertices()
    _, _, fo = o.vertices()
    assert np.array_equal(fg, fo)  # the capped forces are written back identically
    g.stage("collide_stream")
    o.stage("collide_stream")
    g.stage("interpolate")
    o.stage("interpolate")
    _, vg, _ = g.vertices()
    _, vo, _ = o.vertices()
    assert cases.rel_l2(vg, vo) < 1e-12


# ----------------------------------------------------------- mechanics stage --
@pytest.mark.parametrize("kind", ["RBC", "PLT", "WBC", "WBC2562"])
def test_mechanics_stage_fp64(kind):
    rng = np.random.default_rng(6)
    u = cases.units()
    g, o = pair(40, 40, 40, tau=u.tau)
    _one_cell(g, o, u, kind, (20.2, 19.7, 20.4), rng, amp=0.08)
    nv = g.num_vertices
    vel = 1e-4 * rng.normal(size=(nv, 3))
    both(g, o, "set_vertices", None, vel, None)
    g.stage("mechanics", 1)
    o.stage("mechanics", 1)
    _, _, fg = g.vertices()
    _, _, fo = o.vertices()
    assert np.abs(fo).max() > 0
    assert cases.rel_l2(fg, fo) < 1e-11


@pytest.mark.parametrize("kind", ["RBC", "PLT", "WBC"])
def test_force_terms_match_single_stiffness_oracle(kind):
    """hc_get_force_terms (the per-term vectors writeOutput separates): each term
    equals the oracle's whole mechanics with every other stiffness set to zero
    (area: k_area; volume: k_volume; bending: k_bend; link and, for platelets, the
    inner links: k_link; viscosity: eta_m; WBC inner links: k_inner_rigid and
    k_cytoskeleton), and the six terms add up to the device's full mechanics."""
    rng = np.random.default_rng(31)
    u = cases.units()
    model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
    x = cases.perturb(cases.place(v, (12.3, 11.7, 12.9), cases.rotation(rng)), rng, 0.05)
    vel = 1e-4 * rng.uniform(-1, 1, x.shape)
    g = hc.Simulation(0)
    g.lattice_init(26, 24, 26, u.tau)
    tid = g.add_cell_type(model, v, t, p, inner)
    g.add_cells(tid, x[None])
    g.set_vertices(x, vel, None)
    terms = g.force_terms()
    g.stage("mechanics", 1)
    full = g.vertices()[2]
    total = sum(terms.values())
    assert cases.rel_l2(total, full) < 1e-13
    fields = [f[0] for f in type(p)._fields_]

    def oracle_with(keep):
        q = type(p).from_buffer_copy(p)
        for name in ("k_link", "k_bend", "k_volume", "k_area", "eta_m", "k_inner_rigid", "k_cytoskeleton"):
            if name in fields and name not in keep:
                setattr(q, name, 0.0)
        o = OracleSim(26, 24, 26, u.tau)
        ot = o.add_cell_type(model, v, t, q, inner)
        o.add_cells(ot, x[None])
        o.set_vertices(x, vel, None)
        o.stage("mechanics", 1)
        return o.vertices()[2]

    checks = [("area", ("k_area",)), ("volume", ("k_volume",)), ("bending", ("k_bend",))]
    if kind == "PLT":
        checks.append(("link+inner_link", ("k_link",)))
    else:
        checks.append(("link", ("k_link",)))
    if p.eta_m != 0.0:  # the test materials give only the WBC membrane viscosity
        checks.append(("visc", ("eta_m",)))
    if kind == "WBC":
        checks.append(("inner_link", ("k_inner_rigid", "k_cytoskeleton")))
    for name, keep in checks:
        dev = sum(terms[n] for n in name.split("+"))
        ref = oracle_with(keep)
        assert np.abs(ref).max() > 0, name
        assert cases.rel_l2(dev, ref) < 1e-11, name


def test_mechanics_rest_shape_is_force_free():
    u = cases.units()
    g = hc.Simulation(0)
    g.lattice_init(40, 40, 40, u.tau)
    model, v, t, p, inner = cases.mesh("RBC", u, hc.build_mesh)
    tid = g.add_cell_type(model, v, t, p, inner)
    g.add_cells(tid, (v + 20.0)[None])
    g.stage("mechanics", 1)
    _, _, f = g.vertices()
    assert np.abs(f).max() < 1e-12 * max(p.k_link, p.k_bend, p.k_area, p.k_volume)
    vol, area = g.cell_volume_area()
    assert abs(vol[0] * (u.dx * 1e6) ** 3 - 81.116) < 5e-4


# ---------------------------------------------------------------- full step --
def _shear_case(precision, n=32, steps=100, seed=7):
    rng = np.random.default_rng(seed)
    u = cases.units()
    mask = cases.couette_mask(n, n, n)
    g, o = pair(n, n, n, tau=u.tau, periodic=(1, 1, 0), mask=mask, precision=precision)
    shear = 1000.0 * u.dt  # 1000 /s in lattice units
    vhalf = (n - 1) * shear * 0.5
    both(g, o, "set_wall_velocity", 5, [-vhalf, 0, 0])
    both(g, o, "set_wall_velocity", 4, [vhalf, 0, 0])
    both(g, o, "init_equilibrium", 1.0, [0, 0, 0])
    both(g, o, "set_force_limit", u.f_limit)
    _one_cell(g, o, u, "RBC", (n / 2 + 0.3, n / 2 - 0.2, n / 2 + 0.1), rng, amp=0.0)
    x0 = g.vertices()[0]
    for _ in range(steps // 10):
        both(g, o, "step", 10)
    return g, o, x0


def test_full_step_shear_fp64_100_steps():
    g, o, x0 = _shear_case(hc.HC_FP64)
    xg, vg, fg = g.vertices()
    xo, vo, fo = o.vertices()
    print("[full step shear fp64] relative L2 vs oracle: x %.2e, displacement %.2e, v %.2e" % (
        cases.rel_l2(xg, xo), cases.rel_l2(xg - x0, xo - x0), cases.rel_l2(vg, vo)))
    assert cases.rel_l2(xg, xo) < 1e-10
    assert cases.rel_l2(xg - x0, xo - x0) < 1e-10
    assert cases.rel_l2(vg, vo) < 1e-10
    _, ug = g.fluid()
    _, uo = o.fluid()
    assert cases.rel_l2(ug, uo) < 1e-10
    assert g.wall_hits() == o.wall_hits() == 0
    g.check_finite()


def test_full_step_shear_fp32_shifted():
    g, o, x0 = _shear_case(hc.HC_FP32_SHIFTED)
    xg, _, _ = g.vertices()
    xo, _, _ = o.vertices()
    _, ug = g.fluid()
    _, uo = o.fluid()
    assert cases.rel_l2(ug, uo) < 1e-5
    assert cases.rel_l2(xg - x0, xo - x0) < 1e-4


def _mixed(sims, rng_seed=8, nudge=0.0):
    rng = np.random.default_rng(rng_seed)
    u = cases.units()
    for s in sims:
        s.set_body_force([0, 0, 2e-6])
        s.set_force_limit(u.f_limit)
    for kind, centers in (("RBC", [(10, 10, 10), (28, 18, 22)]), ("PLT", [(20, 25, 5), (33, 5, 30)]),
                          ("WBC", [(25, 12, 28)])):
        model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
        xs = np.stack([cases.perturb(cases.place(v, c, cases.rotation(rng)), rng, 0.03) for c in centers])
        if nudge:
            xs = xs * (1 + nudge * np.random.default_rng(1).uniform(-1, 1, xs.shape))
        for s in sims:
            tid = s.add_cell_type(model, v, t, p, inner)
            s.add_cells(tid, xs)


@pytest.mark.parametrize("persistent", [0, 1])
def test_full_step_mixed_suspension_body_force_fp64(persistent, monkeypatch):
    """RBC, platelet and WBC cells stepped by per-type kernels side by side, or by
    the one persistent launch of all types (HC_MIXED=1), against the oracle."""
    monkeypatch.setenv("HC_MIXED", str(persistent))
    u = cases.units()
    nx, ny, nz = 40, 30, 36
    g, o = pair(nx, ny, nz, tau=u.tau)
    _mixed([g, o])
    # the oracle's own response to a 1e-15 relative nudge of the positions: the
    # conditioning of each observable in this strongly coupled case
    o2 = OracleSim(nx, ny, nz, u.tau)
    _mixed([o2], nudge=1e-15)
    x0 = g.vertices()[0]
    for _ in range(5):
        both(g, o, "step", 10)
        o2.step(10)
    xg, vg, _ = g.vertices()
    xo, vo, _ = o.vertices()
    _, v2, _ = o2.vertices()
    print("[mixed fp64] relative L2 vs oracle: x %.2e, displacement %.2e, v %.2e (oracle nudged by 1e-15: v %.2e)" % (
        cases.rel_l2(xg, xo), cases.rel_l2(xg - x0, xo - x0), cases.rel_l2(vg, vo), cases.rel_l2(v2, vo)))
    assert cases.rel_l2(xg, xo) < 1e-10
    assert cases.rel_l2(xg - x0, xo - x0) < 1e-10
    # measured: v 4e-13, the same as the oracle's own response to the 1e-15 nudge
    assert cases.rel_l2(vg, vo) < 1e-10
    _, ug = g.fluid()
    _, uo = o.fluid()
    assert cases.rel_l2(ug, uo) < 1e-10


@pytest.mark.parametrize("mat_ts,vel_ts", [((2, 3, 2), 1), ((1, 1, 1), 2), ((3, 2, 1), 2)])
def test_timescales_match_oracle_fp64(mat_ts, vel_ts):
    """Per-type material timescales (the constitutive model runs only on
    iterations divisible by its type's timescale, hemoCellField.h:63) and the
    particle-velocity timescale (interpolation only on iterations divisible by
    it, hemoCellFields.h:183-202) against the oracle: 30 FP64 steps of a mixed
    suspension, positions, displacements and fluid velocity within 1e-10."""
    u = cases.units()
    nx, ny, nz = 40, 30, 36
    g, o = pair(nx, ny, nz, tau=u.tau)
    rng = np.random.default_rng(12)
    for s in (g, o):
        s.set_body_force([0, 0, 2e-6])
        s.set_force_limit(u.f_limit)
    for (kind, centers), ts in zip((("RBC", [(10, 10, 10), (28, 18, 22)]), ("PLT", [(20, 25, 5), (33, 5, 30)]),
                                    ("WBC", [(25, 12, 28)])), mat_ts):
        model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
        xs = np.stack([cases.perturb(cases.place(v, c, cases.rotation(rng)), rng, 0.03) for c in centers])
        for s in (g, o):
            tid = s.add_cell_type(model, v, t, p, inner, timescale=ts)
            s.add_cells(tid, xs)
    both(g, o, "set_velocity_timescale", vel_ts)
    x0 = g.vertices()[0]
    for _ in range(3):
        both(g, o, "step", 10)
    xg, xo = g.vertices()[0], o.vertices()[0]
    ug, uo = g.fluid()[1], o.fluid()[1]
    e = (cases.rel_l2(xg, xo), cases.rel_l2(xg - x0, xo - x0), cases.rel_l2(ug, uo))
    print("[timescales %s, %d] relative L2 vs oracle: x %.2e, displacement %.2e, u %.2e" % ((mat_ts, vel_ts) + e))
    assert e[0] < 1e-10 and e[1] < 1e-10 and e[2] < 1e-10


def test_stretch_extra_force_fp64():
    rng = np.random.default_rng(9)
    u = cases.units()
    g, o = pair(36, 28, 28, tau=u.tau)
    both(g, o, "set_force_limit", u.f_limit)
    x = _one_cell(g, o, u, "RBC", (18, 14, 14), rng, amp=0.0)
    order = np.argsort(x[:, 0])
    lo, hi = order[:7], order[-7:]
    F = 100e-12 / u.df
    ids = np.concatenate([lo, hi]).astype(np.int64)
    forces = np.zeros((14, 3))
    forces[:7, 0] = -F / 7
    forces[7:, 0] = F / 7
    g.set_extra_vertex_force(ids, forces)
    for _ in range(30):
        # the oracle applies the stretch force the way the reference driver does: before every step
        xo, vo, fo = o.vertices()
        fo[ids] += forces
        o.set_vertices(None, None, fo)
        o.step(1)
    g.step(30)
    xg, _, _ = g.vertices()
    xo, _, _ = o.vertices()
    assert cases.rel_l2(xg, xo) < 1e-10


def test_fused_and_unfused_steps_agree(monkeypatch):
    """The fused vertex kernel (interpolation + advance + mechanics + next spread)
    against the separate kernels, on a mixed periodic case with a stretch force."""
    u = cases.units()
    nx, ny, nz = 40, 30, 36
    sims = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("HC_FUSE", fuse)
        g = hc.Simulation(0)
        g.lattice_init(nx, ny, nz, u.tau)
        _mixed([g])
        g.set_extra_vertex_force(np.array([3, 500], dtype=np.int64), np.array([[1e-4, 0, 0], [-1e-4, 0, 0]]))
        sims.append(g)
    for _ in range(4):
        for g in sims:
            g.step(5)
    (xa, va, fa), (xb, vb, fb) = sims[0].vertices(), sims[1].vertices()
    assert cases.rel_l2(xa, xb) < 1e-12
    assert cases.rel_l2(va, vb) < 1e-8
    _, ua = sims[0].fluid()
    _, ub = sims[1].fluid()
    assert cases.rel_l2(ua, ub) < 1e-10
    assert np.array_equal(sims[0].force_field(), np.zeros_like(ub.T))


# ------------------------------------------------------ binning and repulsion --
def _crowded_pair(shape, periodic, seed, n_plt, n_rbc, precision=hc.HC_FP64):
    rng = np.random.default_rng(seed)
    u = cases.units()
    g, o = pair(*shape, tau=u.tau, periodic=periodic, precision=precision)
    for kind, n in (("PLT", n_plt), ("RBC", n_rbc)):
        if n:
            model, v, t, p, inner, xs = cases.crowd(kind, n, shape, rng, u, hc.build_mesh)
            for s in (g, o):
                tid = s.add_cell_type(model, v, t, p, inner)
                s.add_cells(tid, xs)
    return g, o, u


@pytest.mark.parametrize("periodic", [(1, 1, 1), (1, 0, 1)])
def test_vertex_bins_bit_exact(periodic):
    g, o, _ = _crowded_pair((14, 12, 16), periodic, seed=11, n_plt=40, n_rbc=2)
    bg, bo = g.vertex_bins(), o.vertex_bins()
    assert np.array_equal(bg, bo)
    assert (bo >= 0).any() and (periodic == (1, 1, 1)) == bool((bo >= 0).all())


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 1)])
def test_repulsion_stage_fp64(periodic):
    g, o, _ = _crowded_pair((12, 10, 13), periodic, seed=12, n_plt=40, n_rbc=1)
    both(g, o, "set_repulsion", 2.5e-3, 0.8, 1)
    g.stage("repulsion")
    o.stage("repulsion")
    fg, fo = g.repulsion_force(), o.repulsion_force()
    assert np.abs(fo).max() > 0
    assert cases.rel_l2(fg, fo) < 1e-14
    # a second call after moving the cells reuses the (cleared) bins
    xg = g.vertices()[0] + 0.37
    both(g, o, "set_vertices", xg, None, None)
    g.stage("repulsion")
    o.stage("repulsion")
    assert cases.rel_l2(g.repulsion_force(), o.repulsion_force()) < 1e-14


def test_full_step_with_repulsion_fp64():
    g, o, u = _crowded_pair((18, 16, 20), (1, 1, 1), seed=13, n_plt=30, n_rbc=1)
    both(g, o, "set_body_force", [0, 0, 1e-6])
    both(g, o, "set_force_limit", u.f_limit)
    both(g, o, "set_repulsion", 2e-5, 0.9, 2)
    for _ in range(3):
        both(g, o, "step", 5)
    assert np.abs(o.repulsion_force()).max() > 0
    xg, vg, _ = g.vertices()
    xo, vo, _ = o.vertices()
    assert cases.rel_l2(xg, xo) < 1e-10
    assert cases.rel_l2(g.repulsion_force(), o.repulsion_force()) < 1e-9


def _near_walls_pair(shape, seed):
    """Platelets and one RBC in a channel (bounce-back planes y = 0 and y = ny-1),
    several platelets within a lattice spacing of the walls."""
    rng = np.random.default_rng(seed)
    u = cases.units()
    mask = cases.channel_mask(*shape, axis=1)
    g, o = pair(*shape, tau=u.tau, periodic=(1, 0, 1), mask=mask)
    for kind, centers in (("PLT", [(5.0, 3.1, 6.0), (13.0, shape[1] - 4.2, 9.0), (8.0, 3.3, 15.0)]),
                          ("RBC", [(10.0, shape[1] / 2.0, 10.0)])):
        model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
        xs = np.stack([cases.place(v, c, cases.rotation(rng)) for c in centers])
        for s in (g, o):
            tid = s.add_cell_type(model, v, t, p, inner)
            s.add_cells(tid, xs)
    return g, o, u


def test_wall_repulsion_stage_and_steps_fp64():
    """Wall (boundary particle) repulsion, hemoCellParticleField.cpp:865-918: the
    per-vertex force against the oracle after one application, then whole steps
    with it on every other step (positions within 1e-10, the force within 1e-12)."""
    g, o, u = _near_walls_pair((18, 24, 20), seed=17)
    both(g, o, "set_force_limit", u.f_limit)
    both(g, o, "set_boundary_repulsion", 2e-4, 1.3, 1)
    both(g, o, "stage", "boundary_repulsion")
    fo = o.boundary_repulsion_force()
    assert np.abs(fo).max() > 0 and np.count_nonzero(np.abs(fo).sum(axis=1)) >= 2
    assert cases.rel_l2(g.boundary_repulsion_force(), fo) < 1e-14
    both(g, o, "set_boundary_repulsion", 2e-4, 1.3, 2)
    both(g, o, "set_body_force", [0, 0, 1e-6])
    for _ in range(3):
        both(g, o, "step", 5)
    assert cases.rel_l2(g.vertices()[0], o.vertices()[0]) < 1e-10
    assert cases.rel_l2(g.boundary_repulsion_force(), o.boundary_repulsion_force()) < 1e-12
    assert g.wall_hits() == o.wall_hits() == 0


# ---------------------------------------------------------- cell statistics --
def test_cell_stats_match_reference_formulas():
    u = cases.units()
    g = hc.Simulation(0)
    g.lattice_init(40, 30, 36, u.tau)
    _mixed([g])
    g.step(20)
    x, v, _ = g.vertices()
    st = g.cell_stats()
    base, c = 0, 0
    # cells in the order _mixed adds them
    for kind, count in (("RBC", 2), ("PLT", 2), ("WBC", 1)):
        _, mv, mt, _, _ = cases.mesh(kind, u, hc.build_mesh)
        for _ in range(count):
            sl = slice(base, base + len(mv))
            ref = cases.cell_stats_reference(x[sl], v[sl], mt)
            assert abs(st["volume"][c] - ref["volume"]) <= 1e-12 * abs(ref["volume"])
            assert abs(st["area"][c] - ref["area"]) <= 1e-12 * ref["area"]
            assert np.allclose(st["centroid"][c], ref["centroid"], rtol=1e-13, atol=0)
            assert np.allclose(st["velocity"][c], ref["velocity"], rtol=1e-10, atol=1e-20)
            assert np.array_equal(st["bbox"][c], ref["bbox"])
            assert abs(st["stretch"][c] - ref["stretch"]) <= 1e-15 * ref["stretch"]
            base += len(mv)
            c += 1
    assert c == len(st["volume"])


@pytest.mark.parametrize("precision", [hc.HC_FP64, hc.HC_FP32_SHIFTED])
def test_fluid_statistics(precision):
    """FluidInfo statistics (helper/fluidInfo.cpp:27-141) over the non-bounce-back
    nodes: |u| against numpy on the device's own velocity field, |F| of the
    vertex forces spread into a reset field against the FP64 oracle's spread of
    the same positions and forces."""
    u = cases.units()
    nx, ny, nz = 40, 30, 36
    mask = cases.channel_mask(nx, ny, nz, axis=1)
    g = hc.Simulation(0)
    g.lattice_init(nx, ny, nz, u.tau, periodic=(1, 0, 1), node_type=mask, precision=precision)
    g.init_equilibrium(1.0, [0, 0, 0])
    _mixed([g])
    g.step(7)
    fluid = mask == 0
    mag = np.linalg.norm(g.fluid()[1][fluid], axis=1)
    st = g.fluid_stats("velocity")
    assert st["ncells"] == int(fluid.sum())
    assert st["min"] == pytest.approx(mag.min(), rel=1e-14, abs=1e-300)
    assert st["max"] == pytest.approx(mag.max(), rel=1e-14)
    assert st["avg"] == pytest.approx(mag.mean(), rel=1e-12)
    x, _, f = g.vertices()
    sf = g.fluid_stats("force")
    o = OracleSim(nx, ny, nz, u.tau, periodic=(1, 0, 1), node_type=mask)
    o.set_force_limit(u.f_limit)
    _mixed([o])
    o.set_vertices(x, None, f)
    o.set_force_field(np.zeros((3, nx * ny * nz)))
    o.stage("spread")
    fm = np.linalg.norm(o.force_field()[:, fluid], axis=0)
    tol = 1e-12 if precision == hc.HC_FP64 else 1e-6
    assert sf["ncells"] == int(fluid.sum())
    assert sf["max"] == pytest.approx(fm.max(), rel=tol)
    assert sf["avg"] == pytest.approx(fm.mean(), rel=tol)


# ------------------------------------------------------------- step graphs --
def test_graph_replays_match_eager_steps(monkeypatch):
    """Steady fused steps replayed from a captured two-step graph against the eager
    launches, replay after replay, with host reads and a changed body force (which
    forces a new capture) in between.  The spread's float atomics commit in no fixed
    order, so two eager runs already differ at round-off: the graphed run must stay
    within the fused-step tolerances of the eager one, and as close to it as a second
    eager run is, up to a factor.  Two eager runs often commit in the same order and
    agree to the last bit, while the graph's parallel branches commit in another; one
    force sum differing in its last bit grows over the coupled steps to ~1e-10 in the
    vertex velocities and the populations (8e-11 and 1.1e-10 observed), hence 1e-9
    for both; a graph replaying a stale argument would be off by far more."""
    u = cases.units()
    sims = []
    for graph in ("0", "0", "1"):
        monkeypatch.setenv("HC_GRAPH", graph)
        g = hc.Simulation(0)
        g.lattice_init(40, 30, 36, u.tau)
        _mixed([g])
        g.set_extra_vertex_force(np.array([3, 500], dtype=np.int64), np.array([[1e-4, 0, 0], [-1e-4, 0, 0]]))
        sims.append(g)
    eager, eager2, graphed = sims
    for rnd, n in enumerate((1, 6, 6, 7, 4, 9)):
        if rnd == 3:
            for g in sims:
                g.set_body_force([0, 3e-6, 1e-6])
        for g in sims:
            g.step(n)
        (xa, va, _), (x2, v2, _), (xb, vb, _) = eager.vertices(), eager2.vertices(), graphed.vertices()
        assert cases.rel_l2(xa, xb) < 1e-12, rnd
        assert cases.rel_l2(va, vb) < 1e-8, rnd
        assert cases.rel_l2(va, vb) <= max(100 * cases.rel_l2(va, v2), 1e-9), rnd
        fa, fb = eager.populations(), graphed.populations()
        assert cases.rel_l2(fa, fb) < 1e-9, rnd
    assert eager.iteration == graphed.iteration == 33


@pytest.mark.parametrize("precision", [hc.HC_FP64, hc.HC_FP32_SHIFTED])
def test_wbc_beside_rbc_with_positions_in_global_memory(monkeypatch, precision):
    """Beside the RBC kernel, the WBC kernel keeps its positions in global memory
    (256 threads per cell) instead of LDS.  Only where the positions live differs,
    so a step of each way agrees to the last bit; the bound absorbs only a different
    commit order of the first spread's float atomics at nodes shared by three or
    more cells (relative L2 1e-13 in FP64, 1e-6 in FP32)."""
    u = cases.units()
    out = []
    for gpos in ("1", "0"):
        monkeypatch.setenv("HC_SIDE_GPOS", gpos)
        g = hc.Simulation(0)
        g.lattice_init(40, 30, 36, u.tau, precision=precision)
        _mixed([g])
        g.step(3)
        out.append(g.vertices())
    tol = 1e-13 if precision == hc.HC_FP64 else 1e-6
    for a, b in zip(out[0], out[1]):
        assert cases.rel_l2(a, b) <= tol


def test_autotune_steps_like_plain_steps():
    """hc_autotune times real steps under each launch shape and keeps the fastest:
    the run it leaves behind is the same simulation as plain steps (the shapes
    change only the commit order of float atomics, hence the round-off bounds of
    the graph test), and it reports a choice for every parameter it tried."""
    u = cases.units()
    sims = []
    for _ in range(2):
        g = hc.Simulation(0)
        g.lattice_init(40, 30, 36, u.tau)
        _mixed([g])
        sims.append(g)
    tuned, plain = sims
    rep = tuned.autotune(1)
    assert set(rep) >= {"spread_box_bytes", "rbc_threads", "wbc_threads"}
    assert all("chosen" in v for v in rep.values())
    n = tuned.iteration
    plain.step(n)
    tuned.step(5)
    plain.step(5)
    assert tuned.iteration == plain.iteration == n + 5
    (xa, va, _), (xb, vb, _) = tuned.vertices(), plain.vertices()
    assert cases.rel_l2(xa, xb) < 1e-12
    assert cases.rel_l2(va, vb) < 1e-8


@pytest.mark.parametrize("block", ["256", "512"])
def test_threads_per_cell_agree_with_the_oracle(monkeypatch, block):
    """The fused cell kernel at an explicit number of threads per RBC (the
    suspension's setting) against the oracle; small cases otherwise run with the
    widest block."""
    monkeypatch.setenv("HC_CELL_BLOCK", block)
    u = cases.units()
    g, o = pair(40, 30, 36, tau=u.tau)
    _mixed([g, o])
    for _ in range(3):
        both(g, o, "step", 10)
    xg, vg, _ = g.vertices()
    xo, vo, _ = o.vertices()
    assert cases.rel_l2(xg, xo) < 1e-10
    _, ug = g.fluid()
    _, uo = o.fluid()
    assert cases.rel_l2(ug, uo) < 1e-10


def test_velocity_walls_in_their_own_kernel(monkeypatch):
    """Velocity-wall nodes reconstructed by the separate kernel (HC_FUSE_VEL=0)
    against the oracle, as the in-kernel path is by the Couette tests."""
    monkeypatch.setenv("HC_FUSE_VEL", "0")
    g, o, x0 = _shear_case(hc.HC_FP64, steps=20)
    xg, _, _ = g.vertices()
    xo, _, _ = o.vertices()
    assert cases.rel_l2(xg, xo) < 1e-10
    _, ug = g.fluid()
    _, uo = o.fluid()
    assert cases.rel_l2(ug, uo) < 1e-10


def test_periodicity_changed_after_binding():
    """hc_set_periodicity on a bound lattice (setSystemPeriodicity after the lattice
    exists, core/hemoCell.cpp:168-180) steps as a lattice created with that
    periodicity, cells included."""
    print("[parity] periodicity", flush=True)
    u = cases.units()
    shape = (24, 20, 22)
    mask = cases.channel_mask(*shape, axis=1)
    rng = np.random.default_rng(31)
    f0 = cases.random_populations(int(np.prod(shape)), rng, amp=1e-4)
    sims = []
    for start in ((0, 0, 0), (1, 0, 1)):
        g = hc.Simulation(0)
        g.lattice_init(*shape, u.tau, periodic=start, node_type=mask)
        g.set_populations(f0)
        g.set_body_force([0, 0, 1e-6])
        g.set_force_limit(u.f_limit)
        model, v, t, p, inner = cases.mesh("PLT", u, hc.build_mesh)
        tid = g.add_cell_type(model, v, t, p, inner)
        g.add_cells(tid, cases.place(v, (1.2, 10.3, 20.9), cases.rotation(np.random.default_rng(2)))[None])
        sims.append(g)
    sims[0].set_periodicity((1, 0, 1))
    for g in sims:
        g.step(7)
    assert cases.rel_l2(sims[0].vertices()[0], sims[1].vertices()[0]) < 1e-12
    assert cases.rel_l2(sims[0].populations(), sims[1].populations()) < 1e-12
