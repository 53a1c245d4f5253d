"""Pins of the oracle's immersed-boundary coupling and membrane mechanics (CPU).

IBM (core/immersedBoundaryMethod.h:37-138, hemoCellParticleField.cpp:819-863):
  phi2 weights are the trilinear hat weights (partition of unity, exact for
  linear fields), the spread conserves the total force, spread and interpolation
  are adjoint, wall nodes never receive weight, and forces above f_limit are
  capped to f_limit in place.
Mechanics (mechanics/*Model.cpp): zero force on the rest shape, zero net force,
zero net torque for the area/volume/link terms, translation and rotation
covariance, and the link force equals the negative gradient of its energy
E = k L0 (s^2/2 - ln|9 - s^2|/2), s = (L - L0)/L0.
"""
import numpy as np
import pytest

import cases
import mesh_oracle
from oracle import OracleSim, feq


def rbc_sim(n=40, kind="RBC", **over):
    u = cases.units()
    o = OracleSim(n, n, n, u.tau)
    model, v, t, p, inner = cases.mesh(kind, u, mesh_oracle.build_mesh)
    for k, val in over.items():
        setattr(p, k, val)
    tid = o.add_cell_type(model, v, t, p, inner)
    return o, tid, v, t, p


# ------------------------------------------------------------------- IBM ----
def test_phi2_weights_are_trilinear():
    rng = np.random.default_rng(1)
    o = OracleSim(8, 9, 10, 1.0)
    for _ in range(50):
        p = rng.uniform(0, [8, 9, 10])
        idx, w = o.ibm_kernel(p)
        assert len(idx) <= 8 and abs(w.sum() - 1) < 1e-14
        lo = np.floor(p)
        fr = p - lo
        for n, wk in zip(idx, w):
            z = n % 10
            y = (n // 10) % 9
            x = n // 90
            d = [(x - lo[0]) % 8, (y - lo[1]) % 9, (z - lo[2]) % 10]
            expect = np.prod([fr[a] if d[a] == 1 else 1 - fr[a] for a in range(3)])
            assert abs(wk - expect) < 1e-14


def test_phi2_skips_wall_nodes_and_renormalises():
    mask = cases.channel_mask(6, 6, 6, axis=1)
    o = OracleSim(6, 6, 6, 1.0, periodic=(1, 0, 1), node_type=mask)
    idx, w = o.ibm_kernel([2.3, 0.8, 3.6])  # the y = 0 plane is a wall
    ys = (np.asarray(idx) // 6) % 6
    assert np.all(ys >= 1) and abs(w.sum() - 1) < 1e-14
    assert len(idx) == 4  # only the y = 1 layer survives


def _eq_field(o, ufun):
    nx, ny, nz = o.shape
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    U = ufun(X.reshape(-1), Y.reshape(-1), Z.reshape(-1))
    f = np.zeros((19, o.N))
    for n in range(o.N):
        for i in range(19):
            f[i, n] = feq(i, 0.0, U[n])
    o.set_populations(f)
    return U


def test_interpolation_exact_for_linear_field():
    u = cases.units()
    o = OracleSim(20, 20, 20, u.tau, periodic=(0, 0, 0))
    A = np.array([[1e-5, 2e-5, -1e-5], [3e-6, -4e-6, 2e-5], [-2e-5, 1e-6, 5e-6]])
    _eq_field(o, lambda x, y, z: (np.stack([x, y, z], -1) - 10.0) @ A.T + np.array([1e-4, -2e-4, 5e-5]))
    model, v, t, p, inner = cases.mesh("RBC", u, mesh_oracle.build_mesh)
    tid = o.add_cell_type(model, v, t, p, inner)
    x = cases.place(v, (10.3, 9.6, 10.1), cases.rotation(np.random.default_rng(2)))
    o.add_cells(tid, x[None])
    o.stage("interpolate")
    _, vel, _ = o.vertices()
    # u = j/rho at equilibrium; tiny deviation from rho = 1 + O(u^2) terms is absent (rhoBar = 0)
    expect = (x - 10.0) @ A.T + np.array([1e-4, -2e-4, 5e-5])
    assert np.max(np.abs(vel - expect)) < 1e-17 + 1e-12 * np.abs(expect).max()


def test_spread_conserves_force_and_is_adjoint_to_interpolation():
    rng = np.random.default_rng(3)
    o, tid, v, t, p = rbc_sim(24)
    x = cases.perturb(cases.place(v, (12.2, 11.7, 12.4), cases.rotation(rng)), rng, 0.05)
    o.add_cells(tid, x[None])
    F = 1e-3 * rng.normal(size=x.shape)
    o.set_vertices(None, None, F)
    U = _eq_field(o, lambda X, Y, Z: 1e-4 * np.stack([np.sin(X / 3.0), np.cos(Y / 4.0), np.sin(Z / 5.0 + 1)], -1))
    o.stage("spread")
    Fn = o.force_field()
    assert np.allclose(Fn.sum(axis=1), F.sum(axis=0), rtol=1e-12, atol=1e-18)
    o.stage("interpolate")
    _, vel, _ = o.vertices()
    # <spread F, u> = <F, interp u>; the node velocity includes +F/2 of the spread itself
    rho, unode = o.fluid()
    lhs = np.sum(Fn.T * unode)
    rhs = np.sum(F * vel)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_force_cap_rescales_in_place():
    u = cases.units()
    o, tid, v, t, p = rbc_sim(24)
    o.add_cells(tid, (v + 12.0)[None])
    F = np.zeros((len(v), 3))
    F[0] = [3 * u.f_limit, 4 * u.f_limit, 0]
    F[1] = [0.5 * u.f_limit, 0, 0]
    o.set_vertices(None, None, F)
    o.set_force_limit(u.f_limit)
    o.stage("spread")
    _, _, F2 = o.vertices()
    assert np.allclose(F2[0], [0.6 * u.f_limit, 0.8 * u.f_limit, 0], rtol=1e-15)
    assert np.array_equal(F2[1], F[1])
    assert abs(o.force_field().sum(axis=1)[0] - (0.6 + 0.5) * u.f_limit) < 1e-16


# ------------------------------------------------------------- mechanics ----
def mech_forces(o, x, vel=None):
    o.set_vertices(x, vel, None)
    o.stage("mechanics", 1)
    return o.vertices()[2]


@pytest.mark.parametrize("kind", ["RBC", "PLT", "WBC"])
@pytest.mark.parametrize("shift", [0.0, 300.0])
def test_rest_shape_is_force_free(kind, shift):
    # also 300 spacings from the origin: the volume's triple products are taken
    # relative to the cell's vertex 0, so they do not lose ~|x|^3/V digits to
    # cancellation as absolute coordinates (the reference's form) would; there
    # the volume term alone would leave forces ~1e-6 of the stiffness.  The bound
    # is 1e-11 there because storing the shifted positions rounds the shape by
    # ulp(300) ~ 6e-14 spacings.
    o, tid, v, t, p = rbc_sim(40, kind)
    x = v + shift
    o.add_cells(tid, x[None])
    F = mech_forces(o, x)
    bound = 1e-12 if shift == 0.0 else 1e-11
    assert np.abs(F).max() < bound * max(p.k_link, p.k_bend, p.k_area, p.k_volume)


@pytest.mark.parametrize("kind", ["RBC", "PLT", "WBC"])
def test_internal_forces_sum_to_zero(kind):
    rng = np.random.default_rng(4)
    o, tid, v, t, p = rbc_sim(40, kind)
    x = cases.perturb(v + 20.0, rng, 0.1)
    o.add_cells(tid, x[None])
    vel = 1e-4 * rng.normal(size=x.shape)
    F = mech_forces(o, x, vel)
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).sum()


@pytest.mark.parametrize("kind", ["RBC", "PLT", "WBC"])
def test_torque_free_without_bending(kind):
    rng = np.random.default_rng(5)
    o, tid, v, t, p = rbc_sim(40, kind, k_bend=0.0, eta_m=0.0)
    x = cases.perturb(v + 20.0, rng, 0.1)
    o.add_cells(tid, x[None])
    F = mech_forces(o, x)
    c = x.mean(axis=0)
    tau = np.cross(x - c, F).sum(axis=0)
    assert np.abs(tau).max() < 1e-12 * np.abs(np.cross(x - c, F)).sum()


@pytest.mark.parametrize("kind", ["RBC", "PLT", "WBC"])
def test_translation_and_rotation_covariance(kind):
    rng = np.random.default_rng(6)
    o, tid, v, t, p = rbc_sim(40, kind)
    x = cases.perturb(v, rng, 0.08)
    o.add_cells(tid, (x + 20.0)[None])
    F0 = mech_forces(o, x + 20.0)
    F1 = mech_forces(o, x + np.array([20.7, 19.1, 20.3]))
    assert np.allclose(F0, F1, rtol=1e-9, atol=1e-13 * np.abs(F0).max())
    R = cases.rotation(rng)
    F2 = mech_forces(o, x @ R.T + 20.0)
    assert np.allclose(F2, F0 @ R.T, rtol=1e-8, atol=1e-11 * np.abs(F0).max())


def test_link_force_is_energy_gradient():
    rng = np.random.default_rng(7)
    o, tid, v, t, p = rbc_sim(40, "RBC", k_bend=0.0, k_area=0.0, k_volume=0.0)
    x = cases.perturb(v + 20.0, rng, 0.15)
    o.add_cells(tid, x[None])
    tab = o.type_tables(tid)
    e, L0 = tab["edge"], tab["edge_len"]

    def energy(xx):
        L = np.linalg.norm(xx[e[:, 1]] - xx[e[:, 0]], axis=1)
        s = (L - L0) / L0
        return np.sum(p.k_link * L0 * (s * s / 2 - 0.5 * np.log(np.abs(9 - s * s))))

    F = mech_forces(o, x)
    h = 1e-6
    for k in rng.choice(len(v), 6, replace=False):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k, a] += h
            xm[k, a] -= h
            grad = (energy(xp) - energy(xm)) / (2 * h)
            assert abs(F[k, a] + grad) < 1e-6 * np.abs(F).max()


def test_volume_and_area_forces_restore_the_rest_shape():
    o, tid, v, t, p = rbc_sim(40, "RBC", k_bend=0.0, k_link=0.0)
    c = v.mean(axis=0)
    x = c + 1.02 * (v - c) + 20.0
    o.add_cells(tid, x[None])
    F = mech_forces(o, x)
    # an inflated cell is pushed back inwards everywhere on average
    assert np.sum(F * (x - x.mean(axis=0))) < 0


# ------------------------------------------------------------------- mesh ----
def test_rbc_rest_volume_matches_reference_template():
    u = cases.units()
    v, t = mesh_oracle.build_mesh(1, 3.91e-6 / u.dx, 600)
    assert (len(v), len(t)) == (642, 1280)
    vol_um3 = mesh_oracle.mesh_volume(v, t) * (u.dx * 1e6) ** 3
    assert abs(vol_um3 - 81.116) < 5e-4  # tools/cell_templates/RBC_template.xml:16


@pytest.mark.parametrize("shape,r,n,a,counts", [(1, 7.82, 600, 0.3, (642, 1280)), (6, 2.0, 66, 0.25, (66, 128)),
                                               (0, 8.0, 600, 0.3, (642, 1280)), (0, 8.0, 2000, 0.3, (2562, 5120))])
def test_mesh_topology(shape, r, n, a, counts):
    v, t = mesh_oracle.build_mesh(shape, r, n, a)
    assert (len(v), len(t)) == counts
    edges = {tuple(sorted((f[k], f[(k + 1) % 3]))) for f in t for k in range(3)}
    assert len(v) - len(edges) + len(t) == 2  # closed genus-0 surface
    assert mesh_oracle.mesh_volume(v, t) > 0  # outward orientation


def test_tables_consistent():
    o, tid, v, t, p = rbc_sim(40)
    tab = o.type_tables(tid)
    assert len(tab["edge"]) == 1920
    assert np.all(tab["edge"][:, 0] < tab["edge"][:, 1])
    tri = {tuple(sorted(f)) for f in t}
    for vid in range(0, len(v), 37):
        r = tab["ring"][vid, :tab["nring"][vid]]
        for j in range(len(r)):
            assert tuple(sorted((vid, r[j], r[(j + 1) % len(r)]))) in tri
    assert abs(tab["volume_eq"] - mesh_oracle.mesh_volume(v, t)) < 1e-10 * tab["volume_eq"]
