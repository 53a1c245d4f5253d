"""Pins of the oracle's lattice arithmetic (CPU).

Palabos, which holds the reference's LBM arithmetic, is not vendored, so the
oracle's Guo-forced BGK, bounce-back and regularized velocity boundary are pinned
by what the mathematics fixes, with expectations derived independently of the
oracle's code:
  * moments of one collision from D3Q19 isotropy: rhoBar' = rhoBar,
    j' = j + rho F, Pi' = (1-w) Pi + w Pi_eq(rho u) + rho (1 - w/2)(u F + F u)
    with u = j/rho + F/2 and Pi_eq(J) = rhoBar/3 I + J J / rho;
  * rest state, mass conservation, global momentum gain under a uniform force;
  * analytic flows: decaying shear wave (nu = (tau - 1/2)/3), body-force
    Poiseuille channel between bounce-back walls, Couette flow between
    velocity walls.
"""
import numpy as np
import pytest

import cases
from oracle import C, W, OracleSim, collide_one, feq, velocity_bc_one


def moments(f):
    rb = f.sum(axis=0)
    j = np.einsum("id,i...->d...", C.astype(float), f)
    Pi = np.einsum("ia,ib,i...->ab...", C.astype(float), C.astype(float), f)
    return rb, j, Pi


def test_lattice_constants():
    assert np.isclose(W.sum(), 1.0)
    assert np.allclose(np.einsum("i,ia->a", W, C), 0)
    assert np.allclose(np.einsum("i,ia,ib->ab", W, C, C), np.eye(3) / 3)
    for i in range(1, 10):
        assert np.array_equal(C[i + 9], -C[i])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_collision_moments_closed_form(seed):
    rng = np.random.default_rng(seed)
    f = 1e-2 * rng.uniform(-1, 1, 19)
    F = 1e-3 * rng.uniform(-1, 1, 3)
    omega = 1.0 / 0.83
    rb, j, Pi = moments(f)
    rho = 1 + rb
    u = j / rho + F / 2
    J = rho * u
    Pi_eq = rb / 3 * np.eye(3) + np.outer(J, J) / rho
    mu = rho * (1 - omega / 2)
    out = collide_one(f, F, omega)
    rb2, j2, Pi2 = moments(out)
    assert abs(rb2 - rb) < 1e-16
    assert np.allclose(j2, j + rho * F, atol=1e-16, rtol=0)
    expect = (1 - omega) * Pi + omega * Pi_eq + mu * (np.outer(u, F) + np.outer(F, u))
    assert np.allclose(Pi2, expect, atol=1e-15, rtol=0)


def test_equilibrium_is_a_fixed_point():
    j = np.array([1e-3, -2e-3, 5e-4])
    rb = 3e-3
    f = np.array([feq(i, rb, j) for i in range(19)])
    out = collide_one(f, np.zeros(3), 1.3)
    assert np.allclose(out, f, atol=1e-17, rtol=0)
    # and its moments are the equilibrium ones
    r2, j2, Pi2 = moments(f)
    assert abs(r2 - rb) < 1e-17
    assert np.allclose(j2, j, atol=1e-17)
    assert np.allclose(Pi2, rb / 3 * np.eye(3) + np.outer(j, j) / (1 + rb), atol=1e-17)


def test_velocity_boundary_imposes_wall_velocity():
    rng = np.random.default_rng(3)
    uw = np.array([2e-3, -1e-3, 0.0])
    for axis in range(3):
        for orient in (-1, 1):
            f = 1e-3 * rng.uniform(-1, 1, 19)
            out = velocity_bc_one(f, axis, orient, uw.copy() if axis != 0 else np.array([0.0, 2e-3, -1e-3]))
            w = uw if axis != 0 else np.array([0.0, 2e-3, -1e-3])
            rb, j, _ = moments(out)
            assert np.allclose(j, (1 + rb) * w, atol=1e-17)
            # density from the known populations only (Zou-He mass balance)
            known = [i for i in range(19) if C[i][axis] in (0, orient)]
            on = sum(f[i] for i in known if C[i][axis] == 0)
            nrm = sum(f[i] for i in known if C[i][axis] == orient)
            vn = orient * w[axis]
            assert abs(rb - (2 * nrm + on - vn) / (1 + vn)) < 1e-15


def test_rest_state_is_exact():
    o = OracleSim(6, 5, 4, 0.9)
    o.step(20)
    assert np.count_nonzero(o.populations()) == 0


def test_mass_conservation_periodic_with_force():
    rng = np.random.default_rng(4)
    o = OracleSim(6, 7, 5, 0.7)
    o.set_populations(cases.random_populations(o.N, rng))
    o.set_body_force([1e-5, 2e-5, -1e-5])
    m0 = o.populations().sum()
    o.step(30)
    assert abs(o.populations().sum() - m0) < 1e-13


def test_global_momentum_gain_equals_force_times_mass():
    rng = np.random.default_rng(5)
    o = OracleSim(5, 6, 7, 1.3)
    o.set_populations(cases.random_populations(o.N, rng))
    g = np.array([2e-5, -1e-5, 3e-5])
    o.set_body_force(g)
    for _ in range(3):
        f = o.populations()
        mass = (1 + f.sum(axis=0)).sum()
        j0 = np.einsum("ia,in->a", C.astype(float), f)
        o.step(1)
        j1 = np.einsum("ia,in->a", C.astype(float), o.populations())
        assert np.allclose(j1 - j0, g * mass, rtol=1e-11, atol=0)


def test_bounceback_box_conserves_mass():
    rng = np.random.default_rng(6)
    n = 7
    t = np.ones((n, n, n), dtype=np.uint8)
    t[1:-1, 1:-1, 1:-1] = 0
    mask = t.reshape(-1)
    o = OracleSim(n, n, n, 0.8, periodic=(0, 0, 0), node_type=mask)
    f = cases.random_populations(o.N, rng)
    f[:, mask == 1] = 0
    o.set_populations(f)
    m0 = f.sum()
    o.step(25)
    assert abs(o.populations().sum() - m0) < 1e-14


def test_shear_wave_decay_matches_viscosity():
    L, tau, u0 = 32, 0.8, 1e-4
    nu = (tau - 0.5) / 3
    k = 2 * np.pi / L
    o = OracleSim(1, 1, L, tau)
    z = np.arange(L)
    ux = u0 * np.sin(k * z)
    f = np.zeros((19, L))
    for i in range(19):
        for n in range(L):
            f[i, n] = feq(i, 0.0, np.array([ux[n], 0, 0]))
    # start from the equilibrium plus the first-order non-equilibrium part so the
    # wave decays as one exponential from step 0
    o.set_populations(f)
    T = 200
    o.step(T)
    _, u = o.fluid()
    amp = 2 / L * np.sum(u[:, 0] * np.sin(k * z))
    expect = u0 * np.exp(-nu * k * k * T)
    assert abs(amp / expect - 1) < 0.01


def test_poiseuille_channel_between_bounceback_walls():
    ny, tau, g = 12, 1.0, 1e-6
    nu = (tau - 0.5) / 3
    mask = cases.channel_mask(1, ny, 1, axis=1)
    o = OracleSim(1, ny, 1, tau, periodic=(1, 0, 1), node_type=mask)
    o.set_body_force([g, 0, 0])
    o.step(6000)
    _, u = o.fluid()
    y = np.arange(ny)
    fl = (y >= 1) & (y <= ny - 2)
    # full-way bounce-back: walls half way between wall and first fluid node
    expect = g / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)
    ux = u[:, 0]
    assert np.allclose(ux[fl], ux[fl][::-1], rtol=1e-9)
    assert np.max(np.abs(ux[fl] - expect[fl])) < 0.02 * expect[fl].max()


def test_couette_between_velocity_walls_is_linear():
    nz, tau, U = 9, 1.0, 1e-3
    mask = cases.couette_mask(1, 1, nz)
    o = OracleSim(1, 1, nz, tau, periodic=(1, 1, 0), node_type=mask)
    o.set_wall_velocity(4, [U, 0, 0])
    o.set_wall_velocity(5, [-U, 0, 0])
    o.init_equilibrium(1.0, [0, 0, 0])
    o.step(4000)
    _, u = o.fluid()
    z = np.arange(nz)
    expect = U * (1 - 2 * z / (nz - 1))
    assert np.max(np.abs(u[:, 0] - expect)) < 1e-9
    assert np.max(np.abs(u[:, 1:])) < 1e-12
