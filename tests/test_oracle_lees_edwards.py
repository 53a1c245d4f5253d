"""Pins of the oracle's Lees-Edwards planes (CPU), helper/leesEdwardsBC.h.

After each collide-and-stream the planes z = nz-1 and z = 0 relax once more
towards the equilibrium of (-+ vHalf, 0, 0), vHalf = (nz-1) shear / 2, and the
five populations that entered across the plane come from the plane at x + d
(top) or x - d (bottom).  Checked here:

* the steady state of a cell-free box is the exact linear profile from +vHalf at
  z = 0 to -vHalf at z = nz-1 (the planes act as moving walls), at any
  displacement, with uniform density;
* a fluid at rest stays at rest bit for bit;
* an integer displacement m moves the entering populations of the top plane by
  -m and of the bottom plane by +m along x (np.roll of the d = 0 result) and
  changes nothing else; a fractional one interpolates linearly between the two
  neighbouring integer results.
"""
import numpy as np
import pytest

import cases
from oracle import OracleSim

TOP_IN = [3, 6, 8, 16, 18]
BOTTOM_IN = [7, 9, 12, 15, 17]


@pytest.mark.parametrize("disp", [0.0, 0.37])
def test_steady_state_is_the_linear_shear_profile(disp):
    nx, ny, nz, shear = 6, 4, 21, 1e-3
    o = OracleSim(nx, ny, nz, 1.0)
    o.set_lees_edwards(shear, disp)
    o.step(3000)
    rho, u = o.fluid()
    u = u.reshape(nx, ny, nz, 3)
    vh = (nz - 1) * shear / 2
    z = np.arange(nz)
    lin = vh * (1 - 2 * z / (nz - 1))
    assert np.abs(u[..., 0] - lin[None, None, :]).max() < 1e-9 * vh
    assert np.abs(u[..., 1:]).max() < 1e-12 * vh
    assert np.abs(rho - 1.0).max() < 1e-12


def test_rest_state_stays_at_rest():
    o = OracleSim(5, 4, 6, 0.8)
    o.set_lees_edwards(0.0, 0.6)
    o.step(5)
    assert np.all(o.populations() == 0.0)


def _stage(f0, shape, disp, tau=0.9):
    o = OracleSim(*shape, tau)
    o.set_populations(f0)
    o.set_lees_edwards(1.5e-3, disp)
    o.set_iteration(1)  # d = fmod(disp x 1, nx)
    o.stage_lees_edwards()
    return o.populations().reshape(19, *shape)


def test_displacement_shifts_the_entering_populations():
    shape = (7, 3, 5)
    rng = np.random.default_rng(4)
    f0 = cases.random_populations(int(np.prod(shape)), rng)
    a = _stage(f0, shape, 0.0)
    for m in (1, 3):
        b = _stage(f0, shape, float(m))
        for i in range(19):
            top_in, bot_in = i in TOP_IN, i in BOTTOM_IN
            # top plane: entering populations taken from x + m
            want = np.roll(a[i, :, :, -1], -m, axis=0) if top_in else a[i, :, :, -1]
            assert np.array_equal(b[i, :, :, -1], want), (m, i)
            want = np.roll(a[i, :, :, 0], m, axis=0) if bot_in else a[i, :, :, 0]
            assert np.array_equal(b[i, :, :, 0], want), (m, i)
            assert np.array_equal(b[i, :, :, 1:-1], a[i, :, :, 1:-1])
    # a fractional displacement interpolates between the neighbouring integers
    b1, b2, bf = _stage(f0, shape, 1.0), _stage(f0, shape, 2.0), _stage(f0, shape, 1.25)
    for i in TOP_IN:
        assert np.allclose(bf[i, :, :, -1], 0.25 * b2[i, :, :, -1] + 0.75 * b1[i, :, :, -1], rtol=0, atol=1e-17)
    for i in BOTTOM_IN:
        assert np.allclose(bf[i, :, :, 0], 0.25 * b2[i, :, :, 0] + 0.75 * b1[i, :, :, 0], rtol=0, atol=1e-17)
