"""Pins of the oracle's wall-contact removal (CPU).

The reference removes a particle whose nearest node is a wall node right after
its advance (core/hemoCellParticleField.cpp:566-588); its cell is then
incomplete: applyConstitutiveModel (:633-675) resets the forces of the type's
vertices and runs the model on complete cells only, so the cell's remaining
vertices carry no membrane force but keep drifting with the fluid; an
interior-viscosity run deletes incomplete cells whole before each entire-grid
pass (core/hemoCell.cpp:347-351, deleteIncompleteCells :510-553).

Case (cases.wall_contact_case): walls y = 0 and y = 19, an RBC lying flat with
its lowest vertex 0.3 spacings above the y = 0 wall, a platelet mid-channel.
"""
import numpy as np

import cases
import mesh_oracle
from oracle import OracleSim


def _case(iv=False):
    u = cases.units()
    o = OracleSim(*cases.WALL_SHAPE, u.tau, periodic=(1, 0, 1),
                  node_type=cases.channel_mask(*cases.WALL_SHAPE, axis=1))
    rbc, plt = cases.wall_contact_case([o], u, mesh_oracle.build_mesh)
    if iv:
        o.set_interior_viscosity(0, 2.0 * u.tau)
        o.set_interior_viscosity_timescales(1, 5)
    return o, rbc, plt


def test_vertices_on_wall_nodes_are_removed_once_and_stay():
    o, (r0, r1), (p0, p1) = _case()
    x0 = o.vertices()[0]
    o.step(1)
    x1 = o.vertices()[0]
    rm = o.removed_vertices()
    # exactly the vertices whose advanced position rounds to the y = 0 wall nodes
    expect = np.zeros_like(rm)
    expect[r0:r1] = x1[r0:r1, 1] < 0.5
    assert rm.sum() > 0 and np.array_equal(rm, expect)
    assert not rm[p0:p1].any()
    assert o.wall_hits() == rm.sum()
    assert np.all(x1[rm, 1] != x0[rm, 1])  # they advanced once, then were removed
    o.step(20)
    x2 = o.vertices()[0]
    assert np.array_equal(o.removed_vertices(), rm)  # nothing else reached a wall
    assert o.wall_hits() == rm.sum()  # a removed vertex is counted once
    assert np.array_equal(x2[rm], x1[rm])  # removed vertices do not move
    keep = ~rm
    keep[p0:] = False
    assert np.abs(x2[keep] - x1[keep]).max() > 1e-4  # the rest drift with the fluid
    assert list(o.incomplete_cells()) == [True, False]


def test_incomplete_cell_carries_no_force():
    o, (r0, r1), (p0, p1) = _case()
    o.step(3)
    _, _, f = o.vertices()
    assert np.all(f[r0:r1] == 0.0)
    assert np.abs(f[p0:p1]).max() > 0.0
    # the next spread holds the platelet's contributions only: nothing below y = 8
    # (the platelet's support lies at 9 <= y <= 15)
    o.stage("spread")
    F = o.force_field().reshape(3, *cases.WALL_SHAPE)
    assert np.all(F[:, :, :8, :] == 0.0)
    assert np.abs(F).max() > 0.0


def test_complete_cells_are_untouched_by_the_removal():
    """The same case with the RBC 1.3 spacings above the wall: nothing is
    removed and its mechanics run (non-zero forces)."""
    u = cases.units()
    o = OracleSim(*cases.WALL_SHAPE, u.tau, periodic=(1, 0, 1),
                  node_type=cases.channel_mask(*cases.WALL_SHAPE, axis=1))
    (r0, r1), _ = cases.wall_contact_case([o], u, mesh_oracle.build_mesh, gap=1.3)
    o.step(3)
    assert not o.removed_vertices().any() and o.wall_hits() == 0
    assert np.abs(o.vertices()[2][r0:r1]).max() > 0.0


def test_interior_viscosity_grid_pass_deletes_incomplete_cells():
    o, (r0, r1), (p0, p1) = _case(iv=True)
    o.set_iteration(1)
    o.step(1)  # iteration 1: vertices removed, no entire-grid pass
    rm = o.removed_vertices()
    assert 0 < rm[r0:r1].sum() < r1 - r0
    o.step(4)  # iterations 2..5: the entire-grid pass at iteration 5
    rm = o.removed_vertices()
    assert rm[r0:r1].all() and not rm[p0:p1].any()
    # the deleted cell has no inner nodes; the platelet's class is not set (it has
    # no interior viscosity), so no node is inner
    assert not np.any(o.interior_viscosity() > 0)
