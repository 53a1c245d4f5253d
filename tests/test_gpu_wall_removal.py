"""Wall-contact removal on the GPU against the oracle.

A vertex whose nearest node is a wall node after its advance is removed
(core/hemoCellParticleField.cpp:566-588); its cell is then incomplete: no
mechanics, zero force (applyConstitutiveModel, :633-675), the other vertices
drift on; an interior-viscosity run deletes the cell whole at its next
entire-grid pass (core/hemoCell.cpp:347-351).  Case: cases.wall_contact_case
(walls y = 0 and y = 19, an RBC 0.3 spacings above the y = 0 wall, a platelet
mid-channel).

* FP64, fused and unfused step kernels: after 1 and 31 steps the removed
  vertices, the incomplete cells and the wall-hit count equal the oracle's;
  positions, displacements, vertex velocities and forces and the populations
  within 1e-10 relative L2 (the incomplete RBC's forces exactly zero).
* FP32-shifted: the same vertices removed, positions within 1e-6.
* Interior viscosity (grid pass every 5 steps, from iteration 1): the cell
  loses its bottom vertices at iteration 1 and is deleted whole at iteration 5,
  as in the oracle.
* A run checkpointed after the removal (removed flags saved and restored with
  the rest of the state) and resumed in a fresh context ends bit for bit where
  the uninterrupted run ends (exact FP64 spread on both).
* Two such cases stepped together through hc_step_batch remove the same
  vertices as each stepped alone and end within 1e-12 of them.
"""
import numpy as np
import pytest

import cases
import hemocell_amd as hc
from oracle import OracleSim

pytestmark = pytest.mark.gpu


def _gpu(precision, u):
    g = hc.Simulation(0)
    g.lattice_init(*cases.WALL_SHAPE, u.tau, periodic=(1, 0, 1),
                   node_type=cases.channel_mask(*cases.WALL_SHAPE, axis=1), precision=precision)
    return g


def _pair(precision, monkeypatch, fuse=True, iv=False):
    monkeypatch.setenv("HC_FUSE", "1" if fuse else "0")
    u = cases.units()
    g = _gpu(precision, u)
    o = OracleSim(*cases.WALL_SHAPE, u.tau, periodic=(1, 0, 1),
                  node_type=cases.channel_mask(*cases.WALL_SHAPE, axis=1))
    ranges = cases.wall_contact_case([g, o], u, hc.build_mesh)
    if iv:
        for s in (g, o):
            s.set_interior_viscosity(0, 2.0 * u.tau)
            s.set_interior_viscosity_timescales(1, 5)
            s.set_iteration(1)
    return g, o, ranges


def _compare(g, o, x0):
    xg, vg, fg = g.vertices()
    xo, vo, fo = o.vertices()
    e = dict(x=cases.rel_l2(xg, xo), disp=cases.rel_l2(xg - x0, xo - x0), v=cases.rel_l2(vg, vo),
             f=cases.rel_l2(fg, fo), pop=cases.rel_l2(g.populations(), o.populations()))
    return e, fg


@pytest.mark.parametrize("fuse", [True, False])
def test_removal_matches_the_oracle_fp64(monkeypatch, fuse):
    print("[wall removal] fp64 fuse=%s" % fuse, flush=True)
    g, o, ((r0, r1), (p0, p1)) = _pair(hc.HC_FP64, monkeypatch, fuse)
    x0 = g.vertices()[0]
    g.step(1)
    o.step(1)
    rm = o.removed_vertices()
    assert rm.sum() > 0 and np.array_equal(g.removed_vertices(), rm)
    assert g.wall_hits() == o.wall_hits() == rm.sum()
    for _ in range(3):
        g.step(10)
        o.step(10)
    assert np.array_equal(g.removed_vertices(), o.removed_vertices())
    assert g.wall_hits() == o.wall_hits()
    assert list(g.incomplete_cells()) == list(o.incomplete_cells()) == [True, False]
    e, fg = _compare(g, o, x0)
    print("[wall removal] fp64 fuse=%s relative L2 vs oracle after 31 steps:" % fuse, e)
    assert all(v < 1e-10 for v in e.values()), e
    assert np.all(fg[r0:r1] == 0.0) and np.abs(fg[p0:p1]).max() > 0.0


def test_removal_fp32(monkeypatch):
    print("[wall removal] fp32", flush=True)
    g, o, _ = _pair(hc.HC_FP32_SHIFTED, monkeypatch)
    g.step(31)
    o.step(31)
    assert np.array_equal(g.removed_vertices(), o.removed_vertices())
    assert g.wall_hits() == o.wall_hits() > 0
    e = cases.rel_l2(g.vertices()[0], o.vertices()[0])
    print("[wall removal] fp32 positions relative L2 vs oracle after 31 steps: %.2e" % e)
    assert e < 1e-6


def test_interior_viscosity_deletes_the_incomplete_cell(monkeypatch):
    print("[wall removal] interior viscosity", flush=True)
    g, o, ((r0, r1), (p0, p1)) = _pair(hc.HC_FP64, monkeypatch, iv=True)
    x0 = g.vertices()[0]
    g.step(1)
    o.step(1)
    rg = g.removed_vertices()
    assert np.array_equal(rg, o.removed_vertices()) and 0 < rg[r0:r1].sum() < r1 - r0
    g.step(4)
    o.step(4)
    rg = g.removed_vertices()
    assert np.array_equal(rg, o.removed_vertices()) and rg[r0:r1].all() and not rg[p0:p1].any()
    assert np.array_equal(g.interior_viscosity() > 0, o.interior_viscosity() > 0)
    g.step(5)
    o.step(5)
    e, _ = _compare(g, o, x0)
    print("[wall removal] interior viscosity relative L2 vs oracle after 10 steps:", e)
    assert all(v < 1e-10 for v in e.values()), e


def test_resume_after_removal_is_bit_identical(monkeypatch):
    print("[wall removal] checkpoint resume", flush=True)
    monkeypatch.setenv("HC_FUSE", "1")
    u = cases.units()

    def fresh():
        g = _gpu(hc.HC_FP64, u)
        cases.wall_contact_case([g], u, hc.build_mesh)
        g.set_exact_spread(True)
        return g

    whole, first = fresh(), fresh()
    whole.step(12)
    first.step(5)
    x, v, f = first.vertices()
    st = dict(x=x, v=v, f=f, pop=first.populations(), it=first.iteration, rm=first.removed_vertices())
    first.close()
    assert st["rm"].any()
    resumed = fresh()
    resumed.set_vertices(st["x"], st["v"], st["f"])
    resumed.set_populations(st["pop"])
    resumed.set_iteration(st["it"])
    resumed.set_removed_vertices(st["rm"])
    resumed.step(7)
    for a, b in zip(whole.vertices(), resumed.vertices()):
        assert np.array_equal(a, b)
    assert np.array_equal(whole.populations(), resumed.populations())
    assert np.array_equal(whole.removed_vertices(), resumed.removed_vertices())


def test_batched_cases_remove_like_alone(monkeypatch):
    """Two wall-contact cases (fluid set moving at different speeds) stepped
    together through hc_step_batch remove the same vertices as each stepped
    alone, and end within 1e-12 of them (the batch's cell kernels use other block
    sizes, so the volume sums round differently)."""
    print("[wall removal] batch", flush=True)
    monkeypatch.setenv("HC_FUSE", "1")
    u = cases.units()

    def fresh(u0):
        g = _gpu(hc.HC_FP64, u)
        cases.wall_contact_case([g], u, hc.build_mesh, u0=u0)
        return g

    alone = [fresh(-0.004), fresh(-0.002)]
    batch = [fresh(-0.004), fresh(-0.002)]
    for s in alone:
        s.step(25)
    hc.step_batch(batch, 10)
    hc.step_batch(batch, 15)
    for a, b in zip(alone, batch):
        assert np.array_equal(a.removed_vertices(), b.removed_vertices()) and a.removed_vertices().any()
        assert a.wall_hits() == b.wall_hits()
        assert cases.rel_l2(b.vertices()[0], a.vertices()[0]) <= 1e-12
        assert cases.rel_l2(b.populations(), a.populations()) <= 1e-12
