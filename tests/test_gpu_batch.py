"""Batched steps of several independent cases (hc_step_batch) against each case
stepped alone (hc_step): every case computes what its own fused step computes,
so FP64 positions and populations agree to round-off (the spread's float
atomics commit in no fixed order): relative L2 <= 1e-12, FP32 <= 1e-5.  The
cases differ in cell kind, box size, boundary set-up (velocity walls on two or
six faces, stretch forces) and AA parity (some have taken an odd number of
steps before the batch)."""
import numpy as np
import pytest

import cases as tcases
import hemocell_amd as hc
from hemocell_amd import serving

pytestmark = pytest.mark.gpu


def _cases():
    return [serving.Case(ident=0, kind="shear", cell="RBC", edge=32, steps=0, shear_rate=300.0),
            serving.Case(ident=1, kind="stretch", cell="PLT", edge=24, steps=0, force_pN=80.0),
            serving.Case(ident=2, kind="shear", cell="WBC", edge=36, steps=0, shear_rate=40.0),
            serving.Case(ident=3, kind="stretch", cell="RBC", edge=34, steps=0, force_pN=150.0),
            serving.Case(ident=4, kind="shear", cell="PLT", edge=20, steps=0, shear_rate=900.0)]


@pytest.mark.parametrize("precision,tol", [(hc.HC_FP64, 1e-12), (hc.HC_FP32_SHIFTED, 1e-5)])
def test_batched_steps_match_single_steps(precision, tol):
    print("[batch] precision", precision, flush=True)
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    pre = [0, 1, 2, 3, 0]  # steps taken alone before the batch: mixed AA parities
    alone = [serving.build(c, units, precision) for c in _cases()]
    batch = [serving.build(c, units, precision) for c in _cases()]
    for s, k in zip(alone + batch, pre + pre):
        if k:
            s.step(k)
    for s in alone:
        s.step(57)
    hc.step_batch(batch, 20)
    hc.step_batch(batch, 37)
    for a, b in zip(alone, batch):
        a.synchronize()
        b.synchronize()
        assert a.iteration == b.iteration
        xa, xb = a.vertices()[0], b.vertices()[0]
        assert tcases.rel_l2(xb, xa) <= tol
        assert tcases.rel_l2(b.populations(), a.populations()) <= tol
        assert np.all(np.isfinite(xb))
    for s in alone + batch:
        s.close()


def test_batch_with_a_case_it_cannot_batch():
    print("[batch] fallback", flush=True)
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    cs = _cases()[:3]
    alone = [serving.build(c, units, hc.HC_FP64) for c in cs]
    batch = [serving.build(c, units, hc.HC_FP64) for c in cs]
    for s in (alone[1], batch[1]):
        s.set_repulsion(1e-4, 1.0, 1)  # repulsion: stepped on its own inside the batch call
    for s in alone:
        s.step(15)
    hc.step_batch(batch, 15)
    for a, b in zip(alone, batch):
        assert tcases.rel_l2(b.vertices()[0], a.vertices()[0]) <= 1e-12
    for s in alone + batch:
        s.close()


def test_batch_arguments():
    print("[batch] arguments", flush=True)
    units = hc.lbm_base_parameters(5e-7, 1e-7)
    c = _cases()[0]
    a = serving.build(c, units, hc.HC_FP64)
    b = serving.build(c, units, hc.HC_FP32_SHIFTED)
    with pytest.raises(hc.HcError):
        hc.step_batch([a, a], 1)
    with pytest.raises(hc.HcError):
        hc.step_batch([a, b], 1)
    hc.step_batch([a], 0)
    assert a.iteration == 0
    a.close()
    b.close()


def test_empty_and_ragged_inputs():
    """A cell-free context and a context whose cell type has no cells step in a
    batch as alone; interior viscosity, the stages and the statistics accept an
    empty suspension; zero steps and an empty batch are no-ops."""
    print("[batch] empty", flush=True)
    u = tcases.units()
    shape = (16, 12, 10)
    f0 = tcases.random_populations(int(np.prod(shape)), np.random.default_rng(3))
    sims = []
    for _ in range(2):
        pair = []
        for k in range(2):
            g = hc.Simulation(0)
            g.lattice_init(*shape, u.tau)
            g.set_populations(f0)
            g.set_body_force([1e-6, 0, 0])
            if k == 1:  # a registered type without cells
                model, v, t, p, inner = tcases.mesh("PLT", u, hc.build_mesh)
                tid = g.add_cell_type(model, v, t, p, inner)
                g.set_interior_viscosity(tid, 2.0)
            pair.append(g)
        sims.append(pair)
    for g in sims[0]:
        g.step(9)
    hc.step_batch(sims[1], 9)
    hc.step_batch([], 5)
    for a, b in zip(sims[0], sims[1]):
        assert b.iteration == a.iteration == 9
        assert np.array_equal(b.populations(), a.populations())
    g = sims[1][1]
    g.stage("interior_viscosity", 0)
    assert np.all(g.interior_viscosity() == 0.0)
    assert g.cells()[0].shape == (0,)
    st = g.fluid_stats()
    assert st["ncells"] == int(np.prod(shape))
    for pair in sims:
        for s in pair:
            s.close()

