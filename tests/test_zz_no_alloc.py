"""hc_get_cell_stats_noalloc through the C ABI: the statistics it stages in the
library's static device array equal, bit for bit, those of hc_get_cell_stats
(which allocates a staging buffer per call), on two calls in a row on two
contexts, the second with fewer cells (so the array keeps the first call's
larger contents past the second call's cells); its refusals carry their reason."""
import ctypes

import numpy as np
import pytest

import cases
import hemocell_amd as hc

pytestmark = pytest.mark.gpu


def _suspension(n_rbc, n_plt, seed):
    u = cases.units()
    sim = hc.Simulation(0)
    sim.lattice_init(48, 40, 44, u.tau, periodic=(1, 1, 1), precision=hc.HC_FP64)
    sim.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    sim.set_body_force((0.0, 0.0, 2e-6))
    sim.set_force_limit(u.f_limit)
    rng = np.random.default_rng(seed)
    for kind, n, centres in (("RBC", n_rbc, [(12, 10, 11), (36, 10, 11), (12, 30, 33), (36, 30, 33)]),
                             ("PLT", n_plt, [(24, 20, 22), (4, 4, 40), (44, 36, 4)])):
        model, v, t, p, inner = cases.mesh(kind, u, hc.build_mesh)
        tid = sim.add_cell_type(model, v, t, p, inner)
        x = np.stack([cases.perturb(cases.place(v, centres[k], cases.rotation(rng)), rng, 0.02) for k in range(n)])
        sim.add_cells(tid, x)
    sim.step(6)
    sim.synchronize()
    return sim


def _packed(st):
    return np.concatenate([st["volume"][:, None], st["area"][:, None], st["centroid"], st["velocity"], st["bbox"],
                           st["stretch"][:, None]], axis=1)


def test_noalloc_cell_stats_equal_the_allocating_entry_point():
    lib = ctypes.CDLL(hc.LIB_PATH)
    lib.hc_get_cell_stats_noalloc.restype = ctypes.c_int
    lib.hc_get_cell_stats_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64]
    for label, (n_rbc, n_plt, seed) in (("first call, 7 cells", (4, 3, 1)), ("second call, 3 cells", (2, 1, 2))):
        print("[no-alloc]", label, flush=True)
        sim = _suspension(n_rbc, n_plt, seed)
        nc = n_rbc + n_plt
        want = _packed(sim.cell_stats())
        got = np.full((nc, hc.CELL_STATS_QUANTITIES), np.nan)
        st = lib.hc_get_cell_stats_noalloc(sim.h, got.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           nc)
        assert st == hc.HC_OK, sim.lib.hc_last_error(sim.h)
        assert np.array_equal(got, want), (label, np.abs(got - want).max())
        assert np.all(got[:, 0] > 0) and np.all(got[:, 14] > 0)
        # the same through the Python method, after more steps
        sim.step(3)
        assert np.array_equal(sim.cell_stats_noalloc(), _packed(sim.cell_stats()))
        # an output that holds fewer cells than the context: refused with its reason
        small = np.zeros((nc - 1, hc.CELL_STATS_QUANTITIES))
        st = lib.hc_get_cell_stats_noalloc(sim.h, small.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           nc - 1)
        assert st == hc.HC_INVALID_ARG and b"fewer cells" in sim.lib.hc_last_error(sim.h)
        sim.close()


def test_noalloc_cell_stats_refuses_cells_without_device_tables():
    u = cases.units()
    sim = hc.Simulation(0)
    sim.lattice_init(32, 32, 32, u.tau, periodic=(1, 1, 1), precision=hc.HC_FP64)
    model, v, t, p, inner = cases.mesh("RBC", u, hc.build_mesh)
    tid = sim.add_cell_type(model, v, t, p, inner)
    sim.add_cells(tid, cases.place(v, (16, 16, 16))[None])
    with pytest.raises(hc.HcError, match="call hc_step or hc_get_cell_stats first"):
        sim.cell_stats_noalloc()
    sim.cell_stats()  # builds the cell tables
    out = sim.cell_stats_noalloc()
    assert out.shape == (1, 15) and out[0, 0] > 0
    sim.close()
