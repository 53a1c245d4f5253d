"""The FP32 RBC mechanics over its three table forms, on the GPU.

The FP32-shifted RBC kernel reads the ring-ordered tables by default (per
vertex its 16-bit 1-ring, the rest areas of triangles (k, ring[j], ring[j+1])
and rest lengths of edges (k, ring[j]), summed around the ring), the per-slot
compact tables with HC_MECH_TABLES=compact (one record per triangle/edge slot,
summed in the reference's triangle, contributor and edge order), and the
variable-length lists with HC_MECH_TABLES=csr.  Case: one RBC, perturbed by
0.05 spacings and rotated, in a 24^3 sheared box (the smoke case), one step.

* compact and csr compute the same terms in the same order (the compiler fuses
  other multiply-adds where an edge is taken from its other end): forces
  within FP32 rounding (relative L2 <= 1e-6);
* ring-ordered: the same forces up to FP32 rounding (relative L2 <= 1e-6);
* all three against the FP64 oracle of the same step (relative L2 <= 1e-5).
"""
import numpy as np
import pytest

import cases
import hemocell_amd as hc
from oracle import OracleSim

pytestmark = pytest.mark.gpu

N = 24


def _setup(s, u, x, model, v, t, p, inner):
    mask = cases.couette_mask(N, N, N)
    vhalf = (N - 1) * 1000.0 * u.dt * 0.5
    s.set_wall_velocity(5, [-vhalf, 0, 0])
    s.set_wall_velocity(4, [vhalf, 0, 0])
    s.init_equilibrium(1.0, [0, 0, 0])
    s.set_force_limit(u.f_limit)
    tid = s.add_cell_type(model, v, t, p, inner)
    s.add_cells(tid, x[None])
    return mask


def test_ring_compact_and_list_tables_agree(monkeypatch):
    u = cases.units()
    rng = np.random.default_rng(3)
    model, v, t, p, inner = cases.mesh("RBC", u, hc.build_mesh)
    x = cases.perturb(cases.place(v, (12.2, 11.9, 12.1), cases.rotation(rng)), rng, 0.05)
    mask = cases.couette_mask(N, N, N)
    forces = {}
    for mode in ("", "compact", "csr"):
        print("[mech tables] fp32 mode %r" % (mode or "ring"), flush=True)
        monkeypatch.setenv("HC_MECH_TABLES", mode)
        g = hc.Simulation(0)
        g.lattice_init(N, N, N, u.tau, periodic=(1, 1, 0), node_type=mask, precision=hc.HC_FP32_SHIFTED)
        _setup(g, u, x, model, v, t, p, inner)
        g.step(1)
        forces[mode or "ring"] = g.vertices()[2]
        g.close()
    o = OracleSim(N, N, N, u.tau, periodic=(1, 1, 0), node_type=mask)
    _setup(o, u, x, model, v, t, p, inner)
    o.step(1)
    fo = o.vertices()[2]
    assert np.abs(fo).max() > 0
    e_cc = cases.rel_l2(forces["compact"], forces["csr"])
    e_ring = cases.rel_l2(forces["ring"], forces["compact"])
    errs = {k: cases.rel_l2(f, fo) for k, f in forces.items()}
    print("[mech tables] compact vs csr %.2e, ring vs compact %.2e; vs the FP64 oracle %s" % (e_cc, e_ring, errs))
    assert e_cc <= 1e-6
    assert 0 < e_ring <= 1e-6
    assert all(e <= 1e-5 for e in errs.values()), errs
