"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

Each fixture records the seed and the input it was made from; the tests that use
them state their tolerance.  Regenerate with `python tests/golden/make_golden.py`
(from the repository root, after building oracle/) only when the oracle changes on
purpose, and say why in the commit.

  collide_one.npz   one Guo-BGK node update for 3 seeded states (tau 1.82)
  periodic8.npz     8x8x8 periodic lattice, random populations, body force, 10 steps
  rbc_tables.npz    RBC rest-mesh tables (edges, rest lengths/areas/angles, V0, A0, e0)
  rbc_forces.npz    RBC mechanics forces for a seeded +-5% perturbation of the rest mesh
  ibm_weights.npz   phi2 support and weights of 16 seeded positions, some next to a wall
  rbc_shear32.npz   one RBC in a 32^3 Couette box at 1000 1/s (velocity walls z = 0, 31), 100 steps
  c3_crop_64x64x128.npz  an input, not an oracle output: a 64x64x128 periodic crop of the
                    Ht 30% suspension packed by hemocell_amd.packing (random positions and
                    orientations): 242 RBC, 24 PLT and one 2562-vertex WBC
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]

import cases  # noqa: E402
import mesh_oracle  # noqa: E402
from oracle import OracleSim, collide_one  # noqa: E402

SEED = 20261202


def collide():
    rng = np.random.default_rng(SEED)
    f = 1e-3 * rng.uniform(-1, 1, size=(3, 19))
    F = 1e-5 * rng.uniform(-1, 1, size=(3, 3))
    out = np.stack([collide_one(f[k], F[k], 1.0 / 1.82) for k in range(3)])
    np.savez_compressed(os.path.join(HERE, "collide_one.npz"), f=f, F=F, omega=1.0 / 1.82, out=out, seed=SEED)


def periodic8():
    rng = np.random.default_rng(SEED + 1)
    f0 = cases.random_populations(512, rng, 1e-3)
    g = np.array([1e-6, -2e-6, 3e-6])
    o = OracleSim(8, 8, 8, 1.82)
    o.set_populations(f0)
    o.set_body_force(g)
    o.step(10)
    np.savez_compressed(os.path.join(HERE, "periodic8.npz"), f0=f0, g=g, tau=1.82, steps=10, f=o.populations(),
                        seed=SEED + 1)


def rbc():
    u = cases.units()
    model, v, t, p, inner = cases.mesh("RBC", u, mesh_oracle.build_mesh)
    o = OracleSim(40, 40, 40, u.tau)
    tid = o.add_cell_type(model, v, t, p, inner)
    tab = o.type_tables(tid)
    np.savez_compressed(os.path.join(HERE, "rbc_tables.npz"), vertices=v, triangles=t,
                        **{k: tab[k] for k in ("edge", "edge_tris", "edge_outer", "ring", "nring", "edge_len",
                                               "edge_angle", "tri_area", "patch_dist", "volume_eq", "area_mean_eq",
                                               "edge_mean_eq", "angle_mean_eq")})
    rng = np.random.default_rng(SEED + 2)
    x = cases.perturb(cases.place(v, (20.2, 19.7, 20.4), cases.rotation(rng)), rng, 0.05)
    o.add_cells(tid, x[None])
    o.stage("mechanics", 1)
    _, _, f = o.vertices()
    np.savez_compressed(os.path.join(HERE, "rbc_forces.npz"), x=x, f=f, k=np.array(
        [p.k_link, p.k_bend, p.k_volume, p.k_area, p.eta_m]), seed=SEED + 2)


def ibm():
    nx, ny, nz = 12, 10, 14
    mask = cases.channel_mask(nx, ny, nz, axis=1)
    o = OracleSim(nx, ny, nz, 1.0, periodic=(1, 0, 1), node_type=mask)
    rng = np.random.default_rng(SEED + 3)
    pos = rng.uniform([0, 0.6, 0], [nx, ny - 1.6, nz], size=(16, 3))
    pos[:4, 1] = [0.7, 1.1, ny - 2.2, ny - 1.8]  # next to the walls y = 0 and y = ny-1
    idx = np.full((16, 27), -1, dtype=np.int64)
    w = np.zeros((16, 27))
    for k, p in enumerate(pos):
        i, ww = o.ibm_kernel(p)
        idx[k, :len(i)] = i
        w[k, :len(ww)] = ww
    np.savez_compressed(os.path.join(HERE, "ibm_weights.npz"), shape=(nx, ny, nz), mask=mask, pos=pos, idx=idx, w=w,
                        seed=SEED + 3)


def shear32():
    u = cases.units()
    n = 32
    mask = cases.couette_mask(n, n, n)
    o = OracleSim(n, n, n, u.tau, periodic=(1, 1, 0), node_type=mask)
    uw = (n - 1) * 1000.0 * u.dt * 0.5  # shear rate 1000 1/s (helper/hemocellInit.hh:75-77)
    o.set_wall_velocity(4, (uw, 0, 0))
    o.set_wall_velocity(5, (-uw, 0, 0))
    o.init_equilibrium(1.0, (0, 0, 0))
    o.set_force_limit(u.f_limit)
    model, v, t, p, inner = cases.mesh("RBC", u, mesh_oracle.build_mesh)
    tid = o.add_cell_type(model, v, t, p, inner)
    x0 = cases.place(v, (15.7, 16.2, 15.4))
    o.add_cells(tid, x0[None])
    o.step(100)
    x, vel, _ = o.vertices()
    _, uf = o.fluid()
    uslice = uf.reshape(n, n, n, 3)[:, 16, :, :]
    np.savez_compressed(os.path.join(HERE, "rbc_shear32.npz"), mask=mask, wall_u=uw, x0=x0, x=x, v=vel,
                        u_slice_y16=uslice, tau=u.tau, steps=100)


def c3_crop():
    import hemocell_amd as hc
    from hemocell_amd import packing
    from hemocell_amd import suspension as susp
    types = susp.cell_types(hc.lbm_base_parameters(5e-7, 1e-7))
    shape = (64, 64, 128)
    counts = susp.cell_counts(types, float(np.prod(shape)), 0.30)
    counts["WBC"] = 1  # 1:500 rounds to none at this size; the crop keeps one
    pk = packing.pack(shape, types, counts, seed=SEED + 5)
    assert packing.overlap_free(types, pk, shape, 0.49)
    packing.save(os.path.join(HERE, "c3_crop_64x64x128.npz"), pk, shape=shape, hematocrit=0.30, seed=SEED + 5,
                 counts=[counts["RBC"], counts["PLT"], counts["WBC"]], walls=-1, tube=0)


if __name__ == "__main__":
    if sys.argv[1:] == ["c3_crop"]:
        c3_crop()
        sys.exit(0)
    collide()
    periodic8()
    rbc()
    ibm()
    shear32()
    c3_crop()
    print("fixtures written to", HERE)
