"""Pins of the oracle's interior viscosity (CPU).

The reference's INTERIOR_VISCOSITY build marks the fluid nodes inside a membrane
and relaxes them with tau_inner (core/hemoCellField.cpp:99-110):

* the entire-grid pass (core/hemoCellParticleField.cpp:776-808) counts the
  crossings of a ray from each node of a cell's bounding box along (-1,-1,-1)
  with the cell's triangles (helper/octree.h:107-150, helper/mollerTrumbore.h);
  checked here against an independent point-in-polyhedron rule, the winding
  number from the triangles' solid angles (Van Oosterom and Strackee), and
  against the closed form of a sphere;
* the membrane pass (:745-774) re-classifies the support nodes of each vertex
  by the sign of (node - x) . normal; checked against the radial direction of a
  sphere, with the untouched nodes keeping their class;
* the collision of a node with class k uses 1/tau_inner: a lattice whose nodes
  are all inner steps exactly as a plain lattice with tau = tau_inner.
"""
import numpy as np
import pytest

import cases
import mesh_oracle
from oracle import OracleSim, ray_crosses


def test_ray_crossing_known_answers():
    tri = ([-5.0, -5.0, 0.0], [5.0, -5.0, 0.0], [0.0, 5.0, 0.0])
    # the ray from (0,0,3) meets z = 0 at (-3,-3,0), inside the triangle
    assert ray_crosses(*tri, [0, 0, 3]) == 1
    # from below the plane the crossing is behind the origin
    assert ray_crosses(*tri, [0, 0, -3]) == 0
    # from (0,0,9) the ray meets z = 0 at (-9,-9,0), outside
    assert ray_crosses(*tri, [0, 0, 9]) == 0
    # reversed orientation crosses as well (the count ignores the side)
    assert ray_crosses(tri[0], tri[2], tri[1], [0, 0, 3]) == 1
    # a triangle parallel to the ray (its plane holds the direction (1,1,1)): det = 0
    assert ray_crosses([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0, -1.0, 0.0], [3, 2, 2]) == 0
    # degenerate triangle
    assert ray_crosses([1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [4, 4, 4]) == 0


def _winding(points, x, tris):
    """Sum of the solid angles of the triangles seen from each point / 4 pi."""
    a = x[tris[:, 0]][None] - points[:, None]
    b = x[tris[:, 1]][None] - points[:, None]
    c = x[tris[:, 2]][None] - points[:, None]
    la, lb, lc = (np.linalg.norm(q, axis=2) for q in (a, b, c))
    num = np.einsum("pti,pti->pt", a, np.cross(b, c))
    den = la * lb * lc + np.einsum("pti,pti->pt", a, b) * lc + np.einsum("pti,pti->pt", b, c) * la + \
        np.einsum("pti,pti->pt", c, a) * lb
    return np.abs((2.0 * np.arctan2(num, den)).sum(axis=1) / (4.0 * np.pi))


def _surface_distance(points, x, tris):
    """Lower bound of the distance to the surface: to the nearest triangle's plane
    within its vertices' reach, here bounded by the nearest vertex or plane."""
    d = np.full(len(points), np.inf)
    for t in tris:
        p0, p1, p2 = x[t]
        n = np.cross(p1 - p0, p2 - p0)
        n /= np.linalg.norm(n)
        dp = np.abs((points - p0) @ n)
        r = max(np.linalg.norm(p1 - p0), np.linalg.norm(p2 - p0))
        near = np.linalg.norm(points - p0, axis=1) <= r + 1.0
        d = np.where(near, np.minimum(d, dp), d)
    return d


def _sphere_sim(kind, center, shape=(28, 26, 27), periodic=(1, 1, 1), seed=5, mask=None):
    u = cases.units()
    o = OracleSim(*shape, u.tau, periodic=periodic, node_type=mask)
    model, v, t, p, inner = cases.mesh(kind, u, mesh_oracle.build_mesh)
    tid = o.add_cell_type(model, v, t, p, inner)
    rng = np.random.default_rng(seed)
    x = cases.place(v, center, cases.rotation(rng))
    o.add_cells(tid, x[None])
    return o, tid, x, np.asarray(t), u


@pytest.mark.parametrize("kind,center", [("RBC", (12.37, 11.81, 12.53)), ("WBC", (13.6, 12.2, 12.9)),
                                         ("RBC", (0.73, 25.4, 13.1))])
def test_entire_grid_pass_matches_winding_number(kind, center):
    o, tid, x, tris, u = _sphere_sim(kind, center)
    o.set_interior_viscosity(tid, 2.5)
    o.stage_interior_viscosity(0)
    inner = o.interior_viscosity() != 0.0
    assert set(np.unique(o.interior_viscosity())) <= {0.0, 2.5}
    # every node of the unwrapped bounding box, classified independently
    lo, hi = np.floor(x.min(axis=0)) - 1, np.ceil(x.max(axis=0)) + 1
    g = np.stack(np.meshgrid(*[np.arange(lo[a], hi[a] + 1) for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
    w = _winding(g, x, tris)
    clear = _surface_distance(g, x, tris) > 1e-6
    shape = np.array(o.shape)
    wr = np.mod(g.astype(np.int64), shape)
    got = inner[wr[:, 0], wr[:, 1], wr[:, 2]]
    want = w > 0.5
    assert np.array_equal(got[clear], want[clear])
    # nothing outside the bounding box is marked
    assert inner.sum() == got.sum()
    # the count approaches the enclosed volume in lattice units
    vol = abs(np.einsum("ti,ti->t", x[tris[:, 0]], np.cross(x[tris[:, 1]], x[tris[:, 2]])).sum() / 6.0)
    assert abs(inner.sum() - vol) < 0.2 * vol


def test_entire_grid_pass_leaves_walls_and_clamps_to_the_domain():
    # bounce-back walls on y = 0 and y = ny-1, the cell pressed against the lower one
    shape = (26, 24, 26)
    mask = cases.channel_mask(*shape, axis=1)
    o, tid, x, tris, u = _sphere_sim("RBC", (12.4, 4.6, 12.7), shape=shape, periodic=(1, 0, 1), mask=mask)
    o.set_interior_viscosity(tid, 3.0)
    o.stage_interior_viscosity(0)
    tau = o.interior_viscosity()
    m = mask.reshape(shape)
    assert np.all(tau[m != 0] == 0.0)
    g = np.argwhere(m == 0).astype(np.float64)
    w = _winding(g, x, tris)
    clear = _surface_distance(g, x, tris) > 1e-6
    got = tau[m == 0] != 0.0
    assert np.array_equal(got[clear], (w > 0.5)[clear])


@pytest.mark.parametrize("model", [0, 1])
def test_membrane_pass_follows_the_outward_normal(model):
    # a sphere (the WBC icosphere, radius 8) registered with the RBC model (0),
    # whose mechanics provide the normal, or with the PLT model (1), whose do not
    center = np.array([12.37, 11.81, 12.53])
    u = cases.units()
    o = OracleSim(28, 26, 27, u.tau)
    _, v, t, _, _ = cases.mesh("WBC", u, mesh_oracle.build_mesh)
    t = np.asarray(t)
    p = cases.params(cases.RBC_MAT if model == 0 else cases.PLT_MAT, len(t), u, model)
    tid = o.add_cell_type(model, v, t, p, None)
    x = cases.place(v, center, cases.rotation(np.random.default_rng(5)))
    o.add_cells(tid, x[None])
    o.set_interior_viscosity(tid, 2.5)
    # every node starts inner: the pass must rewrite exactly the nodes it visits
    o.set_interior_viscosity_classes(np.full(o.N, tid + 1, dtype=np.uint8))
    o.stage_interior_viscosity(1)
    flat = o.interior_viscosity().reshape(-1)
    shape = np.array(o.shape)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).reshape(-1, 3)
    d = idx - center
    d -= shape * np.round(d / shape)
    r = np.linalg.norm(d, axis=1)
    R = np.linalg.norm(x - center, axis=1)
    # a vertex's support nodes lie within sqrt(3) of it: farther nodes keep their class
    far = (r < R.min() - 2.0) | (r > R.max() + 2.0)
    assert np.all(flat[far] == 2.5)
    changed = flat == 0.0
    # nodes farther than the mean rest edge from every vertex are skipped
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    e = np.unique(e, axis=0)
    edge_mean = np.linalg.norm(v[e[:, 0]] - v[e[:, 1]], axis=1).mean()
    beyond = r > R.max() + edge_mean
    assert np.all(flat[beyond] == 2.5) and np.any(beyond & (r < R.max() + 1.7))
    if model == 0:
        # visited nodes outside the sphere become background, those inside stay
        # inner (margin: the normal is radial only up to the mesh's facets)
        assert not np.any(changed & (r < R.min() - 0.25))
        shell = (r > R.max() + 0.25) & (r < R.max() + 0.75)
        assert shell.sum() > 50 and np.all(changed[shell])
    else:
        # a zero normal: every visited node becomes background, inside or out
        assert np.any(changed & (r < R.min() - 0.25)) and np.any(changed & (r > R.max() + 0.25))


def test_inner_nodes_relax_with_tau_inner():
    shape = (10, 9, 8)
    u = cases.units()
    rng = np.random.default_rng(2)
    f0 = cases.random_populations(int(np.prod(shape)), rng)
    tau_in = 1.37
    runs = []
    for tau, iv in ((u.tau, True), (tau_in, False)):
        o = OracleSim(*shape, tau)
        model, v, t, p, inner = cases.mesh("PLT", u, mesh_oracle.build_mesh)
        tid = o.add_cell_type(model, v, t, p, inner)
        o.set_body_force([1e-5, -2e-5, 3e-5])
        o.set_populations(f0)
        if iv:
            o.set_interior_viscosity(tid, tau_in)
            o.set_interior_viscosity_timescales(1000, 1000)
            o.set_iteration(1)  # no pass runs before iteration 1000
            o.set_interior_viscosity_classes(np.full(o.N, tid + 1, dtype=np.uint8))
        o.step(4)
        runs.append(o.populations())
    assert np.array_equal(runs[0], runs[1])
