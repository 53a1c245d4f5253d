"""Pins of the oracle's vertex binning and cell-cell repulsion (CPU).

update_pg (core/hemoCellParticleField.cpp:137-168) puts every vertex in the bin
of its nearest node; applyRepulsionForce (:677-743) pairs each node with itself
and 13 forward neighbours.  The oracle's half-stencil walk is checked against an
all-pairs sum in which each pair's multiplicity follows from the bin relation
alone (same bin: both orders, neighbouring bins: once), so a missing or repeated
stencil offset, a sign error or a wrong bin rule fails it.
"""
import numpy as np
import pytest

import cases
import mesh_oracle
from oracle import OracleSim


def _crowded(shape, periodic, seed, n_plt=24, n_rbc=0):
    rng = np.random.default_rng(seed)
    u = cases.units()
    o = OracleSim(*shape, u.tau, periodic=periodic)
    cells = []
    for kind, n in (("PLT", n_plt), ("RBC", n_rbc)):
        if n == 0:
            continue
        model, v, t, p, inner, xs = cases.crowd(kind, n, shape, rng, u, mesh_oracle.build_mesh)
        tid = o.add_cell_type(model, v, t, p, inner)
        o.add_cells(tid, xs)
        cells.append(xs.reshape(-1, 3))
    x = np.concatenate(cells)
    return o, x


def _cell_index(o_counts):
    return np.concatenate([np.full(nv, c) for c, nv in enumerate(o_counts)])


@pytest.mark.parametrize("periodic", [(1, 1, 1), (1, 0, 1)])
def test_bins_are_nearest_nodes(periodic):
    shape = (12, 11, 13)
    o, x = _crowded(shape, periodic, seed=3)
    b = o.vertex_bins()
    n = np.asarray(shape)
    inside = b >= 0
    node = np.stack([b // (n[1] * n[2]), (b // n[2]) % n[1], b % n[2]], axis=1)
    d = x - node
    for a in range(3):
        if periodic[a]:
            d[:, a] -= n[a] * np.round(d[:, a] / n[a])
    # the bin node is within half a spacing of the vertex on every axis
    assert np.all(np.abs(d[inside]) <= 0.5 + 1e-12)
    # a vertex has no bin only when it lies beyond half a spacing outside a non-periodic face
    np_axes = [a for a in range(3) if not periodic[a]]
    beyond = np.zeros(len(x), dtype=bool)
    for a in np_axes:
        beyond |= (x[:, a] < -0.5) | (x[:, a] >= n[a] - 0.5)
    assert np.array_equal(~inside, beyond)
    if periodic == (1, 1, 1):
        assert inside.all()


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 1)])
@pytest.mark.parametrize("r_c", [0.45, 0.9])
def test_repulsion_matches_all_pairs(periodic, r_c):
    shape = (10, 9, 11)
    o, x = _crowded(shape, periodic, seed=4, n_plt=30, n_rbc=1)
    counts = [66] * 30 + [642]
    cell = _cell_index(counts)
    k = 2.5e-3
    o.set_repulsion(k, r_c, 1)
    o.stage("repulsion")
    F = o.repulsion_force()
    Fb = cases.repulsion_brute_force(x, cell, shape, periodic, k, r_c)
    assert np.abs(Fb).max() > 0
    assert cases.rel_l2(F, Fb) < 1e-13
    # pairwise action and reaction: the internal forces sum to zero
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).sum()


def test_repulsion_is_zero_within_one_cell():
    o, x = _crowded((10, 10, 10), (1, 1, 1), seed=5, n_plt=0, n_rbc=1)
    o.set_repulsion(1.0, 0.9, 1)
    o.stage("repulsion")
    assert np.all(o.repulsion_force() == 0.0)


# ------------------------------------------------------------ wall repulsion --
def _wall_sim(shape, periodic, mask, kind="PLT"):
    u = cases.units()
    o = OracleSim(*shape, u.tau, periodic=periodic, node_type=mask)
    model, v, t, p, inner = cases.mesh(kind, u, mesh_oracle.build_mesh)
    tid = o.add_cell_type(model, v, t, p, inner)
    return o, tid, v


def test_wall_repulsion_of_one_vertex_over_a_flat_wall():
    """A vertex 0.9 above the wall plane y = 0, straight over a node: only that
    node is within r_c = 1.2 (the next ones are sqrt(1 + 0.81) away), so the force
    is k (r_c / 0.9) along +y (hemoCellParticleField.cpp:908-911)."""
    shape = (10, 8, 10)
    mask = cases.channel_mask(*shape, axis=1)
    o, tid, v = _wall_sim(shape, (1, 0, 1), mask)
    x = np.tile([[4.0, 0.9, 6.0]], (len(v), 1))
    x[1:] = [5.0, 4.0, 5.0]  # the other vertices far from the walls
    o.add_cells(tid, x[None])
    k, rc = 2e-3, 1.2
    o.set_boundary_repulsion(k, rc)
    o.stage("boundary_repulsion")
    fb = o.boundary_repulsion_force()
    assert np.allclose(fb[0], [0.0, k * (rc / 0.9), 0.0], rtol=1e-15, atol=0)
    assert np.all(fb[1:] == 0.0)


@pytest.mark.parametrize("periodic,axis", [((1, 0, 1), 1), ((0, 1, 1), 0)])
def test_wall_repulsion_matches_all_pairs(periodic, axis):
    """Against every (vertex, boundary particle) pair: boundary particles are the
    bounce-back nodes with a non-bounce-back node among their 26 neighbours
    (hemoCellParticleField.cpp:865-887); a vertex feels those in the 3x3x3
    neighbourhood of its nearest node that are closer than r_c, with the
    separation taken as the minimum image on periodic axes."""
    rng = np.random.default_rng(41)
    shape = (12, 10, 11)
    mask = cases.channel_mask(*shape, axis=axis)
    o, tid, v = _wall_sim(shape, periodic, mask)
    # vertices scattered through the whole box, many next to the walls
    x = rng.uniform(0.0, 1.0, size=(len(v), 3)) * (np.array(shape) - 1)
    x[:, axis] = rng.uniform(0.3, 2.5, len(v))
    x[::2, axis] = shape[axis] - 1 - x[::2, axis]
    o.add_cells(tid, x[None])
    k, rc = 3e-3, 1.4
    o.set_boundary_repulsion(k, rc)
    o.stage("boundary_repulsion")
    fb = o.boundary_repulsion_force()
    m = mask.reshape(shape)
    dims = np.array(shape)
    per = np.array(periodic, dtype=bool)
    bp = []
    for node in np.argwhere(m == 1):
        found = False
        for d in np.ndindex(3, 3, 3):
            q = node + np.array(d) - 1
            if np.any((q < 0) | (q >= dims)) and np.any(((q < 0) | (q >= dims)) & ~per):
                continue
            q = np.where(per, q % dims, q)
            if m[tuple(q)] != 1:
                found = True
                break
        if found:
            bp.append(node)
    bp = np.array(bp)
    ref = np.zeros_like(x)
    for i, p in enumerate(x):
        c = np.floor(p + 0.5).astype(int)
        for b in bp:
            d = b - c
            d = np.where(per, (d + dims // 2) % dims - dims // 2, d)  # nearest periodic image of b
            if np.abs(d).max() > 1:
                continue
            r = p - (c + d)
            dist = np.sqrt((r * r).sum())
            if dist < rc:
                ref[i] += k * (1 / (dist / rc)) * (r / dist)
    assert np.abs(ref).max() > 0
    assert np.allclose(fb, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
