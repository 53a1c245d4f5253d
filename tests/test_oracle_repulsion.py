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
