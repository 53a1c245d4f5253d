"""Seeded random packings of cells: uniform random positions and orientations,
no two cells overlapping.

The reference writes such inputs with tools/packCells (packCells.cpp:99-274,
an iterative ellipsoid packer writing RBC.pos / PLT.pos).  This module packs
by force-biased growth instead: every cell is a rigid union of spheres that
covers its membrane mesh (an RBC a ring of 16 spheres round the rim, 8 inside
it and one at the centre; a platelet or a WBC its bounding sphere).  The cells
start at seeded uniform random positions and orientations, shrunk by a common
factor; each iteration pushes the spheres of every overlapping pair of cells
apart (translating both cells and turning them by the torque of the push),
and once no pair overlaps the factor grows, until the cells have their full
size with a clearance `gap` between any two of them and from the walls.

Lattice units throughout.  Result: {name: (centres [n, 3], rotations [n, 3, 3])};
instances(types, packing) gives the vertex positions {name: [n, nv, 3]}.
"""
import math

import numpy as np
from scipy.spatial import cKDTree


def _quat_to_matrix(q):
    a, b, c, d = q.T
    return np.stack([np.stack([a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)], -1),
                     np.stack([2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)], -1),
                     np.stack([2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d], -1)], 1)


def _quat_mul(p, q):
    a1, b1, c1, d1 = p.T
    a2, b2, c2, d2 = q.T
    return np.stack([a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2, a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
                     a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2, a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2], -1)


def sphere_cover(verts, kind, margin=0.15):
    """Local sphere centres [k, 3] and radii [k] whose union covers the cell's
    vertices (centred on their mean) with `margin` to spare.  An RBC (disc axis
    z, as cell_types builds it) gets three rings on its mid-plane; anything
    else its bounding sphere."""
    v = verts - verts.mean(axis=0)
    if kind != "RBC":
        return np.zeros((1, 3)), np.array([np.linalg.norm(v, axis=1).max() + margin])
    rho = np.hypot(v[:, 0], v[:, 1])
    rmax = rho.max()
    rings = [(0.0, 1), (0.40 * rmax, 8), (0.71 * rmax, 16)]
    cen, rad = [], []
    for r0, n in rings:
        # the radius that reaches every vertex of the band this ring is nearest to
        for k in range(n):
            a = 2 * math.pi * (k + 0.5 * (n == 8)) / n
            cen.append((r0 * math.cos(a), r0 * math.sin(a), 0.0))
    cen = np.array(cen)
    d = np.linalg.norm(v[:, None, :] - cen[None, :, :], axis=2)
    owner = np.argmin(d, axis=1)
    rad = np.zeros(len(cen))
    np.maximum.at(rad, owner, d[np.arange(len(v)), owner])
    # a ring shares its largest radius; every vertex then lies inside
    start = 0
    for _, n in rings:
        rad[start:start + n] = rad[start:start + n].max()
        start += n
    rad += margin
    assert np.all(np.min(np.linalg.norm(v[:, None, :] - cen[None], axis=2) - rad[None], axis=1) <= -margin + 1e-9)
    return cen, rad


def _random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def pack(shape, types, counts, seed=20261202, gap=0.5, walls=None, tube=None, max_iter=20000, log=None, tol=1e-3):
    """Random positions and orientations of counts[name] cells of each type in a
    box of `shape` nodes.  walls: an axis (0, 1 or 2) bounded by wall planes at
    0 and n-1; tube: (axis, centre, radius) of a cylinder the cells stay inside;
    other axes periodic.  Cells keep `gap` (less `tol`) from each other and from
    the walls.  Returns {name: (centres, rotations)}."""
    rng = np.random.default_rng(seed)
    shape = np.asarray(shape, dtype=np.float64)
    names = [n for n in ("RBC", "PLT", "WBC") if counts.get(n, 0) > 0]
    body_type, loc, rad, sb = [], [], [], []
    nb = 0
    for ti, name in enumerate(names):
        cen, r = sphere_cover(types[name].verts, name)
        for _ in range(counts[name]):
            body_type.append(ti)
            loc.append(cen)
            rad.append(r)
            sb.append(np.full(len(r), nb))
            nb += 1
    body_type = np.array(body_type)
    loc = np.concatenate(loc)
    rad = np.concatenate(rad)
    sb = np.concatenate(sb)
    big = rad > 4.0  # the WBCs' spheres: paired with everything by their own query
    small = ~big
    # moment of inertia per body (about its centre, sphere centres only), for turning
    inertia = np.maximum(np.bincount(sb, weights=np.sum(loc * loc, axis=1) + rad * rad, minlength=nb), 1.0)
    periodic = np.ones(3, dtype=bool)
    lo = np.zeros(3)
    hi = shape.copy()
    if walls is not None:
        periodic[walls] = False
        lo[walls], hi[walls] = gap, shape[walls] - 1.0 - gap
    if tube is not None:
        ax, tc, tr = tube
        others = [a for a in range(3) if a != ax]
        periodic[others] = False
    # start positions inside the allowed region, random orientations
    c = rng.uniform(0, 1, size=(nb, 3)) * shape
    if walls is not None:
        c[:, walls] = rng.uniform(lo[walls] + 8.0, hi[walls] - 8.0, size=nb)
    if tube is not None:
        ax, tc, tr = tube
        r = (tr - 8.0) * np.sqrt(rng.uniform(0, 1, nb))
        a = rng.uniform(0, 2 * np.pi, nb)
        o0, o1 = [x for x in range(3) if x != ax]
        c[:, o0] = tc[0] + r * np.cos(a)
        c[:, o1] = tc[1] + r * np.sin(a)
    q = _random_quats(rng, nb)
    box = np.where(periodic, shape, 8.0 * shape.max())
    lam = 0.5
    it = 0
    rsmall = rad[small].max() if small.any() else 0.0
    while it < max_iter:
        it += 1
        R = _quat_to_matrix(q)
        arm = lam * np.einsum("mij,mj->mi", R[sb], loc)
        S = c[sb] + arm
        r = lam * rad
        Sw = np.where(periodic, np.mod(S, shape), S + 4.0 * shape.max())
        tree = cKDTree(Sw[small], boxsize=box)
        idx_small = np.flatnonzero(small)
        pr = tree.query_pairs(2 * lam * rsmall + gap, output_type="ndarray")
        i, j = idx_small[pr[:, 0]], idx_small[pr[:, 1]]
        if big.any():
            ib = np.flatnonzero(big)
            hits = tree.query_ball_point(Sw[ib], lam * (rad[ib] + rsmall) + gap)
            bi = np.concatenate([np.full(len(h), b) for b, h in zip(ib, hits)]) if len(ib) else np.zeros(0, int)
            bj = idx_small[np.concatenate(hits).astype(int)] if len(ib) else np.zeros(0, int)
            gi, gj = np.meshgrid(ib, ib, indexing="ij")
            m = gi < gj
            i = np.concatenate([i, bi, gi[m]])
            j = np.concatenate([j, bj, gj[m]])
        keep = sb[i] != sb[j]
        i, j = i[keep], j[keep]
        d = S[j] - S[i]
        d -= np.where(periodic, np.round(d / shape) * shape, 0.0)
        dist = np.linalg.norm(d, axis=1)
        over = r[i] + r[j] + gap - dist
        hit = over > tol
        i, j, d, dist, over = i[hit], j[hit], d[hit], dist[hit], over[hit]
        n = d / np.maximum(dist, 1e-12)[:, None]
        push = 0.5 * (over + 2 * tol)[:, None] * n  # moves sphere j's cell by +push, sphere i's by -push
        F = np.zeros((len(S), 3))
        np.add.at(F, i, -push)
        np.add.at(F, j, push)
        # walls and tube: push the sphere back inside
        wall_over = 0.0
        if walls is not None:
            a = walls
            low = lo[a] - (S[:, a] - r)
            high = (S[:, a] + r) - hi[a]
            F[:, a] += np.where(low > tol, low + tol, 0.0) - np.where(high > tol, high + tol, 0.0)
            wall_over = max(low.max(), high.max(), 0.0)
        if tube is not None:
            ax, tc, tr = tube
            o0, o1 = [x for x in range(3) if x != ax]
            dy, dz = S[:, o0] - tc[0], S[:, o1] - tc[1]
            rr = np.hypot(dy, dz)
            out = rr + r - (tr - gap)
            m = out > tol
            F[m, o0] -= (out[m] + tol) * dy[m] / np.maximum(rr[m], 1e-12)
            F[m, o1] -= (out[m] + tol) * dz[m] / np.maximum(rr[m], 1e-12)
            wall_over = max(wall_over, out.max(initial=0.0))
        worst = max(over.max(initial=0.0), wall_over)
        if worst <= tol:
            if lam >= 1.0:
                break
            lam = min(1.0, lam + (0.01 if lam < 0.9 else 0.004))
            continue
        # rigid-body response: mean push per cell (as many spheres as push it)
        moved = np.bincount(sb, weights=np.any(F != 0, axis=1), minlength=nb)
        dc = np.stack([np.bincount(sb, weights=F[:, k], minlength=nb) for k in range(3)], -1)
        dc /= np.maximum(moved, 1.0)[:, None]
        tq = np.cross(arm, F)
        w = np.stack([np.bincount(sb, weights=tq[:, k], minlength=nb) for k in range(3)], -1)
        w /= (lam * lam * inertia)[:, None]
        ang = np.linalg.norm(w, axis=1)
        cap = np.minimum(1.0, 0.05 / np.maximum(ang, 1e-30))
        w *= cap[:, None]
        ang *= cap
        c += 1.1 * dc
        half = 0.5 * ang
        dq = np.concatenate([np.cos(half)[:, None], (np.sin(half) / np.maximum(ang, 1e-30))[:, None] * w], 1)
        q = _quat_mul(dq, q)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        c = np.where(periodic, np.mod(c, shape), c)
        if log and it % 200 == 0:
            log("pack: iteration %d, scale %.3f, %d overlapping pairs, worst %.3f" % (it, lam, len(i), worst))
    else:
        raise RuntimeError("packing did not converge in %d iterations (scale %.3f)" % (max_iter, lam))
    if log:
        log("pack: %d cells in %d iterations" % (nb, it))
    out = {}
    R = _quat_to_matrix(q)
    for ti, name in enumerate(names):
        m = body_type == ti
        out[name] = (c[m], R[m])
    return out


def instances(types, packing):
    """Vertex positions {name: [n, nv, 3]} of a packing (each mesh centred on its
    vertex mean, rotated, then moved to its centre)."""
    out = {}
    for name, (cen, rot) in packing.items():
        v = types[name].verts - types[name].verts.mean(axis=0)
        out[name] = np.einsum("cij,vj->cvi", rot, v) + cen[:, None, :]
    return out


def overlap_free(types, packing, shape, gap, periodic=(True, True, True)):
    """Whether the sphere covers of a packing keep `gap` between every two cells."""
    S, r, b = [], [], []
    k = 0
    for name, (cen, rot) in packing.items():
        loc, rad = sphere_cover(types[name].verts, name)
        for ci in range(len(cen)):
            S.append(cen[ci] + loc @ rot[ci].T)
            r.append(rad)
            b.append(np.full(len(rad), k))
            k += 1
    S, r, b = np.concatenate(S), np.concatenate(r), np.concatenate(b)
    shape = np.asarray(shape, dtype=np.float64)
    per = np.asarray(periodic)
    Sw = np.where(per, np.mod(S, shape), S + 4.0 * shape.max())
    tree = cKDTree(Sw, boxsize=np.where(per, shape, 8.0 * shape.max()))
    pr = tree.query_pairs(2 * r.max() + gap, output_type="ndarray")
    i, j = pr[:, 0], pr[:, 1]
    keep = b[i] != b[j]
    i, j = i[keep], j[keep]
    d = S[j] - S[i]
    d -= np.where(per, np.round(d / shape) * shape, 0.0)
    return bool(np.all(np.linalg.norm(d, axis=1) >= r[i] + r[j] + gap - 1e-9))
