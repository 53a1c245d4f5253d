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
import os

import numpy as np
from scipy.spatial import cKDTree

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def data_path(name):
    """A committed packing of hemocell_amd/data."""
    return os.path.join(DATA, name)


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


def pack(shape, types, counts, seed=20261202, gap=0.5, walls=None, tube=None, max_iter=20000, log=None, tol=1e-3,
         lam0=0.5, wall_margin=1.5):
    """Random positions and orientations of counts[name] cells of each type in a
    box of `shape` nodes.  walls: an axis (0, 1 or 2) bounded by wall planes at
    0 and n-1, which the cells keep wall_margin + gap away from; tube: (axis,
    centre, radius) of a cylinder the cells stay inside; other axes periodic.  Cells keep `gap` (less `tol`) from each other and from
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
        lo[walls], hi[walls] = wall_margin + gap, shape[walls] - 1.0 - wall_margin - gap
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
    lam = lam0
    it = 0
    rsmall = rad[small].max() if small.any() else 0.0
    idx_small = np.flatnonzero(small)
    ib = np.flatnonzero(big)
    gi, gj = np.meshgrid(ib, ib, indexing="ij")
    bb = gi < gj
    skin = 1.0
    S_ref, lam_ref, cand = None, None, None
    while it < max_iter:
        it += 1
        R = _quat_to_matrix(q)
        arm = lam * np.einsum("mij,mj->mi", R[sb], loc)
        S = c[sb] + arm
        r = lam * rad
        # candidate pairs within the reach plus a skin, rebuilt when a sphere has
        # moved (or all have grown) by more than the skin allows
        if S_ref is not None:
            dS = S - S_ref
            dS -= np.where(periodic, np.round(dS / shape) * shape, 0.0)
            drift = 2 * np.sqrt(np.max(np.sum(dS * dS, axis=1))) + 2 * (lam - lam_ref) * rad.max()
        if S_ref is None or drift > skin:
            Sw = np.where(periodic, np.mod(S, shape), S + 4.0 * shape.max())
            tree = cKDTree(Sw[small], boxsize=box)
            pr = tree.query_pairs(2 * lam * rsmall + gap + skin, output_type="ndarray")
            ci, cj = idx_small[pr[:, 0]], idx_small[pr[:, 1]]
            if len(ib):
                hits = tree.query_ball_point(Sw[ib], lam * (rad[ib] + rsmall) + gap + skin)
                bi = np.concatenate([np.full(len(h), b) for b, h in zip(ib, hits)])
                bj = idx_small[np.concatenate(hits).astype(int)]
                ci = np.concatenate([ci, bi, gi[bb]])
                cj = np.concatenate([cj, bj, gj[bb]])
            other = sb[ci] != sb[cj]
            cand = (ci[other], cj[other])
            S_ref, lam_ref = S.copy(), lam
        i, j = cand
        d = S[j] - S[i]
        d -= np.where(periodic, np.round(d / shape) * shape, 0.0)
        dist = np.linalg.norm(d, axis=1)
        over = r[i] + r[j] + gap - dist
        hit = over > tol
        i, j, d, dist, over = i[hit], j[hit], d[hit], dist[hit], over[hit]
        n = d / np.maximum(dist, 1e-12)[:, None]
        push = 0.5 * (over + 2 * tol)[:, None] * n  # moves sphere j's cell by +push, sphere i's by -push
        F = np.stack([np.bincount(j, weights=push[:, k], minlength=len(S)) -
                      np.bincount(i, weights=push[:, k], minlength=len(S)) for k in range(3)], -1).astype(np.float64)
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
        # rigid-body response: the mean push over the cell's contacts (a cell
        # pressed from many sides by one big sphere's neighbours moves by their
        # mean, not their sum), at most a lattice spacing per iteration
        contacts = np.bincount(sb[i], minlength=nb) + np.bincount(sb[j], minlength=nb)
        if walls is not None or tube is not None:
            contacts = contacts + np.bincount(sb, weights=np.any(F != 0, axis=1), minlength=nb)
        dc = np.stack([np.bincount(sb, weights=F[:, k], minlength=nb) for k in range(3)], -1)
        dc /= np.maximum(contacts, 1.0)[:, None]
        step = np.linalg.norm(dc, axis=1)
        dc *= np.minimum(1.0, 1.0 / np.maximum(step, 1e-30))[:, None]
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


def tile(packing, tile_shape, reps):
    """A packing of a box periodic along the tiled axes, repeated reps[a] times
    along each axis a (the tiles join without overlap because the packing is
    periodic across those faces)."""
    offs = np.stack(np.meshgrid(*[np.arange(r) for r in reps], indexing="ij"), -1).reshape(-1, 3)
    offs = offs * np.asarray(tile_shape, dtype=np.float64)
    out = {}
    for name, (cen, rot) in packing.items():
        out[name] = ((cen[None, :, :] + offs[:, None, :]).reshape(-1, 3),
                     np.broadcast_to(rot[None], (len(offs),) + rot.shape).reshape(-1, 3, 3).copy())
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


def save(path, packing, **meta):
    """A packing as an .npz file: per type its centres and unit quaternions
    (float32; rounding moves a vertex by well under 1e-3 lattice units), plus
    the metadata given (shape, seed, hematocrit, ...)."""
    arrays = {}
    for name, (cen, rot) in packing.items():
        arrays[name + "_centres"] = cen.astype(np.float32)
        arrays[name + "_quats"] = _matrix_to_quat(rot).astype(np.float32)
    for k, v in meta.items():
        arrays["meta_" + k] = np.asarray(v)
    np.savez_compressed(path, **arrays)


def load(path):
    """(packing, meta) of a file written by save()."""
    d = np.load(path, allow_pickle=False)
    out, meta = {}, {}
    for k in d.files:
        if k.startswith("meta_"):
            meta[k[5:]] = d[k]
        elif k.endswith("_centres"):
            name = k[:-8]
            q = d[name + "_quats"].astype(np.float64)
            q /= np.linalg.norm(q, axis=1, keepdims=True)
            out[name] = (d[k].astype(np.float64), _quat_to_matrix(q))
    return out, meta


def _matrix_to_quat(R):
    """Unit quaternions (w, x, y, z) of rotation matrices [n, 3, 3]."""
    q = np.zeros((len(R), 4))
    tr = np.trace(R, axis1=1, axis2=2)
    for k in range(len(R)):
        m = R[k]
        if tr[k] > 0:
            s = 2.0 * math.sqrt(1.0 + tr[k])
            q[k] = (0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s)
        else:
            a = int(np.argmax(np.diag(m)))
            b, c = (a + 1) % 3, (a + 2) % 3
            s = 2.0 * math.sqrt(max(1.0 + m[a, a] - m[b, b] - m[c, c], 1e-300))
            q[k, 0] = (m[c, b] - m[b, c]) / s
            q[k, 1 + a] = 0.25 * s
            q[k, 1 + b] = (m[b, a] + m[a, b]) / s
            q[k, 1 + c] = (m[c, a] + m[a, c]) / s
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def main(argv=None):
    """python -m hemocell_amd.packing OUT.npz --shape nx,ny,nz --hematocrit H
    [--walls AXIS] [--tube] [--plt-per-rbc R] [--wbc-per-rbc R] [--seed S]"""
    import argparse
    import sys
    import time

    from . import lbm_base_parameters
    from . import suspension as susp
    ap = argparse.ArgumentParser(prog="python -m hemocell_amd.packing")
    ap.add_argument("out")
    ap.add_argument("--shape", required=True)
    ap.add_argument("--hematocrit", type=float, required=True)
    ap.add_argument("--plt-per-rbc", type=float, default=0.1)
    ap.add_argument("--wbc-per-rbc", type=float, default=1.0 / 500)
    ap.add_argument("--walls", type=int, default=None, help="axis bounded by wall planes at 0 and n-1")
    ap.add_argument("--tube", action="store_true", help="cells inside the tube along x of tube_fluid()")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--max-iter", type=int, default=20000)
    a = ap.parse_args(argv)
    shape = tuple(int(v) for v in a.shape.split(","))
    types = susp.cell_types(lbm_base_parameters(5e-7, 1e-7))
    tube = None
    if a.tube:
        if shape[1] != shape[2]:
            raise SystemExit("--tube needs a square y-z cross-section")
        fluid = float(susp.tube_fluid(1, shape[1]).sum()) * shape[0]
        c = (shape[1] - 1) / 2.0
        # the supports (one node each way) stay off the wall nodes
        tube = (0, (c, c), shape[1] / 2.0 - 0.5 - 1.5)
    else:
        fluid = float(np.prod(shape)) if a.walls is None else float(np.prod(shape)) * (shape[a.walls] - 2) / shape[a.walls]
    counts = susp.cell_counts(types, fluid, a.hematocrit, a.plt_per_rbc, a.wbc_per_rbc)
    t0 = time.time()
    log = lambda m: print("%s (%.0f s)" % (m, time.time() - t0), file=sys.stderr, flush=True)  # noqa: E731
    pk = pack(shape, types, counts, seed=a.seed, walls=a.walls, tube=tube, log=log, max_iter=a.max_iter)
    periodic = [True, True, True]
    if a.walls is not None:
        periodic[a.walls] = False
    if a.tube:
        periodic[1] = periodic[2] = False
    if not overlap_free(types, pk, shape, 0.49, periodic):
        raise SystemExit("packing check failed")
    save(a.out, pk, shape=shape, hematocrit=a.hematocrit, seed=a.seed, counts=[counts.get(n, 0) for n in
                                                                             ("RBC", "PLT", "WBC")],
         walls=-1 if a.walls is None else a.walls, tube=int(a.tube))
    log("wrote %s: %s" % (a.out, counts))


if __name__ == "__main__":
    main()
