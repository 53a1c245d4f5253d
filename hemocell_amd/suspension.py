"""Seeded, non-overlapping cell suspensions for periodic boxes and channels.

Stands in for the reference's cell packing inputs (the RBC.pos / PLT.pos files
written by tools/packCells, read by io/readPositionsBloodCells.cpp:186-285):
positions are generated here, directly in lattice units.

Packing: the box is cut into layers along z and 16x16-node columns in x, y.
RBCs lie with their disc axis along z, one per slot, rotated by a uniform random
angle about that axis and jittered inside the slot margin.  Platelets take one
slot each with a uniform random orientation (their bounding sphere fits the
slot).  WBCs take a block of consecutive layers in one column, never next to
another WBC.  No two cells overlap by construction.
"""
import math

import numpy as np

from . import build_mesh, model_constants, MODEL_RBC, MODEL_PLT, MODEL_WBC

RBC_XML = dict(kLink=15.0, kBend=80.0, kVolume=20.0, kArea=5.0, eta_m=0.0)
PLT_XML = dict(kLink=25.0, kBend=250.0, kVolume=100.0, kArea=8.0, eta_m=0.0)
WBC_XML = dict(kLink=60.0, kBend=200.0, kVolume=20.0, kArea=20.0, eta_m=1.0e-9, kInnerRigid=6.40625e-12,
               kCytoskeleton=6.40625e-15, coreRadius=2.5e-6, radius=4.0e-6)


def _rot_z(a):
    c, s = np.cos(a), np.sin(a)
    R = np.zeros((len(a), 3, 3))
    R[:, 0, 0], R[:, 0, 1], R[:, 1, 0], R[:, 1, 1], R[:, 2, 2] = c, -s, s, c, 1.0
    return R


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    a, b, c, d = q.T
    return np.stack([np.stack([a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)], -1),
                     np.stack([2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)], -1),
                     np.stack([2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d], -1)], 1)


def antipodal_pairs(verts, count=None):
    """Inner edges of a centrally symmetric mesh: each vertex with its opposite."""
    v = verts - verts.mean(axis=0)
    pairs, used = [], set()
    for i in range(len(v)):
        if i in used:
            continue
        d = np.linalg.norm(v + v[i], axis=1)
        j = int(np.argmin(d))
        if j != i and j not in used and d[j] < 1e-6 * np.linalg.norm(v[i]) + 1e-9:
            pairs.append((i, j))
            used.update((i, j))
    pairs = np.array(pairs, dtype=np.int32).reshape(-1, 2)
    return pairs if count is None else pairs[:count]


class CellType:
    def __init__(self, name, model, verts, tris, params, inner=None):
        self.name, self.model, self.verts, self.tris, self.params, self.inner = name, model, verts, tris, params, inner

    @property
    def nv(self):
        return len(self.verts)


def cell_types(units, wbc_triangles=2000):
    """RBC (642 vertices), PLT (66-vertex 2x2x0.5 um ellipsoid) and WBC (8 um sphere) at these units."""
    dx = units.dx
    out = {}
    v, t = build_mesh(1, 3.91e-6 / dx, 600)
    # disc axis of the built RBC is y; turn it to z for layer packing
    v = v[:, [0, 2, 1]] * np.array([1.0, -1.0, 1.0])
    out["RBC"] = CellType("RBC", MODEL_RBC, v, t, model_constants(MODEL_RBC, len(t), units, **RBC_XML))
    v, t = build_mesh(6, 1.0e-6 / dx, 66, 0.25)
    out["PLT"] = CellType("PLT", MODEL_PLT, v, t, model_constants(MODEL_PLT, len(t), units, **PLT_XML),
                          antipodal_pairs(v, 21))
    v, t = build_mesh(0, 4.0e-6 / dx, wbc_triangles)
    out["WBC"] = CellType("WBC", MODEL_WBC, v, t, model_constants(MODEL_WBC, len(t), units, **WBC_XML),
                          antipodal_pairs(v))
    return out


def pack(shape, types, hematocrit=0.30, plt_per_rbc=0.1, wbc_per_rbc=1.0 / 500, seed=20261202, layer=5.6,
         wall_axis=None):
    """Vertex positions per type: {name: array[ncells, nv, 3]} in lattice units.

    wall_axis 0 or 1: that axis is bounded by wall planes at 0 and n-1, and the
    first and last columns along it stay empty, so no vertex's support reaches a
    wall node."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    rbc = types["RBC"]
    vol_rbc = _volume(rbc.verts, rbc.tris)
    n_rbc = int(round(hematocrit * nx * ny * nz / vol_rbc))
    n_plt = int(round(plt_per_rbc * n_rbc))
    n_wbc = int(round(wbc_per_rbc * n_rbc))
    ci, cj = nx // 16, ny // 16
    nl = int(nz // layer)
    h = nz / nl
    ext = {k: np.ptp(t.verts, axis=0) for k, t in types.items()}
    if ext["RBC"][2] > h or max(ext["RBC"][0], ext["RBC"][1]) > 16.0:
        raise ValueError("RBC does not fit a 16 x 16 x %.2f slot" % h)
    wl = int(math.ceil((ext["WBC"][2] + 0.2) / h)) if n_wbc else 0
    free = np.ones((ci, cj, nl), dtype=bool)
    if wall_axis == 0:
        free[0, :, :] = free[ci - 1, :, :] = False
    elif wall_axis == 1:
        free[:, 0, :] = free[:, cj - 1, :] = False
    elif wall_axis is not None:
        raise ValueError("wall_axis must be 0, 1 or None")
    wbc_centres = []
    cols = [(i, j) for i in range(0, ci, 2) for j in range(0, cj, 2) if free[i, j, 0]]
    rng.shuffle(cols)
    if n_wbc > len(cols):
        raise ValueError("too many WBCs for this box")
    for k in range(n_wbc):
        i, j = cols[k]
        l0 = int(rng.integers(0, nl - wl + 1))
        free[i, j, l0:l0 + wl] = False
        wbc_centres.append((16 * i + 8.0, 16 * j + 8.0, (l0 + wl / 2.0) * h))
    slots = np.argwhere(free)
    if n_rbc + n_plt > len(slots):
        raise ValueError("hematocrit %.2f does not fit this packing (%d slots for %d cells)" %
                         (hematocrit, len(slots), n_rbc + n_plt))
    pick = rng.permutation(len(slots))[:n_rbc + n_plt]
    centres = np.stack([16 * slots[pick, 0] + 8.0, 16 * slots[pick, 1] + 8.0, (slots[pick, 2] + 0.5) * h], -1)
    out = {}
    # RBC: rotation about z, jitter within the margins
    c = centres[:n_rbc]
    mxy = max(0.0, (16.0 - max(ext["RBC"][0], ext["RBC"][1])) / 2 - 0.02)
    mz = max(0.0, (h - ext["RBC"][2]) / 2 - 0.05)
    c = c + np.stack([rng.uniform(-mxy, mxy, n_rbc), rng.uniform(-mxy, mxy, n_rbc), rng.uniform(-mz, mz, n_rbc)], -1)
    out["RBC"] = _instances(rbc.verts, _rot_z(rng.uniform(0, 2 * np.pi, n_rbc)), c)
    p = centres[n_rbc:]
    r = np.linalg.norm(types["PLT"].verts - types["PLT"].verts.mean(0), axis=1).max()
    m = max(0.0, min(8.0, h / 2) - r - 0.05)
    p = p + rng.uniform(-m, m, size=p.shape) * np.array([1.0, 1.0, 1.0])
    out["PLT"] = _instances(types["PLT"].verts, _random_rotations(rng, n_plt), p)
    if n_wbc:
        out["WBC"] = _instances(types["WBC"].verts, _random_rotations(rng, n_wbc), np.array(wbc_centres))
    return out


def cell_counts(types, fluid_nodes, hematocrit, plt_per_rbc=0.1, wbc_per_rbc=1.0 / 500):
    """Cells of each type for a hematocrit (RBC volume over the fluid volume)."""
    rbc = types["RBC"]
    n_rbc = int(round(hematocrit * fluid_nodes / _volume(rbc.verts, rbc.tris)))
    return {"RBC": n_rbc, "PLT": int(round(plt_per_rbc * n_rbc)), "WBC": int(round(wbc_per_rbc * n_rbc))}


def tube_fluid(length, diameter):
    """Fluid mask of a tube along x in a length x diameter x diameter box: fluid
    iff (y - c)^2 + (z - c)^2 < r^2 with c = (diameter - 1) / 2, r = diameter / 2 - 0.5;
    the rest is wall.  Indexed [x, y, z]."""
    c = (diameter - 1) / 2.0
    r = diameter / 2.0 - 0.5
    y, z = np.meshgrid(np.arange(diameter), np.arange(diameter), indexing="ij")
    inside = (y - c) ** 2 + (z - c) ** 2 < r * r
    return np.broadcast_to(inside, (length, diameter, diameter))


def pack_tube(length, diameter, types, hematocrit=0.20, plt_per_rbc=0.1, seed=20261202, layer=5.6, margin=1.6):
    """Seeded, non-overlapping cells inside the tube of tube_fluid(length, diameter).

    Slots are 16 (x) by 16 (y) by `layer` (z) boxes whose corners in the y-z
    cross-section stay `margin` inside the wall radius; RBCs lie with their disc
    axis along z, rotated by a random angle about it; platelets take a slot each
    with a random orientation.  The hematocrit counts RBC volume over the tube's
    fluid volume.  Returns {name: array[ncells, nv, 3]} in lattice units."""
    rng = np.random.default_rng(seed)
    c = (diameter - 1) / 2.0
    r_in = diameter / 2.0 - 0.5 - margin
    rbc = types["RBC"]
    ext = np.ptp(rbc.verts, axis=0)
    if ext[2] > layer or max(ext[0], ext[1]) > 16.0:
        raise ValueError("RBC does not fit a 16 x 16 x %.2f slot" % layer)
    # the cross-section's slot grid: the y offset, z offset and row stagger that
    # fit the most slots inside the circle
    def grid(yoff, zoff, stagger):
        out = []
        nrow = int(diameter // layer) + 2
        for i in range(-nrow, nrow + 1):
            zc = c + zoff + layer * i
            for yc in np.arange(c - diameter, c + diameter + 0.1, 16.0) + yoff + (stagger if i % 2 else 0.0):
                if all(np.hypot(yc + dy - c, zc + dz - c) < r_in for dy in (-8.0, 8.0)
                       for dz in (-layer / 2, layer / 2)):
                    out.append((yc, zc))
        return out
    cross = max((grid(yo, zo, st) for yo in np.arange(0.0, 8.01, 0.5) for zo in np.arange(0.0, layer, 0.4)
                 for st in (0.0, 8.0)), key=len)
    slots = [(8.0 + 16.0 * k, yc, zc) for k in range(length // 16) for (yc, zc) in cross]
    fluid = float(tube_fluid(1, diameter).sum()) * length
    n_rbc = int(round(hematocrit * fluid / _volume(rbc.verts, rbc.tris)))
    n_plt = int(round(plt_per_rbc * n_rbc))
    if n_rbc + n_plt > len(slots):
        raise ValueError("hematocrit %.2f does not fit this tube (%d slots for %d cells)" %
                         (hematocrit, len(slots), n_rbc + n_plt))
    pick = rng.permutation(len(slots))[:n_rbc + n_plt]
    centres = np.array(slots)[pick]
    out = {"RBC": _instances(rbc.verts, _rot_z(rng.uniform(0, 2 * np.pi, n_rbc)), centres[:n_rbc])}
    if n_plt:
        out["PLT"] = _instances(types["PLT"].verts, _random_rotations(rng, n_plt), centres[n_rbc:])
    return out


def _instances(verts, rots, centres):
    v = verts - verts.mean(axis=0)
    return np.einsum("cij,vj->cvi", rots, v) + centres[:, None, :]


def _volume(v, t):
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def hematocrit(types, cells, shape):
    vol = _volume(types["RBC"].verts, types["RBC"].tris)
    return len(cells["RBC"]) * vol / float(np.prod(shape))
