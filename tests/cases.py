"""Shared case builders for the tests (lattices, cell types, seeded placements).

The unit conversion here is the test side's own restatement of
mechanics/constantConversion.cpp:36-59 and mechanics/cellMechanics.h:50-78;
tests compare it with the product's hc_lbm_base_parameters / hc_model_constants
and with the closed-form values tabulated from the reference formulas.
"""
import math
from types import SimpleNamespace

import numpy as np

KBT = 4.100531391e-21

# <MaterialModel> defaults of the reference templates (tools/cell_templates/*.xml)
RBC_MAT = dict(kLink=15.0, kBend=80.0, kVolume=20.0, kArea=5.0, eta_m=0.0, radius=3.91e-6, minNumTriangles=600)
PLT_MAT = dict(kLink=25.0, kBend=250.0, kVolume=100.0, kArea=8.0, eta_m=0.0, radius=1.0e-6, aspectRatio=0.25,
               minNumTriangles=66)
WBC_MAT = dict(kLink=60.0, kBend=200.0, kVolume=20.0, kArea=20.0, eta_m=1.0e-9, kInnerRigid=6.40625e-12,
               kCytoskeleton=6.40625e-15, coreRadius=2.5e-6, radius=4.0e-6, minNumTriangles=600)


def units(dx=5e-7, dt=1e-7, nu_p=1.1e-6, rho_p=1025.0, kBT=KBT):
    if dt < 0:
        tau = 1.0
        nu_lbm = (tau - 0.5) / 3.0
        dt = nu_lbm / nu_p * dx * dx
    else:
        nu_lbm = nu_p * dt / (dx * dx)
        tau = 3.0 * nu_lbm + 0.5
    dm = rho_p * dx ** 3
    df = dm * dx / dt ** 2
    return SimpleNamespace(dx=dx, dt=dt, tau=tau, nu_lbm=nu_lbm, dm=dm, df=df, f_limit=50e-12 / df,
                           kBT_lbm=kBT / (df * dx))


def params(mat, ntri, u, model):
    import hemocell_amd as hc  # ModelParams struct only (no device call)
    plc = 7.5e-9 / u.dx
    eq = 5e-7 / u.dx
    s = 1280.0 / ntri
    p = hc.ModelParams()
    p.k_link = mat["kLink"] * u.kBT_lbm / plc
    p.k_bend = mat["kBend"] * u.kBT_lbm / eq
    p.k_volume = mat["kVolume"] * s * u.kBT_lbm / eq
    p.k_area = mat["kArea"] * s * u.kBT_lbm / eq
    p.eta_m = mat.get("eta_m", 0.0) * u.dx / u.dt / u.df
    if model == 2:
        p.k_inner_rigid = mat["kInnerRigid"] / u.df
        p.k_cytoskeleton = mat["kCytoskeleton"] / u.df
        p.core_radius = mat["coreRadius"] / u.dx
        p.radius = mat["radius"] / u.dx
    return p


def antipodal_pairs(verts, count=None):
    """Inner edges: vertex pairs opposite through the centre (the mesh is centred at 0)."""
    v = verts - verts.mean(axis=0)
    pairs = []
    used = set()
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


def mesh(kind, u, builder):
    """(model, vertices, triangles, params, inner) for kind in RBC/PLT/WBC/WBC2562."""
    if kind == "RBC":
        v, t = builder(1, RBC_MAT["radius"] / u.dx, RBC_MAT["minNumTriangles"])
        return 0, v, t, params(RBC_MAT, len(t), u, 0), None
    if kind == "PLT":
        v, t = builder(6, PLT_MAT["radius"] / u.dx, PLT_MAT["minNumTriangles"], PLT_MAT["aspectRatio"])
        return 1, v, t, params(PLT_MAT, len(t), u, 1), antipodal_pairs(v, 21)
    if kind in ("WBC", "WBC2562"):
        ntri = 600 if kind == "WBC" else 2000
        v, t = builder(0, WBC_MAT["radius"] / u.dx, ntri)
        return 2, v, t, params(WBC_MAT, len(t), u, 2), antipodal_pairs(v)
    raise ValueError(kind)


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def place(verts, center, R=None):
    v = verts - verts.mean(axis=0)
    if R is not None:
        v = v @ R.T
    return v + np.asarray(center, dtype=np.float64)


def perturb(x, rng, amp):
    return x + amp * rng.uniform(-1, 1, size=x.shape)


def random_populations(N, rng, amp=1e-3):
    return amp * rng.uniform(-1, 1, size=(19, N))


def couette_mask(nx, ny, nz):
    """Velocity walls on z = 0 (face 4) and z = nz-1 (face 5)."""
    t = np.zeros((nx, ny, nz), dtype=np.uint8)
    t[:, :, 0] = 2 + 4
    t[:, :, nz - 1] = 2 + 5
    return t.reshape(-1)


def channel_mask(nx, ny, nz, axis=1):
    """Bounce-back walls on the two faces normal to `axis`."""
    t = np.zeros((nx, ny, nz), dtype=np.uint8)
    sl = [slice(None)] * 3
    sl[axis] = 0
    t[tuple(sl)] = 1
    sl[axis] = t.shape[axis] - 1
    t[tuple(sl)] = 1
    return t.reshape(-1)


def tube_mask(nx, ny, nz, axis=0):
    """Bounce-back everywhere outside a cylinder along `axis` (analytic mask)."""
    idx = np.indices((nx, ny, nz)).astype(np.float64)
    others = [a for a in range(3) if a != axis]
    dims = [nx, ny, nz]
    ca = (dims[others[0]] - 1) / 2.0
    cb = (dims[others[1]] - 1) / 2.0
    r = min(dims[others[0]], dims[others[1]]) / 2.0 - 0.5
    d2 = (idx[others[0]] - ca) ** 2 + (idx[others[1]] - cb) ** 2
    return np.where(d2 < r * r, 0, 1).astype(np.uint8).reshape(-1)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
