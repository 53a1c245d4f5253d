"""Cell-mesh construction restated in Python -- TEST INFRASTRUCTURE ONLY.

This is the oracle for the product's C++ mesh builder (hemocell_amd/csrc/mesh.cpp).
Only tests/, smoke() and bench.py's cpu_baseline may import it.

Restated from the HemoCell reference:
  * icosahedron + midpoint refinement      helper/meshGeneratingFunctions.hh:32-151
  * Evans-Fung RBC map                     helper/meshGeneratingFunctions.hh:154-168
  * ellipsoid map                          helper/meshGeneratingFunctions.hh:171-183
  * RBC / ellipsoid / sphere construction  helper/meshGeneratingFunctions.hh:214-271,
                                           helper/meshGeneratingFunctions.h:69-94
The Palabos parts (octahedron sphere, TriangleSet::rotate, vertex numbering of
the de-duplicated mesh, inflate) are not vendored: they are restated here as
(1) octahedron (+x,-x,+y,-y,+z,-z) with the same refinement rule,
(2) the z-x-z Euler rotation matrix, (3) vertices numbered in order of first
appearance over the triangle soup (exact-coordinate match), and (4) each vertex
moved 1e-3 lattice units along its unit vertex normal (sum of unit triangle
normals).  Their exact reproduction of Palabos is parity unpinned; the rest
volume of the inflated RBC (81.116 um^3, RBC_template.xml:16) pins (4).
"""
import math

import numpy as np

# shapes (config/constant_defaults.h:80-84)
WBC_SPHERE = 0
RBC_FROM_SPHERE = 1
ELLIPSOID_FROM_SPHERE = 6

_T = -0.8506508084
_O = -0.5257311121


def _icosahedron():
    v = [None,
         (_T, _O, 0.0), (-_T, _O, 0.0), (-_T, -_O, 0.0), (_T, -_O, 0.0),
         (_O, 0.0, _T), (_O, 0.0, -_T), (-_O, 0.0, -_T), (-_O, 0.0, _T),
         (0.0, _T, _O), (0.0, -_T, _O), (0.0, -_T, -_O), (0.0, _T, -_O)]
    faces = [(5, 8, 9), (5, 10, 8), (6, 12, 7), (6, 7, 11), (1, 4, 5), (1, 6, 4),
             (3, 2, 8), (3, 7, 2), (9, 12, 1), (9, 2, 12), (10, 4, 11), (10, 11, 3),
             (9, 1, 5), (12, 6, 1), (5, 4, 10), (6, 11, 4), (8, 2, 9), (7, 12, 2),
             (8, 10, 3), (7, 3, 11)]
    return [[list(v[a]), list(v[b]), list(v[c])] for a, b, c in faces]


def _octahedron():
    px, mx = [1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]
    py, my = [0.0, 1.0, 0.0], [0.0, -1.0, 0.0]
    pz, mz = [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]
    faces = [(px, py, pz), (py, mx, pz), (mx, my, pz), (my, px, pz),
             (py, px, mz), (mx, py, mz), (my, mx, mz), (px, my, mz)]
    return [[list(a), list(b), list(c)] for a, b, c in faces]


def _norm(a):
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _refine(tris, min_tri):
    while len(tris) < min_tri:
        size = len(tris)
        for i in range(size):
            va, vb, vc = tris[i]
            vd = [0.5 * (va[k] + vb[k]) for k in range(3)]
            ve = [0.5 * (vb[k] + vc[k]) for k in range(3)]
            vf = [0.5 * (vc[k] + va[k]) for k in range(3)]
            nd, ne, nf = _norm(vd), _norm(ve), _norm(vf)
            vd = [c / nd for c in vd]
            ve = [c / ne for c in ve]
            vf = [c / nf for c in vf]
            tris[i] = [vd, ve, vf]
            tris.append([va, vd, vf])
            tris.append([vd, vb, ve])
            tris.append([vf, ve, vc])
    return tris


def _rotation(phi, theta, psi):
    cf, sf = math.cos(phi), math.sin(phi)
    ct, st = math.cos(theta), math.sin(theta)
    cp, sp = math.cos(psi), math.sin(psi)
    return [[cp * cf - ct * sf * sp, cp * sf + ct * cf * sp, sp * st],
            [-sp * cf - ct * sf * cp, -sp * sf + ct * cf * cp, cp * st],
            [st * sf, -st * cf, ct]]


def _rotate(tris, a):
    out = []
    for t in tris:
        nt = []
        for p in t:
            q = [0.0, 0.0, 0.0]
            for i in range(3):
                s = 0.0
                for j in range(3):
                    s += a[i][j] * p[j]
                q[i] = s
            nt.append(q)
        out.append(nt)
    return out


def _sign(v):
    return (1 if 0.0 < v else 0) - (1 if v < 0.0 else 0)


def _rbc_point(p, R=1.0):
    x, y, z = p
    r2 = x * x + y * y
    s = _sign(z)
    x *= R
    y *= R
    if 1 - r2 < 0:
        r2 = 1
    C0, C2, C4 = 0.054322, 1.001279, -0.561381
    z = s * R * math.sqrt(1 - r2) * (C0 + C2 * r2 + C4 * r2 * r2)
    return [x, y, z]


def _ellipsoid_point(p, R, aspect):
    x, y, z = p
    r2 = x * x + y * y
    s = _sign(z)
    if 1 - r2 < 0:
        r2 = 1
    x *= R
    y *= R
    z = s * aspect * R * math.sqrt(1 - r2)
    return [x, y, z]


def _dedup(tris):
    index = {}
    verts = []
    faces = []
    for t in tris:
        f = []
        for p in t:
            key = (p[0], p[1], p[2])
            if key not in index:
                index[key] = len(verts)
                verts.append(list(p))
            f.append(index[key])
        faces.append(f)
    return np.array(verts, dtype=np.float64), np.array(faces, dtype=np.int32)


def _inflate(verts, faces, amount=1e-3):
    acc = np.zeros_like(verts)
    for f in faces:
        v0, v1, v2 = verts[f[0]], verts[f[1]], verts[f[2]]
        e1 = [v1[k] - v0[k] for k in range(3)]
        e2 = [v2[k] - v0[k] for k in range(3)]
        n = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
        nn = _norm(n)
        n = [c / nn for c in n]
        for k in f:
            for c in range(3):
                acc[k][c] += n[c]
    out = verts.copy()
    for i in range(len(verts)):
        nv = _norm(acc[i])
        for c in range(3):
            out[i][c] = verts[i][c] + amount * (acc[i][c] / nv)
    return out


def build_mesh(shape, radius_lu, min_triangles, aspect_ratio=0.3):
    """Rest mesh (vertices [nv,3] in lattice units, triangles [nt,3]) centred near 0."""
    half = math.pi / 2.0
    if shape == WBC_SPHERE:
        tris = _refine(_icosahedron(), min_triangles)
        tris = [[[c * radius_lu for c in p] for p in t] for t in tris]
    elif shape == RBC_FROM_SPHERE:
        rot = _rotation(half, half, 0.0)
        tris = _rotate(_refine(_icosahedron(), min_triangles), rot)
        tris = [[_rbc_point(p) for p in t] for t in tris]
        tris = [[[c * radius_lu for c in p] for p in t] for t in tris]
        tris = _rotate(tris, rot)
    elif shape == ELLIPSOID_FROM_SPHERE:
        rot = _rotation(half, half, 0.0)
        tris = _rotate(_refine(_octahedron(), min_triangles), rot)
        tris = [[_ellipsoid_point(p, radius_lu, aspect_ratio) for p in t] for t in tris]
        tris = _rotate(tris, rot)
    else:
        raise ValueError("unsupported shape %r" % shape)
    verts, faces = _dedup(tris)
    return _inflate(verts, faces), faces


def mesh_volume(verts, faces):
    v = 0.0
    for f in faces:
        a, b, c = verts[f[0]], verts[f[1]], verts[f[2]]
        v += np.dot(a, np.cross(b, c)) / 6.0
    return float(v)


def mesh_area(verts, faces):
    s = 0.0
    for f in faces:
        a, b, c = verts[f[0]], verts[f[1]], verts[f[2]]
        s += 0.5 * np.linalg.norm(np.cross(b - a, c - a))
    return float(s)
