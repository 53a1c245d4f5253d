"""Closed forms of the membrane force terms on a regular octahedron.

The octahedron (vertices +-r e_x, +-r e_y, +-r e_z; 8 faces, 12 edges; every
vertex has 4 neighbours) is registered as the rest shape of a cell type and
then loaded uniformly scaled by s about its centre.  By symmetry every term's
force on every vertex is radial, and its size follows from the reference
formulas by hand:

  link   (rbcHighOrderModel.cpp:169-183)  edge L = s r sqrt2, ef = s - 1,
         fs = k_link (ef + ef / |9 - ef^2|); the 4 edges at a vertex pull it by
         sum_b (x_b - x_a)/L = -2 sqrt2 e_a:            F = -2 sqrt2 fs e_a
  area   (:62-84)   ar = s^2 - 1, afm = k_area (ar + ar / |0.09 - ar^2|); the 4
         faces at a vertex give sum (c_t - x_a) = -(8/3) s r e_a:
                                                        F = -(8/3) s r afm e_a
  volume (:100-112) vfrac = s^3 - 1, vf = -k_vol vfrac / |0.01 - vfrac^2|;
         each face adds vf n_t A_t / A_eq with A_t / A_eq = s^2 and the 4 unit
         normals at a vertex sum to (4 / sqrt3) e_a:     F = vf s^2 (4/sqrt3) e_a
  bending (RBC/WBC, :127-166) the ring mean of a vertex is the centre, so
         ndev = -s r, ndev_eq = -r, mean rest edge r sqrt2, dDev = (1 - s)/sqrt2,
         sb = k_bend (dDev + dDev / |0.0555 - dDev^2|); the vertex takes +sb e_a
         and each of its 4 neighbours -sb/4 of its own normal, which cancel:
                                                        F = sb e_a
  inner links along the three diameters (len = 2 s r):
         PLT (pltSimpleModel.cpp:185-205): ef = s - 1, F = 5 k_link ef (-e_a)
         WBC (wbcHighOrderModel.cpp:196-222): cytoskeleton below 2 R and core
         below 2 R_core, each pushing apart with k (1 - len / (2 R)):
                               F = [k_cyto (1 - s r / R) + k_core (1 - s r / R_core)] e_a
  membrane viscosity (:185-199) with only vertex 0 (+r e_x) moving at w e_x:
         each of its 4 edges gives eta (w / sqrt2) along (e_b - e_x)/sqrt2,
         F_0 = -2 eta w e_x; capped per edge at 50/4 (FORCE_LIMIT/4):
         F_0 = -sqrt2 (50/4) * 4 / 2 ... = -(50/4) (4 / sqrt2) e_x.

Nothing here reuses the oracle's or the library's code: the expected forces
are numbers from the formulas above.
"""
import numpy as np

SQ2, SQ3 = np.sqrt(2.0), np.sqrt(3.0)
# config/constant_defaults.h:157-173
MAX_VOL, MAX_AREA, MAX_BEND, MAX_PERS, FORCE_LIMIT = 0.01, 0.09, 0.0555, 9.0, 50.0

AXES = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)


def octahedron(r):
    """Vertices (6, 3) and outward triangles (8, 3) of the regular octahedron."""
    tris = []
    for ix, sx in ((0, 1), (1, -1)):
        for iy, sy in ((2, 1), (3, -1)):
            for iz, sz in ((4, 1), (5, -1)):
                tris.append((ix, iy, iz) if sx * sy * sz > 0 else (ix, iz, iy))
    return r * AXES.copy(), np.array(tris, dtype=np.int32)


def diameters():
    return np.array([[0, 1], [2, 3], [4, 5]], dtype=np.int32)


def g(x, cap):
    return x + x / abs(cap - x * x)


def radial(mag):
    """Force (6, 3): the same signed size along every vertex's outward axis."""
    return mag * AXES


def link(s, r, k):
    return radial(-2 * SQ2 * k * g(s - 1.0, MAX_PERS))


def area(s, r, k):
    return radial(-(8.0 / 3.0) * s * r * k * g(s * s - 1.0, MAX_AREA))


def volume(s, r, k):
    vf = -k * (s ** 3 - 1.0) / abs(MAX_VOL - (s ** 3 - 1.0) ** 2)
    return radial(vf * s * s * 4.0 / SQ3)


def bending(s, r, k):
    return radial(k * g((1.0 - s) / SQ2, MAX_BEND))


def plt_inner(s, r, k_link):
    return radial(-5.0 * k_link * (s - 1.0))


def wbc_inner(s, r, k_cyto, radius, k_core, core_radius):
    m = 0.0
    if 2 * s * r < 2 * radius:
        m += k_cyto * (1.0 - s * r / radius)
    if 2 * s * r < 2 * core_radius:
        m += k_core * (1.0 - s * r / core_radius)
    return radial(m)


def viscosity_vertex0(w, eta):
    """Force on vertex 0 when only it moves, at w e_x (no other term)."""
    per_edge = eta * abs(w) / SQ2
    if per_edge > FORCE_LIMIT / 4.0:
        return -np.sign(w) * (FORCE_LIMIT / 4.0) * (4.0 / SQ2) * AXES[0]
    return -2.0 * eta * w * AXES[0]
