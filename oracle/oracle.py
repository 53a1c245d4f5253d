"""ctypes binding of the FP64 CPU oracle (libhc_oracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  Its interface mirrors hemocell_amd.Simulation so parity tests can
drive both sides with the same calls.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhc_oracle.so")

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_L = ctypes.POINTER(ctypes.c_long)
_INT = ctypes.c_int

_SIG = {
    "ho_create": (_P, [_INT, _INT, _INT, ctypes.POINTER(ctypes.c_uint8), ctypes.c_double, ctypes.POINTER(_INT)]),
    "ho_destroy": (None, [_P]),
    "ho_set_body_force": (None, [_P, _D]),
    "ho_set_wall_velocity": (None, [_P, _INT, _D]),
    "ho_set_f_limit": (None, [_P, ctypes.c_double]),
    "ho_set_timescales": (None, [_P, _INT]),
    "ho_set_repulsion": (None, [_P, ctypes.c_double, ctypes.c_double, _INT]),
    "ho_set_iter": (None, [_P, ctypes.c_long]),
    "ho_init_equilibrium": (None, [_P, ctypes.c_double, _D]),
    "ho_add_cell_type": (_INT, [_P, _INT, _INT, _INT, _I32, _D, _INT, _I32, _D, _INT]),
    "ho_add_cells": (_INT, [_P, _INT, _INT, _D]),
    "ho_type_sizes": (None, [_P, _INT, _I32]),
    "ho_type_tables": (None, [_P, _INT, _I32, _I32, _I32, _I32, _I32, _D, _D, _D, _D, _D, _D]),
    "ho_num_threads": (_INT, []),
    "ho_num_vertices": (ctypes.c_long, [_P]),
    "ho_wall_hits": (ctypes.c_long, [_P]),
    "ho_get_removed": (None, [_P, ctypes.c_void_p]),
    "ho_set_removed": (None, [_P, ctypes.c_void_p]),
    "ho_get_populations": (None, [_P, _D]),
    "ho_set_populations": (None, [_P, _D]),
    "ho_get_force_field": (None, [_P, _D]),
    "ho_set_force_field": (None, [_P, _D]),
    "ho_get_fluid": (None, [_P, _D, _D]),
    "ho_get_vertices": (None, [_P, _D, _D, _D]),
    "ho_set_vertices": (None, [_P, _D, _D, _D]),
    "ho_get_repulsion": (None, [_P, _D]),
    "ho_stage_repulsion": (None, [_P]),
    "ho_set_boundary_repulsion": (None, [_P, ctypes.c_double, ctypes.c_double, _INT]),
    "ho_get_boundary_repulsion": (None, [_P, _D]),
    "ho_stage_boundary_repulsion": (None, [_P]),
    "ho_set_interior_viscosity": (None, [_P, _INT, ctypes.c_double]),
    "ho_set_interior_viscosity_timescales": (None, [_P, _INT, _INT]),
    "ho_get_interior_viscosity": (None, [_P, _D]),
    "ho_set_interior_viscosity_classes": (None, [_P, ctypes.POINTER(ctypes.c_uint8)]),
    "ho_stage_interior_viscosity": (None, [_P, _INT]),
    "ho_ray_crosses": (_INT, [_D, _D, _D, _L]),
    "ho_set_lees_edwards": (None, [_P, ctypes.c_double, ctypes.c_double]),
    "ho_stage_lees_edwards": (None, [_P]),
    "ho_vertex_bins": (None, [_P, ctypes.POINTER(ctypes.c_long)]),
    "ho_stage_spread": (None, [_P]),
    "ho_stage_collide_stream": (None, [_P]),
    "ho_stage_interpolate": (None, [_P]),
    "ho_stage_advance": (None, [_P]),
    "ho_stage_mechanics": (None, [_P, _INT]),
    "ho_stage_reset_force": (None, [_P]),
    "ho_step": (None, [_P, _INT]),
    "ho_ibm_kernel": (_INT, [_P, _D, _L, _D]),
    "ho_collide_one": (None, [_D, _D, ctypes.c_double]),
    "ho_velocity_bc_one": (None, [_D, _INT, _INT, _D]),
    "ho_feq": (ctypes.c_double, [_INT, ctypes.c_double, _D]),
}

_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        for k, (r, a) in _SIG.items():
            fn = getattr(lib, k)
            fn.restype = r
            fn.argtypes = a
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


# D3Q19 in Palabos order (lattice constants, not derived from the oracle library)
C = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1], [-1, 0, 1],
              [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, -1, 0], [1, 0, 1],
              [1, 0, -1], [0, 1, 1], [0, 1, -1]], dtype=np.int64)
W = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)


class OracleSim:
    def __init__(self, nx, ny, nz, tau, periodic=(1, 1, 1), node_type=None):
        self.lib = library()
        per = (_INT * 3)(*[int(p) for p in periodic])
        nt = None
        if node_type is not None:
            self._types = np.ascontiguousarray(node_type, dtype=np.uint8).reshape(-1)
            nt = self._types.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        self.h = self.lib.ho_create(nx, ny, nz, nt, tau, per)
        self.shape = (nx, ny, nz)
        self.N = nx * ny * nz
        self.type_nv = []

    def close(self):
        if self.h:
            self.lib.ho_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_body_force(self, g):
        self.lib.ho_set_body_force(self.h, _dp(np.asarray(g, dtype=np.float64)))

    def set_wall_velocity(self, face, u):
        self.lib.ho_set_wall_velocity(self.h, face, _dp(np.asarray(u, dtype=np.float64)))

    def set_force_limit(self, fl):
        self.lib.ho_set_f_limit(self.h, fl)

    def set_repulsion(self, k_rep, cutoff_lu, timescale=1):
        self.lib.ho_set_repulsion(self.h, k_rep, cutoff_lu, timescale)

    def set_boundary_repulsion(self, k_rep, cutoff_lu, timescale=1):
        self.lib.ho_set_boundary_repulsion(self.h, k_rep, cutoff_lu, timescale)

    def boundary_repulsion_force(self):
        fr = np.zeros((self.num_vertices, 3))
        self.lib.ho_get_boundary_repulsion(self.h, _dp(fr))
        return fr

    def set_iteration(self, it):
        self.lib.ho_set_iter(self.h, int(it))

    def set_lees_edwards(self, shear_rate, displacement_per_step):
        self.lib.ho_set_lees_edwards(self.h, shear_rate, displacement_per_step)

    def stage_lees_edwards(self):
        self.lib.ho_stage_lees_edwards(self.h)

    def set_interior_viscosity(self, type_id, tau_inner):
        self.lib.ho_set_interior_viscosity(self.h, type_id, tau_inner)

    def set_interior_viscosity_timescales(self, membrane, entire_grid):
        self.lib.ho_set_interior_viscosity_timescales(self.h, membrane, entire_grid)

    def interior_viscosity(self):
        """Relaxation time per node (nx, ny, nz), 0 where the background applies."""
        t = np.zeros(self.N)
        self.lib.ho_get_interior_viscosity(self.h, _dp(t))
        return t.reshape(self.shape)

    def set_interior_viscosity_classes(self, cls):
        c = np.ascontiguousarray(cls, dtype=np.uint8).reshape(-1)
        assert c.size == self.N
        self.lib.ho_set_interior_viscosity_classes(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))

    def stage_interior_viscosity(self, which):
        """0: entire-grid ray-casting pass; 1: membrane pass (support at the current positions)."""
        self.lib.ho_stage_interior_viscosity(self.h, which)

    def repulsion_force(self):
        fr = np.zeros((self.num_vertices, 3))
        self.lib.ho_get_repulsion(self.h, _dp(fr))
        return fr

    def vertex_bins(self):
        b = np.zeros(self.num_vertices, dtype=np.int64)
        self.lib.ho_vertex_bins(self.h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        return b

    def set_velocity_timescale(self, ts):
        self.lib.ho_set_timescales(self.h, ts)

    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        self.lib.ho_init_equilibrium(self.h, rho, _dp(np.asarray(u, dtype=np.float64)))

    def add_cell_type(self, model, vertices, triangles, params, inner_edges=None, timescale=1):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        t = np.ascontiguousarray(triangles, dtype=np.int32)
        ie = None if inner_edges is None or len(inner_edges) == 0 else np.ascontiguousarray(inner_edges, np.int32)
        p = np.array([params.k_link, params.k_bend, params.k_volume, params.k_area, params.eta_m,
                      params.k_inner_rigid, params.k_cytoskeleton, params.core_radius, params.radius])
        tid = self.lib.ho_add_cell_type(self.h, model, v.shape[0], t.shape[0], t.ctypes.data_as(_I32), _dp(v),
                                        0 if ie is None else ie.shape[0],
                                        None if ie is None else ie.ctypes.data_as(_I32), _dp(p), timescale)
        if tid < 0:
            raise ValueError("oracle rejected the mesh")
        self.type_nv.append(v.shape[0])
        return tid

    def add_cells(self, type_id, xyz):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        ncells = x.shape[0] if x.ndim == 3 else 1
        self.lib.ho_add_cells(self.h, type_id, ncells, _dp(x))
        if not hasattr(self, "_cell_nv"):
            self._cell_nv = []
        self._cell_nv += [x.shape[-2]] * ncells

    def type_tables(self, t):
        s = np.zeros(4, dtype=np.int32)
        self.lib.ho_type_sizes(self.h, t, s.ctypes.data_as(_I32))
        nv, nt, ne, ni = [int(q) for q in s]
        out = dict(edge=np.zeros((ne, 2), np.int32), edge_tris=np.zeros((ne, 2), np.int32),
                   edge_outer=np.zeros((ne, 2), np.int32), ring=np.zeros((nv, 6), np.int32),
                   nring=np.zeros(nv, np.int32), edge_len=np.zeros(ne), edge_angle=np.zeros(ne),
                   tri_area=np.zeros(nt), patch_dist=np.zeros(nv), inner_len=np.zeros(max(ni, 1)),
                   scalars=np.zeros(4))
        i32 = lambda a: a.ctypes.data_as(_I32)  # noqa: E731
        self.lib.ho_type_tables(self.h, t, i32(out["edge"]), i32(out["edge_tris"]), i32(out["edge_outer"]),
                                i32(out["ring"]), i32(out["nring"]), _dp(out["edge_len"]), _dp(out["edge_angle"]),
                                _dp(out["tri_area"]), _dp(out["patch_dist"]), _dp(out["inner_len"]),
                                _dp(out["scalars"]))
        out["volume_eq"], out["area_mean_eq"], out["edge_mean_eq"], out["angle_mean_eq"] = out["scalars"]
        return out

    @property
    def num_vertices(self):
        return self.lib.ho_num_vertices(self.h)

    def step(self, n=1):
        self.lib.ho_step(self.h, n)

    def stage(self, name, *args):
        getattr(self.lib, "ho_stage_" + name)(self.h, *args)

    def populations(self):
        f = np.zeros(19 * self.N)
        self.lib.ho_get_populations(self.h, _dp(f))
        return f.reshape(19, self.N)

    def set_populations(self, f):
        self.lib.ho_set_populations(self.h, _dp(np.ascontiguousarray(f, dtype=np.float64).reshape(-1)))

    def force_field(self):
        F = np.zeros(3 * self.N)
        self.lib.ho_get_force_field(self.h, _dp(F))
        return F.reshape(3, self.N)

    def set_force_field(self, F):
        self.lib.ho_set_force_field(self.h, _dp(np.ascontiguousarray(F, dtype=np.float64).reshape(-1)))

    def fluid(self):
        rho = np.zeros(self.N)
        u = np.zeros(3 * self.N)
        self.lib.ho_get_fluid(self.h, _dp(rho), _dp(u))
        return rho, u.reshape(self.N, 3)

    def vertices(self):
        nv = self.num_vertices
        x, v, f = np.zeros((nv, 3)), np.zeros((nv, 3)), np.zeros((nv, 3))
        self.lib.ho_get_vertices(self.h, _dp(x), _dp(v), _dp(f))
        return x, v, f

    def set_vertices(self, x=None, v=None, f=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, v, f)]
        self.lib.ho_set_vertices(self.h, *[_dp(a) for a in arrs])

    def wall_hits(self):
        return self.lib.ho_wall_hits(self.h)

    def removed_vertices(self):
        """One bool per vertex: removed by wall contact (advanceParticles)."""
        r = np.zeros(self.num_vertices, dtype=np.uint8)
        self.lib.ho_get_removed(self.h, r.ctypes.data)
        return r.astype(bool)

    def set_removed_vertices(self, removed):
        r = np.ascontiguousarray(np.asarray(removed, dtype=bool).astype(np.uint8))
        assert r.shape == (self.num_vertices,)
        self.lib.ho_set_removed(self.h, r.ctypes.data)

    def incomplete_cells(self):
        """One bool per cell: a vertex of it was removed."""
        r = self.removed_vertices()
        nv = getattr(self, "_cell_nv", [])
        starts = np.concatenate([[0], np.cumsum(nv)])[:-1].astype(np.int64)
        return np.array([r[s:s + n].any() for s, n in zip(starts, nv)], dtype=bool)

    def ibm_kernel(self, p):
        idx = np.zeros(27, dtype=np.int64)
        w = np.zeros(27)
        n = self.lib.ho_ibm_kernel(self.h, _dp(np.asarray(p, dtype=np.float64)), idx.ctypes.data_as(_L), _dp(w))
        return idx[:n], w[:n]


def collide_one(f, F, omega):
    f = np.array(f, dtype=np.float64)
    library().ho_collide_one(_dp(f), _dp(np.asarray(F, dtype=np.float64)), omega)
    return f


def velocity_bc_one(f, axis, orient, uw):
    f = np.array(f, dtype=np.float64)
    library().ho_velocity_bc_one(_dp(f), axis, orient, _dp(np.asarray(uw, dtype=np.float64)))
    return f


def feq(i, rho_bar, j):
    return library().ho_feq(i, rho_bar, _dp(np.asarray(j, dtype=np.float64)))


def ray_crosses(v0, v1, v2, site):
    """1 when the ray from the integer site along (-1,-1,-1) crosses triangle (v0, v1, v2)."""
    s = np.asarray(site, dtype=np.int64)
    return library().ho_ray_crosses(_dp(np.asarray(v0, dtype=np.float64)), _dp(np.asarray(v1, dtype=np.float64)),
                                    _dp(np.asarray(v2, dtype=np.float64)), s.ctypes.data_as(_L))
