"""hemocell_amd -- MI355X-native HemoCell IB-LBM time step.

The compute path is libhcgpu.so (hand-written HIP kernels for gfx950 behind the
C ABI declared in include/hemocell_gpu.h).  This module is only a ctypes
binding over that ABI, for Python drivers, tests and the benchmark; it never
computes anything itself and raises if the shared library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhcgpu.so")

HC_OK, HC_INVALID_ARG, HC_OOM, HC_DEVICE, HC_NUMERIC, HC_STATE = range(6)
HC_FP64, HC_FP32_SHIFTED = 0, 1
NODE_FLUID, NODE_BOUNCEBACK, NODE_VELOCITY = 0, 1, 2
MODEL_RBC, MODEL_PLT, MODEL_WBC = 0, 1, 2
SHAPE_WBC_SPHERE, SHAPE_RBC_FROM_SPHERE, SHAPE_ELLIPSOID_FROM_SPHERE = 0, 1, 6
PROF_SPREAD, PROF_COLLIDE, PROF_INTERP, PROF_MECH, PROF_REPULSION, PROF_STEP = range(6)

_STATUS = {1: "invalid argument", 2: "out of device memory", 3: "device error", 4: "non-finite values",
           5: "invalid state"}


class HcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (_STATUS.get(status, "error %d" % status), msg))
        self.status = status


class ModelParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("k_link", "k_bend", "k_volume", "k_area", "eta_m",
                                               "k_inner_rigid", "k_cytoskeleton", "core_radius", "radius")]


class CaseSpec(ctypes.Structure):
    """hc_case_spec: one single-cell case for the request loop (lattice units)."""
    _fields_ = [("kind", ctypes.c_int), ("edge", ctypes.c_int), ("steps", ctypes.c_int64),
                ("precision", ctypes.c_int), ("tau", ctypes.c_double), ("f_limit", ctypes.c_double),
                ("wall_u", ctypes.c_double), ("stretch_force", ctypes.c_double), ("stretch_n", ctypes.c_int),
                ("model", ctypes.c_int), ("nv", ctypes.c_int), ("nt", ctypes.c_int), ("n_inner", ctypes.c_int),
                ("vertices", ctypes.POINTER(ctypes.c_double)), ("triangles", ctypes.POINTER(ctypes.c_int32)),
                ("inner_edges", ctypes.POINTER(ctypes.c_int32)), ("params", ModelParams)]


class Units(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dx", "dt", "nu_p", "rho_p", "kBT_p",
                                               "tau", "nu_lbm", "dm", "df", "f_limit", "kBT_lbm")]


class Material(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("kLink", "kBend", "kVolume", "kArea", "eta_m",
                                               "kInnerRigid", "kCytoskeleton", "coreRadius", "radius")]


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_INT = ctypes.c_int
_ST = ctypes.c_int

# name: (restype, argtypes) -- exactly the functions of include/hemocell_gpu.h
SIGNATURES = {
    "hc_lbm_base_parameters": (_ST, [ctypes.POINTER(Units)]),
    "hc_model_constants": (_ST, [_INT, ctypes.POINTER(Material), _INT, ctypes.POINTER(Units),
                                 ctypes.POINTER(ModelParams)]),
    "hc_create": (_ST, [_INT, ctypes.POINTER(_P)]),
    "hc_destroy": (None, [_P]),
    "hc_last_error": (ctypes.c_char_p, [_P]),
    "hc_lattice_init": (_ST, [_P, _INT, _INT, _INT, _U8, ctypes.c_double, ctypes.POINTER(_INT), _INT]),
    "hc_set_wall_velocity": (_ST, [_P, _INT, _D]),
    "hc_set_body_force": (_ST, [_P, _D]),
    "hc_init_equilibrium": (_ST, [_P, ctypes.c_double, _D]),
    "hc_build_mesh": (_ST, [_INT, ctypes.c_double, _INT, ctypes.c_double, ctypes.POINTER(_INT),
                            ctypes.POINTER(_INT), _D, _I32]),
    "hc_add_cell_type": (_ST, [_P, _INT, _INT, _INT, _I32, _D, _INT, _I32, ctypes.POINTER(ModelParams), _INT,
                               ctypes.POINTER(_INT)]),
    "hc_add_cells": (_ST, [_P, _INT, _INT, _D, _I32]),
    "hc_set_velocity_timescale": (_ST, [_P, _INT]),
    "hc_set_material_timescale": (_ST, [_P, _INT, _INT]),
    "hc_set_force_limit": (_ST, [_P, ctypes.c_double]),
    "hc_set_exact_spread": (_ST, [_P, ctypes.c_int]),
    "hc_get_removed_vertices": (_ST, [_P, ctypes.c_void_p]),
    "hc_set_removed_vertices": (_ST, [_P, ctypes.c_void_p]),
    "hc_get_incomplete_cells": (_ST, [_P, ctypes.c_void_p]),
    "hc_set_repulsion": (_ST, [_P, ctypes.c_double, ctypes.c_double, _INT]),
    "hc_set_extra_vertex_force": (_ST, [_P, _INT, _I64, _D]),
    "hc_step": (_ST, [_P, _INT]),
    "hc_synchronize": (_ST, [_P]),
    "hc_iteration": (ctypes.c_int64, [_P]),
    "hc_set_iteration": (_ST, [_P, ctypes.c_int64]),
    "hc_autotune": (_ST, [_P, _INT, ctypes.c_char_p, _INT]),
    "hc_stage_spread": (_ST, [_P]),
    "hc_stage_collide_stream": (_ST, [_P]),
    "hc_stage_interpolate": (_ST, [_P]),
    "hc_stage_advance": (_ST, [_P]),
    "hc_stage_mechanics": (_ST, [_P, _INT]),
    "hc_stage_repulsion": (_ST, [_P]),
    "hc_busy": (_INT, [_P]),
    "hc_stage_end": (_ST, [_P]),
    "hc_num_nodes": (ctypes.c_int64, [_P]),
    "hc_num_vertices": (ctypes.c_int64, [_P]),
    "hc_num_cells": (_INT, [_P]),
    "hc_get_populations": (_ST, [_P, _D]),
    "hc_set_populations": (_ST, [_P, _D]),
    "hc_get_force_field": (_ST, [_P, _D]),
    "hc_get_fluid": (_ST, [_P, _D, _D]),
    "hc_get_vertices": (_ST, [_P, _D, _D, _D]),
    "hc_set_vertices": (_ST, [_P, _D, _D, _D]),
    "hc_get_repulsion_force": (_ST, [_P, _D]),
    "hc_get_boundary_repulsion_force": (_ST, [_P, _D]),
    "hc_set_boundary_repulsion": (_ST, [_P, ctypes.c_double, ctypes.c_double, _INT]),
    "hc_stage_boundary_repulsion": (_ST, [_P]),
    "hc_set_repulsion_force": (_ST, [_P, _D]),
    "hc_step_batch": (_ST, [ctypes.POINTER(_P), _INT, _INT]),
    "hc_server_create": (_ST, [_INT, _INT, ctypes.c_double, ctypes.POINTER(_P)]),
    "hc_server_destroy": (None, [_P]),
    "hc_server_last_error": (ctypes.c_char_p, [_P]),
    "hc_case_submit": (_ST, [_P, ctypes.POINTER(CaseSpec), ctypes.POINTER(ctypes.c_int64)]),
    "hc_server_poll": (_ST, [_P, _INT, ctypes.POINTER(_INT)]),
    "hc_case_status": (_INT, [_P, ctypes.c_int64]),
    "hc_case_result": (_ST, [_P, ctypes.c_int64, _D, _D]),
    "hc_server_inject_failure": (_ST, [_P, ctypes.c_int64, _INT]),
    "hc_get_cell_stats_noalloc": (_ST, [_P, _D, ctypes.c_int64]),
    "hc_set_periodicity": (_ST, [_P, ctypes.POINTER(_INT)]),
    "hc_set_lees_edwards": (_ST, [_P, ctypes.c_double, ctypes.c_double]),
    "hc_stage_lees_edwards": (_ST, [_P]),
    "hc_set_interior_viscosity": (_ST, [_P, _INT, ctypes.c_double]),
    "hc_set_interior_viscosity_timescales": (_ST, [_P, _INT, _INT]),
    "hc_get_interior_viscosity": (_ST, [_P, _D]),
    "hc_set_interior_viscosity_field": (_ST, [_P, _D]),
    "hc_stage_interior_viscosity": (_ST, [_P, _INT]),
    "hc_get_cells": (_ST, [_P, _I32, _I32]),
    "hc_get_vertex_bins": (_ST, [_P, _I64]),
    "hc_get_cell_volume_area": (_ST, [_P, _D, _D]),
    "hc_get_cell_stats": (_ST, [_P, _D, _D, _D, _D, _D, _D]),
    "hc_get_fluid_stats": (_ST, [_P, _INT, _D]),
    "hc_get_force_terms": (_ST, [_P, _D, _D, _D, _D, _D, _D]),
    "hc_wall_hits": (ctypes.c_int64, [_P]),
    "hc_check_finite": (_ST, [_P]),
    "hc_profile_enable": (_ST, [_P, _INT]),
    "hc_profile_read": (_ST, [_P, _INT, _D, _I64]),
}

_lib = None


def library():
    """Load libhcgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhcgpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def lbm_base_parameters(dx, dt, nu_p=1.1e-6, rho_p=1025.0, kBT=4.100531391e-21):
    """SI -> lattice conversion (constantConversion.cpp:36-59); dt < 0 selects tau = 1."""
    u = Units(dx=dx, dt=dt, nu_p=nu_p, rho_p=rho_p, kBT_p=kBT)
    st = library().hc_lbm_base_parameters(ctypes.byref(u))
    if st:
        raise HcError(st, "bad unit inputs")
    return u


def model_constants(model, n_triangles, units, kLink, kBend, kVolume, kArea, eta_m=0.0, kInnerRigid=0.0,
                    kCytoskeleton=0.0, coreRadius=0.0, radius=0.0):
    m = Material(kLink=kLink, kBend=kBend, kVolume=kVolume, kArea=kArea, eta_m=eta_m, kInnerRigid=kInnerRigid,
                 kCytoskeleton=kCytoskeleton, coreRadius=coreRadius, radius=radius)
    out = ModelParams()
    st = library().hc_model_constants(model, ctypes.byref(m), n_triangles, ctypes.byref(units), ctypes.byref(out))
    if st:
        raise HcError(st, "bad material inputs")
    return out


def build_mesh(shape, radius_lu, min_triangles, aspect_ratio=0.3):
    """Rest mesh of a cell type: (vertices [nv,3] float64, triangles [nt,3] int32)."""
    lib = library()
    nv, nt = _INT(), _INT()
    st = lib.hc_build_mesh(shape, radius_lu, min_triangles, aspect_ratio, ctypes.byref(nv), ctypes.byref(nt),
                           None, None)
    if st:
        raise HcError(st, "bad mesh arguments")
    v = np.zeros((nv.value, 3), dtype=np.float64)
    t = np.zeros((nt.value, 3), dtype=np.int32)
    lib.hc_build_mesh(shape, radius_lu, min_triangles, aspect_ratio, ctypes.byref(nv), ctypes.byref(nt),
                      _dp(v), t.ctypes.data_as(_I32))
    return v, t


CASE_SHEAR, CASE_STRETCH = 0, 1
CASE_QUEUED, CASE_RUNNING, CASE_DONE, CASE_FAILED = 0, 1, 2, 3
INJECT_NONE, INJECT_RETIRE, INJECT_STREAM_ERROR = 0, 1, 2
CELL_STATS_QUANTITIES, CELL_STATS_NOALLOC_MAX = 15, 262144


class Server:
    """The native request loop (hc_server_*): submit single-cell cases, poll to
    admit, step (batched) and retire them, read the finished ones."""

    def __init__(self, device=0, max_active=32, hbm_budget=64 << 30):
        self.lib = library()
        h = _P()
        st = self.lib.hc_server_create(device, max_active, float(hbm_budget), ctypes.byref(h))
        if st:
            raise HcError(st, "hc_server_create failed")
        self.h = h
        self._nv = {}

    def _chk(self, st):
        if st:
            raise HcError(st, self.lib.hc_server_last_error(self.h).decode())

    def submit(self, kind, edge, steps, model, vertices, triangles, params, tau, f_limit, wall_u=0.0,
               stretch_force=0.0, stretch_n=7, inner_edges=None, precision=HC_FP32_SHIFTED):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        t = np.ascontiguousarray(triangles, dtype=np.int32)
        ie = None if inner_edges is None else np.ascontiguousarray(inner_edges, dtype=np.int32)
        spec = CaseSpec(kind=kind, edge=edge, steps=steps, precision=precision, tau=tau, f_limit=f_limit,
                        wall_u=wall_u, stretch_force=stretch_force, stretch_n=stretch_n, model=model,
                        nv=len(v), nt=len(t), n_inner=0 if ie is None else len(ie),
                        vertices=v.ctypes.data_as(_D), triangles=t.ctypes.data_as(_I32),
                        inner_edges=None if ie is None else ie.ctypes.data_as(_I32), params=params)
        cid = ctypes.c_int64()
        self._chk(self.lib.hc_case_submit(self.h, ctypes.byref(spec), ctypes.byref(cid)))
        self._nv[cid.value] = len(v)
        return cid.value

    def poll(self, chunk=250):
        n = _INT()
        self._chk(self.lib.hc_server_poll(self.h, int(chunk), ctypes.byref(n)))
        return n.value

    def status(self, case_id):
        return self.lib.hc_case_status(self.h, case_id)

    def inject_failure(self, case_id, what):
        """Test hook: the case fails at its next poll, as a failed read-back at
        retirement (INJECT_RETIRE) or a stream error (INJECT_STREAM_ERROR)."""
        self._chk(self.lib.hc_server_inject_failure(self.h, case_id, int(what)))

    def error(self):
        return self.lib.hc_server_last_error(self.h).decode()

    def result(self, case_id):
        x = np.zeros((self._nv[case_id], 3))
        st = np.zeros(3)
        self._chk(self.lib.hc_case_result(self.h, case_id, _dp(x), _dp(st)))
        return x, dict(volume=st[0], area=st[1], stretch=st[2])

    def close(self):
        if self.h:
            self.lib.hc_server_destroy(self.h)
            self.h = None


def step_batch(sims, n=1):
    """n steps of every simulation, stepped together (hc_step_batch): a few kernel
    launches per step for all of them; asynchronous like Simulation.step."""
    sims = list(sims)
    if not sims:
        return
    lib = sims[0].lib
    arr = (_P * len(sims))(*[s.h for s in sims])
    st = lib.hc_step_batch(arr, len(sims), int(n))
    if st:
        msgs = [s.lib.hc_last_error(s.h).decode() for s in sims]
        raise HcError(st, next((m for m in msgs if m), "hc_step_batch failed"))


class Simulation:
    """One device context: lattice, cell types, cells, and the IB-LBM step."""

    def __init__(self, device=0):
        self.lib = library()
        h = _P()
        st = self.lib.hc_create(device, ctypes.byref(h))
        if st:
            raise HcError(st, "hc_create(%d) failed" % device)
        self.h = h
        self.N = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.hc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st:
            raise HcError(st, self.lib.hc_last_error(self.h).decode())

    # lattice ---------------------------------------------------------------
    def lattice_init(self, nx, ny, nz, tau, periodic=(1, 1, 1), node_type=None, precision=HC_FP64):
        per = (_INT * 3)(*[int(p) for p in periodic])
        nt = None
        if node_type is not None:
            self._types = np.ascontiguousarray(node_type, dtype=np.uint8).reshape(-1)
            if self._types.size != nx * ny * nz:
                raise ValueError("node_type must have nx*ny*nz entries")
            nt = self._types.ctypes.data_as(_U8)
        self._chk(self.lib.hc_lattice_init(self.h, nx, ny, nz, nt, tau, per, precision))
        self.shape = (nx, ny, nz)
        self.N = nx * ny * nz
        self.precision = precision

    def set_wall_velocity(self, face, u):
        self._chk(self.lib.hc_set_wall_velocity(self.h, face, _dp(np.asarray(u, dtype=np.float64))))

    def set_body_force(self, g):
        self._chk(self.lib.hc_set_body_force(self.h, _dp(np.asarray(g, dtype=np.float64))))

    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        self._chk(self.lib.hc_init_equilibrium(self.h, rho, _dp(np.asarray(u, dtype=np.float64))))

    # cells -------------------------------------------------------------------
    def add_cell_type(self, model, vertices, triangles, params, inner_edges=None, timescale=1):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        t = np.ascontiguousarray(triangles, dtype=np.int32)
        ie = None if inner_edges is None or len(inner_edges) == 0 else np.ascontiguousarray(inner_edges, np.int32)
        tid = _INT()
        self._chk(self.lib.hc_add_cell_type(self.h, model, v.shape[0], t.shape[0], t.ctypes.data_as(_I32), _dp(v),
                                            0 if ie is None else ie.shape[0],
                                            None if ie is None else ie.ctypes.data_as(_I32),
                                            ctypes.byref(params), timescale, ctypes.byref(tid)))
        return tid.value

    def add_cells(self, type_id, xyz, cell_ids=None):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        ncells = x.shape[0] if x.ndim == 3 else 1
        ids = None
        if cell_ids is not None:
            ids = np.ascontiguousarray(cell_ids, dtype=np.int32)
            if ids.shape != (ncells,):
                raise ValueError("one cell id per cell")
            ids = ids.ctypes.data_as(_I32)
        self._chk(self.lib.hc_add_cells(self.h, type_id, ncells, _dp(x), ids))

    def set_velocity_timescale(self, ts):
        self._chk(self.lib.hc_set_velocity_timescale(self.h, ts))

    def set_material_timescale(self, type_id, ts):
        self._chk(self.lib.hc_set_material_timescale(self.h, type_id, ts))

    def set_force_limit(self, fl):
        self._chk(self.lib.hc_set_force_limit(self.h, fl))

    def set_exact_spread(self, on=True):
        """FP64 only, after lattice_init and set_force_limit: spread contributions
        rounded to a power-of-two quantum so repeated runs agree bit for bit."""
        self._chk(self.lib.hc_set_exact_spread(self.h, 1 if on else 0))

    def set_repulsion(self, k_rep, cutoff_lu, timescale=1):
        self._chk(self.lib.hc_set_repulsion(self.h, k_rep, cutoff_lu, timescale))

    def set_boundary_repulsion(self, k_rep, cutoff_lu, timescale=1):
        self._chk(self.lib.hc_set_boundary_repulsion(self.h, k_rep, cutoff_lu, timescale))

    def boundary_repulsion_force(self):
        fr = np.zeros((self.num_vertices, 3))
        self._chk(self.lib.hc_get_boundary_repulsion_force(self.h, _dp(fr)))
        return fr

    def set_periodicity(self, periodic):
        per = (_INT * 3)(*[int(p) for p in periodic])
        self._chk(self.lib.hc_set_periodicity(self.h, per))

    def set_lees_edwards(self, shear_rate, displacement_per_step):
        """Lees-Edwards planes z = 0 and z = nz-1 driven at -+(nz-1) shear_rate / 2."""
        self._chk(self.lib.hc_set_lees_edwards(self.h, float(shear_rate), float(displacement_per_step)))

    def set_interior_viscosity(self, type_id, tau_inner):
        """Fluid nodes inside the membranes of type_id relax with tau_inner (<= 0: off)."""
        self._chk(self.lib.hc_set_interior_viscosity(self.h, type_id, float(tau_inner)))

    def set_interior_viscosity_timescales(self, membrane, entire_grid):
        self._chk(self.lib.hc_set_interior_viscosity_timescales(self.h, int(membrane), int(entire_grid)))

    def interior_viscosity(self):
        """Relaxation time per node, shape (nx, ny, nz): 0 where the background applies."""
        t = np.zeros(self.shape)
        self._chk(self.lib.hc_get_interior_viscosity(self.h, _dp(t)))
        return t

    def set_interior_viscosity_field(self, tau):
        t = np.ascontiguousarray(tau, dtype=np.float64)
        if t.size != int(np.prod(self.shape)):
            raise ValueError("one value per node expected")
        self._chk(self.lib.hc_set_interior_viscosity_field(self.h, _dp(t)))

    def set_extra_vertex_force(self, ids, forces):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        f = np.ascontiguousarray(forces, dtype=np.float64)
        self._chk(self.lib.hc_set_extra_vertex_force(self.h, ids.size, ids.ctypes.data_as(_I64), _dp(f)))

    # stepping ----------------------------------------------------------------
    def step(self, n=1):
        self._chk(self.lib.hc_step(self.h, n))

    def synchronize(self):
        self._chk(self.lib.hc_synchronize(self.h))

    def busy(self):
        r = self.lib.hc_busy(self.h)
        if r < 0:
            raise HcError(HC_DEVICE, self.lib.hc_last_error(self.h).decode())
        return r == 1

    def stage(self, name, *args):
        self._chk(getattr(self.lib, "hc_stage_" + name)(self.h, *args))

    @property
    def iteration(self):
        return self.lib.hc_iteration(self.h)

    def autotune(self, steps_per_trial=3):
        """Time the per-cell kernels' launch shapes on real steps and keep the fastest;
        returns the timings (ms per step) and the choices."""
        import json
        buf = ctypes.create_string_buffer(4096)
        self._chk(self.lib.hc_autotune(self.h, int(steps_per_trial), buf, len(buf)))
        return json.loads(buf.value.decode() or "{}")

    def set_iteration(self, it):
        self._chk(self.lib.hc_set_iteration(self.h, int(it)))

    # state -------------------------------------------------------------------
    @property
    def num_vertices(self):
        return self.lib.hc_num_vertices(self.h)

    def populations(self):
        f = np.zeros(19 * self.N)
        self._chk(self.lib.hc_get_populations(self.h, _dp(f)))
        return f.reshape(19, self.N)

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        self._chk(self.lib.hc_set_populations(self.h, _dp(f)))

    def force_field(self):
        F = np.zeros(3 * self.N)
        self._chk(self.lib.hc_get_force_field(self.h, _dp(F)))
        return F.reshape(3, self.N)

    def fluid(self, velocity=True):
        """(rho, u) per node; u is None with velocity=False (only rho is read back)."""
        rho = np.zeros(self.N)
        u = np.zeros(3 * self.N) if velocity else None
        self._chk(self.lib.hc_get_fluid(self.h, _dp(rho), _dp(u) if velocity else None))
        return rho, (u.reshape(self.N, 3) if velocity else None)

    def vertices(self):
        nv = self.num_vertices
        x, v, f = np.zeros((nv, 3)), np.zeros((nv, 3)), np.zeros((nv, 3))
        self._chk(self.lib.hc_get_vertices(self.h, _dp(x), _dp(v), _dp(f)))
        return x, v, f

    def set_vertices(self, x=None, v=None, f=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, v, f)]
        self._chk(self.lib.hc_set_vertices(self.h, *[_dp(a) for a in arrs]))

    def repulsion_force(self):
        fr = np.zeros((self.num_vertices, 3))
        self._chk(self.lib.hc_get_repulsion_force(self.h, _dp(fr)))
        return fr

    def set_repulsion_force(self, fr):
        fr = np.ascontiguousarray(fr, dtype=np.float64)
        self._chk(self.lib.hc_set_repulsion_force(self.h, _dp(fr)))

    def removed_vertices(self):
        """One bool per vertex: removed by wall contact."""
        r = np.zeros(self.num_vertices, dtype=np.uint8)
        self._chk(self.lib.hc_get_removed_vertices(self.h, r.ctypes.data))
        return r.astype(bool)

    def set_removed_vertices(self, removed):
        r = np.ascontiguousarray(np.asarray(removed, dtype=bool).astype(np.uint8))
        assert r.shape == (self.num_vertices,)
        self._chk(self.lib.hc_set_removed_vertices(self.h, r.ctypes.data))

    def incomplete_cells(self):
        """One bool per cell (cell order): a vertex of it was removed by wall contact."""
        r = np.zeros(self.lib.hc_num_cells(self.h), dtype=np.uint8)
        self._chk(self.lib.hc_get_incomplete_cells(self.h, r.ctypes.data))
        return r.astype(bool)

    def cells(self):
        """(type id, cell id) of every cell, in cell order."""
        nc = self.lib.hc_num_cells(self.h)
        t = np.zeros(nc, dtype=np.int32)
        c = np.zeros(nc, dtype=np.int32)
        self._chk(self.lib.hc_get_cells(self.h, t.ctypes.data_as(_I32), c.ctypes.data_as(_I32)))
        return t, c

    def vertex_bins(self):
        b = np.zeros(self.num_vertices, dtype=np.int64)
        self._chk(self.lib.hc_get_vertex_bins(self.h, b.ctypes.data_as(_I64)))
        return b

    def cell_volume_area(self):
        nc = self.lib.hc_num_cells(self.h)
        vol, area = np.zeros(nc), np.zeros(nc)
        self._chk(self.lib.hc_get_cell_volume_area(self.h, _dp(vol), _dp(area)))
        return vol, area

    def cell_stats(self):
        """Per-cell dict of arrays: volume, area, centroid, velocity, bbox, stretch (lattice units)."""
        nc = self.lib.hc_num_cells(self.h)
        out = dict(volume=np.zeros(nc), area=np.zeros(nc), centroid=np.zeros((nc, 3)), velocity=np.zeros((nc, 3)),
                   bbox=np.zeros((nc, 6)), stretch=np.zeros(nc))
        self._chk(self.lib.hc_get_cell_stats(self.h, *[_dp(out[k]) for k in
                                                       ("volume", "area", "centroid", "velocity", "bbox", "stretch")]))
        return out

    def cell_stats_noalloc(self, out=None):
        """hc_get_cell_stats_noalloc: the statistics of cell_stats as one (ncells, 15)
        array (volume, area, centroid[3], velocity[3], bbox[6], stretch), staged on the
        device without allocating; `out` may be a preallocated C-contiguous array."""
        nc = self.lib.hc_num_cells(self.h)
        if out is None:
            out = np.zeros((nc, CELL_STATS_QUANTITIES))
        if out.dtype != np.float64 or not out.flags["C_CONTIGUOUS"] or out.ndim != 2 or \
                out.shape[1] != CELL_STATS_QUANTITIES:
            raise ValueError("out must be a C-contiguous float64 array of shape (cells, 15)")
        self._chk(self.lib.hc_get_cell_stats_noalloc(self.h, _dp(out), out.shape[0]))
        return out

    def force_terms(self):
        """Per-vertex force terms at the current state: dict of (nv, 3) arrays
        area, volume, bending, link, visc, inner_link."""
        names = ("area", "volume", "bending", "link", "visc", "inner_link")
        out = {k: np.zeros((self.num_vertices, 3)) for k in names}
        self._chk(self.lib.hc_get_force_terms(self.h, *[_dp(out[k]) for k in names]))
        return out

    def fluid_stats(self, what="velocity"):
        """FluidInfo statistics: dict(min, max, avg, ncells) of |u| ("velocity") or of the
        spread vertex forces |F| ("force") over the nodes that are not bounce-back."""
        out = np.zeros(4)
        self._chk(self.lib.hc_get_fluid_stats(self.h, {"velocity": 0, "force": 1}[what], _dp(out)))
        return dict(min=out[0], max=out[1], avg=out[2], ncells=int(out[3]))

    def wall_hits(self):
        return self.lib.hc_wall_hits(self.h)

    def check_finite(self):
        self._chk(self.lib.hc_check_finite(self.h))

    def profile_enable(self, on=True):
        self._chk(self.lib.hc_profile_enable(self.h, 1 if on else 0))

    def profile_read(self, stage):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        self._chk(self.lib.hc_profile_read(self.h, stage, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value
