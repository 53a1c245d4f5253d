"""The C ABI library on the CPU: it loads, exports exactly what include/hemocell_gpu.h
declares, and its host-only functions (mesh construction, unit conversion) agree
with the oracle and with closed forms.  No device call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

import cases
import hemocell_amd as hc
import mesh_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "hemocell_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hc_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(hc.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_exactly_the_header():
    assert sorted(hc.SIGNATURES) == declared_functions()


@pytest.mark.parametrize("shape,r,n,a", [(1, 7.82, 600, 0.3), (6, 2.0, 66, 0.25), (0, 8.0, 600, 0.3),
                                         (0, 8.0, 2000, 0.3), (1, 5.0, 100, 0.3)])
def test_product_mesh_equals_oracle_mesh(shape, r, n, a):
    v, t = hc.build_mesh(shape, r, n, a)
    vo, to = mesh_oracle.build_mesh(shape, r, n, a)
    assert np.array_equal(t, to)
    assert np.array_equal(v, vo)  # same arithmetic order: bit-exact


def test_build_mesh_rejects_bad_input():
    with pytest.raises(hc.HcError):
        hc.build_mesh(3, 7.82, 600)
    with pytest.raises(hc.HcError):
        hc.build_mesh(1, -1.0, 600)


def test_unit_conversion_closed_forms():
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    ref = cases.units(5e-7, 1e-7)
    # values of mechanics/constantConversion.cpp:36-59 at dx = 0.5 um, dt = 1e-7 s
    assert np.isclose(u.tau, 1.82, rtol=1e-14)
    assert np.isclose(u.df, 6.40625e-9, rtol=1e-14)
    assert np.isclose(u.kBT_lbm, 1.2801658977e-6, rtol=1e-10)
    assert np.isclose(u.f_limit, 50e-12 / 6.40625e-9, rtol=1e-14)
    for k in ("tau", "nu_lbm", "dm", "df", "f_limit", "kBT_lbm"):
        assert np.isclose(getattr(u, k), getattr(ref, k), rtol=1e-15)
    u1 = hc.lbm_base_parameters(5e-7, -1.0)  # tau = 1 branch
    assert u1.tau == 1.0 and np.isclose(u1.dt, (1 / 6) / 1.1e-6 * 25e-14, rtol=1e-14)


def test_model_constants_closed_forms():
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    p = hc.model_constants(hc.MODEL_RBC, 1280, u, 15.0, 80.0, 20.0, 5.0)
    # k_link = kLink kBT/(7.5e-9/dx); k_bend = kBend kBT/(5e-7/dx); area/volume x 1280/#triangles
    assert np.isclose(p.k_link, 1.280166e-3, rtol=1e-6)
    assert np.isclose(p.k_bend, 1.024133e-4, rtol=1e-6)
    assert np.isclose(p.k_volume, 2.560332e-5, rtol=1e-6)
    assert np.isclose(p.k_area, 6.400829e-6, rtol=1e-6)
    w = hc.model_constants(hc.MODEL_WBC, 5120, u, 60.0, 200.0, 20.0, 20.0, 1e-9, 6.40625e-12, 6.40625e-15,
                           2.5e-6, 4e-6)
    assert np.isclose(w.k_inner_rigid, 1e-3, rtol=1e-14) and np.isclose(w.k_cytoskeleton, 1e-6, rtol=1e-14)
    assert np.isclose(w.k_volume, 20.0 * 0.25 * u.kBT_lbm, rtol=1e-14)
    assert np.isclose(w.core_radius, 5.0, rtol=1e-14) and np.isclose(w.radius, 8.0, rtol=1e-14)
    ref = cases.params(cases.WBC_MAT, 5120, cases.units(), 2)
    for k in ("k_link", "k_bend", "k_volume", "k_area", "eta_m", "k_inner_rigid", "k_cytoskeleton"):
        assert np.isclose(getattr(w, k), getattr(ref, k), rtol=1e-14)


def test_suspension_packing_is_seeded_and_non_overlapping():
    from scipy.spatial import cKDTree
    from hemocell_amd import suspension as susp
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    types = susp.cell_types(u)
    shape = (64, 64, 128)
    a = susp.pack(shape, types, seed=3)
    b = susp.pack(shape, types, seed=3)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert abs(susp.hematocrit(types, a, shape) - 0.30) < 0.01
    allv = np.concatenate([v.reshape(-1, 3) for v in a.values()])
    cid = np.concatenate([np.repeat(np.arange(len(v)) + 100000 * k, v.shape[1]) for k, v in enumerate(a.values())])
    tree = cKDTree(allv % np.array(shape), boxsize=shape)
    pairs = tree.query_pairs(0.05, output_type="ndarray")
    assert np.all(cid[pairs[:, 0]] == cid[pairs[:, 1]])
