"""The random-orientation packer (hemocell_amd/packing.py) and the packings the
GPU workloads load: every cell's mesh lies inside its sphere cover, the
packings keep their clearance, are reproducible from their seed, and their
orientations and positions are uniform (an aligned lattice packing would fail
the orientation check by far)."""
import os

import numpy as np
import pytest

import hemocell_amd as hc
from hemocell_amd import packing
from hemocell_amd import suspension as susp

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "hemocell_amd", "data")


@pytest.fixture(scope="module")
def types():
    return susp.cell_types(hc.lbm_base_parameters(5e-7, 1e-7))


def _axes(rot):
    return np.abs(rot[:, :, 2])  # each RBC's disc axis (body z) in the lab frame


def check_packing(types, pk, shape, periodic=(True, True, True)):
    assert packing.overlap_free(types, pk, shape, 0.49, periodic)
    cen, rot = pk["RBC"]
    assert np.allclose(np.einsum("nij,nkj->nik", rot, rot), np.eye(3), atol=1e-6)
    # |axis component| of a uniformly random direction: uniform on [0, 1]
    a = _axes(rot)
    tol = 4 * 0.29 / np.sqrt(len(a))
    assert np.all(np.abs(a.mean(axis=0) - 0.5) < tol), a.mean(axis=0)
    for d in range(3):
        if periodic[d]:
            frac = np.mod(cen[:, d], shape[d]) / shape[d]
            assert abs(frac.mean() - 0.5) < 4 * 0.29 / np.sqrt(len(frac))


@pytest.mark.parametrize("name", ["RBC", "PLT", "WBC"])
def test_sphere_cover_contains_the_mesh(types, name):
    v = types[name].verts - types[name].verts.mean(axis=0)
    cen, rad = packing.sphere_cover(types[name].verts, name)
    gap = np.min(np.linalg.norm(v[:, None, :] - cen[None], axis=2) - rad[None], axis=1)
    assert np.all(gap <= -0.15 + 1e-9)


def test_packing_is_seeded(types):
    shape = (40, 40, 48)
    counts = dict(RBC=18, PLT=3)
    a = packing.pack(shape, types, counts, seed=3)
    b = packing.pack(shape, types, counts, seed=3)
    c = packing.pack(shape, types, counts, seed=4)
    for k in a:
        assert np.array_equal(a[k][0], b[k][0]) and np.array_equal(a[k][1], b[k][1])
    assert not np.allclose(a["RBC"][0], c["RBC"][0])
    assert packing.overlap_free(types, a, shape, 0.49)


def test_save_load_round_trip(types, tmp_path):
    shape = (40, 40, 48)
    pk = packing.pack(shape, types, dict(RBC=12, PLT=2), seed=5)
    packing.save(tmp_path / "p.npz", pk, shape=shape, seed=5)
    back, meta = packing.load(tmp_path / "p.npz")
    assert tuple(meta["shape"]) == shape and int(meta["seed"]) == 5
    for k in pk:
        assert np.abs(back[k][0] - pk[k][0]).max() < 1e-4
        assert np.abs(back[k][1] - pk[k][1]).max() < 1e-6
    x0 = packing.instances(types, pk)["RBC"]
    x1 = packing.instances(types, back)["RBC"]
    assert np.abs(x1 - x0).max() < 1e-4


def test_crop_fixture(types):
    pk, meta = packing.load(os.path.join(HERE, "golden", "c3_crop_64x64x128.npz"))
    shape = tuple(int(v) for v in meta["shape"])
    assert shape == (64, 64, 128)
    assert {k: len(v[0]) for k, v in pk.items()} == {"RBC": 242, "PLT": 24, "WBC": 1}
    check_packing(types, pk, shape)


def test_headline_suspension_packing(types):
    path = os.path.join(DATA, "suspension_256x256x512_ht30.npz")
    pk, meta = packing.load(path)
    shape = tuple(int(v) for v in meta["shape"])
    assert shape == (256, 256, 512) and float(meta["hematocrit"]) == 0.30
    counts = {k: len(v[0]) for k, v in pk.items()}
    assert counts == susp.cell_counts(types, float(np.prod(shape)), 0.30)
    rbc_volume = susp._volume(types["RBC"].verts, types["RBC"].tris)
    assert abs(counts["RBC"] * rbc_volume / np.prod(shape) - 0.30) < 1e-3
    check_packing(types, pk, shape)


def test_tube_packing(types):
    """The Ht 20% tube of the pipe-flow workload: 244 RBC and 24 platelets, every
    sphere of their covers inside radius 30 of the tube axis (the supports stay
    off the wall nodes of the 64-node tube), clear of each other."""
    pk, meta = packing.load(os.path.join(DATA, "tube_256x64x64_ht20.npz"))
    shape = tuple(int(v) for v in meta["shape"])
    assert shape == (256, 64, 64) and int(meta["tube"]) == 1
    assert {k: len(v[0]) for k, v in pk.items()} == {"RBC": 244, "PLT": 24}
    assert packing.overlap_free(types, pk, shape, 0.49, (True, False, False))
    x = packing.instances(types, pk)
    c = (64 - 1) / 2.0
    for name in x:
        r = np.hypot(x[name][..., 1] - c, x[name][..., 2] - c)
        assert r.max() < 30.0, name
    a = _axes(pk["RBC"][1])
    assert np.all(np.abs(a.mean(axis=0) - 0.5) < 4 * 0.29 / np.sqrt(len(a)))


def test_channel_tile_packing(types):
    """The 128x512x128 tile of the largest domain (bounce-back planes y = 0 and
    511, periodic in x and z): every vertex at least 1.5 spacings from the wall
    planes, clear of each other, also once tiled 4 x 8 over 512x512x1024."""
    pk, meta = packing.load(os.path.join(DATA, "channel_tile_128x512x128_ht35.npz"))
    shape = tuple(int(v) for v in meta["shape"])
    assert shape == (128, 512, 128) and int(meta["walls"]) == 1
    assert {k: len(v[0]) for k, v in pk.items()} == {"RBC": 4507, "PLT": 451}
    assert packing.overlap_free(types, pk, shape, 0.49, (True, False, True))
    x = packing.instances(types, pk)
    for name in x:
        y = x[name][..., 1]
        assert y.min() > 1.5 and y.max() < 511 - 1.5, name
    a = _axes(pk["RBC"][1])
    assert np.all(np.abs(a.mean(axis=0) - 0.5) < 4 * 0.29 / np.sqrt(len(a)))
    full = packing.tile(pk, shape, (4, 1, 8))
    assert len(full["RBC"][0]) == 32 * 4507
    assert packing.overlap_free(types, full, (512, 512, 1024), 0.49, (True, False, True))
