"""The per-cell kernels' LDS layouts (hemocell_amd/csrc/hc_lds_plan.h), checked on
the CPU over every mesh, block size, precision, spread box and device limit the
launchers can meet: a layout the plan accepts fits the workgroup limit together
with the kernel's static LDS, the spread box and the vertex list after it stay
inside the dynamic allocation, the staged layout holds the positions,
velocities and bending forces, and a layout that cannot fit is refused (the
launcher then tries a smaller block) instead of launched.  Among the refused:
the 2562-vertex WBC with its positions in LDS on a 64 KB device."""
import ctypes
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RBC, PLT, WBC = 0, 1, 2
MARGIN = 256


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("lds") / "libldsplan.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", out,
                           os.path.join(HERE, "native", "lds_plan_shim.cpp")])
    L = ctypes.CDLL(out)
    i64 = ctypes.c_int64
    L.plan_cell.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i64, i64, i64, i64, i64, i64, ctypes.c_int,
                            ctypes.POINTER(i64)]
    L.plan_cell.restype = None
    L.plan_mixed.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int), i64, i64, i64, i64, ctypes.POINTER(ctypes.c_int)]
    L.plan_mixed.restype = i64
    L.fixed_bytes.argtypes = [ctypes.c_int, ctypes.c_int, i64, ctypes.c_int, ctypes.c_int]
    L.fixed_bytes.restype = i64
    return L


def static_lds(block):
    """Static LDS of a k_cell_step instantiation: 128 doubles of volume chunk
    sums, the volume force, vertex 0's held position, the box bounds and the
    spread's plane counters, rounded up to the 512-byte allocation granule
    (2048 B at every block size)."""
    raw = 8 * 128 + 8 + 24 + 8 * 4 + 2 * 65 * 4
    return (raw + 511) // 512 * 512


def plan(L, model, nv, visc, psize, box, dev, block, gpos):
    out = (ctypes.c_int64 * 7)()
    st = static_lds(block)
    L.plan_cell(model, nv, visc, psize, box, dev, st, st, st, gpos, out)
    ok, staged, fixed, region, lists, dyn, box_nodes = (int(v) for v in out)
    return dict(ok=bool(ok), staged=bool(staged), fixed=fixed, region=region, lists=lists, dyn=dyn,
                box_nodes=box_nodes, st=st)


MESHES = [(RBC, 642, 0), (RBC, 642, 1), (PLT, 66, 1), (WBC, 642, 1), (WBC, 2562, 1), (RBC, 2562, 0)]


@pytest.mark.parametrize("dev", [65536, 163840])
@pytest.mark.parametrize("psize", [4, 8])
def test_every_accepted_layout_fits(lib, dev, psize):
    n = 0
    for model, nv, visc in MESHES:
        for block in (64, 256, 512, 1024):
            for box in (8192, 12288, 16384, 24576, 32768, 49152):
                for gpos in (0, 1):
                    p = plan(lib, model, nv, visc, psize, box, dev, block, gpos)
                    assert p["dyn"] == p["fixed"] + p["region"] + p["lists"]
                    if not p["ok"]:
                        assert p["dyn"] + p["st"] + MARGIN > dev
                        continue
                    n += 1
                    assert p["dyn"] + p["st"] + MARGIN <= dev
                    # the region offset the kernel computes for the instantiation launched
                    kpsize = psize if p["staged"] else 8
                    assert p["fixed"] == lib.fixed_bytes(nv, visc, kpsize, int(p["staged"]), gpos)
                    # box nodes (3 doubles) inside the region, the vertex list right after them
                    assert 24 * p["box_nodes"] <= p["region"]
                    assert p["fixed"] + 24 * p["box_nodes"] + 2 * nv <= p["dyn"]
                    assert p["lists"] >= 2 * nv and p["fixed"] % 16 == 0 and p["lists"] % 16 == 0
                    if gpos:
                        assert not p["staged"] and p["fixed"] == 0 and p["region"] == box
                    elif p["staged"]:
                        assert p["fixed"] >= (6 if visc else 3) * nv * psize
                        if model != PLT:
                            assert p["region"] >= 3 * nv * psize  # bending forces
                        assert p["region"] >= box
                    else:
                        assert p["fixed"] == 24 * nv and p["region"] <= box
    assert n > 100


@pytest.mark.parametrize("psize", [4, 8])
def test_a_launch_always_finds_a_block(lib, psize):
    """The launcher tries 1024, 512, then 256 threads (64 for platelets), and
    then 256 threads with the positions in global memory: the last try fits
    every mesh at every box size on a 64 KB device, and on a 160 KB device
    every mesh fits with its positions in LDS."""
    for model, nv, visc in MESHES:
        last = 64 if model == PLT else 256
        for box in (8192, 12288, 16384, 24576):
            assert plan(lib, model, nv, visc, psize, box, 65536, 256, 1)["ok"], (model, nv, box)
            assert plan(lib, model, nv, visc, psize, box, 163840, last, 0)["ok"], (model, nv, box)
            if nv <= 642:
                assert plan(lib, model, nv, visc, psize, box, 65536, last, 0)["ok"], (model, nv, box)


def test_wbc_1024_threads_on_a_64k_device_is_refused(lib):
    p = plan(lib, WBC, 2562, 1, 8, 12288, 65536, 1024, 0)
    assert not p["ok"]  # 2562 x 24 B positions + 1 KB static + list leave no room
    assert p["fixed"] + p["lists"] + p["st"] + MARGIN > 65536
    q = plan(lib, WBC, 2562, 1, 8, 12288, 163840, 1024, 0)
    assert q["ok"] and not q["staged"] and q["box_nodes"] == 512


def test_mixed_layout(lib):
    model = (ctypes.c_int * 3)(RBC, PLT, WBC)
    nv = (ctypes.c_int * 3)(642, 66, 2562)
    visc = (ctypes.c_int * 3)(0, 1, 1)
    for dev in (65536, 163840):
        for psize in (4, 8):
            for box in (8192, 12288, 16384, 24576, 49152):
                nodes = (ctypes.c_int * 3)()
                st = static_lds(256)
                dyn = lib.plan_mixed(3, model, nv, visc, psize, box, dev, st, nodes)
                if dyn == 0:
                    assert box >= 49152 and dev == 65536
                    continue
                assert dyn % 16 == 0 and dyn + st + MARGIN <= dev
                for t in range(3):
                    gpos = int(model[t] == WBC)
                    fixed = lib.fixed_bytes(nv[t], visc[t], psize, 1 - gpos, gpos)
                    lists = (2 * nv[t] + 15) // 16 * 16
                    assert fixed + 24 * nodes[t] + 2 * nv[t] <= dyn
                    assert 24 * nodes[t] >= box - 24  # every type gets at least the requested box
                    if not gpos and model[t] != PLT:
                        assert dyn - fixed - lists >= 3 * nv[t] * psize  # bending forces
