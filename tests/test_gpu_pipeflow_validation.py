"""The reference's pipe-flow validation (tests/validation/pipeflow/test_pipeflow.cpp)
through the host layer, with the reference's configuration, materials and
position files (tests/golden/pipeflow/) and an analytic tube in place of the
voxelised tube.stl: the number of cells stays constant after every iteration up
to 1000 (no vertex ever lands on a wall node, which would remove it and make
its cell incomplete), and after 100 iterations, every 10th iteration, the
apparent viscosity 0.5 u_max / mean |u| lies in the reference's band (1.03,
3.0) and the mean vertex force stays below 4 pN (test_pipeflow.cpp:84-106).
The reference's count of 42 cells depends on Palabos' voxelisation of the STL
surface; the analytic tube keeps its own count (113: the .pos files' cells
whose vertices, or their periodic images across x, avoid the wall nodes),
which is reported and must stay constant."""
import json
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pipeflow_bands(tmp_path):
    src = os.path.join(ROOT, "tests", "golden", "pipeflow")
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), tmp_path / f)
        os.chmod(tmp_path / f, 0o644)
    drv = os.path.join(ROOT, "hemocell_amd", "bin", "pipeflow")
    print("[pipeflow] running", flush=True)
    r = subprocess.run([drv, "config_pipeflow.xml", "1000", "10"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print("[pipeflow]", res, flush=True)
    assert res["iteration"] == 1000 and res["checks"] == 91
    assert res["cells_loaded"] > 0 and res["cells_min"] == res["cells_max"] == res["cells_loaded"]
    assert res["removed_vertices"] == 0 and res["last_loss_iteration"] == -1
    assert abs(res["pipe_radius"] - 25.0) < 0.5
    assert 1.03 < res["viscosity_min"] and res["viscosity_max"] < 3.0
    assert res["mean_force_pN_max"] < 4.0
