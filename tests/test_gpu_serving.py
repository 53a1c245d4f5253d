"""The serving loop (hemocell_amd/serving.py) on the GPU, with the active cases
stepped together (hc_step_batch) or each on its own: a case gives the same
result whether it runs alone or inside a stream of other cases -- as close as
two runs of it alone are to each other (the spread's float atomics commit in
no fixed order, so repeated runs agree to round-off, not bit for bit) -- and
every case of a stream completes."""
import numpy as np
import pytest

import cases
import hemocell_amd as hc
from hemocell_amd import serving

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batched", [True, False])
def test_alone_and_in_a_mix_are_identical(batched):
    print("[serving] batched", batched, flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = serving.make_cases(6, rate=200.0, seed=7, edge=(32, 40), steps=(150, 400))
    kinds = {(c.kind, c.cell) for c in stream}
    assert len(kinds) >= 3
    done, wall = serving.serve(stream, u, max_active=6, chunk=64, keep_vertices=True, batched=batched)
    assert sorted(c.ident for c in done) == list(range(6))
    for c in done:
        alone, x = serving.run_alone(c, u)
        _, x2 = serving.run_alone(c, u)
        noise = cases.rel_l2(x2, x)
        mix = cases.rel_l2(c.result["x"], x)
        assert mix < 1e-9 and mix <= 100 * max(noise, 1e-13), (c, mix, noise)
        assert abs(c.result["volume"] - alone["volume"]) <= 1e-9 * alone["volume"]
        assert c.latency > 0 and c.finished >= c.started
    st = serving.latency_stats(done, wall)
    assert st["cases"] == 6 and st["p99_s"] >= st["p50_s"] > 0
