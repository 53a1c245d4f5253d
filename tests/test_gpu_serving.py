"""The serving loop (hemocell_amd/serving.py) on the GPU, with the active cases
stepped together (hc_step_batch) or each on its own: a case gives the same
result whether it runs alone or inside a stream of other cases -- as close as
two runs of it alone are to each other (the spread's float atomics commit in
no fixed order, so repeated runs agree to round-off, not bit for bit) -- and
every case of a stream completes."""
import time

import numpy as np
import pytest

import cases
import hemocell_amd as hc
from hemocell_amd import serving
from oracle import OracleSim

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


def test_native_request_loop_matches_alone():
    """The library's request loop (hc_server_*): every case of a stream finishes,
    each as close to the case run alone as two lone runs are to each other."""
    print("[serving] native", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = serving.make_cases(5, rate=500.0, seed=11, edge=(24, 36), steps=(120, 300))
    done, wall = serving.serve_native(stream, u, max_active=3, chunk=50, keep_vertices=True)
    assert sorted(c.ident for c in done) == list(range(5))
    for c in done:
        alone, x = serving.run_alone(c, u)
        _, x2 = serving.run_alone(c, u)
        noise = cases.rel_l2(x2, x)
        mix = cases.rel_l2(c.result["x"], x)
        assert mix < 1e-9 and mix <= 100 * max(noise, 1e-13), (c, mix, noise)
        assert abs(c.result["volume"] - alone["volume"]) <= 1e-9 * alone["volume"]
    st = serving.latency_stats(done, wall)
    assert st["cases"] == 5 and st["p99_s"] >= st["p50_s"] > 0


def test_native_server_arguments():
    print("[serving] native arguments", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    s = hc.Server(0, 2, 1 << 30)
    c = serving.make_cases(1, rate=1.0, seed=3)[0]
    kw = serving.spec(c, u)
    with pytest.raises(hc.HcError):
        s.submit(**dict(kw, edge=2))
    cid = s.submit(**dict(kw, steps=3))
    assert s.status(cid) == hc.CASE_QUEUED and s.status(cid + 99) == -1
    with pytest.raises(hc.HcError):
        s.result(cid)
    while s.status(cid) != hc.CASE_DONE:
        s.poll(10)
    x, st = s.result(cid)
    assert x.shape == (len(kw["vertices"]), 3) and st["volume"] > 0
    s.close()


def test_native_server_failed_case_does_not_stop_the_others():
    """A case whose mesh the library rejects at admission fails alone: poll reports
    the error, the case reads FAILED and the case admitted beside it finishes."""
    print("[serving] native failed case", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    s = hc.Server(0, 4, 1 << 30)
    c = serving.make_cases(1, rate=1.0, seed=5)[0]
    kw = serving.spec(c, u)
    good = s.submit(**dict(kw, steps=20))
    tris = np.array(kw["triangles"], dtype=np.int32).copy()
    tris.flat[0] = len(kw["vertices"]) + 7  # a vertex index past the mesh
    bad = s.submit(**dict(kw, steps=20, triangles=tris))
    with pytest.raises(hc.HcError):
        s.poll(10)
    assert s.status(bad) == hc.CASE_FAILED
    t0 = time.time()
    while s.status(good) != hc.CASE_DONE and time.time() - t0 < 60:  # poll does not block
        s.poll(10)
        time.sleep(1e-3)
    assert s.status(good) == hc.CASE_DONE
    x, st = s.result(good)
    assert np.isfinite(x).all() and st["volume"] > 0
    s.close()


def test_native_server_retire_and_stream_failures():
    """Failures forced from the host (no device error): a case whose result
    read-back fails at retirement and a case whose stream reports an error are
    marked FAILED and release their slots; poll reports each error once, the
    third case finishes with the result it has alone, and the loop drains."""
    print("[serving] native injected failures", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = serving.make_cases(3, rate=1.0, seed=13, edge=(24, 28), steps=(60, 90))
    s = hc.Server(0, 3, 1 << 30)
    ids = [s.submit(**serving.spec(c, u)) for c in stream]
    assert s.poll(20) == 3  # all three admitted and stepping
    s.inject_failure(ids[1], hc.INJECT_RETIRE)
    s.inject_failure(ids[2], hc.INJECT_STREAM_ERROR)
    errors = []
    running, t0 = 3, time.time()
    while running and time.time() - t0 < 60:
        try:
            running = s.poll(20)
        except hc.HcError as e:
            errors.append(str(e))
            running = sum(s.status(i) in (hc.CASE_QUEUED, hc.CASE_RUNNING) for i in ids)
        time.sleep(1e-3)
    assert running == 0
    assert [s.status(i) for i in ids] == [hc.CASE_DONE, hc.CASE_FAILED, hc.CASE_FAILED]
    assert len(errors) == 2 and any("stream" in e for e in errors) and any("retire" in e for e in errors)
    with pytest.raises(hc.HcError):
        s.result(ids[1])
    x, st = s.result(ids[0])
    alone, xa = serving.run_alone(stream[0], u)
    assert cases.rel_l2(x, xa) < 1e-9
    s.close()


def test_serve_native_keeps_serving_past_a_failed_case():
    """serve_native with one case whose mesh the library rejects at admission:
    that case ends failed with its error, every other case of the stream
    finishes, and the latency statistics count it apart."""
    print("[serving] serve_native with a failed case", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = serving.make_cases(4, rate=500.0, seed=17, edge=(24, 30), steps=(80, 160))

    def spec_of(c, units):
        kw = serving.spec(c, units)
        if c.ident == 1:
            tris = np.array(kw["triangles"], dtype=np.int32).copy()
            tris.flat[0] = len(kw["vertices"]) + 7  # a vertex index past the mesh
            kw["triangles"] = tris
        return kw

    done, wall = serving.serve_native(stream, u, max_active=4, chunk=40, spec_of=spec_of)
    assert sorted(c.ident for c in done) == [0, 1, 2, 3]
    bad = [c for c in done if c.ident == 1][0]
    assert bad.result is None and bad.error
    for c in done:
        if c.ident != 1:
            assert c.result and c.result["volume"] > 0 and not c.error
    st = serving.latency_stats(done, wall)
    assert st["cases"] == 3 and st["failed"] == 1


def test_stream_at_full_ranges_matches_alone():
    """The bench's stream generator at its full ranges (edge 32-96, 2000-20000
    steps, shear 10-1000 1/s or stretch 0-200 pN, RBC/PLT/WBC): every case of an
    8-case stream served by the batched loop ends as close to the case run alone
    as two lone runs are to each other, or within 1e-8.  The FP32 form's spread
    box adds doubles in the order the wavefronts' LDS atomics land, so two lone
    runs of the stretched RBC (case 4, 6925 steps) differ by 0 to 3e-9 from run
    to run (measured); a batching mistake moves a case by far more.  The exact
    spread removes that freedom: see the FP64 stream below."""
    print("[serving] full ranges", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = serving.make_cases(8, rate=100.0, seed=23)
    assert max(c.edge for c in stream) > 64 and max(c.steps for c in stream) > 10000
    done, wall = serving.serve(stream, u, max_active=8, chunk=250, keep_vertices=True)
    assert sorted(c.ident for c in done) == list(range(8))
    for c in done:
        alone, x = serving.run_alone(c, u)
        _, x2 = serving.run_alone(c, u)
        noise = cases.rel_l2(x2, x)
        mix = cases.rel_l2(c.result["x"], x)
        print("[serving] case %d (%s %s, edge %d, %d steps): mix %.2e, lone-vs-lone %.2e" % (
            c.ident, c.kind, c.cell, c.edge, c.steps, mix, noise))
        assert mix <= max(100 * noise, 1e-8), (c, mix, noise)


def test_exact_spread_stream_is_bit_identical_to_alone():
    """The same stream in the FP64 form with the exact spread, steps capped at
    1200: every case served by the batched loop ends bit for bit where it ends
    when run alone (the exact spread's sums do not depend on the order of the
    adds, and the batch runs the same kernels per case)."""
    import dataclasses
    print("[serving] exact spread stream", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = [dataclasses.replace(c, steps=min(c.steps, 1200)) for c in serving.make_cases(8, rate=100.0, seed=23)]
    done, _ = serving.serve(stream, u, max_active=8, chunk=250, keep_vertices=True, precision=hc.HC_FP64,
                            exact_spread=True)
    assert sorted(c.ident for c in done) == list(range(8))
    for c in done:
        _, x = serving.run_alone(c, u, precision=hc.HC_FP64, exact_spread=True)
        mix = cases.rel_l2(c.result["x"], x)
        print("[serving] exact case %d (%s %s, edge %d, %d steps): mix %.2e" % (
            c.ident, c.kind, c.cell, c.edge, c.steps, mix))
        assert np.array_equal(c.result["x"], x), (c.ident, mix)


def _oracle_case(case, units):
    """The oracle set up as serving.build sets up the case (FP64)."""
    E = case.edge
    t = serving._cell_types(units)[case.cell]
    mask = np.zeros((E, E, E), dtype=np.uint8)
    if case.kind == "shear":
        mask[:, :, 0] = hc.NODE_VELOCITY + 4
        mask[:, :, E - 1] = hc.NODE_VELOCITY + 5
        o = OracleSim(E, E, E, units.tau, periodic=(1, 1, 0), node_type=mask.reshape(-1))
        vh = (E - 1) * case.shear_rate * units.dt * 0.5
        o.set_wall_velocity(4, (vh, 0.0, 0.0))
        o.set_wall_velocity(5, (-vh, 0.0, 0.0))
    else:
        for a in range(3):
            for side, face in ((0, 2 * a), (E - 1, 2 * a + 1)):
                sl = [slice(None)] * 3
                sl[a] = side
                fm = mask[tuple(sl)]
                fm[fm == 0] = hc.NODE_VELOCITY + face
        o = OracleSim(E, E, E, units.tau, periodic=(0, 0, 0), node_type=mask.reshape(-1))
    o.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    o.set_force_limit(units.f_limit)
    tid = o.add_cell_type(t.model, t.verts, t.tris, t.params, t.inner)
    x = t.verts - 0.5 * (t.verts.min(axis=0) + t.verts.max(axis=0)) + (E - 1) / 2.0
    o.add_cells(tid, x[None])
    extra = None
    if case.kind == "stretch":
        order = np.argsort(x[:, 0], kind="stable")
        n = 7
        f = case.force_pN * 1e-12 / units.df / n
        ids = np.concatenate([order[:n], order[::-1][:n]])
        forces = np.zeros((2 * n, 3))
        forces[:n, 0] = -f
        forces[n:, 0] = f
        extra = (ids, forces)
    return o, extra, x


def test_stream_cases_match_the_oracle():
    """Three cases of the stream (shear and stretch; edge and steps cut to 32-40
    and 300-500 so the oracle finishes), served together in FP64 by the batched
    loop, each against the FP64 oracle: positions and displacements within 1e-10."""
    print("[serving] stream cases vs oracle", flush=True)
    u = hc.lbm_base_parameters(5e-7, 1e-7)
    stream = serving.make_cases(3, rate=100.0, seed=31, edge=(32, 40), steps=(300, 500))
    assert {c.kind for c in stream} == {"shear", "stretch"}
    done, _ = serving.serve(stream, u, max_active=3, chunk=100, keep_vertices=True, precision=hc.HC_FP64)
    for c in done:
        o, extra, x0 = _oracle_case(c, u)
        for _ in range(c.steps):
            if extra is not None:
                _, _, fo = o.vertices()
                fo[extra[0]] += extra[1]
                o.set_vertices(None, None, fo)
            o.step(1)
        xo = o.vertices()[0]
        xg = c.result["x"]
        e = (cases.rel_l2(xg, xo), cases.rel_l2(xg - x0, xo - x0))
        print("[serving] case %d (%s %s, edge %d, %d steps) vs oracle: x %.2e, displacement %.2e" % (
            (c.ident, c.kind, c.cell, c.edge, c.steps) + e))
        assert e[0] < 1e-10 and e[1] < 1e-10
