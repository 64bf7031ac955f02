"""Batched scenes (mgf_batch_step) and the served loop (mgf_server_*) through
the C ABI.  Needs an MI355X.

The bar: every scene's fields after a batched or served run are bit-identical
to the same scene stepped alone with mgf_step (each scene runs the same
kernels in the same order on its own buffers and stream), and a scene of the
batch agrees with the FP32 oracle as a scene stepped alone does.
"""
import numpy as np
import pytest

import mgf
import scenes

pytestmark = pytest.mark.gpu

FIELDS = (mgf.FIELD["vel_x"], mgf.FIELD["vel_y"], mgf.FIELD["vel_z"], mgf.FIELD["fluid_domain"],
          mgf.FIELD["pressure"])
SIZES = (8, 12, 16, 20)


@pytest.fixture(scope="module")
def ctx():
    c = mgf.Context(0)
    yield c
    c.close()


def desc_kw(precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_MULTIGRID):
    return dict(rk=mgf.RK1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=precision, precond=precond)


def specs(n, seed0=100):
    return [scenes.ragged_scene(seed0 + i, sizes=SIZES, h=1.0 / 16, dt=2e-3) for i in range(n)]


def snapshot(s):
    return [s.download(f).view(np.uint32).copy() for f in FIELDS]


def assert_same(a, b, what):
    for k, (x, y) in enumerate(zip(a, b)):
        bad = np.count_nonzero(x != y)
        assert bad == 0, "%s: field %d differs in %d of %d values" % (what, FIELDS[k], bad, x.size)


@pytest.mark.parametrize("precision,precond", [(mgf.PRECISION_MIXED, mgf.PRECOND_MULTIGRID),
                                               (mgf.PRECISION_MIXED, mgf.PRECOND_IDENTITY),
                                               (mgf.PRECISION_PLAIN, mgf.PRECOND_IDENTITY)])
def test_batch_matches_scenes_stepped_alone(ctx, precision, precond):
    sp = specs(7)
    kw = desc_kw(precision, precond)
    solo = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    batch = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    nsteps = 3
    ref_stats = [s.step(nsteps) for s in solo]
    status, stats, ms = mgf.batch_step(ctx, batch, nsteps, workers=4, step_latency=True)
    assert status == [mgf.OK] * len(sp), status
    assert ms.shape == (len(sp), nsteps) and np.all(ms > 0)
    for i, p in enumerate(sp):
        assert stats[i].steps == nsteps and stats[i].converged == 1
        assert stats[i].cg_iterations == ref_stats[i]["cg_iterations"], (p.name, stats[i].cg_iterations)
        assert_same(snapshot(batch[i]), snapshot(solo[i]), p.name)
    for s in solo + batch:
        s.close()


def test_batch_without_latency_and_repeat(ctx):
    """Two batched calls (no per-step latency) equal one solo call of the summed
    steps; the worker count does not change the results."""
    sp = specs(5, seed0=200)
    kw = desc_kw()
    solo = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    b1 = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    b2 = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    for s in solo:
        s.step(4)
    st, _, ms = mgf.batch_step(ctx, b1, 2, workers=1)
    assert ms is None and st == [mgf.OK] * 5
    mgf.batch_step(ctx, b1, 2, workers=8)
    mgf.batch_step(ctx, b2, 4, workers=3)
    for i in range(len(sp)):
        ref = snapshot(solo[i])
        assert_same(snapshot(b1[i]), ref, "b1 %d" % i)
        assert_same(snapshot(b2[i]), ref, "b2 %d" % i)
    for s in solo + b1 + b2:
        s.close()


def test_batched_scene_matches_fp32_oracle(ctx, orc):
    """A non-cubic scene stepped inside a batch agrees with the FP32 oracle as
    an unbatched scene does (velocity within 1e-4 relative L2)."""
    sp = specs(4, seed0=300)
    kw = desc_kw()
    batch = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    mgf.batch_step(ctx, batch, 2, workers=4)
    spec = sp[0]
    o = scenes.build_oracle_scene(orc, spec)
    prm = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-6)
    o.step(prm)
    o.step(prm)
    for f in (0, 1, 2):
        g = batch[0].download(f).astype(np.float64)
        e = o.download(f).astype(np.float64)
        fin = (np.abs(e) < 1e30) & (np.abs(g) < 1e30)
        assert np.array_equal(np.abs(e) < 1e30, np.abs(g) < 1e30)
        d = np.linalg.norm(e[fin])
        assert np.linalg.norm(g[fin] - e[fin]) <= 1e-4 * max(d, 1e-30), (spec.name, f)
    o.close()
    for s in batch:
        s.close()


def test_batch_rejects_bad_arguments(ctx):
    sp = specs(2, seed0=400)
    a = scenes.build_gpu_scene(mgf, ctx, sp[0], **desc_kw())
    L = mgf.lib()
    import ctypes as C
    arr = (C.c_void_p * 2)(a.h.value, a.h.value)
    assert L.mgf_batch_step(ctx.h, arr, 2, 1, 0, None, None, None) == mgf.ERR_INVALID
    assert L.mgf_batch_step(ctx.h, arr, -1, 1, 0, None, None, None) == mgf.ERR_INVALID
    assert L.mgf_batch_step(ctx.h, arr, 1, -1, 0, None, None, None) == mgf.ERR_INVALID
    assert L.mgf_batch_step(ctx.h, None, 0, 1, 0, None, None, None) == mgf.OK
    a.close()


def elongated(nxyz):
    """A scene of nx x ny x nz interior cells (spacing 1/64) with a fluid
    block over the lower part of x and y and the full z extent."""
    h = 1.0 / 64
    ext = [v * float(np.float32(h)) for v in nxyz]
    return scenes.SceneSpec(nxyz, 2e-3, (0.0, 0.0, 0.0), (0.4 * ext[0], 0.5 * ext[1], ext[2]), [],
                            "elongated_%dx%dx%d" % tuple(nxyz), h=h)


def test_batch_long_rows_and_lines_match_scenes_stepped_alone(ctx):
    """Scenes whose x rows (303 faces) and y and z lines (533 and 903 faces)
    select the distance transform's larger kernel classes: stepped together
    in lockstep (multi-scene records of those classes), their fields equal the
    same scenes stepped alone, bit for bit."""
    sp = [elongated((300, 10, 8)), elongated((9, 530, 8)), elongated((8, 9, 900))]
    kw = desc_kw()
    solo = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    batch = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    nsteps = 2
    ref = [s.step(nsteps) for s in solo]
    status, stats, _ = mgf.batch_step(ctx, batch, nsteps, workers=2)
    assert status == [mgf.OK] * len(sp), status
    for i, p in enumerate(sp):
        assert stats[i].steps == nsteps and stats[i].converged == 1, p.name
        assert stats[i].cg_iterations == ref[i]["cg_iterations"], (p.name, stats[i].cg_iterations)
        assert_same(snapshot(batch[i]), snapshot(solo[i]), p.name)
    for s in solo + batch:
        s.close()


def test_scene_bytes_for_matches_a_stepped_scene(ctx):
    for p in specs(3, seed0=500):
        d = mgf.make_desc(p.res, p.origin, p.spacing, p.dt, **desc_kw())
        need = mgf.scene_bytes_for(ctx, d)
        s = mgf.Scene(ctx, d)
        s.fluid_box(p.fluid_min, p.fluid_max)
        s.step(1)
        assert s.device_bytes() == need, (p.name, s.device_bytes(), need)
        s.close()


def test_server_results_match_scenes_stepped_alone(ctx):
    """Requests for several scenes, submitted while others run, overlap on the
    device; each scene's result equals the same steps run alone."""
    sp = specs(6, seed0=600)
    kw = desc_kw()
    solo = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    with mgf.Server(ctx, budget_bytes=8 << 30, slots=3) as srv:
        served = []
        for p in sp:
            s = srv.open_scene(mgf.make_desc(p.res, p.origin, p.spacing, p.dt, **kw))
            assert s is not None
            scenes.setup_gpu_scene(s, p)
            served.append(s)
        tickets = {}
        plan = [(i, k) for k in range(3) for i in range(len(sp))]  # 3 requests of 1-2 steps per scene
        steps = [0] * len(sp)
        for i, k in plan:
            n = 1 + (i + k) % 2
            tickets[srv.submit(served[i], n)] = (i, k, n)
            steps[i] += n
        got = []
        while len(got) < len(plan):
            got += srv.poll(cap=64, wait_ms=10000)
        assert sorted(r.ticket for r in got) == sorted(tickets)
        last_end = {}
        for r in sorted(got, key=lambda r: r.ticket):
            i, k, n = tickets[r.ticket]
            assert r.status == mgf.OK and r.steps == n and r.stats.steps == n
            assert r.t_submit_ms <= r.t_start_ms <= r.t_end_ms
            # one scene's requests run in submission order, one at a time
            if i in last_end:
                assert r.t_start_ms >= last_end[i]
            last_end[i] = r.t_end_ms
        for i, s in enumerate(solo):
            s.step(steps[i])
            assert_same(snapshot(served[i]), snapshot(s), sp[i].name)
        b, q, run = srv.state()
        assert q == 0 and run == 0 and b > 0
    for s in solo:
        s.close()


def test_server_lockstep_rounds_match_scenes_stepped_alone(ctx):
    """With at least 8 ready scenes that can step in lockstep, a worker steps
    them together, one step per round (requests of two steps span two
    rounds, or finish on their scene's stream); every scene's result equals
    the same steps run alone, and the rounds show as groups of at least 8
    results ending at the same instant."""
    sp = specs(12, seed0=700)
    kw = desc_kw()
    solo = [scenes.build_gpu_scene(mgf, ctx, p, **kw) for p in sp]
    with mgf.Server(ctx, budget_bytes=8 << 30, slots=2) as srv:
        served = []
        for p in sp:
            s = srv.open_scene(mgf.make_desc(p.res, p.origin, p.spacing, p.dt, **kw))
            assert s is not None
            scenes.setup_gpu_scene(s, p)
            served.append(s)
        tickets = {}
        steps = [0] * len(sp)
        for k in range(3):
            for i in range(len(sp)):
                n = 1 + (i + k) % 2
                tickets[srv.submit(served[i], n)] = (i, n)
                steps[i] += n
        got = []
        while len(got) < len(tickets):
            got += srv.poll(cap=64, wait_ms=10000)
        assert sorted(r.ticket for r in got) == sorted(tickets)
        ends = {}
        for r in got:
            i, n = tickets[r.ticket]
            assert r.status == mgf.OK and r.steps == n and r.stats.steps == n, (i, r.status, r.stats.steps)
            assert r.stats.true_converged == 1
            assert r.t_submit_ms <= r.t_start_ms <= r.t_end_ms
            ends[r.t_end_ms] = ends.get(r.t_end_ms, 0) + 1
        assert max(ends.values()) >= 8, sorted(ends.values())
        for i, s in enumerate(solo):
            s.step(steps[i])
            assert_same(snapshot(served[i]), snapshot(s), sp[i].name)
        b, q, run = srv.state()
        assert q == 0 and run == 0 and b > 0
    for s in solo:
        s.close()


def test_server_admission_by_budget(ctx):
    sp = specs(2, seed0=700)
    kw = desc_kw()
    d0 = mgf.make_desc(sp[0].res, sp[0].origin, sp[0].spacing, sp[0].dt, **kw)
    d1 = mgf.make_desc(sp[1].res, sp[1].origin, sp[1].spacing, sp[1].dt, **kw)
    b0, b1 = mgf.scene_bytes_for(ctx, d0), mgf.scene_bytes_for(ctx, d1)
    with mgf.Server(ctx, budget_bytes=max(b0, b1) + min(b0, b1) // 2, slots=2) as srv:
        s0 = srv.open_scene(d0)
        assert s0 is not None
        assert srv.state()[0] == b0
        assert srv.open_scene(d1) is None            # refused: the budget is full
        srv.close_scene(s0)
        assert srv.state()[0] == 0
        s1 = srv.open_scene(d1)
        assert s1 is not None
        scenes.setup_gpu_scene(s1, sp[1])
        srv.submit(s1, 1)
        r = srv.poll(cap=4, wait_ms=-1)
        assert len(r) == 1 and r[0].status == mgf.OK
        # solids added after admission count against the budget from then on
        assert srv.state()[0] == s1.device_bytes() > b1
