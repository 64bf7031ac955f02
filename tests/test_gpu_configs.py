"""The batched, served and large-grid workloads at their own sizes, and the
divergence condition of the metric.  Needs an MI355X.

  * 256 batched scenes with n_x, n_y, n_z from {32,48,64,96,128} and 0-3
    spheres (mgf_batch_step, 2 steps): every scene meets the 1e-6 FP64
    residual; 8 sampled scenes are bit-identical to the same scenes stepped
    alone; one sampled scene agrees with the FP32 oracle within 1e-4;
  * served scenes with n per axis log-uniform in [32,256] and 0-4 spheres
    under an HBM budget that binds: admission is refused, then granted once a
    scene closes; every served result equals the scene stepped alone;
  * the obstacle scene at 256^3: the divergence of the stored FP32 velocity
    after each of three steps, against the FP64 residual of the solve;
  * the blocks-without-fluid shortcut changes no bit;
  * the 16-bit face-distance limit of the extrapolation.
"""
import numpy as np
import pytest

import mgf
import scenes

pytestmark = pytest.mark.gpu

FIELDS = (mgf.FIELD["vel_x"], mgf.FIELD["vel_y"], mgf.FIELD["vel_z"], mgf.FIELD["fluid_domain"],
          mgf.FIELD["pressure"])
MIXED = dict(rk=mgf.RK1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
             precond=mgf.PRECOND_MULTIGRID)


@pytest.fixture(scope="module")
def ctx():
    c = mgf.Context(0)
    yield c
    c.close()


def snapshot(s):
    return [s.download(f).view(np.uint32).copy() for f in FIELDS]


def assert_same(a, b, what):
    for k, (x, y) in enumerate(zip(a, b)):
        bad = np.count_nonzero(x != y)
        assert bad == 0, "%s: field %d differs in %d of %d values" % (what, FIELDS[k], bad, x.size)


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_config4_batched_256_ragged_scenes(ctx, orc):
    specs = [scenes.ragged_scene(i) for i in range(256)]
    assert {v for p in specs for v in p.nxyz} <= set(scenes.BATCH_SIZES)
    batch = [scenes.build_gpu_scene(mgf, ctx, p, **MIXED) for p in specs]
    nsteps = 2
    status, stats, _ = mgf.batch_step(ctx, batch, nsteps)
    assert status == [mgf.OK] * len(specs), [i for i, v in enumerate(status) if v != mgf.OK][:8]
    for i, st in enumerate(stats):
        assert st.converged == 1 and st.true_converged == 1 and st.steps == nsteps, (i, specs[i].name)
        assert st.residual_rel <= 1e-6, (i, specs[i].name, st.residual_rel)
    cells = sum(p.interior_cells for p in specs)
    print("config 4: %d scenes, %.3g interior cells, max FP64 residual %.2e"
          % (len(specs), cells, max(st.residual_rel for st in stats)))
    # a sample of the batch stepped alone: bit-identical
    sample = list(range(0, 256, 32))
    for i in sample:
        solo = scenes.build_gpu_scene(mgf, ctx, specs[i], **MIXED)
        solo.step(nsteps)
        assert_same(snapshot(batch[i]), snapshot(solo), specs[i].name)
        solo.close()
    # one sampled scene against the FP32 oracle (the smallest of the sample)
    i = min(sample, key=lambda k: specs[k].interior_cells)
    o = scenes.build_oracle_scene(orc, specs[i])
    p = orc.params(specs[i].dt, rk=1, tol_mode=1, tol=1e-6)
    o.step(p, n=nsteps)
    g = np.concatenate([batch[i].download(f) for f in (0, 1, 2)])
    e = np.concatenate([o.download(f) for f in (0, 1, 2)])
    d = rel_l2(g, e)
    print("config 4: %s vs the FP32 oracle after %d steps: rel L2 %.2e" % (specs[i].name, nsteps, d))
    assert d <= 1e-4, (specs[i].name, d)
    o.close()
    for s in batch:
        s.close()


def test_config5_served_budget_binds(ctx):
    """Served scenes (n per axis log-uniform in [32,256], 0-4 spheres) under a
    budget of the first three scenes' bytes: the fourth is refused, admitted
    once a scene closes, and each result equals the same steps run alone."""
    specs = [scenes.ragged_scene(1000 + i, log_uniform=(32, 256), max_spheres=4) for i in range(4)]
    descs = [mgf.make_desc(p.res, p.origin, p.spacing, p.dt, **MIXED) for p in specs]
    need = [mgf.scene_bytes_for(ctx, d) for d in descs]
    budget = sum(need[:3]) + min(need) // 4
    assert budget < sum(need)
    with mgf.Server(ctx, budget_bytes=budget, slots=3) as srv:
        served = []
        for k in range(3):
            s = srv.open_scene(descs[k])
            assert s is not None, specs[k].name
            scenes.setup_gpu_scene(s, specs[k])
            served.append(s)
        assert srv.open_scene(descs[3]) is None          # refused: the budget binds
        tickets = {}
        for rep in range(2):
            for k, s in enumerate(served):
                tickets[srv.submit(s, 1)] = k
        got = []
        while len(got) < len(tickets):
            got += srv.poll(cap=16, wait_ms=20000)
        assert all(r.status == mgf.OK and r.stats.true_converged == 1 for r in got)
        # each result equals the scene stepped alone
        for k, s in enumerate(served):
            solo = scenes.build_gpu_scene(mgf, ctx, specs[k], **MIXED)
            solo.step(2)
            assert_same(snapshot(s), snapshot(solo), specs[k].name)
            solo.close()
        # a scene closes: the refused one is admitted now
        srv.close_scene(served[0])
        s3 = srv.open_scene(descs[3])
        assert s3 is not None
        scenes.setup_gpu_scene(s3, specs[3])
        srv.submit(s3, 1)
        r = []
        while not r:
            r = srv.poll(cap=4, wait_ms=20000)
        assert r[0].status == mgf.OK and r[0].stats.true_converged == 1
        solo = scenes.build_gpu_scene(mgf, ctx, specs[3], **MIXED)
        solo.step(1)
        assert_same(snapshot(s3), snapshot(solo), specs[3].name)
        solo.close()
        print("config 5: budget %.3g GB, scene bytes %s" % (budget / 1e9, ["%.3g GB" % (b / 1e9) for b in need]))


def test_obstacles_256_divergence_of_stored_velocity(ctx):
    """The headline scene, three steps with the divergence check on: each solve
    meets the 1e-6 FP64 residual, and ||D u|| / ||b|| of the stored FP32
    velocity stays within the FP32 rounding floor measured on the config-1
    trajectory (at most a few 1e-6)."""
    spec = scenes.obstacles(256, seed=0)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **MIXED)
    s.set_divergence_check(True)
    for step in range(3):
        st = s.step(1)
        assert st["converged"] == 1 and st["residual_rel"] <= 1e-6, st
        print("256^3 step %d: FP64 residual %.2e, divergence of the stored velocity %.2e"
              % (step, st["residual_rel"], st["div_residual"]))
        assert 0 < st["div_residual"] <= 1e-5, st
    s.close()


def test_block_skip_changes_no_bit(ctx):
    """Blocks without fluid skipped (the default) or run: the same bits, the
    same iteration counts, on a scene with large dry regions."""
    spec = scenes.obstacles(64, seed=4)
    kw = dict(MIXED, mg_kernels=mgf.MG_FUSED)
    a = scenes.build_gpu_scene(mgf, ctx, spec, **kw)
    b = scenes.build_gpu_scene(mgf, ctx, spec, **kw)
    b.set_block_skip(False)
    for step in range(3):
        sa, sb = a.step(1), b.step(1)
        assert sa["cg_iterations"] == sb["cg_iterations"] and sa["converged"] == sb["converged"] == 1
        assert_same(snapshot(a), snapshot(b), "step %d" % step)
    a.close(); b.close()


def test_extrapolation_16bit_limit(ctx):
    """A grid with nx + ny + nz + 3 >= 65535 is created and steps without
    extrapolation (one layer); extrapolating over more layers is refused."""
    res = (65534, 2, 1)
    d = mgf.make_desc(res, (0, 0, 0), (1.0, 1.0, 1.0), 1e-3, layers=1, **MIXED)
    s = mgf.Scene(ctx, d)
    s.fluid_box((0.0, 0.0, 0.0), (100.0, 1.0, 1.0))
    st = s.step(1)
    assert st["converged"] == 1
    d2 = mgf.make_desc(res, (0, 0, 0), (1.0, 1.0, 1.0), 1e-3, layers=1000, **MIXED)
    s.set_params(d2)
    with pytest.raises(mgf.MgfError) as e:
        s.extrapolate()
    assert e.value.status == mgf.ERR_UNSUPPORTED
    s.close()
