"""Fused multigrid descent / ascent (k_mg_march, forced on every 3-D level
with MGF_MG_FUSED) against one kernel per sweep (MGF_MG_PER_SWEEP).

The fused kernels form every value of the V-cycle with the same operations in
the same order as the per-sweep kernels (k_mg_smooth2_first, k_mg_smooth,
the restriction kernels, k_mg_smooth_prolong), so z = M^-1 r must be bitwise
equal.  r . z is summed over other blocks, so it agrees to FP64 rounding.
Grids: even and odd extents, tiles cut at the grid's edges, chunks that end
past the last plane, extents of 2, 2-D grids (one axis of one cell: those
levels keep the per-sweep kernels), random solids and air.  Needs an MI355X.
"""
import numpy as np
import pytest

import mgf
import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mgf.Context(0)
    yield c
    c.close()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def random_scene(ctx, res, seed, kernels, psolid=0.1, pfluid=0.7):
    rng = np.random.default_rng(seed)
    n = int(np.prod(res))
    solid = np.where(rng.random(n) < psolid, -1.0, 1.0).astype(np.float32)
    fluid = np.where(rng.random(n) < pfluid, -1.0, 1.0).astype(np.float32)
    d = mgf.make_desc(res, (0, 0, 0), (0.1, 0.2, 0.3), 0.01, tol_mode=mgf.TOL_RELATIVE, tol=1e-6,
                      precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_MULTIGRID, mg_kernels=kernels)
    s = mgf.Scene(ctx, d)
    s.add_solid(solid)
    s.upload(3, fluid)
    return s


def both(ctx, res, seed, **kw):
    return (random_scene(ctx, res, seed, mgf.MG_FUSED, **kw),
            random_scene(ctx, res, seed, mgf.MG_PER_SWEEP, **kw))


def check_same(fused, ref, r, what):
    zf, rzf = fused.apply_preconditioner(r)
    zr, rzr = ref.apply_preconditioner(r)
    bad = np.nonzero(bits(zf) != bits(zr))[0]
    print(what, "cells", zf.size, "differing", bad.size, flush=True)
    assert bad.size == 0, "%s: %d of %d differ, first %s: fused %r per-sweep %r" % (
        what, bad.size, zf.size, bad[:5], zf[bad[:5]], zr[bad[:5]])
    assert np.all(np.isfinite(zf))
    assert abs(rzf - rzr) <= 1e-12 * max(abs(rzr), 1e-300), (what, rzf, rzr)
    return zf


@pytest.mark.parametrize("res", [(8, 8, 8), (2, 2, 2), (3, 5, 7), (21, 17, 13), (64, 11, 40), (59, 12, 35),
                                 (117, 21, 66), (130, 130, 9), (9, 9, 130)])
def test_fused_vcycle_bitwise_per_sweep(ctx, res):
    """Random r on random markers (10% solid, 70% of the rest fluid)."""
    f, p = both(ctx, res, 5 + sum(res))
    rng = np.random.default_rng(sum(res))
    r = rng.standard_normal(int(np.prod(res))).astype(np.float32)
    z = check_same(f, p, r, "res %s" % (res,))
    assert np.any(z != 0.0)
    f.close(); p.close()


@pytest.mark.parametrize("res", [(1, 40, 33), (37, 1, 20), (33, 25, 1)])
def test_fused_vcycle_2d_grids(ctx, res):
    """A grid with an axis of one cell: no level is fused (every axis must
    halve), so both settings run the same kernels."""
    f, p = both(ctx, res, 3 + sum(res))
    r = np.random.default_rng(1).standard_normal(int(np.prod(res))).astype(np.float32)
    check_same(f, p, r, "res %s" % (res,))
    f.close(); p.close()


def test_fused_vcycle_all_fluid_and_all_air(ctx):
    """Edge markers: every cell fluid (no Dirichlet cell) and every cell air
    (z = 0)."""
    for pf, what in ((1.0, "all fluid"), (0.0, "all air")):
        f, p = both(ctx, (34, 30, 26), 9, psolid=0.0, pfluid=pf)
        r = np.random.default_rng(2).standard_normal(34 * 30 * 26).astype(np.float32)
        z = check_same(f, p, r, what)
        if pf == 0.0:
            assert np.all(z == 0.0)
        f.close(); p.close()


def test_fused_vcycle_obstacles_130(ctx):
    """The obstacle scene at 128^3 (130^3 padded): the residual of its first
    pressure right-hand side through both V-cycles."""
    spec = scenes.obstacles(128, seed=1)
    kw = dict(rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
              precond=mgf.PRECOND_MULTIGRID)
    f = scenes.build_gpu_scene(mgf, ctx, spec, mg_kernels=mgf.MG_FUSED, **kw)
    p = scenes.build_gpu_scene(mgf, ctx, spec, mg_kernels=mgf.MG_PER_SWEEP, **kw)
    n = int(np.prod(spec.res))
    r = np.random.default_rng(3).standard_normal(n).astype(np.float32)
    check_same(f, p, r, "obstacles 130^3")
    f.close(); p.close()


def test_fused_solve_matches_per_sweep(ctx):
    """Whole steps: the fused setting takes the same number of CG iterations
    and gives the same velocities to FP64-summation rounding (r . z, hence
    beta, is summed in another order)."""
    spec = scenes.obstacles(96, seed=4)
    kw = dict(rk=3, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
              precond=mgf.PRECOND_MULTIGRID)
    f = scenes.build_gpu_scene(mgf, ctx, spec, mg_kernels=mgf.MG_FUSED, **kw)
    p = scenes.build_gpu_scene(mgf, ctx, spec, mg_kernels=mgf.MG_PER_SWEEP, **kw)
    for step in range(3):
        sf, sp = f.step(1), p.step(1)
        print("step", step, sf["cg_iterations"], sp["cg_iterations"], sf["residual_rel"], sp["residual_rel"],
              flush=True)
        assert sf["converged"] == 1 and sp["converged"] == 1
        assert abs(sf["cg_iterations"] - sp["cg_iterations"]) <= 1
        assert sf["residual_rel"] <= 1e-6
    for fid in (0, 1, 2):
        a, b = f.download(fid), p.download(fid)
        m = b != np.finfo(np.float32).max
        assert np.array_equal(a != np.finfo(np.float32).max, m)
        d = np.linalg.norm((a[m] - b[m]).astype(np.float64)) / max(np.linalg.norm(b[m].astype(np.float64)), 1e-30)
        assert d <= 1e-5, (fid, d)
    f.close(); p.close()


def test_auto_fuses_large_levels_only(ctx):
    """MGF_MG_AUTO: the 256^3 obstacle scene (17M cells) runs the fused level-0
    kernels (the probes see their launches), a 64^3 scene none; z is bitwise
    the per-sweep V-cycle's in both."""
    for n, fused in ((256, True), (64, False)):
        spec = scenes.obstacles(n, seed=1)
        kw = dict(rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
                  precond=mgf.PRECOND_MULTIGRID)
        a = scenes.build_gpu_scene(mgf, ctx, spec, **kw)
        p = scenes.build_gpu_scene(mgf, ctx, spec, mg_kernels=mgf.MG_PER_SWEEP, **kw)
        a.set_kernel_probes(True)
        a.kernel_probes(reset=True)
        r = np.random.default_rng(n).standard_normal(int(np.prod(spec.res))).astype(np.float32)
        check_same(a, p, r, "auto %d^3" % n)
        pr = a.kernel_probes(reset=True)
        assert ("k_mg_march" in pr and pr["k_mg_march"]["launches"] > 0) == fused, pr.get("k_mg_march")
        a.close(); p.close()


@pytest.mark.parametrize("first,then", [(1, 2), (0, 2), (2, 1)])
def test_switch_mg_kernels_after_a_step(ctx, first, then):
    """mg_kernels changed by mgf_scene_set_params after the hierarchy exists
    (a level that becomes fused may have no wet-block flags: it marches every
    block): the next step is bitwise that of a scene created with the new
    setting and stepped the same way."""
    spec = scenes.obstacles(96, seed=6)
    kw = dict(rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
              precond=mgf.PRECOND_MULTIGRID)
    a = scenes.build_gpu_scene(mgf, ctx, spec, mg_kernels=first, **kw)
    b = scenes.build_gpu_scene(mgf, ctx, spec, mg_kernels=then, **kw)
    b.set_params(mgf.make_desc(spec.res, spec.origin, spec.spacing, spec.dt, mg_kernels=first, **kw))
    a.step(1)
    b.step(1)
    a.set_params(mgf.make_desc(spec.res, spec.origin, spec.spacing, spec.dt, mg_kernels=then, **kw))
    b.set_params(mgf.make_desc(spec.res, spec.origin, spec.spacing, spec.dt, mg_kernels=then, **kw))
    sa, sb = a.step(1), b.step(1)
    assert sa["converged"] == 1 and sb["converged"] == 1 and sa["cg_iterations"] == sb["cg_iterations"]
    for fid in (0, 1, 2, 9):
        x, y = a.download(fid), b.download(fid)
        assert np.array_equal(bits(x), bits(y)), fid
    a.close(); b.close()


def test_apply_preconditioner_arguments(ctx):
    d = mgf.make_desc((8, 8, 8), precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_IDENTITY)
    s = mgf.Scene(ctx, d)
    with pytest.raises(mgf.MgfError) as ei:
        s.apply_preconditioner(np.zeros(512, np.float32))
    assert ei.value.status == mgf.ERR_UNSUPPORTED
    s.close()
    d = mgf.make_desc((8, 8, 8), precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_MULTIGRID)
    d.mg_kernels = 3
    with pytest.raises(mgf.MgfError):
        mgf.Scene(ctx, d)
