"""Particle sets, the particle advection solvers (PIC / FLIP / mixed) and the
grid step with a particle solver, on the GPU through the C ABI, against the
CPU oracle.  Needs an MI355X.

Bars:
  * generated particles: bit-exact (same seeded generator, same order);
  * particle positions and velocities after a solver's first call: bit-exact
    (grid-to-particle sums and the advection are order-free);
  * grid outputs of the solvers (particle-to-grid splats): the sums are FP32
    atomics, as in the reference, so |got - oracle| <= 2e-5 (1 + |oracle|);
    later calls carry that rounding into the particles: positions within
    1e-5 of the grid extent, velocities within 1e-4 (1 + |v|);
  * the step with a particle solver: particle markers of the first step
    bit-exact, face velocities within 1e-4 relative L2 of the oracle over
    three steps (plain FP32 CG, relative tolerance 1e-6 on both sides);
  * the projection with particle markers equals the fluid-domain projection
    bitwise when the particles sit one per fluid cell.
"""
import numpy as np
import pytest

import mgf
import scenes

pytestmark = pytest.mark.gpu
RES = (12, 10, 9)
ORIGIN = (-0.2, 0.1, 0.0)
SPACING = (0.25, 0.25, 0.25)


@pytest.fixture(scope="module")
def ctx():
    c = mgf.Context(0)
    yield c
    c.close()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(got, exp, what):
    got, exp = np.asarray(got, np.float32).ravel(), np.asarray(exp, np.float32).ravel()
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    bad = np.nonzero(bits(got) != bits(exp))[0]
    assert bad.size == 0, "%s: %d of %d differ, first %s: got %r exp %r" % (
        what, bad.size, got.size, bad[:5], got[bad[:5]], exp[bad[:5]])


def assert_close(got, exp, what, rel=2e-5):
    got, exp = np.asarray(got, np.float64).ravel(), np.asarray(exp, np.float64).ravel()
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    err = np.abs(got - exp) / (1.0 + np.abs(exp))
    i = int(np.argmax(err)) if err.size else 0
    assert err.size == 0 or err[i] <= rel, "%s: worst %d got %r exp %r (err %g)" % (what, i, got[i], exp[i], err[i])


def face_sizes(res):
    nx, ny, nz = res
    return [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]


def fields(rng, res):
    n = int(np.prod(res))
    return rng.uniform(-1, 1, n).astype(np.float32), rng.uniform(-0.3, 1, n).astype(np.float32)


@pytest.mark.parametrize("per_cell,with_solid", [(1, True), (3, True), (2, False), (0, True)])
def test_generate_bitexact(ctx, orc, per_cell, with_solid):
    rng = np.random.default_rng(10 + per_cell)
    fluid, solid = fields(rng, RES)
    solid = solid if with_solid else None
    P = mgf.Particles.generate(ctx, RES, ORIGIN, SPACING, fluid, solid, per_cell, seed=1234)
    exp = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, per_cell, seed=1234)
    assert len(P) == len(exp)
    assert_bitexact(P.get("pos"), exp, "generated positions")
    assert not P.get("vel").any() and not P.get("color").any() and not P.get("scalar").any()
    P.close()


def test_generate_many_tiles(ctx, orc):
    """A dam-break domain of 130^3 cells: more tiles than the scan's work-items."""
    spec = scenes.dam_break(128)
    s = scenes.build_oracle_scene(orc, spec)
    phi = s.download(3)
    walls = next(spec.solids()).ravel()
    P = mgf.Particles.generate(ctx, spec.res, spec.origin, spec.spacing, phi, walls, 2, seed=99)
    exp = orc.generate_particles(spec.res, spec.origin, spec.spacing, phi, walls, 2, seed=99)
    assert len(P) == len(exp) > 0
    assert_bitexact(P.get("pos"), exp, "generated positions 130^3")
    P.close(); s.close()


def test_particles_arrays_roundtrip(ctx):
    rng = np.random.default_rng(2)
    pos = rng.standard_normal((50, 3)).astype(np.float32)
    vel = rng.standard_normal((50, 3)).astype(np.float32)
    P = mgf.Particles(ctx, pos, vel, None)
    assert len(P) == 50
    assert_bitexact(P.get("pos"), pos, "pos")
    assert_bitexact(P.get("vel"), vel, "vel")
    assert not P.get("color").any()
    col = rng.standard_normal(150).astype(np.float32)
    P.set("color", col)
    assert_bitexact(P.get("color"), col, "color")
    assert P.data("pos").value and P.data("vector").value
    with pytest.raises(mgf.MgfError):
        P.set("scalar", np.zeros(49, np.float32))
    with pytest.raises(mgf.MgfError):
        mgf.check(mgf.lib().mgf_particles_download(P.h, 9, None, 0), "bad array")
    P.close()
    empty = mgf.Particles(ctx, n=0)
    assert len(empty) == 0
    empty.close()


def swirl(res, origin, spacing):
    o, h = np.asarray(origin, np.float64), np.asarray(spacing, np.float64)
    c = o + 0.5 * np.asarray(res) * h
    out = []
    for comp in range(3):
        shape = list(res)
        shape[comp] += 1
        idx = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
        x = [o[a] + (idx[a] + (0.0 if a == comp else 0.5)) * h[a] for a in range(3)]
        f = [-(x[1] - c[1]) + 0.3, (x[0] - c[0]) - 0.1 * (x[2] - c[2]), 0.2 * np.sin(x[0]) + 0.05][comp]
        out.append(f.astype(np.float32).transpose(2, 1, 0).ravel())
    return out


def test_particles_advect_matches_op(ctx, orc):
    rng = np.random.default_rng(4)
    fluid, solid = fields(rng, RES)
    pos = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, 2, seed=3)
    u, v, w = swirl(RES, ORIGIN, SPACING)
    P = mgf.Particles(ctx, pos)
    P.advect(RES, ORIGIN, SPACING, u, v, w, 0.9, 0.5, 3, True)
    exp = orc.advect_particles(RES, ORIGIN, SPACING, u, v, w, pos, 0.9, 0.5, 3, True)
    assert_bitexact(P.get("pos"), exp, "particle set advection")
    P.close()


@pytest.mark.parametrize("kind", ["pic", "flip", "mix"])
@pytest.mark.parametrize("alg", ["linear", "quadratic", "cubic"])
def test_solver_velocity_matches_oracle(ctx, orc, kind, alg):
    rng = np.random.default_rng(20)
    fluid, solid = fields(rng, RES)
    pos = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, 2, seed=21)
    ext = float(max(np.asarray(RES) * np.asarray(SPACING)))
    P = mgf.Particles(ctx, pos)
    g = mgf.ParticleSolver(ctx, RES, ORIGIN, SPACING, kind, P, cfl=0.7, alg=alg, mixing=0.3)
    o = orc.ParticleSolver(RES, ORIGIN, SPACING, kind, pos, cfl=0.7, alg=alg, mixing=0.3)
    vel = swirl(RES, ORIGIN, SPACING)
    for call in range(3):
        inp = [a + np.float32(0.1 * call) * rng.standard_normal(a.size).astype(np.float32) for a in vel]
        got = g.advect_velocity(inp, vel, 0.4, 3, True)
        exp = o.advect_velocity(inp, vel, 0.4, 3, True)
        what = "%s %s call %d" % (kind, alg, call)
        if call == 0:
            assert_bitexact(P.get("pos"), o.get(0), what + " positions")
            assert_bitexact(P.get("vel"), o.get(1), what + " velocities")
            for a, b in zip(got, exp):
                assert_close(a, b, what)
        else:
            assert np.abs(P.get("pos") - o.get(0)).max() <= 1e-5 * ext, what
            assert_close(P.get("vel"), o.get(1), what + " velocities", rel=1e-4)
            for a, b in zip(got, exp):
                assert_close(a, b, what, rel=1e-4)
    g.close(); P.close(); o.close()


def test_solver_scalar_vector_match_oracle(ctx, orc):
    rng = np.random.default_rng(30)
    fluid, solid = fields(rng, RES)
    pos = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, 3, seed=31)
    n = int(np.prod(RES))
    P = mgf.Particles(ctx, pos)
    g = mgf.ParticleSolver(ctx, RES, ORIGIN, SPACING, "flip", P, alg="quadratic")
    o = orc.ParticleSolver(RES, ORIGIN, SPACING, "flip", pos, alg="quadratic")
    for call in range(2):
        phi = rng.standard_normal(n).astype(np.float32)
        col = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
        assert_close(g.advect_scalar(phi), o.advect_scalar(phi), "scalar")
        for a, b in zip(g.advect_vector(col), o.advect_vector(col)):
            assert_close(a, b, "vector")
        assert_bitexact(P.get("scalar"), o.get(3), "particle scalars")
        assert_bitexact(P.get("vector"), o.get(4), "particle vectors")
    g.close(); P.close(); o.close()


def test_solver_rejects_bad_arguments(ctx):
    P = mgf.Particles(ctx, np.zeros((4, 3), np.float32))
    with pytest.raises(mgf.MgfError):
        mgf.ParticleSolver(ctx, RES, ORIGIN, SPACING, "pic", P, alg="p2gsum")
    with pytest.raises(mgf.MgfError):
        h = mgf.C.c_void_p()
        mgf.check(mgf.lib().mgf_psolver_create(ctx.h, mgf.grid(RES, ORIGIN, SPACING), 5, P.h, 1.0, 2, 0.1,
                                               mgf.C.byref(h)), "bad kind")
    g = mgf.ParticleSolver(ctx, RES, ORIGIN, SPACING, "pic", P)
    U = [ctx.alloc(4 * m) for m in face_sizes(RES)]
    try:
        with pytest.raises(mgf.MgfError):   # output aliasing the input
            mgf.check(mgf.lib().mgf_psolver_advect_velocity(g.h, *U, *U, 0.1, 1, 0, *U), "alias")
    finally:
        for p in U:
            ctx.free(p)
    g.close(); P.close()


def cell_centres(spec, cells):
    nx, ny = spec.res[0], spec.res[1]
    h, o = np.float64(spec.h), np.asarray(spec.origin, np.float64)
    ijk = np.stack([cells % nx, (cells // nx) % ny, cells // (nx * ny)], 1)
    return (o + (ijk + 0.5) * h).astype(np.float32)


def plain_desc():
    return dict(rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_PLAIN,
                precond=mgf.PRECOND_IDENTITY)


def test_projection_with_particle_markers_equals_domain(ctx):
    spec = scenes.dam_break(32)
    a = scenes.build_gpu_scene(mgf, ctx, spec, **plain_desc())
    b = scenes.build_gpu_scene(mgf, ctx, spec, **plain_desc())
    for s in (a, b):
        s.update_boundaries()
        s.advect()
        s.apply_forces()
    phi, solid = a.download(3), a.download(4)
    cells = np.nonzero((phi < 0) & ~(solid < 0))[0]
    P = mgf.Particles(ctx, cell_centres(spec, cells))
    sa = a.pressure_solve()
    sb = b.pressure_solve_particles(P)
    assert sa["cg_iterations"] == sb["cg_iterations"]
    for f in (8, 0, 1, 2, 9):
        assert_bitexact(b.download(f), a.download(f), "field %d" % f)
    P.close(); a.close(); b.close()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("kind", ["pic", "flip", "mix"])
def test_step_particles_matches_oracle(ctx, orc, kind):
    spec = scenes.dam_break(32)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **plain_desc())
    o = scenes.build_oracle_scene(orc, spec)
    phi = o.download(3)
    walls = next(spec.solids()).ravel()
    pos = orc.generate_particles(spec.res, spec.origin, spec.spacing, phi, walls, 4, seed=8)
    P = mgf.Particles(ctx, pos)
    g = mgf.ParticleSolver(ctx, spec.res, spec.origin, spec.spacing, kind, P, cfl=1.0)
    ops = orc.ParticleSolver(spec.res, spec.origin, spec.spacing, kind, pos, cfl=1.0)
    p = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-6)
    for step in range(3):
        st = s.step_particles(g, 1)
        ost = o.step_particles(ops, p, 1)
        assert st["converged"] == 1 and ost["converged"] == 1
        if step == 0:
            assert_bitexact(P.get("pos"), ops.get(0), "positions after step 1")
            assert_bitexact(s.download(8), o.download(8), "particle markers")
        vg = np.concatenate([s.download(f) for f in (0, 1, 2)])
        vo = np.concatenate([o.download(f) for f in (0, 1, 2)])
        assert rel_l2(vg, vo) <= 1e-4, (kind, step, rel_l2(vg, vo))
    g.close(); P.close(); ops.close(); s.close(); o.close()


def test_scene_generate_particles_matches_oracle(ctx, orc):
    """generateFluid's particle branch: the scene's fluid domain and the marker
    of its solids (built before the first step) give the oracle's particles."""
    spec = scenes.obstacles(24, seed=3, nspheres=2)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **plain_desc())
    o = scenes.build_oracle_scene(orc, spec)
    o.update_boundaries(orc.params(spec.dt))       # static solids: the marker only
    P = s.generate_particles(3, seed=41)
    exp = orc.generate_particles(spec.res, spec.origin, spec.spacing, o.download(3), o.download(4), 3, seed=41)
    assert len(P) == len(exp) > 0
    assert_bitexact(P.get("pos"), exp, "scene particles")
    # the scene's own step still starts from the unmoved solids
    g = mgf.ParticleSolver(ctx, spec.res, spec.origin, spec.spacing, "pic", P)
    st = s.step_particles(g, 1)
    assert st["converged"] == 1
    assert_bitexact(s.download(4), o.download(4), "solid marker after the step")
    g.close(); P.close(); s.close(); o.close()
