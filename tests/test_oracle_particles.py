"""Pins the oracle's particle generation, particle advection solvers
(PIC / FLIP / mixed) and the grid step with a particle solver before the GPU
is compared with them.  CPU only.

What fixes each piece independently of the oracle's code:
  * the jitter generator: SplitMix64's published output sequence from state 0;
  * generation: particle counts, the cell each particle lies in and the order
    of the cells, determinism per seed;
  * the solvers: identities between them that a dropped term, a wrong sign or
    swapped blend weights break: FLIP's first call is PIC; MIX with mixing 1
    is PIC and with mixing 0 is FLIP; FLIP adds exactly the grid change
    (a uniform change c raises every particle velocity by c, a zero change
    keeps them bitwise), PIC does not keep them; a constant field survives a
    PIC transfer; scalar and vector fields are sampled onto the particles on
    the first call only;
  * the step with particle markers: with one particle per fluid cell the
    markers, and so the whole projection, equal the fluid-domain ones.
"""
import numpy as np
import pytest

import scenes

RES = (10, 9, 8)
ORIGIN = (-0.2, 0.1, 0.0)
SPACING = (0.25, 0.25, 0.25)


def face_sizes(res):
    nx, ny, nz = res
    return [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]


def fluid_fields(rng, res):
    n = int(np.prod(res))
    fluid = rng.uniform(-1, 1, n).astype(np.float32)
    solid = rng.uniform(-0.3, 1, n).astype(np.float32)
    return fluid, solid


def test_splitmix64_published_sequence(orc):
    # the first two outputs of SplitMix64 from state 0 (the state advances by
    # the golden-ratio increment before each output)
    L = orc.lib()
    assert L.orc_splitmix64(0) == 0xE220A8397B1DCDAF
    assert L.orc_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_generate_particles_cells_and_order(orc):
    rng = np.random.default_rng(1)
    fluid, solid = fluid_fields(rng, RES)
    per = 3
    pos = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, per, seed=7)
    cells = np.nonzero((fluid < 0) & ~(solid < 0))[0]
    assert pos.shape == (cells.size * per, 3)
    nx, ny, nz = RES
    coords = np.stack([cells % nx, (cells // nx) % ny, cells // (nx * ny)], 1)
    rel = (pos.astype(np.float64) - np.asarray(ORIGIN)) / np.asarray(SPACING)
    own = np.repeat(coords, per, axis=0)
    # inside its own cell (the far face included: x + r rounds to x + 1 when r is near 1)
    assert np.all(rel >= own - 1e-6) and np.all(rel <= own + 1 + 1e-6)
    frac = rel - own
    assert 0.4 < frac.mean() < 0.6 and frac.std() > 0.2
    again = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, per, seed=7)
    assert np.array_equal(pos, again)
    other = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, per, seed=8)
    assert not np.array_equal(pos, other)
    no_solid = orc.generate_particles(RES, ORIGIN, SPACING, fluid, None, 1, seed=7)
    assert len(no_solid) == int(np.sum(fluid < 0))


def velocity(rng, res, scale=1.0):
    return [(scale * rng.standard_normal(m)).astype(np.float32) for m in face_sizes(res)]


def solver(orc, kind, pos, mixing=0.1, alg="linear"):
    return orc.ParticleSolver(RES, ORIGIN, SPACING, kind, pos, cfl=0.7, alg=alg, mixing=mixing)


@pytest.fixture
def setup(orc):
    rng = np.random.default_rng(3)
    fluid, solid = fluid_fields(rng, RES)
    pos = orc.generate_particles(RES, ORIGIN, SPACING, fluid, solid, 2, seed=11)
    return rng, pos


@pytest.mark.parametrize("alg", ["linear", "quadratic", "cubic"])
def test_flip_first_call_is_pic(orc, setup, alg):
    rng, pos = setup
    u = velocity(rng, RES)
    a, b = solver(orc, "pic", pos, alg=alg), solver(orc, "flip", pos, alg=alg)
    oa = a.advect_velocity(u, u, 0.3, 3, True)
    ob = b.advect_velocity(u, u, 0.3, 3, True)
    for x, y in zip(oa, ob):
        assert np.array_equal(x, y)
    assert np.array_equal(a.get(0), b.get(0)) and np.array_equal(a.get(1), b.get(1))


def test_mix_endpoints_are_pic_and_flip(orc, setup):
    rng, pos = setup
    u1, u2 = velocity(rng, RES), velocity(rng, RES)
    pic, flip = solver(orc, "pic", pos), solver(orc, "flip", pos)
    m1, m0 = solver(orc, "mix", pos, mixing=1.0), solver(orc, "mix", pos, mixing=0.0)
    for u in (u1, u2):
        o_pic, o_m1 = pic.advect_velocity(u, u, 0.2, 2, True), m1.advect_velocity(u, u, 0.2, 2, True)
        o_flip, o_m0 = flip.advect_velocity(u, u, 0.2, 2, True), m0.advect_velocity(u, u, 0.2, 2, True)
        for x, y in zip(o_pic, o_m1):
            assert np.array_equal(x, y)
        for x, y in zip(o_flip, o_m0):
            assert np.array_equal(x, y)
    # the two endpoints differ once the grid has changed between calls
    assert not np.array_equal(pic.get(1), flip.get(1))
    mid = solver(orc, "mix", pos, mixing=0.5)
    for u in (u1, u2):
        mid.advect_velocity(u, u, 0.2, 2, True)
    # the blend lies between the two (per particle, up to rounding)
    a, b, c = pic.get(1), flip.get(1), mid.get(1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    assert np.all(c >= lo - 1e-5) and np.all(c <= hi + 1e-5)


def test_flip_adds_exactly_the_grid_change(orc, setup):
    rng, pos = setup
    u = velocity(rng, RES)
    zero = [np.zeros(m, np.float32) for m in face_sizes(RES)]
    flip, pic = solver(orc, "flip", pos), solver(orc, "pic", pos)
    out = flip.advect_velocity(u, zero, 0.0, 1, False)     # no motion: dt = 0
    pic_out = pic.advect_velocity(u, zero, 0.0, 1, False)
    v1 = flip.get(1)
    # no grid change: the particle velocities stay bitwise (PIC resamples them)
    flip.advect_velocity(out, zero, 0.0, 1, False)
    pic.advect_velocity(pic_out, zero, 0.0, 1, False)
    assert np.array_equal(flip.get(1), v1)
    assert not np.array_equal(pic.get(1), v1)
    # a uniform change c on the grid raises every particle velocity by c
    c = np.float32(0.75)
    out2 = flip.advect_velocity([o + c for o in out], zero, 0.0, 1, False)
    assert np.allclose(flip.get(1) - v1, c, atol=1e-5)
    # and the splat of that is the previous output plus c where weights exist
    for o2, o in zip(out2, out):
        w = o2 != 0
        assert np.allclose(o2[w], o[w] + c, atol=1e-4)


def test_pic_keeps_a_constant_field(orc, setup):
    rng, pos = setup
    zero = [np.zeros(m, np.float32) for m in face_sizes(RES)]
    const = [np.full(m, 1.5, np.float32) for m in face_sizes(RES)]
    pic = solver(orc, "pic", pos, alg="quadratic")
    out = pic.advect_velocity(const, zero, 0.0, 3, False)
    assert np.allclose(pic.get(1), 1.5, rtol=1e-6)
    for o in out:
        touched = o != 0
        assert touched.any()
        assert np.allclose(o[touched], 1.5, rtol=1e-6)


def test_scalar_and_vector_seed_once(orc, setup):
    rng, pos = setup
    n = int(np.prod(RES))
    pic = solver(orc, "pic", pos)
    phi1, phi2 = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    a = pic.advect_scalar(phi1)
    b = pic.advect_scalar(phi2)             # the particles keep their first values
    assert np.array_equal(a, b)
    assert np.array_equal(pic.get(3), orc.g2p(RES, ORIGIN, SPACING, 0, pos, phi1, "linear"))
    c1 = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    c2 = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    x = pic.advect_vector(c1)
    y = pic.advect_vector(c2)
    for p, q in zip(x, y):
        assert np.array_equal(p, q)
    for k in range(3):
        exp = orc.g2p(RES, ORIGIN, SPACING, 0, pos, c1[k], "linear", span=3, off=k)
        assert np.array_equal(pic.get(4)[k::3], exp[k::3])


def cell_centres(spec, cells):
    nx, ny = spec.res[0], spec.res[1]
    h, o = np.float64(spec.h), np.asarray(spec.origin, np.float64)
    ijk = np.stack([cells % nx, (cells // nx) % ny, cells // (nx * ny)], 1)
    return (o + (ijk + 0.5) * h).astype(np.float32)


def test_particle_markers_projection_equals_domain_projection(orc):
    spec = scenes.dam_break(16)
    a = scenes.build_oracle_scene(orc, spec)
    b = scenes.build_oracle_scene(orc, spec)
    p = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-6)
    for s in (a, b):
        s.update_boundaries(p)
        s.advect(p)
        s.apply_forces(p)
    phi = a.download(3)
    solid = a.download(4)
    fluid_cells = np.nonzero((phi < 0) & ~(solid < 0))[0]
    ps = orc.ParticleSolver(spec.res, spec.origin, spec.spacing, "pic", cell_centres(spec, fluid_cells))
    sa = a.pressure_solve(p)
    sb = b.pressure_solve_particles(ps, p)
    assert sa["cg_iterations"] == sb["cg_iterations"]
    assert np.array_equal(a.download(8), b.download(8))
    for f in (0, 1, 2, 9):
        assert np.array_equal(a.download(f), b.download(f)), f
    ps.close(); a.close(); b.close()


def test_step_with_particles_is_divergence_free(orc):
    spec = scenes.dam_break(16)
    s = scenes.build_oracle_scene(orc, spec)
    phi = s.download(3)
    solid = s.download(4)
    # the solid marker appears at the first boundary update: use the walls' sign
    walls = next(spec.solids()).ravel()
    pos = orc.generate_particles(spec.res, spec.origin, spec.spacing, phi, walls, 4, seed=5)
    assert solid.size == phi.size
    p = orc.params(spec.dt, rk=3, tol_mode=1, tol=1e-7)
    for kind in ("pic", "flip", "mix"):
        t = scenes.build_oracle_scene(orc, spec)
        ps = orc.ParticleSolver(spec.res, spec.origin, spec.spacing, kind, pos, cfl=1.0)
        st = t.step_particles(ps, p, 2)
        assert st["converged"] == 1
        markers = t.download(8)
        assert np.sum(markers == 1) > 0
        # the projection makes the velocity divergence-free on fluid cells (before
        # extrapolation changes air faces only)
        div = orc.divergence(spec.res, spec.origin, spec.spacing, t.download(0), t.download(1),
                             t.download(2)).ravel()
        fl = markers == 1
        vmax = max(np.abs(t.download(f)).max() for f in (0, 1, 2))
        assert np.abs(div[fl]).max() <= 1e-3 * vmax / spec.h, kind
        # particles moved and stay in the grid box
        q = ps.get(0).reshape(-1, 3)
        lo = np.asarray(spec.origin, np.float32)
        hi = lo + np.asarray(spec.res, np.float32) * np.float32(spec.h)
        assert np.all(q[: int(np.prod(spec.res))] >= lo) and np.all(q[: int(np.prod(spec.res))] <= hi)
        assert not np.array_equal(q, pos)
        ps.close(); t.close()
    s.close()
