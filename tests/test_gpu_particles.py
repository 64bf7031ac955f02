"""Particle operators of the HIP library (through the C ABI) against the CPU
oracle and the reference's particle goldens.  Needs an MI355X.

Bars:
  * grid-to-particle transfers, particle counts and fluid markers, particle
    advection (RK1/2/3 with CFL substeps, with and without the clamp to the
    grid box): bit-exact against the FP32 oracle;
  * particle-to-grid transfers: the sums are FP32 atomics (as in the
    reference, EulerParticlesTool.cu:443-711), so their rounding depends on
    the order the adds land in; |got - oracle| <= 2e-5 * (1 + |oracle|) per
    value, and the weights, which are sums of the same kernel values, within
    the same bound;
  * reference goldens (UnitTest-ParticlesAdvection.cu): the tolerance stored
    with each case.
"""
import numpy as np
import pytest

import mgf

pytestmark = pytest.mark.gpu
ALGS_P2G = ("p2gsum", "linear", "quadratic", "cubic")
ALGS_G2P = ("g2pnearest", "linear", "quadratic", "cubic")


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


def check_map(arr, exp_map, tol, label):
    bad = [(label, int(k), float(arr[int(k)]), v) for k, v in exp_map.items() if not abs(float(arr[int(k)]) - v) < tol]
    assert not bad, bad[:10]


def particles(rng, res, origin, spacing, n, outside=0.1):
    """n particles spread over the grid box, a fraction of them outside it."""
    lo = np.asarray(origin, np.float64)
    ext = np.asarray(res, np.float64) * np.asarray(spacing, np.float64)
    p = lo + rng.uniform(-outside, 1.0 + outside, (n, 3)) * ext
    return p.astype(np.float32)


SETUPS = [
    # (res, origin, spacing): a power-of-two spacing, a general one, a flat one
    ((12, 9, 7), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
    ((10, 13, 6), (-1.3, 0.4, 2.0), (0.3, 0.45, 0.7)),
    ((17, 1, 9), (0.5, -0.25, 0.0), (0.125, 0.125, 0.125)),
]


def test_reference_particle_goldens(ctx, cases):
    for c in cases["particles"]:
        if c["kind"] == "p2g":
            f, _ = mgf.particles_to_grid(ctx, c["res"], c["origin"], c["spacing"], c["field"], c["pos"], c["val"],
                                         c["alg"], span=c["span"])
            check_map(f, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "g2p":
            v = mgf.grid_to_particles(ctx, c["res"], c["origin"], c["spacing"], c["field"], c["pos"], c["grid"],
                                      c["alg"], span=c["span"])
            check_map(v, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "p2g_fcv":
            for k, comp in enumerate("xyz"):
                f, _ = mgf.particles_to_grid(ctx, c["res"], c["origin"], c["spacing"], k + 1, c["pos"], c["val"],
                                             c["alg"], span=3, off=k)
                if comp in c["expect"]:
                    check_map(f, c["expect"][comp], c["tol"], c["name"] + ":" + comp)
        elif c["kind"] == "advect":
            p = mgf.advect_particles(ctx, c["res"], c["origin"], c["spacing"], c["u"], c["v"], c["w"], c["pos"],
                                     c["dt"], c["cfl"], c["order"])
            check_map(p, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "markers":
            m = mgf.particle_markers(ctx, c["res"], c["origin"], c["spacing"], c["solid"], c["pos"])
            check_map(m, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "density":
            P = mgf.Particles.generate(ctx, c["res"], c["origin"], c["spacing"], c["fluid"], c["solid"],
                                       c["per_cell"], seed=2024)
            d = mgf.particle_markers(ctx, c["res"], c["origin"], c["spacing"], None, P.get("pos"))
            P.close()
            check_map(d, c["expect"], c["tol"], c["name"])
        else:
            raise AssertionError(c["kind"])


@pytest.mark.parametrize("setup", range(len(SETUPS)))
@pytest.mark.parametrize("alg", ALGS_P2G)
def test_p2g_matches_oracle(ctx, orc, setup, alg):
    res, origin, spacing = SETUPS[setup]
    rng = np.random.default_rng(100 + setup)
    pos = particles(rng, res, origin, spacing, 3000)
    val = rng.standard_normal(3 * len(pos)).astype(np.float32)
    for kind in range(4):
        off = kind - 1 if kind else 0
        for norm in (True, False):
            f, w = mgf.particles_to_grid(ctx, res, origin, spacing, kind, pos, val, alg, span=3, off=off,
                                         normalize=norm)
            ef, ew = orc.p2g(res, origin, spacing, kind, pos, val, alg, span=3, off=off, normalize=norm)
            what = "p2g %s kind %d norm %d" % (alg, kind, norm)
            assert_close(w, ew, what + " weight")
            # normalised values: where the weight is a tiny sum of far kernel
            # tails the quotient amplifies the order-dependent rounding
            # (zero weight: the plain sum is kept, as P2GSUM always does)
            keep = ((ew > 1e-3) | (ew <= 0)) if norm else np.ones(ew.shape, bool)
            assert_close(f[keep], ef[keep], what)


@pytest.mark.parametrize("setup", range(len(SETUPS)))
@pytest.mark.parametrize("alg", ALGS_G2P)
def test_g2p_bitexact(ctx, orc, setup, alg):
    res, origin, spacing = SETUPS[setup]
    rng = np.random.default_rng(200 + setup)
    pos = particles(rng, res, origin, spacing, 4000)
    for kind in range(4):
        field = rng.standard_normal(mgf.field_extent(res, kind)).astype(np.float32)
        prior = rng.standard_normal(3 * len(pos)).astype(np.float32)
        off = (kind + 1) % 3
        got = mgf.grid_to_particles(ctx, res, origin, spacing, kind, pos, field, alg, span=3, off=off, out=prior)
        exp = orc.g2p(res, origin, spacing, kind, pos, field, alg, span=3, off=off, out=prior)
        assert_bitexact(got, exp, "g2p %s kind %d" % (alg, kind))
        # the other slots of each particle are left as they were
        other = np.ones(3 * len(pos), bool)
        other[off::3] = False
        assert_bitexact(got[other], prior[other], "g2p untouched slots")


@pytest.mark.parametrize("setup", range(len(SETUPS)))
def test_counts_and_markers_bitexact(ctx, orc, setup):
    res, origin, spacing = SETUPS[setup]
    rng = np.random.default_rng(300 + setup)
    pos = particles(rng, res, origin, spacing, 5000, outside=0.3)
    got = mgf.particle_markers(ctx, res, origin, spacing, None, pos)
    exp = orc.particle_markers(res, origin, spacing, None, pos, counts_only=True)
    assert_bitexact(got, exp, "particle counts")
    assert got.sum() == len(pos)    # every particle lands in a (clamped) cell
    solid = rng.uniform(-1, 1, int(np.prod(res))).astype(np.float32)
    got = mgf.particle_markers(ctx, res, origin, spacing, solid, pos)
    exp = orc.particle_markers(res, origin, spacing, solid, pos)
    assert_bitexact(got, exp, "fluid markers")
    assert set(np.unique(got)) <= {0.0, 1.0, 2.0}


def swirl(res, origin, spacing, scale=1.0):
    """A smooth face-centred velocity: a rotation about the box centre plus a
    drift, sampled at the face positions of each component."""
    nx, ny, nz = res
    o = np.asarray(origin, np.float64)
    h = np.asarray(spacing, np.float64)
    c = o + 0.5 * np.asarray(res) * h
    out = []
    for comp in range(3):
        shape = [nx, ny, nz]
        shape[comp] += 1
        idx = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
        x = [o[a] + (idx[a] + (0.0 if a == comp else 0.5)) * h[a] for a in range(3)]
        if comp == 0:
            f = -(x[1] - c[1]) + 0.3
        elif comp == 1:
            f = (x[0] - c[0]) - 0.1 * (x[2] - c[2])
        else:
            f = 0.2 * np.sin(x[0]) + 0.05
        out.append((scale * f).astype(np.float32).transpose(2, 1, 0).ravel())
    return out


@pytest.mark.parametrize("setup", range(len(SETUPS)))
@pytest.mark.parametrize("order", (1, 2, 3))
@pytest.mark.parametrize("fix", (False, True))
def test_advect_particles_bitexact(ctx, orc, setup, order, fix):
    res, origin, spacing = SETUPS[setup]
    rng = np.random.default_rng(400 + 10 * setup + order)
    pos = particles(rng, res, origin, spacing, 2500, outside=0.0)
    u, v, w = swirl(res, origin, spacing)
    dt = 2.7 * min(spacing)   # several CFL substeps at cfl 0.5
    got = mgf.advect_particles(ctx, res, origin, spacing, u, v, w, pos, dt, 0.5, order, fix)
    exp = orc.advect_particles(res, origin, spacing, u, v, w, pos, dt, 0.5, order, fix)
    assert_bitexact(got, exp, "advect order %d fix %d" % (order, fix))
    assert not np.array_equal(got, pos.ravel())


def test_advect_particles_still_field(ctx, orc):
    res, origin, spacing = SETUPS[0]
    pos = particles(np.random.default_rng(5), res, origin, spacing, 100)
    z = [np.zeros(mgf.field_extent(res, k), np.float32) for k in (1, 2, 3)]
    got = mgf.advect_particles(ctx, res, origin, spacing, *z, pos, 0.25, 0.5, 3, False)
    assert_bitexact(got, pos, "no velocity, no motion")


def test_axpy_scale_fma(ctx):
    rng = np.random.default_rng(6)
    n = 100003
    x, y = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    a = np.float32(0.37)
    X, Y = ctx.to_device(x), ctx.to_device(y)
    try:
        mgf.check(mgf.lib().mgf_op_axpy(ctx.h, n, a, X, Y), "axpy")
        got = ctx.to_host(Y, n)
        # a*x is exact in 80-bit long double; the sum then rounds once more
        # only in bits far below FP32's, so this is fma(a, x, y) rounded once
        L = np.longdouble
        exp = (L(a) * x.astype(L) + y.astype(L)).astype(np.float32)
        assert_bitexact(got, exp, "axpy")
        mgf.check(mgf.lib().mgf_op_scale(ctx.h, n, a, X), "scale")
        assert_bitexact(ctx.to_host(X, n), (a * x).astype(np.float32), "scale")
    finally:
        ctx.free(X)
        ctx.free(Y)


def test_empty_particle_sets(ctx, orc):
    res, origin, spacing = SETUPS[1]
    pos = np.zeros((0, 3), np.float32)
    f, w = mgf.particles_to_grid(ctx, res, origin, spacing, 0, pos, np.zeros(0, np.float32), "linear")
    assert not f.any() and not w.any()
    assert mgf.grid_to_particles(ctx, res, origin, spacing, 2, pos, np.ones(mgf.field_extent(res, 2)),
                                 "cubic").size == 0
    assert not mgf.particle_markers(ctx, res, origin, spacing, None, pos).any()
    u, v, w3 = swirl(res, origin, spacing)
    assert mgf.advect_particles(ctx, res, origin, spacing, u, v, w3, pos, 0.5, 0.5, 2).size == 0


def test_bad_particle_arguments(ctx):
    res, origin, spacing = SETUPS[0]
    pos = np.zeros((4, 3), np.float32)
    val = np.zeros(4, np.float32)
    with pytest.raises(mgf.MgfError):
        mgf.particles_to_grid(ctx, res, origin, spacing, 0, pos, val, "g2pnearest")
    with pytest.raises(mgf.MgfError):
        mgf.grid_to_particles(ctx, res, origin, spacing, 0, pos, np.zeros(mgf.field_extent(res, 0)), "p2gsum")
    with pytest.raises(mgf.MgfError):
        mgf.particles_to_grid(ctx, res, origin, spacing, 4, pos, val, "linear")
    with pytest.raises(mgf.MgfError):
        mgf.particles_to_grid(ctx, res, origin, spacing, 0, pos, val, "linear", span=2, off=2)
    u, v, w = swirl(res, origin, spacing)
    with pytest.raises(mgf.MgfError):
        mgf.advect_particles(ctx, res, origin, spacing, u, v, w, pos, float("nan"), 0.5, 1)
