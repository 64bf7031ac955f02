"""Parity of the HIP library (through the C ABI) with the CPU oracle and with
the reference's own golden values.  Needs an MI355X.

Bars (the same ones the oracle was pinned to):
  * markers, matrix rows, right-hand side, back-traced positions, advected
    fields, extrapolated faces: bit-exact against the FP32 oracle;
  * pressure / projected velocity: plain mode within 1e-6 relative L2 of the
    FP32 oracle (its dot products differ only in FP64 summation order); mixed
    mode within 1e-4 relative L2 of the FP32 oracle, and its true residual
    ||b - A p|| / ||b|| <= 1e-6 in FP64;
  * reference goldens: |got - expected| < 1e-4 (GRID_SOLVER_EPSILON).
"""
import numpy as np
import pytest

import mgf
import scenes

pytestmark = pytest.mark.gpu
FLT_MAX = float(np.finfo(np.float32).max)


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


def rel_l2(a, b, mask=None):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    if mask is not None:
        a, b = a[mask], b[mask]
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def finite_faces(*arrs):
    return [np.asarray(a) != FLT_MAX for a in arrs]


# ---------------------------------------------------------------------------
# Operators, bit-exact against the oracle.
def rand_vel(rng, res, scale):
    nx, ny, nz = res
    return [(rng.standard_normal(n) * scale).astype(np.float32)
            for n in ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1))]


@pytest.mark.parametrize("order", [1, 2, 3])
def test_backtrace_bitexact(ctx, orc, order):
    rng = np.random.default_rng(order)
    res, o, h = (9, 7, 5), (0.25, -0.5, 1.0), (0.5, 0.25, 0.75)
    u, v, w = rand_vel(rng, res, 1.5)
    n = 4000
    lo, hi = np.array(o) - 1.0, np.array(o) + np.array(res) * np.array(h) + 1.0
    P = [rng.uniform(lo[k], hi[k], n).astype(np.float32) for k in range(3)]
    dt = 0.37
    e = orc.backtrace(res, o, h, u, v, w, dt, order, *P)
    dev = [ctx.to_device(a) for a in (u, v, w)] + [ctx.to_device(a) for a in P]
    g = mgf.grid(res, o, h)
    mgf.check(mgf.lib().mgf_op_backtrace(ctx.h, g, dev[0], dev[1], dev[2], np.float32(dt), order,
                                         dev[3], dev[4], dev[5], n), "backtrace")
    for k in range(3):
        assert_bitexact(ctx.to_host(dev[3 + k], n), e[k], "backtrace p%d" % k)
    for p in dev:
        ctx.free(p)


def test_sample_cc_bitexact(ctx, orc):
    rng = np.random.default_rng(7)
    res, o, h = (11, 6, 4), (-1.0, 0.5, 0.25), (0.3, 0.7, 1.1)
    f = rng.standard_normal(np.prod(res)).astype(np.float32)
    n = 5000
    P = [rng.uniform(o[k] - 2, o[k] + res[k] * h[k] + 2, n).astype(np.float32) for k in range(3)]
    e = orc.sample_cc(res, o, h, f, *P)
    df, *dp = [ctx.to_device(a) for a in [f] + P]
    out = ctx.alloc(4 * n)
    mgf.check(mgf.lib().mgf_op_sample_cc(ctx.h, mgf.grid(res, o, h), df, dp[0], dp[1], dp[2], out, n),
              "sample_cc")
    assert_bitexact(ctx.to_host(out, n), e, "sample_cc")
    for p in [df, out] + dp:
        ctx.free(p)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_advect_bitexact(ctx, orc, order):
    rng = np.random.default_rng(10 + order)
    res, o, h = (13, 10, 8), (-0.1, 0.0, 0.2), (0.1, 0.1, 0.1)
    u, v, w = rand_vel(rng, res, 0.8)
    iu, iv, iw = rand_vel(rng, res, 1.0)
    phi = rng.standard_normal(np.prod(res)).astype(np.float32)
    dt = 0.05
    eu, ev, ew = orc.advect_fc(res, o, h, iu, iv, iw, u, v, w, dt, order)
    ephi = orc.advect_cc(res, o, h, phi, u, v, w, dt, order)
    g = mgf.grid(res, o, h)
    D = {k: ctx.to_device(a) for k, a in dict(u=u, v=v, w=w, iu=iu, iv=iv, iw=iw, phi=phi).items()}
    O = [ctx.alloc(4 * a.size) for a in (iu, iv, iw)]
    op = ctx.alloc(4 * phi.size)
    mgf.check(mgf.lib().mgf_op_advect_fc(ctx.h, g, D["iu"], D["iv"], D["iw"], D["u"], D["v"], D["w"],
                                         np.float32(dt), order, O[0], O[1], O[2]), "advect_fc")
    mgf.check(mgf.lib().mgf_op_advect_cc(ctx.h, g, D["phi"], D["u"], D["v"], D["w"], np.float32(dt),
                                         order, op), "advect_cc")
    for k, e in enumerate((eu, ev, ew)):
        assert_bitexact(ctx.to_host(O[k], e.size), e, "advect_fc c%d" % k)
    assert_bitexact(ctx.to_host(op, phi.size), ephi, "advect_cc")
    for p in list(D.values()) + O + [op]:
        ctx.free(p)


@pytest.mark.parametrize("seed", [0, 1])
def test_matvec_bitexact(ctx, orc, seed):
    rng = np.random.default_rng(seed)
    res = (12, 9, 7)
    n = int(np.prod(res))
    markers = rng.choice([0, 1, 1, 1, 2], n).astype(np.uint8)
    scale = np.array([8.19, 4.1, 2.05], np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    A = orc.build_matrix(res, markers.astype(np.float32), scale)
    e = orc.matvec(res, A, x)
    dm, dx, dy = ctx.to_device(markers), ctx.to_device(x), ctx.alloc(4 * n)
    mgf.check(mgf.lib().mgf_op_pressure_matvec(ctx.h, mgf.grid(res), mgf.fp(scale), dm, dx, dy), "matvec")
    assert_bitexact(ctx.to_host(dy, n), e, "matvec")
    for p in (dm, dx, dy):
        ctx.free(p)


@pytest.mark.parametrize("layers", [0, 1, 3, 1000])
def test_extrapolate_bitexact(ctx, orc, layers):
    rng = np.random.default_rng(layers)
    res = (10, 8, 6)
    u, v, w = rand_vel(rng, res, 1.0)
    for a in (u, v, w):
        a[rng.random(a.size) < 0.93] = FLT_MAX
    eu, ev, ew, used = orc.extrapolate_fc(res, u, v, w, layers)
    D = [ctx.to_device(a) for a in (u, v, w)]
    got_used = mgf.C.c_int32()
    mgf.check(mgf.lib().mgf_op_extrapolate_fc(ctx.h, mgf.grid(res), D[0], D[1], D[2], layers,
                                              mgf.C.byref(got_used)), "extrapolate")
    for k, e in enumerate((eu, ev, ew)):
        assert_bitexact(ctx.to_host(D[k], e.size), e, "extrapolate c%d" % k)
    assert got_used.value == used
    for p in D:
        ctx.free(p)


@pytest.mark.parametrize("res", [(300, 20, 18), (600, 9, 7), (1100, 4, 3), (8, 530, 6), (7, 6, 520),
                                 (6, 300, 5), (6, 900, 5), (5, 4, 1030)])
def test_extrapolate_bitexact_long_lines(ctx, orc, res):
    """Grids whose rows and lines select the other distance-transform
    kernels: rows of 301 and 601 faces (8 and 16 register chunks per lane), of
    1101 (the two-pass row kernel); y and z lines of 301, 531, 521 and 901
    faces (the segmented line kernel's classes up to 320, 576 and 1024 faces)
    and z lines of 1031 (the two-sweep line kernel)."""
    rng = np.random.default_rng(sum(res))
    u, v, w = rand_vel(rng, res, 1.0)
    for a in (u, v, w):
        a[rng.random(a.size) < 0.97] = FLT_MAX
    eu, ev, ew, used = orc.extrapolate_fc(res, u, v, w, 1000)
    D = [ctx.to_device(a) for a in (u, v, w)]
    got_used = mgf.C.c_int32()
    mgf.check(mgf.lib().mgf_op_extrapolate_fc(ctx.h, mgf.grid(res), D[0], D[1], D[2], 1000,
                                              mgf.C.byref(got_used)), "extrapolate")
    for k, e in enumerate((eu, ev, ew)):
        assert_bitexact(ctx.to_host(D[k], e.size), e, "extrapolate c%d" % k)
    assert got_used.value == used
    for p in D:
        ctx.free(p)


def test_extrapolate_bitexact_16bit_distances(ctx, orc):
    """Face distances are stored in 16 bits: a row of 65001 u faces with one
    known face at its start has distances up to 65000 (the sorted layers stop
    at 1000); a grid with nx + ny + nz + 3 >= 65535 is refused before any
    launch."""
    res = (65000, 2, 1)
    rng = np.random.default_rng(7)
    u, v, w = rand_vel(rng, res, 1.0)
    u[:] = FLT_MAX
    u[0] = 0.5
    for a in (v, w):
        a[rng.random(a.size) < 0.97] = FLT_MAX
    eu, ev, ew, used = orc.extrapolate_fc(res, u, v, w, 1000)
    D = [ctx.to_device(a) for a in (u, v, w)]
    got_used = mgf.C.c_int32()
    mgf.check(mgf.lib().mgf_op_extrapolate_fc(ctx.h, mgf.grid(res), D[0], D[1], D[2], 1000,
                                              mgf.C.byref(got_used)), "extrapolate")
    for k, e in enumerate((eu, ev, ew)):
        assert_bitexact(ctx.to_host(D[k], e.size), e, "extrapolate c%d" % k)
    assert got_used.value == used
    for p in D:
        ctx.free(p)
    big = (65533, 1, 1)
    B = [ctx.to_device(np.full(n, FLT_MAX, np.float32)) for n in (65534, 2 * 65533, 2 * 65533)]
    st = mgf.lib().mgf_op_extrapolate_fc(ctx.h, mgf.grid(big), B[0], B[1], B[2], 1000, mgf.C.byref(got_used))
    assert st == mgf.ERR_UNSUPPORTED
    for p in B:
        ctx.free(p)


# ---------------------------------------------------------------------------
# Reference goldens through the GPU.
def pressure_scene_from_case(ctx, c, precision=mgf.PRECISION_PLAIN):
    d = mgf.make_desc(c["res"], c["origin"], c["spacing"], c["dt"], precision=precision)
    s = mgf.Scene(ctx, d)
    i = c["inputs"]
    for fid, key in ((0, "u"), (1, "v"), (2, "w"), (3, "fluid"), (4, "solid"), (5, "us"), (6, "vs"),
                     (7, "ws")):
        s.upload(fid, i[key])
    return s


@pytest.mark.parametrize("idx", range(11))
def test_reference_pressure_goldens(ctx, cases, orc, idx):
    c = cases["pressure"][idx]
    i = c["inputs"]
    s = pressure_scene_from_case(ctx, c)
    st = s.pressure_solve()
    got = dict(markers=s.download(8), A=s.download(11), b=s.download(10), x=s.download(9),
               u=s.download(0), v=s.download(1), w=s.download(2))
    e = c["expect"]
    ov = c.get("tol_override") or {}
    for key, arr in got.items():
        for k, val in e.get(key, {}).items():
            tol = ov.get(key + ":" + k, c["tol"])
            assert abs(float(arr[int(k)]) - val) < tol or (val == FLT_MAX and arr[int(k)] == FLT_MAX), \
                (c["name"], key, k, arr[int(k)], val)
    o = orc.pressure_solve(c["res"], c["origin"], c["spacing"], c["dt"], i["u"], i["v"], i["w"],
                           i["us"], i["vs"], i["ws"], i["solid"], i["fluid"])
    for key in ("markers", "A", "b"):
        assert_bitexact(got[key], o[key], c["name"] + " " + key)
    assert st["cg_iterations"] == o["stats"]["cg_iterations"]
    for key in ("x", "u", "v", "w"):
        fin = np.asarray(o[key]) != FLT_MAX
        assert np.array_equal(np.asarray(got[key]) != FLT_MAX, fin), key
        assert rel_l2(got[key], o[key], fin) <= 1e-6, (c["name"], key)
    s.close()


def test_reference_semilagrangian_goldens(ctx, cases):
    for c in cases["semilagrangian"]:
        i = c["inputs"]
        g = mgf.grid(c["res"], c["origin"], c["spacing"])
        V = [ctx.to_device(np.asarray(i[k], np.float32)) for k in ("u", "v", "w")]
        if c["kind"] == "backtrace":
            P = [ctx.to_device(np.asarray(i[k], np.float32)) for k in ("px", "py", "pz")]
            n = len(i["px"])
            mgf.check(mgf.lib().mgf_op_backtrace(ctx.h, g, V[0], V[1], V[2], np.float32(c["dt"]),
                                                 c["order"], P[0], P[1], P[2], n), "backtrace")
            for k, key in enumerate(("px", "py", "pz")):
                got = ctx.to_host(P[k], n)
                assert np.max(np.abs(got - np.asarray(c["expect"][key]))) < c["tol"], c["name"]
            for p in P:
                ctx.free(p)
        elif c["kind"] == "advect_cc":
            f = ctx.to_device(np.asarray(i["field"], np.float32))
            n = len(i["field"])
            out = ctx.alloc(4 * n)
            mgf.check(mgf.lib().mgf_op_advect_cc(ctx.h, g, f, V[0], V[1], V[2], np.float32(c["dt"]),
                                                 c["order"], out), "advect_cc")
            got = ctx.to_host(out, n)
            if "field_parts" in c["expect"]:
                for k, (a, b, cc) in c["expect"]["field_parts"].items():
                    d = ((np.float32(got[int(k)]) - np.float32(a)) - np.float32(b)) - np.float32(cc)
                    assert abs(d) < c["tol"], (c["name"], k)
            elif isinstance(c["expect"]["field"], dict):
                for k, val in c["expect"]["field"].items():
                    assert abs(got[int(k)] - val) < c["tol"], (c["name"], k)
            else:
                assert np.max(np.abs(got - np.asarray(c["expect"]["field"]))) < c["tol"], c["name"]
            ctx.free(f)
            ctx.free(out)
        else:
            I = [ctx.to_device(np.asarray(i[k], np.float32)) for k in ("iu", "iv", "iw")]
            O = [ctx.alloc(4 * len(i[k])) for k in ("iu", "iv", "iw")]
            mgf.check(mgf.lib().mgf_op_advect_fc(ctx.h, g, I[0], I[1], I[2], V[0], V[1], V[2],
                                                 np.float32(c["dt"]), c["order"], O[0], O[1], O[2]),
                      "advect_fc")
            for k, key in enumerate("uvw"):
                got = ctx.to_host(O[k], len(c["expect"][key]))
                assert np.max(np.abs(got - np.asarray(c["expect"][key]))) < c["tol"], c["name"]
            for p in I + O:
                ctx.free(p)
        for p in V:
            ctx.free(p)


def test_reference_extrapolation_fc_golden(ctx, cases):
    c = [x for x in cases["extrapolation"] if x["kind"] == "fc"][0]
    D = [ctx.to_device(np.asarray(c["inputs"][k], np.float32)) for k in "uvw"]
    mgf.check(mgf.lib().mgf_op_extrapolate_fc(ctx.h, mgf.grid(c["res"]), D[0], D[1], D[2], c["layers"],
                                              None), "extrapolate")
    for k, key in enumerate("uvw"):
        got = ctx.to_host(D[k], len(c["expect"][key]))
        for j, (g_, e_) in enumerate(zip(got, c["expect"][key])):
            assert (e_ == FLT_MAX and g_ == FLT_MAX) or abs(g_ - e_) < c["tol"], (key, j, g_, e_)
        ctx.free(D[k])


def test_reference_external_forces_golden(ctx, cases):
    c = cases["external_forces"][0]
    d = mgf.make_desc(c["res"], dt=c["dt"], gravity=c["gravity"], force=c["force"])
    s = mgf.Scene(ctx, d)
    for fid, k in ((0, "u"), (1, "v"), (2, "w")):
        s.upload(fid, c["inputs"][k])
    for n in range(9):
        s.apply_forces()
        for fid, k in ((0, "u"), (1, "v"), (2, "w")):
            assert np.max(np.abs(s.download(fid) - np.asarray(c["expect_steps"][n][k]))) < c["tol"]
    s.close()


def test_reference_boundary_goldens(ctx, cases):
    for c in cases["boundary"]:
        if c.get("kind") in ("move", "resample"):
            continue  # translations / per-solid SDFs are internal state here
        d = mgf.make_desc(c["res"], c["origin"], c["spacing"], 0.0)
        s = mgf.Scene(ctx, d)
        for sol in c["solids"]:
            s.add_solid(sol["sdf"], sol["vel"])
        s.update_boundaries()
        if "marker" in c["expect"]:
            m = s.download(4)
            for k, val in c["expect"]["marker"].items():
                assert abs(m[int(k)] - val) < c["tol"], k
        else:
            F = dict(u=s.download(5), v=s.download(6), w=s.download(7))
            for key, val in c["expect"]["abs"].items():
                comp, idx = key.split(":")
                assert abs(abs(F[comp][int(idx)]) - val) < c["tol"], key
            for key, val in c["expect"]["val"].items():
                comp, idx = key.split(":")
                assert abs(F[comp][int(idx)] - val) < c["tol"], key
        s.close()


def test_boundary_matches_oracle(ctx, orc, cases):
    for c in cases["boundary"]:
        if c.get("kind") in ("move", "resample"):
            continue
        d = mgf.make_desc(c["res"], c["origin"], c["spacing"], 0.0)
        s = mgf.Scene(ctx, d)
        o = orc.Scene(c["res"], c["origin"], c["spacing"])
        for sol in c["solids"]:
            s.add_solid(sol["sdf"], sol["vel"])
            o.add_solid(sol["sdf"], sol["vel"])
        s.update_boundaries()
        o.update_boundaries(orc.params(0.0))
        for fid in (4, 5, 6, 7):
            assert_bitexact(s.download(fid), o.download(fid), "boundary field %d" % fid)
        s.close()
        o.close()


def test_moving_solids_match_oracle(ctx, orc):
    """Moving obstacles (CBoundarys::updateBoundarys, Boundary.cpp:115-150):
    each refresh moves every solid by -dt * velocity, re-samples its SDF and
    rebuilds the solid marker and the solid face velocity; bit-exact against
    the oracle over several refreshes, the faces between two moving solids
    included (taken in the sequential order the oracle states)."""
    spec = scenes.obstacles(24, seed=6, nspheres=4)
    dt = 0.05
    s = mgf.Scene(ctx, mgf.make_desc(spec.res, spec.origin, spec.spacing, dt))
    o = orc.Scene(spec.res, spec.origin, spec.spacing)
    vels = [(0, 0, 0), (0.5, 0, 0), (0, -0.3, 0.2), (0.1, 0.1, -0.1), (-0.4, 0.2, 0)]
    for k, sdf in enumerate(spec.solids()):
        s.add_solid(sdf, vels[k])
        o.add_solid(sdf, vels[k])
    p = orc.params(dt)
    for step in range(4):
        s.update_boundaries()
        o.update_boundaries(p)
        for fid in (4, 5, 6, 7):
            assert_bitexact(s.download(fid), o.download(fid), "refresh %d field %d" % (step, fid))
    s.close(); o.close()


@pytest.mark.parametrize("precision", [mgf.PRECISION_PLAIN, mgf.PRECISION_MIXED])
def test_reference_integration_goldens(ctx, cases, precision):
    for c in cases["integration"]:
        d = mgf.make_desc(c["res"], c["origin"], c["spacing"], c["dt"], rk=3, precision=precision)
        s = mgf.Scene(ctx, d)
        s.add_solid(c["inputs"]["solid"])
        s.upload(3, c["inputs"]["fluid"])
        for n in range(c["steps"]):
            s.step(1)
            if c["expect"].get("all_faces_zero"):
                for f in (0, 1, 2):
                    assert np.max(np.abs(s.download(f))) < c["tol"], (c["name"], n, f)
            else:
                v = s.download(1)
                for face in c["expect"]["v_faces"]:
                    assert abs(v[face] - c["expect"]["v_per_step"][n]) < c["tol"], (c["name"], n, face)
        s.close()


# ---------------------------------------------------------------------------
# Whole steps on the dam-break scene (64^3, dx = 1/64, dt = 2e-3, RK1).
def dam_desc_kw(precision, precond=mgf.PRECOND_IDENTITY):
    return dict(rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=precision, precond=precond)


PRECONDS = [pytest.param(mgf.PRECOND_IDENTITY, id="identity"), pytest.param(mgf.PRECOND_JACOBI, id="jacobi"),
            pytest.param(mgf.PRECOND_MULTIGRID, id="multigrid")]


def sync_state(s, o):
    """Start the GPU scene from the oracle's state so each stage is compared
    from identical inputs."""
    for fid in (0, 1, 2, 3):
        s.upload(fid, o.download(fid))


def test_step_stages_match_oracle(ctx, orc):
    spec = scenes.dam_break(32, dt=2e-3)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **dam_desc_kw(mgf.PRECISION_PLAIN))
    o = scenes.build_oracle_scene(orc, spec)
    p = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-6)
    for step in range(3):
        sync_state(s, o)
        s.update_boundaries(); o.update_boundaries(p)
        for fid in (4, 5, 6, 7):
            assert_bitexact(s.download(fid), o.download(fid), "solid %d" % fid)
        s.advect(); o.advect(p)
        for fid in (0, 1, 2, 3):
            assert_bitexact(s.download(fid), o.download(fid), "step %d advected field %d" % (step, fid))
        s.apply_forces(); o.apply_forces(p)
        for fid in (0, 1, 2):
            assert_bitexact(s.download(fid), o.download(fid), "step %d forced field %d" % (step, fid))
        st = s.pressure_solve(); ost = o.pressure_solve(p)
        assert_bitexact(s.download(8), o.download(8), "step %d markers" % step)
        assert_bitexact(s.download(10), o.download(10), "step %d rhs" % step)
        assert abs(st["cg_iterations"] - ost["cg_iterations"]) <= 1, (st, ost)
        for fid in (0, 1, 2):
            g_, e_ = s.download(fid), o.download(fid)
            assert np.array_equal(g_ == FLT_MAX, e_ == FLT_MAX)
            assert rel_l2(g_, e_, e_ != FLT_MAX) <= 1e-6
        # extrapolation from identical projected fields is bit-exact
        for fid in (0, 1, 2):
            s.upload(fid, o.download(fid))
        s.extrapolate(); o.extrapolate(p)
        for fid in (0, 1, 2):
            assert_bitexact(s.download(fid), o.download(fid), "step %d extrapolated %d" % (step, fid))
    s.close(); o.close()


@pytest.mark.parametrize("rk", [1, 3])
def test_colour_advection_bitexact(ctx, orc, rk):
    """The colour field (setColorField, GridFluidSolver.cpp:208-212) is advected
    with the velocity the step starts from, as advect(CCV) does: bit-exact
    against the oracle.  The full step carries the same colour as the stage
    path, and a scene without colour reads zeros."""
    spec = scenes.obstacles(24, seed=4)
    kw = dict(rk=rk, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
              precond=mgf.PRECOND_MULTIGRID)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **kw)
    full = scenes.build_gpu_scene(mgf, ctx, spec, **kw)
    o = scenes.build_oracle_scene(orc, spec)
    assert not np.any(s.download(13))
    rng = np.random.default_rng(rk)
    for f in (0, 1, 2):
        v = (rng.standard_normal(s.count(f)) * 0.5).astype(np.float32)
        s.upload(f, v); full.upload(f, v); o.upload(f, v)
    for k in range(3):
        c = rng.uniform(0, 1, spec.padded_cells).astype(np.float32)
        s.upload(13 + k, c); full.upload(13 + k, c); o.upload(13 + k, c)
    p = orc.params(spec.dt, rk=rk, tol_mode=1, tol=1e-6)
    s.update_boundaries(); o.update_boundaries(p)
    s.advect(); o.advect(p)
    for k in range(3):
        assert_bitexact(s.download(13 + k), o.download(13 + k), "colour %d" % k)
    assert_bitexact(s.download(3), o.download(3), "fluid domain advected beside the colour")
    full.step(1)
    for k in range(3):
        assert_bitexact(full.download(13 + k), s.download(13 + k), "colour after a full step %d" % k)
    s.close(); full.close(); o.close()


@pytest.mark.parametrize("layers", [1000, 4])
def test_fluid_sdf_and_density_bitexact(ctx, orc, layers):
    """generateCurFluidSDF / generateCurFluidDensity (GridFluidSolver.cpp:
    571-613) from the scene's current fluid domain and solid marker, bit-exact
    against the oracle's restatement of the reference's passes."""
    spec = scenes.obstacles(30, seed=5)
    s = scenes.build_gpu_scene(mgf, ctx, spec, rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6,
                               precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_MULTIGRID, layers=layers)
    for step in range(2):
        if step:
            s.step(1)
        phi, solid = s.download(3), s.download(4)
        assert_bitexact(s.download(17), orc.fluid_density(spec.res, phi, solid), "density step %d" % step)
        assert_bitexact(s.download(16), orc.fluid_sdf(spec.res, phi, layers), "sdf step %d" % step)
    s.close()


def compare_trajectory(s, o, step, tol_v, what):
    phi_g, phi_o = s.download(3), o.download(3)
    if step == 0:
        assert_bitexact(phi_g, phi_o, what + " fluid domain step 0")
    # marker-defining sign of the fluid domain
    flips = int(np.sum((phi_g < 0) != (phi_o < 0)))
    assert flips <= max(2, phi_g.size // 100000), (what, step, flips)
    assert rel_l2(phi_g, phi_o) <= 1e-5, (what, step, rel_l2(phi_g, phi_o))
    # relative L2 over the whole face-centred velocity (a near-zero component
    # alone has no meaningful relative error)
    g_ = np.concatenate([s.download(f) for f in (0, 1, 2)])
    e_ = np.concatenate([o.download(f) for f in (0, 1, 2)])
    assert rel_l2(g_, e_) <= tol_v, (what, step, rel_l2(g_, e_))


def host_div_rel(s):
    """||D u|| / ||b|| over the fluid cells (last layer of each axis left out)
    from the downloaded FP32 velocity, markers and right-hand side, in FP64:
    the host restatement of the library's divergence check."""
    nx, ny, nz = [int(v) for v in (s.desc.nx, s.desc.ny, s.desc.nz)]
    h = [float(v) for v in s.desc.spacing]
    u = s.download(0).astype(np.float64).reshape(nz, ny, nx + 1)
    v = s.download(1).astype(np.float64).reshape(nz, ny + 1, nx)
    w = s.download(2).astype(np.float64).reshape(nz + 1, ny, nx)
    m = s.download(8).reshape(nz, ny, nx) == 1
    m[-1, :, :] = False; m[:, -1, :] = False; m[:, :, -1] = False
    with np.errstate(over="ignore", invalid="ignore"):
        d = (u[:, :, 1:] - u[:, :, :-1]) / h[0] + (v[:, 1:, :] - v[:, :-1, :]) / h[1] + (w[1:] - w[:-1]) / h[2]
    b = s.download(10).astype(np.float64).reshape(nz, ny, nx)
    return float(np.linalg.norm(d[m]) / np.linalg.norm(b[m]))


def fp32_rounding_floor(o64, spec):
    """||D round32(u)|| / ||b|| of the FP64 oracle's projected velocity: the
    divergence left by storing an (almost) exactly projected velocity in FP32."""
    nx, ny, nz = spec.res
    h = spec.spacing
    with np.errstate(over="ignore", invalid="ignore"):
        u = o64.download(0, f64=True).astype(np.float32).astype(np.float64).reshape(nz, ny, nx + 1)
        v = o64.download(1, f64=True).astype(np.float32).astype(np.float64).reshape(nz, ny + 1, nx)
        w = o64.download(2, f64=True).astype(np.float32).astype(np.float64).reshape(nz + 1, ny, nx)
        d = (u[:, :, 1:] - u[:, :, :-1]) / h[0] + (v[:, 1:, :] - v[:, :-1, :]) / h[1] + (w[1:] - w[:-1]) / h[2]
    m = o64.download(8).reshape(nz, ny, nx) == 1
    m[-1, :, :] = False; m[:, -1, :] = False; m[:, :, -1] = False
    b = o64.download(10, f64=True).astype(np.float64).reshape(nz, ny, nx)
    return float(np.linalg.norm(d[m]) / np.linalg.norm(b[m]))


def test_dam_break_64_plain_matches_fp32_oracle(ctx, orc):
    """Config 1 with the plain FP32 CG, all 10 steps against the FP32 oracle.
    The reference's stop rule is on the FP32 recurrence residual; the FP64
    residual of the FP32 pressure stagnates above 1e-6 (true_converged 0)."""
    spec = scenes.dam_break(64)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **dam_desc_kw(mgf.PRECISION_PLAIN))
    s.set_divergence_check(True)
    o = scenes.build_oracle_scene(orc, spec)
    p = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-6)
    for step in range(10):
        st = s.step(1)
        ost = o.step(p)
        assert st["converged"] == 1
        assert abs(st["cg_iterations"] - ost["cg_iterations"]) <= 2, (step, st, ost)
        compare_trajectory(s, o, step, 1e-5, "plain")
        g_ = np.concatenate([s.download(f) for f in (0, 1, 2)])
        e_ = np.concatenate([o.download(f) for f in (0, 1, 2)])
        # the divergence of the stored velocity is the FP64 residual of the
        # FP32 pressure (the projection invariant D u = A p - b), to rounding
        assert st["div_residual"] > 0 and abs(st["div_residual"] - st["residual_rel"]) <= 0.05 * st["residual_rel"] + 2e-6, st
        assert abs(host_div_rel(s) - st["div_residual"]) <= 1e-6 * st["div_residual"] + 1e-12
        print("plain step %d: vel drift vs FP32 oracle %.2e, iterations %d/%d, FP64 residual %.2e, "
              "div %.2e, true_converged %d" % (step, rel_l2(g_, e_), st["cg_iterations"], ost["cg_iterations"],
                                               st["residual_rel"], st["div_residual"], st["true_converged"]))
    s.close(); o.close()


@pytest.mark.parametrize("precond", PRECONDS)
def test_dam_break_64_mixed(ctx, orc, precond):
    """Config 1, all 10 steps, mixed precision: true FP64 residual <= 1e-6;
    velocity within 1e-4 relative L2 of both the FP32 and the FP64 oracle
    trajectories; the divergence of the stored FP32 velocity is the residual
    plus the FP32 rounding floor of the FP64 oracle's projected velocity."""
    spec = scenes.dam_break(64)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **dam_desc_kw(mgf.PRECISION_MIXED, precond))
    s.set_divergence_check(True)
    o32 = scenes.build_oracle_scene(orc, spec)
    o64 = scenes.build_oracle_scene(orc, spec, f64=True)
    p = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-6)
    p64 = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-12)
    for step in range(10):
        st = s.step(1)
        o32.step(p)
        # the FP64 oracle by stages, to read its projected velocity
        o64.update_boundaries(p64); o64.advect(p64); o64.apply_forces(p64); o64.pressure_solve(p64)
        floor = fp32_rounding_floor(o64, spec)
        o64.extrapolate(p64)
        assert st["converged"] == 1 and st["true_converged"] == 1 and st["residual_rel"] <= 1e-6, st
        compare_trajectory(s, o32, step, 1e-4, "mixed vs fp32")
        g_ = np.concatenate([s.download(f) for f in (0, 1, 2)])
        e_ = np.concatenate([o64.download(f, f64=True) for f in (0, 1, 2)])
        e32 = np.concatenate([o32.download(f) for f in (0, 1, 2)])
        d64 = rel_l2(g_, e_)
        assert d64 <= 1e-4, (step, d64)
        div = st["div_residual"]
        assert abs(host_div_rel(s) - div) <= 1e-6 * div + 1e-12
        assert 0 < div <= st["residual_rel"] + 2.0 * floor + 1e-7, (step, div, st["residual_rel"], floor)
        print("mixed step %d: vel drift vs FP32 oracle %.2e, vs FP64 oracle %.2e; FP64 residual %.2e, "
              "div %.2e (FP32 rounding floor %.2e)" % (step, rel_l2(g_, e32), d64, st["residual_rel"], div, floor))
    s.close(); o32.close(); o64.close()


@pytest.mark.parametrize("precond", PRECONDS)
def test_single_step_mixed_pressure_vs_fp64_oracle(ctx, orc, precond):
    """From one identical state: the mixed solve's pressure agrees with the
    FP64 oracle's to the 1e-6 residual."""
    spec = scenes.obstacles(40, seed=1)
    s = scenes.build_gpu_scene(mgf, ctx, spec, **dam_desc_kw(mgf.PRECISION_MIXED, precond))
    o64 = scenes.build_oracle_scene(orc, spec, f64=True)
    p64 = orc.params(spec.dt, rk=1, tol_mode=1, tol=1e-13)
    s.update_boundaries(); s.advect(); s.apply_forces()
    o64.update_boundaries(p64); o64.advect(p64); o64.apply_forces(p64)
    st = s.pressure_solve()
    ost = o64.pressure_solve(p64)
    x = s.download(9)
    x64 = o64.download(9, f64=True)
    assert st["residual_rel"] <= 1e-6 and ost["residual_rel"] <= 1e-12
    # ||e||/||x|| <= cond(A) * 1e-6; measured ~1.3e-5 on this 42^3 system
    assert rel_l2(x, x64) <= 1e-4, rel_l2(x, x64)
    s.close(); o64.close()


@pytest.mark.parametrize("precond", PRECONDS)
def test_determinism(ctx, precond):
    spec = scenes.obstacles(48, seed=3)
    outs = []
    for _ in range(2):
        s = scenes.build_gpu_scene(mgf, ctx, spec, **dam_desc_kw(mgf.PRECISION_MIXED, precond))
        s.step(2)
        outs.append([s.download(f) for f in (0, 1, 2, 3)])
        s.close()
    for a, b in zip(*outs):
        assert_bitexact(a, b, "repeat run")


@pytest.mark.parametrize("res", [(1, 1, 1), (5, 7, 3), (1, 9, 4), (17, 2, 1)])
def test_ragged_and_tiny_grids(ctx, orc, res):
    rng = np.random.default_rng(sum(res))
    n = int(np.prod(res))
    solid = np.where(rng.random(n) < 0.2, -1.0, 1.0).astype(np.float32)
    fluid = np.where(rng.random(n) < 0.7, -1.0, 1.0).astype(np.float32)
    u, v, w = rand_vel(rng, res, 0.3)
    d = mgf.make_desc(res, (0, 0, 0), (0.1, 0.2, 0.3), 0.01, rk=2, tol_mode=mgf.TOL_RELATIVE, tol=1e-6)
    s = mgf.Scene(ctx, d)
    o = orc.Scene(res, (0, 0, 0), (0.1, 0.2, 0.3))
    for sc in (s, o):
        sc.add_solid(solid)
        sc.upload(3, fluid)
        sc.upload(0, u); sc.upload(1, v); sc.upload(2, w)
    p = orc.params(0.01, rk=2, tol_mode=1, tol=1e-6)
    for step in range(2):
        # a few dozen cells: the reference's N-iteration cap can stop the FP32
        # CG short of 1e-6 on both sides
        st = s.step(1, allow_unconverged=True)
        ost = o.step(p, 1)
        assert st["converged"] == ost["converged"]
        if step == 0:
            assert_bitexact(s.download(3), o.download(3), "phi")
        for fid in (0, 1, 2):
            g_, e_ = s.download(fid), o.download(fid)
            assert np.array_equal(g_ == FLT_MAX, e_ == FLT_MAX)
            assert rel_l2(g_, e_, e_ != FLT_MAX) <= 1e-5
    s.close(); o.close()


def ragged_scene(ctx, orc, res, psolid, seed, **kw):
    rng = np.random.default_rng(seed)
    n = int(np.prod(res))
    solid = np.where(rng.random(n) < psolid, -1.0, 1.0).astype(np.float32)
    fluid = np.where(rng.random(n) < 0.7, -1.0, 1.0).astype(np.float32)
    u, v, w = rand_vel(rng, res, 0.3)
    d = mgf.make_desc(res, (0, 0, 0), (0.1, 0.2, 0.3), 0.01, rk=2, tol_mode=mgf.TOL_RELATIVE, tol=1e-6,
                      precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_MULTIGRID, **kw)
    s = mgf.Scene(ctx, d)
    o = orc.Scene(res, (0, 0, 0), (0.1, 0.2, 0.3), f64=True)
    for sc in (s, o):
        sc.add_solid(solid)
        sc.upload(3, fluid)
        sc.upload(0, u); sc.upload(1, v); sc.upload(2, w)
    return s, o


@pytest.mark.parametrize("res", [(5, 7, 3), (1, 9, 4), (17, 2, 1), (33, 9, 70)])
def test_multigrid_ragged_grids(ctx, orc, res):
    """Multigrid on grids whose axes are odd, 1 or 2 cells long (axes of one
    cell are not coarsened), 10% random solids: the FP64 residual meets 1e-6
    and the projected velocity matches the FP64 oracle's."""
    s, o = ragged_scene(ctx, orc, res, 0.1, 7 + sum(res), max_it=10000)
    p = orc.params(0.01, rk=2, tol_mode=1, tol=1e-13, max_it=100000)
    st = s.step(1)
    ost = o.step(p, 1)
    assert st["converged"] == 1 and st["residual_rel"] <= 1e-6, st
    assert ost["converged"] == 1
    for fid in (0, 1, 2):
        g_, e_ = s.download(fid), o.download(fid, f64=True)
        assert np.array_equal(g_ == FLT_MAX, np.float32(e_) == FLT_MAX)
        assert rel_l2(g_, e_, g_ != FLT_MAX) <= 1e-4
    s.close(); o.close()


def test_multigrid_inconsistent_system_reports_not_converged(ctx, orc):
    """A 1x1x1 fluid cell has an all-zero matrix row (every neighbour outside
    the grid) and a non-zero right-hand side: no solver can meet the
    tolerance.  The step reports NOT_CONVERGED within its iteration cap and
    leaves finite velocities, as the FP64 oracle's CG does not converge
    either."""
    s, o = ragged_scene(ctx, orc, (1, 1, 1), 0.0, 11, max_it=50)
    p = orc.params(0.01, rk=2, tol_mode=1, tol=1e-13, max_it=1000)
    with pytest.raises(mgf.MgfError) as ei:
        s.step(1)
    assert ei.value.status == mgf.ERR_NOT_CONVERGED
    ost = o.step(p, 1)
    assert ost["converged"] == 0
    for fid in (0, 1, 2):
        assert np.all(np.isfinite(s.download(fid)))
    s.close(); o.close()


def test_multigrid_cuts_iterations(ctx):
    """The V-cycle preconditioner leaves the system and tolerance unchanged
    and needs far fewer iterations than the identity on an obstacle scene."""
    spec = scenes.obstacles(96, seed=2)
    its = {}
    for pc in (mgf.PRECOND_IDENTITY, mgf.PRECOND_MULTIGRID):
        s = scenes.build_gpu_scene(mgf, ctx, spec, **dam_desc_kw(mgf.PRECISION_MIXED, pc))
        st = s.step(1)
        assert st["converged"] == 1 and st["residual_rel"] <= 1e-6, st
        its[pc] = st["cg_iterations"]
        s.close()
    assert its[mgf.PRECOND_MULTIGRID] * 8 <= its[mgf.PRECOND_IDENTITY], its


@pytest.mark.parametrize("precond", PRECONDS)
def test_switch_to_mixed_after_creation(ctx, precond):
    """A scene created with the plain FP32 path and switched to the mixed solver
    by mgf_scene_set_params (as CGridFluidSolver::setPressureOptions does)
    takes the buffers the mixed solver needs on its first solve and gives the
    same step as a scene created mixed."""
    spec = scenes.obstacles(64, seed=5)
    kw = dam_desc_kw(mgf.PRECISION_MIXED, precond)
    made = scenes.build_gpu_scene(mgf, ctx, spec, **kw)
    switched = scenes.build_gpu_scene(mgf, ctx, spec, **dam_desc_kw(mgf.PRECISION_PLAIN))
    switched.set_params(mgf.make_desc(spec.res, spec.origin, spec.spacing, spec.dt, **kw))
    a, b = made.step(1), switched.step(1)
    assert a["converged"] == 1 and b["converged"] == 1
    assert a["cg_iterations"] == b["cg_iterations"]
    for fid in (0, 1, 2):
        assert_bitexact(switched.download(fid), made.download(fid), "field %d" % fid)
    made.close(); switched.close()


def test_preconditioner_options_are_checked(ctx):
    d = mgf.make_desc((8, 8, 8), precision=mgf.PRECISION_PLAIN, precond=mgf.PRECOND_MULTIGRID)
    with pytest.raises(mgf.MgfError):
        mgf.Scene(ctx, d)           # PLAIN is the reference's identity CG only
    d = mgf.make_desc((8, 8, 8), precision=mgf.PRECISION_MIXED, precond=7)
    with pytest.raises(mgf.MgfError):
        mgf.Scene(ctx, d)
    d = mgf.make_desc((8, 8, 8), precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_MULTIGRID, mg_omega=2.5)
    with pytest.raises(mgf.MgfError):
        mgf.Scene(ctx, d)


@pytest.mark.parametrize("precond", PRECONDS)
def test_no_fluid_and_zero_rhs(ctx, precond):
    res = (8, 8, 8)
    d = mgf.make_desc(res, dt=0.01, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
                      precond=precond)
    s = mgf.Scene(ctx, d)
    s.upload(3, np.full(512, 1.0, np.float32))   # all air
    st = s.step(1)
    assert st["cg_iterations"] == 0 and st["converged"] == 1
    for fid in (0, 1, 2):
        assert np.all(s.download(fid) == FLT_MAX)  # no fluid: every face unknown, no layer 0
    s.close()


def test_invalid_arguments(ctx):
    d = mgf.make_desc((4, 4, 4))
    d.rk_order = 5
    with pytest.raises(mgf.MgfError):
        mgf.Scene(ctx, d)
    d = mgf.make_desc((4, 4, 4))
    s = mgf.Scene(ctx, d)
    with pytest.raises(mgf.MgfError):
        s.upload(0, np.zeros(3, np.float32))
    with pytest.raises(mgf.MgfError):
        s.add_solid(np.zeros(5, np.float32))
    s.close()


# ---------------------------------------------------------------------------
# Full size (256^3 with obstacles): size-independent properties.
@pytest.mark.parametrize("precond", PRECONDS)
def test_obstacles_256_one_step_properties(ctx, precond):
    spec = scenes.obstacles(256, seed=0)
    s = scenes.build_gpu_scene(mgf, ctx, spec, rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6,
                               precision=mgf.PRECISION_MIXED, precond=precond)
    s.update_boundaries()
    s.advect()
    s.apply_forces()
    st = s.pressure_solve()
    bn = np.linalg.norm(s.download(10).astype(np.float64))
    assert st["converged"] == 1 and st["residual_rel"] <= 1e-6, st
    m = s.download(8)
    div1 = s.download(12)
    g = spec.res[0]
    idx = np.arange(m.size)
    x, y, z = idx % g, (idx // g) % g, idx // (g * g)
    inner = (m == 1) & (x < g - 1) & (y < g - 1) & (z < g - 1)
    # projection invariant: div(u_new) = A p - b on fluid cells (FP32 storage)
    assert bn > 0
    assert np.linalg.norm(div1[inner].astype(np.float64)) <= 1e-4 * bn
    phi = s.download(3)
    solid = s.download(4) < 0
    exp_m = np.where(solid, 2, np.where(phi < 0, 1, 0))
    assert np.array_equal(m, exp_m)
    used = s.extrapolate()
    assert used > 0
    for fid in (0, 1, 2):
        a = s.download(fid)
        assert np.all(np.isfinite(a)) and not np.any(a == FLT_MAX)
    s.close()


def test_large_512_steps_properties(ctx):
    """The large single scene: 512^3 interior (514^3 padded, 135.8M cells,
    408M faces), a seeded fluid block and 8 spheres, mixed-precision
    multigrid PCG.  One step by stages with the projection invariant checked,
    then a full step: both converge to the 1e-6 FP64 residual, markers follow
    their definition, and every face holds a finite extrapolated value."""
    spec = scenes.obstacles(512, seed=0, nspheres=8)
    s = scenes.build_gpu_scene(mgf, ctx, spec, rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6,
                               precision=mgf.PRECISION_MIXED, precond=mgf.PRECOND_MULTIGRID)
    s.update_boundaries()
    s.advect()
    s.apply_forces()
    st = s.pressure_solve()
    assert st["converged"] == 1 and st["residual_rel"] <= 1e-6, st
    m = s.download(8)
    phi = s.download(3)
    solid = s.download(4) < 0
    assert np.array_equal(m, np.where(solid, 2, np.where(phi < 0, 1, 0)).astype(np.float32))
    del phi, solid
    bn = np.linalg.norm(s.download(10).astype(np.float64))
    div1 = s.download(12)
    g = spec.res[0]
    inner = (m == 1).reshape(g, g, g)
    inner[-1, :, :] = False; inner[:, -1, :] = False; inner[:, :, -1] = False
    assert bn > 0 and np.linalg.norm(div1.reshape(g, g, g)[inner].astype(np.float64)) <= 1e-4 * bn
    del m, div1, inner
    assert s.extrapolate() > 0
    st2 = s.step(1)
    assert st2["converged"] == 1 and st2["residual_rel"] <= 1e-6, st2
    for fid in (0, 1, 2):
        a = s.download(fid)
        assert np.all(np.isfinite(a)) and not np.any(a == FLT_MAX)
    s.close()


def test_warm_start_converges_with_fewer_iterations(ctx):
    """The opt-in warm start (previous pressure as the CG's first iterate)
    meets the same FP64 stop rule in fewer iterations, and its trajectory stays
    within the projected-velocity tolerance of the x = 0 start."""
    spec = scenes.obstacles(64, seed=2, nspheres=3)
    kw = dict(rk=1, tol_mode=mgf.TOL_RELATIVE, tol=1e-6, precision=mgf.PRECISION_MIXED,
              precond=mgf.PRECOND_MULTIGRID)
    cold = scenes.build_gpu_scene(mgf, ctx, spec, **kw)
    warm = scenes.build_gpu_scene(mgf, ctx, spec, warm_start=True, **kw)
    it_cold = it_warm = 0
    for step in range(6):
        a, b = cold.step(1), warm.step(1)
        assert a["converged"] == 1 and b["converged"] == 1, (step, a, b)
        assert b["residual_rel"] <= 1e-6, (step, b)
        it_cold += a["cg_iterations"]
        it_warm += b["cg_iterations"]
        vc = np.concatenate([cold.download(f) for f in (0, 1, 2)])
        vw = np.concatenate([warm.download(f) for f in (0, 1, 2)])
        assert rel_l2(vw, vc) <= 1e-4, (step, rel_l2(vw, vc))
        # the interface may move by a flipped cell where the pressures differ in the last bits
        flips = int(np.sum(cold.download(mgf.FIELD["markers"]) != warm.download(mgf.FIELD["markers"])))
        assert flips <= 2, (step, flips)
    assert it_warm < it_cold, (it_warm, it_cold)
    cold.close(); warm.close()
