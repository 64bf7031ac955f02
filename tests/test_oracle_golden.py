"""Pins the CPU oracle (oracle/mac_oracle.hpp) before it is trusted.

1. Against the golden values of the reference's own tests
   (tests/golden/reference_cases.json.gz, written by make_fixtures.py).
2. Against what the mathematics fixes, independently of the oracle's code:
   dense FP64 solves of the pressure system, the projection invariant
   div(u_new) = A p - b on fluid cells, linearity/constant-field checks of the
   sampler, and BFS distances computed by brute force for the extrapolation.
3. Planted mistakes: each check is shown to fail on a deliberately wrong
   variant computed here on the CPU.
CPU only.
"""
import numpy as np
import pytest

FLT_MAX = float(np.finfo(np.float32).max)


def near(got, exp, tol):
    got = float(got)
    if exp == FLT_MAX:
        return got == FLT_MAX
    return abs(got - exp) < tol


def check_map(arr, exp_map, tol, label, override=None):
    bad = []
    for k, v in exp_map.items():
        t = tol
        if override and (label + ":" + k) in override:
            t = override[label + ":" + k]
        if not near(arr[int(k)], v, t):
            bad.append((label, int(k), float(arr[int(k)]), v))
    assert not bad, bad[:10]


def arr(x):
    return np.asarray(x, np.float32)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("idx", range(11))
def test_pressure_goldens(cases, orc, idx):
    c = cases["pressure"][idx]
    i = c["inputs"]
    out = orc.pressure_solve(c["res"], c["origin"], c["spacing"], c["dt"], i["u"], i["v"], i["w"],
                             i["us"], i["vs"], i["ws"], i["solid"], i["fluid"])
    e = c["expect"]
    ov = c.get("tol_override")
    for key, got in (("markers", out["markers"]), ("A", out["A"]), ("b", out["b"]),
                     ("x", out["x"]), ("u", out["u"]), ("v", out["v"]), ("w", out["w"])):
        if key in e:
            check_map(got, e[key], c["tol"], key, ov)


def test_pressure_reference_stop_rule(cases, orc):
    """sqrt(||r||) < 1e-3, i.e. ||r|| < 1e-6 (CuMatrixFreePCG.cpp:51-52)."""
    c = cases["pressure"][6]  # MatrixFreePCG1
    i = c["inputs"]
    out = orc.pressure_solve(c["res"], c["origin"], c["spacing"], c["dt"], i["u"], i["v"], i["w"],
                             i["us"], i["vs"], i["ws"], i["solid"], i["fluid"])
    st = out["stats"]
    assert st["converged"] == 1
    assert np.sqrt(st["residual_recurrence"]) < 1e-3


# ---------------------------------------------------------------------------
def test_semilagrangian_goldens(cases, orc):
    for c in cases["semilagrangian"]:
        i = c["inputs"]
        if c["kind"] == "backtrace":
            px, py, pz = orc.backtrace(c["res"], c["origin"], c["spacing"], i["u"], i["v"], i["w"],
                                       c["dt"], c["order"], i["px"], i["py"], i["pz"])
            for got, exp in ((px, c["expect"]["px"]), (py, c["expect"]["py"]), (pz, c["expect"]["pz"])):
                assert np.max(np.abs(got - np.asarray(exp))) < c["tol"], c["name"]
        elif c["kind"] == "advect_cc":
            out = orc.advect_cc(c["res"], c["origin"], c["spacing"], i["field"], i["u"], i["v"], i["w"],
                                c["dt"], c["order"])
            if "field_parts" in c["expect"]:
                for k, (a, b, cc) in c["expect"]["field_parts"].items():
                    d = ((np.float32(out[int(k)]) - np.float32(a)) - np.float32(b)) - np.float32(cc)
                    assert abs(d) < c["tol"], (c["name"], k, out[int(k)], a, b, cc)
                continue
            e = c["expect"]["field"]
            if isinstance(e, dict):
                check_map(out, e, c["tol"], c["name"])
            else:
                assert np.max(np.abs(out - np.asarray(e))) < c["tol"], c["name"]
        elif c["kind"] == "advect_fc":
            ou, ov, ow = orc.advect_fc(c["res"], c["origin"], c["spacing"], i["iu"], i["iv"], i["iw"],
                                       i["u"], i["v"], i["w"], c["dt"], c["order"])
            for got, exp in ((ou, c["expect"]["u"]), (ov, c["expect"]["v"]), (ow, c["expect"]["w"])):
                assert np.max(np.abs(got - np.asarray(exp))) < c["tol"], c["name"]


def test_external_forces_goldens(cases, orc):
    c = cases["external_forces"][0]
    u, v, w = (arr(c["inputs"][k]) for k in "uvw")
    for n in range(9):
        u, v, w = orc.external_forces(c["res"], u, v, w, c["gravity"], c["force"], c["dt"])
        e = c["expect_steps"][n]
        for got, k in ((u, "u"), (v, "v"), (w, "w")):
            assert np.max(np.abs(got - np.asarray(e[k]))) < c["tol"]
    # resizeExternalForcesSolver clears the extra force; gravity stays.
    u, v, w = (arr(c["inputs"][k]) for k in "uvw")
    for n in range(9):
        u, v, w = orc.external_forces(c["res"], u, v, w, c["gravity"], [0, 0, 0], c["dt"])
        e = c["expect_after_reset"][n]
        for got, k in ((u, "u"), (v, "v"), (w, "w")):
            assert np.max(np.abs(got - np.asarray(e[k]))) < c["tol"]


def test_extrapolation_goldens(cases, orc):
    for c in cases["extrapolation"]:
        if c["kind"] == "grid":
            d = arr(c["inputs"]["data"]).reshape(c["shape"])
            vals, ids, _ = orc.extrapolate_grid(d, c["layers"])
            lay = np.where(ids == np.iinfo(np.int32).max, FLT_MAX, ids).ravel()
            for k, (g, e) in enumerate(zip(lay, c["expect"]["layers"])):
                assert near(g, e, c["tol"]), (c["name"], "layer", k, g, e)
            for k, (g, e) in enumerate(zip(vals.ravel(), c["expect"]["data"])):
                assert near(g, e, c["tol"]), (c["name"], "data", k, g, e)
        else:
            res = c["res"]
            comps = []
            for comp, key in enumerate("uvw"):
                shape = orc.face_shape(res, comp)
                d = arr(c["inputs"][key]).reshape(shape)
                vals, ids, _ = orc.extrapolate_grid(d, c["layers"])
                lay = np.where(ids == np.iinfo(np.int32).max, FLT_MAX, ids).ravel()
                for k, (g, e) in enumerate(zip(lay, c["expect"]["layers_" + key])):
                    assert near(g, e, c["tol"]), (c["name"], key, k, g, e)
                for k, (g, e) in enumerate(zip(vals.ravel(), c["expect"][key])):
                    assert near(g, e, c["tol"]), (c["name"], key, k, g, e)
                comps.append(vals.ravel())
            # the face-centred entry point gives the same values
            u, v, w, _ = orc.extrapolate_fc(res, c["inputs"]["u"], c["inputs"]["v"], c["inputs"]["w"],
                                            c["layers"])
            for a, b in zip((u, v, w), comps):
                assert np.array_equal(a, b)


def test_boundary_goldens(cases, orc):
    for c in cases["boundary"]:
        s = orc.Scene(c["res"], c["origin"], c["spacing"])
        for sol in c["solids"]:
            s.add_solid(sol["sdf"], sol["vel"])
        s.update_boundaries(orc.params(c.get("dt", 0.0)))
        kind = c.get("kind")
        if kind == "move":
            for k, t in enumerate(c["expect"]["translation"]):
                got, _ = s.solid_state(k)
                assert np.max(np.abs(got - np.asarray(t))) < c["tol"]
        elif kind == "resample":
            for k, e in enumerate(c["expect"]["sdf"]):
                _, got = s.solid_state(k)
                assert np.max(np.abs(got - np.asarray(e))) < c["tol"]
        elif "marker" in c["expect"]:
            check_map(s.download(4), c["expect"]["marker"], c["tol"], "marker")
        else:
            fields = {"u": s.download(5), "v": s.download(6), "w": s.download(7)}
            for key, val in c["expect"]["abs"].items():
                comp, idx = key.split(":")
                assert abs(abs(fields[comp][int(idx)]) - val) < c["tol"], key
            for key, val in c["expect"]["val"].items():
                comp, idx = key.split(":")
                assert abs(fields[comp][int(idx)] - val) < c["tol"], key
        s.close()


def test_cg_goldens(cases, orc):
    for c in cases["cg"]:
        A = np.asarray(c["A"], np.float64)
        b = np.asarray(c["b"], np.float64)
        x = None
        for _ in range(c["repeats"]):
            x, _ = orc.cg_dense(A, b, minv=c["minv"], max_it=c["max_it"], x0=x)
            r = b - A @ x.astype(np.float64)
            assert np.linalg.norm(r) <= c["residual_max"], c["name"]


def test_field_goldens(cases, orc):
    for c in cases["fields"]:
        i = c["inputs"]
        if c["kind"] == "sample_fc":
            for comp, key in enumerate("uvw"):
                ou, ov, ow = orc.sample_fc(c["res"], c["origin"], c["spacing"], i["u"], i["v"], i["w"],
                                           i["p" + key + "x"], i["p" + key + "y"], i["p" + key + "z"])
                got = (ou, ov, ow)[comp]
                assert np.max(np.abs(got - np.asarray(c["expect"][key]))) < c["tol"], (c["name"], key)
        else:
            d = orc.divergence(c["res"], c["origin"], c["spacing"], i["u"], i["v"], i["w"])
            assert np.max(np.abs(d.ravel() - np.asarray(c["expect"]["div"]))) < c["tol"], c["name"]


def test_integration_goldens(cases, orc):
    for c in cases["integration"]:
        s = orc.Scene(c["res"], c["origin"], c["spacing"])
        s.add_solid(c["inputs"]["solid"], (0, 0, 0))
        s.upload(3, c["inputs"]["fluid"])
        p = orc.params(c["dt"], rk=3)
        for n in range(c["steps"]):
            s.step(p)
            if c["expect"].get("all_faces_zero"):
                for f in (0, 1, 2):
                    assert np.max(np.abs(s.download(f))) < c["tol"], (c["name"], n, f)
            else:
                v = s.download(1)
                for face in c["expect"]["v_faces"]:
                    assert abs(v[face] - c["expect"]["v_per_step"][n]) < c["tol"], (c["name"], n, face)
        s.close()


# ---------------------------------------------------------------------------
# Checks fixed by mathematics, not by the oracle's code.
def random_pressure_scene(rng, res, air_frac=0.2, solid_frac=0.15):
    nx, ny, nz = res
    n = nx * ny * nz
    solid = np.where(rng.random(n) < solid_frac, -1.0, 1.0).astype(np.float32)
    fluid = np.where(rng.random(n) < air_frac, 1.0, -1.0).astype(np.float32)
    u = rng.standard_normal((nx + 1) * ny * nz).astype(np.float32)
    v = rng.standard_normal(nx * (ny + 1) * nz).astype(np.float32)
    w = rng.standard_normal(nx * ny * (nz + 1)).astype(np.float32)
    us = rng.standard_normal(u.size).astype(np.float32) * 0.1
    vs = rng.standard_normal(v.size).astype(np.float32) * 0.1
    ws = rng.standard_normal(w.size).astype(np.float32) * 0.1
    return solid, fluid, u, v, w, us, vs, ws


def dense_from_compact(res, A):
    """Dense matrix of fdmMatrixVectorMul's operator, written from the 4-value
    layout [diag,+x,+y,+z] with -d couplings read from the neighbour."""
    nx, ny, nz = res
    n = nx * ny * nz
    M = np.zeros((n, n))
    A = np.asarray(A, np.float64).reshape(n, 4)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                c = (z * ny + y) * nx + x
                M[c, c] = A[c, 0]
                if x + 1 < nx:
                    M[c, c + 1] = A[c, 1]
                    M[c + 1, c] = A[c, 1]
                if y + 1 < ny:
                    M[c, c + nx] = A[c, 2]
                    M[c + nx, c] = A[c, 2]
                if z + 1 < nz:
                    M[c, c + nx * ny] = A[c, 3]
                    M[c + nx * ny, c] = A[c, 3]
    return M


def laplacian_dense(res, markers, scale):
    """The variable-domain Poisson operator written from its definition:
    fluid rows: sum over in-grid non-solid neighbours of s_d (p_c - p_n) with
    p_n = 0 for air; non-fluid rows: identity."""
    nx, ny, nz = res
    n = nx * ny * nz
    M = np.zeros((n, n))
    m = np.asarray(markers).ravel()
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                c = (z * ny + y) * nx + x
                if m[c] != 1:
                    M[c, c] = 1
                    continue
                for dx, dy, dz, s in ((1, 0, 0, scale[0]), (-1, 0, 0, scale[0]), (0, 1, 0, scale[1]),
                                      (0, -1, 0, scale[1]), (0, 0, 1, scale[2]), (0, 0, -1, scale[2])):
                    X, Y, Z = x + dx, y + dy, z + dz
                    if not (0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz):
                        continue
                    q = (Z * ny + Y) * nx + X
                    if m[q] == 2:
                        continue
                    M[c, c] += s
                    if m[q] == 1:
                        M[c, q] -= s
    return M


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matrix_is_the_poisson_operator(orc, seed):
    rng = np.random.default_rng(seed)
    res = (5, 4, 3)
    solid, fluid, u, v, w, us, vs, ws = random_pressure_scene(rng, res)
    out = orc.pressure_solve(res, (0, 0, 0), (0.5, 0.25, 1.0), 0.1, u, v, w, us, vs, ws, solid, fluid)
    scale = [np.float32(0.1) / (np.float32(0.5) ** 2), np.float32(0.1) / (np.float32(0.25) ** 2),
             np.float32(0.1) / np.float32(1.0)]
    M = laplacian_dense(res, out["markers"], scale)
    assert np.allclose(dense_from_compact(res, out["A"]), M, rtol=1e-6, atol=1e-6)
    # matvec == dense product (exact products summed in FP64, rounded once)
    x = rng.standard_normal(M.shape[0]).astype(np.float32)
    y = orc.matvec(res, out["A"], x)
    assert np.allclose(y, M @ x.astype(np.float64), rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_cg_matches_dense_solve(orc, seed):
    rng = np.random.default_rng(seed)
    res = (4, 4, 4)
    solid, fluid, u, v, w, us, vs, ws = random_pressure_scene(rng, res)
    out = orc.pressure_solve(res, (0, 0, 0), (1, 1, 1), 1.0, u, v, w, us, vs, ws, solid, fluid,
                             tol_mode=1, tol=1e-7)
    M = dense_from_compact(res, out["A"])
    xd = np.linalg.solve(M, out["b"].astype(np.float64))
    assert np.allclose(out["x"], xd, rtol=1e-4, atol=1e-4 * max(1.0, np.abs(xd).max()))
    # the FP64 oracle forms A and b in double: it agrees with the FP32-formed
    # system's exact solution to the FP32 rounding of A and b
    o64 = orc.pressure_solve_f64(res, (0, 0, 0), (1, 1, 1), 1.0, u, v, w, us, vs, ws, solid, fluid,
                                 tol_mode=1, tol=1e-13)
    assert np.allclose(o64["x"], xd, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(xd).max()))
    assert o64["stats"]["residual_rel"] < 1e-12


@np.errstate(over="ignore", invalid="ignore")
def div_numpy(res, spacing, u, v, w):
    nx, ny, nz = res
    U = np.asarray(u, np.float64).reshape(nz, ny, nx + 1)
    V = np.asarray(v, np.float64).reshape(nz, ny + 1, nx)
    W = np.asarray(w, np.float64).reshape(nz + 1, ny, nx)
    return ((U[:, :, 1:] - U[:, :, :-1]) / spacing[0] + (V[:, 1:, :] - V[:, :-1, :]) / spacing[1]
            + (W[1:, :, :] - W[:-1, :, :]) / spacing[2]).ravel()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_projection_invariant(orc, seed):
    """After the projection, div(u_new) on fluid cells equals A p - b; with
    the FP64 oracle solved to 1e-13 that is ~0 relative to div(u*)."""
    rng = np.random.default_rng(seed)
    res = (6, 5, 4)
    sp = (0.5, 0.5, 0.5)
    solid, fluid, u, v, w, us, vs, ws = random_pressure_scene(rng, res)
    o32 = orc.pressure_solve(res, (0, 0, 0), sp, 0.05, u, v, w, us, vs, ws, solid, fluid)
    # fluid cells off the max-index planes (whose outer faces the projection
    # sets to UNKNOWN, EulerSolverTool.cu:308-321)
    nx, ny, nz = res
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    inner = ((x < nx - 1) & (y < ny - 1) & (z < nz - 1)).ravel()
    fl = (o32["markers"] == 1) & inner
    o64 = orc.pressure_solve_f64(res, (0, 0, 0), sp, 0.05, u, v, w, us, vs, ws, solid, fluid,
                                 tol_mode=1, tol=1e-13)
    d_new = div_numpy(res, sp, o64["u"], o64["v"], o64["w"])[fl]
    d_old = div_numpy(res, sp, u, v, w)[fl]
    assert np.linalg.norm(d_new) <= 1e-9 * np.linalg.norm(d_old)
    # planted mistake: flipping the sign of the gradient update breaks it
    wrong = div_numpy(res, sp, 2 * np.asarray(u, np.float64) - o64["u"],
                      2 * np.asarray(v, np.float64) - o64["v"], 2 * np.asarray(w, np.float64) - o64["w"])[fl]
    assert np.linalg.norm(wrong) > 1e-3 * np.linalg.norm(d_old)


def test_sampler_reproduces_linear_fields(orc):
    """Trilinear interpolation is exact on affine fields inside the grid and
    clamps outside (values at the nearest cell centre along each axis)."""
    rng = np.random.default_rng(5)
    res, o, h = (7, 6, 5), (0.25, -1.0, 2.0), (0.5, 0.25, 1.0)
    nx, ny, nz = res
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cx, cy, cz = o[0] + (x + 0.5) * h[0], o[1] + (y + 0.5) * h[1], o[2] + (z + 0.5) * h[2]
    f = (1.5 * cx - 2.0 * cy + 0.75 * cz + 3.0).astype(np.float32)
    px = rng.uniform(o[0] + 0.5 * h[0], o[0] + (nx - 0.5) * h[0], 500)
    py = rng.uniform(o[1] + 0.5 * h[1], o[1] + (ny - 0.5) * h[1], 500)
    pz = rng.uniform(o[2] + 0.5 * h[2], o[2] + (nz - 0.5) * h[2], 500)
    got = orc.sample_cc(res, o, h, f, px, py, pz)
    assert np.allclose(got, 1.5 * px - 2.0 * py + 0.75 * pz + 3.0, atol=2e-5)
    # outside the grid the value is that of the clamped point
    qx = np.float32(o[0] - 3.0)
    got = orc.sample_cc(res, o, h, f, [qx], [py[0]], [pz[0]])
    ref = orc.sample_cc(res, o, h, f, [o[0] + 0.5 * h[0]], [py[0]], [pz[0]])
    assert abs(got[0] - ref[0]) < 1e-5


def test_advection_constant_and_zero_velocity(orc):
    rng = np.random.default_rng(3)
    res, o, h = (6, 5, 4), (0, 0, 0), (1, 1, 1)
    nx, ny, nz = res
    f = rng.standard_normal(nx * ny * nz).astype(np.float32)
    zu, zv, zw = (np.zeros((nx + 1) * ny * nz), np.zeros(nx * (ny + 1) * nz), np.zeros(nx * ny * (nz + 1)))
    for order in (1, 2, 3):
        # zero velocity: identity (face and cell positions land on samples)
        assert np.array_equal(orc.advect_cc(res, o, h, f, zu, zv, zw, 0.1, order), f)
        # a constant field stays constant under any velocity
        c = np.full(nx * ny * nz, 2.5, np.float32)
        vel = [rng.standard_normal(a.size) for a in (zu, zv, zw)]
        assert np.allclose(orc.advect_cc(res, o, h, c, *vel, 0.1, order), 2.5, atol=1e-6)


def bfs_layers(valid, layers):
    """Brute-force layers: L1 distance to the nearest known entry (clamped
    6-neighbourhood on an obstacle-free grid), capped below `layers`."""
    nz, ny, nx = valid.shape
    pts = np.argwhere(valid)
    out = np.full(valid.shape, np.iinfo(np.int32).max, np.int64)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                if len(pts) == 0:
                    continue
                d = np.min(np.abs(pts - [z, y, x]).sum(axis=1))
                if d < layers:
                    out[z, y, x] = d
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_extrapolation_layers_are_l1_distance(orc, seed):
    rng = np.random.default_rng(seed)
    d = np.full((5, 6, 7), FLT_MAX, np.float32)
    known = rng.random(d.shape) < 0.05
    known[2, 3, 3] = True
    d[known] = rng.standard_normal(known.sum())
    for layers in (3, 1000):
        vals, ids, _ = orc.extrapolate_grid(d, layers)
        assert np.array_equal(ids, bfs_layers(known, layers))
        # each newly reached entry is the mean of its lower-layer neighbours
        nz, ny, nx = d.shape
        for z, y, x in np.argwhere(ids > 0):
            if ids[z, y, x] == np.iinfo(np.int32).max:
                assert vals[z, y, x] == FLT_MAX
                continue
            L = ids[z, y, x]
            nb = []
            for dz, dy, dx in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)):
                Z, Y, X = min(max(z + dz, 0), nz - 1), min(max(y + dy, 0), ny - 1), min(max(x + dx, 0), nx - 1)
                if ids[Z, Y, X] < L:
                    nb.append(vals[Z, Y, X])
            assert abs(vals[z, y, x] - np.mean(nb)) < 1e-5


def test_scene_colour_is_advected_with_the_old_velocity(orc):
    """The oracle step carries the colour field as CSemiLagrangian::advect(CCV)
    does (SemiLagrangian.cpp:152-167): each channel is the CCS advection of the
    channel by the velocity the step started from; a constant colour stays
    constant, and a scene without colour keeps none."""
    import scenes
    rng = np.random.default_rng(5)
    spec = scenes.obstacles(10, seed=2)
    s = scenes.build_oracle_scene(orc, spec)
    n = int(np.prod(spec.res))
    col = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    for k in range(3):
        s.upload(13 + k, col[k])
    vel = [s.download(f) for f in (0, 1, 2)]
    for order in (1, 3):
        p = orc.params(spec.dt, rk=order, tol_mode=1, tol=1e-6)
        s.update_boundaries(p)
        vel = [s.download(f) for f in (0, 1, 2)]
        exp = [orc.advect_cc(spec.res, spec.origin, spec.spacing, col[k], *vel, spec.dt, order) for k in range(3)]
        s.advect(p)
        got = [s.download(13 + k) for k in range(3)]
        for k in range(3):
            assert np.array_equal(got[k].view(np.uint32), exp[k].view(np.uint32)), (order, k)
        col = got
        # give the next round a non-zero velocity
        for f in (0, 1, 2):
            s.upload(f, rng.standard_normal(vel[f].size).astype(np.float32) * 0.3)
    c = scenes.build_oracle_scene(orc, spec)
    for k in range(3):
        c.upload(13 + k, np.full(n, 0.25 * (k + 1), np.float32))
    p = orc.params(spec.dt, rk=3, tol_mode=1, tol=1e-6)
    c.step(p, 2)
    for k in range(3):
        assert np.allclose(c.download(13 + k), 0.25 * (k + 1), atol=1e-6)
    s.close(); c.close()


def test_fluid_density_closed_form(orc):
    """buildFluidDensity on a fluid block against walls: fluid cells touching
    air are FluidSurfaceDensity (0.2), fluid cells whose neighbours are fluid
    or solid are 1, solid and air cells 0 (a solid neighbour is not air)."""
    n = 10
    res = (n, n, n)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    wall = (x == 0) | (y == 0) | (z == 0) | (x == n - 1) | (y == n - 1) | (z == n - 1)
    solid = np.where(wall, -1.0, np.finfo(np.float32).max).astype(np.float32)
    fluid = (~wall) & (x <= 4) & (y <= 6)
    phi = np.where(fluid, -500.0, 500.0).astype(np.float32)
    got = orc.fluid_density(res, phi.ravel(), solid.ravel()).reshape(res)
    surface = fluid & ((x == 4) | (y == 6))
    exp = np.where(surface, np.float32(0.2), np.where(fluid, 1.0, 0.0)).astype(np.float32)
    assert np.array_equal(got, exp)
    # a fluid cell inside a solid is not fluid
    phi2 = np.full(n ** 3, -500.0, np.float32)
    assert np.all(orc.fluid_density(res, phi2, solid.ravel()).reshape(res)[wall] == 0)


@pytest.mark.parametrize("passes", [1000, 3])
def test_fluid_sdf_closed_form(orc, passes):
    """buildFluidSDF on a fluid slab z < k: inside cells at distance d from the
    last fluid layer hold 0.5-0.2 (d = 0) and -(d - (0.5-0.2)) below; outside
    cells d layers beyond the first air layer hold d + (1.5-0.2), up to
    passes-1 layers (the rest keep +-UNKNOWN).  Around one fluid cell the
    outside value grows with the L1 distance."""
    U = np.finfo(np.float32).max
    nx, ny, nz, k = 4, 3, 16, 6
    phi = np.where(np.arange(nz)[:, None, None] < k, -500.0, 500.0) * np.ones((nz, ny, nx))
    got = orc.fluid_sdf((nx, ny, nz), phi.astype(np.float32).ravel(), passes).reshape(nz, ny, nx)
    for z in range(nz):
        if z < k:
            d = k - 1 - z
            e = np.float32(0.5 - 0.2) if d == 0 else (np.float32(-(d - (0.5 - 0.2))) if d < passes else -U)
        else:
            d = z - k
            e = np.float32(d + (1.5 - 0.2)) if d < passes else U
        assert np.all(got[z] == e), (z, got[z].ravel()[:3], e)
    n = 9
    phi = np.full((n, n, n), 500.0, np.float32)
    phi[4, 4, 4] = -500.0
    got = orc.fluid_sdf((n, n, n), phi.ravel(), 1000).reshape(n, n, n)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    l1 = np.abs(x - 4) + np.abs(y - 4) + np.abs(z - 4)
    exp = (l1 - 1 + (1.5 - 0.2)).astype(np.float32)
    exp[4, 4, 4] = np.float32(0.5 - 0.2)
    assert np.array_equal(got, exp)


def test_particle_transfer_goldens(cases, orc):
    """The reference's particle tests (UnitTest-ParticlesAdvection.cu): P2G
    with P2GSUM / LINEAR / QUADRATIC / CUBIC kernels (normalised by the
    weights), G2P with G2PNEAREST and the same kernels, particle advection
    (RK1 with and without substeps, RK2, RK3), fluid markers from particles and
    the particle count of generated particles."""
    for c in cases["particles"]:
        if c["kind"] == "p2g":
            f, _ = orc.p2g(c["res"], c["origin"], c["spacing"], c["field"], c["pos"], c["val"], c["alg"],
                           span=c["span"])
            check_map(f, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "g2p":
            v = orc.g2p(c["res"], c["origin"], c["spacing"], c["field"], c["pos"], c["grid"], c["alg"],
                        span=c["span"])
            check_map(v, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "p2g_fcv":
            for k, comp in enumerate("xyz"):
                f, _ = orc.p2g(c["res"], c["origin"], c["spacing"], k + 1, c["pos"], c["val"], c["alg"],
                               span=3, off=k)
                if comp in c["expect"]:
                    check_map(f, c["expect"][comp], c["tol"], c["name"] + ":" + comp)
        elif c["kind"] == "advect":
            p = orc.advect_particles(c["res"], c["origin"], c["spacing"], c["u"], c["v"], c["w"], c["pos"], c["dt"],
                                     c["cfl"], c["order"])
            check_map(p, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "markers":
            m = orc.particle_markers(c["res"], c["origin"], c["spacing"], c["solid"], c["pos"])
            check_map(m, c["expect"], c["tol"], c["name"])
        elif c["kind"] == "density":
            pos = orc.generate_particles(c["res"], c["origin"], c["spacing"], c["fluid"], c["solid"],
                                         c["per_cell"], seed=2024)
            d = orc.particle_markers(c["res"], c["origin"], c["spacing"], None, pos, counts_only=True)
            check_map(d, c["expect"], c["tol"], c["name"])
        else:
            raise AssertionError(c["kind"])
