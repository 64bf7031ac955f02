"""mgf_op_extrapolate_fc_noalloc (the entry point that allocates nothing)
against mgf_op_extrapolate_fc through the C ABI: bit for bit, on two calls in
a row, the second on a smaller grid (the shared scratch keeps the first
call's contents), plus the refusal of grids beyond its scratch.  Needs an
MI355X."""
import numpy as np
import pytest

import mgf

pytestmark = pytest.mark.gpu
FLT_MAX = float(np.finfo(np.float32).max)


def faces(rng, res, known):
    nx, ny, nz = res
    out = []
    for n in ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)):
        a = rng.standard_normal(n).astype(np.float32)
        a[rng.random(n) >= known] = FLT_MAX
        out.append(a)
    return out


def run(ctx, entry, res, arrs, layers):
    D = [ctx.to_device(a) for a in arrs]
    used = mgf.C.c_int32(-1)
    st = getattr(mgf.lib(), entry)(ctx.h, mgf.grid(res), D[0], D[1], D[2], layers, mgf.C.byref(used))
    out = [ctx.to_host(d, a.size) for d, a in zip(D, arrs)]
    for d in D:
        ctx.free(d)
    return st, out, used.value


@pytest.fixture(scope="module")
def ctx():
    c = mgf.Context(0)
    yield c
    c.close()


def test_noalloc_matches_allocating_entry_point_twice(ctx):
    rng = np.random.default_rng(11)
    # first call: a grid near the scratch limit (3.9M faces), then a smaller one
    for res, known, layers in (((110, 107, 110), 0.02, 1000), ((40, 24, 31), 0.05, 1000)):
        arrs = faces(rng, res, known)
        st_a, ref, used_a = run(ctx, "mgf_op_extrapolate_fc", res, arrs, layers)
        mgf.check(st_a, "mgf_op_extrapolate_fc %r" % (res,))
        st_b, got, used_b = run(ctx, "mgf_op_extrapolate_fc_noalloc", res, arrs, layers)
        mgf.check(st_b, "mgf_op_extrapolate_fc_noalloc %r" % (res,))
        assert used_a == used_b and used_a > 1, (res, used_a, used_b)
        for k in range(3):
            assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (res, k)
            assert np.any(got[k] != arrs[k]), (res, k)   # something was extrapolated


def test_noalloc_refuses_grids_beyond_its_scratch(ctx):
    used = mgf.C.c_int32()
    # 128^3 has 6.3M faces (the scratch holds 4194304): refused before any launch
    big = (128, 128, 128)
    st = mgf.lib().mgf_op_extrapolate_fc_noalloc(ctx.h, mgf.grid(big), mgf.C.c_void_p(16), mgf.C.c_void_p(16),
                                                 mgf.C.c_void_p(16), 1000, mgf.C.byref(used))
    assert st == mgf.ERR_INVALID
    assert b"scratch" in mgf.lib().mgf_last_error()
    # a long thin grid: few faces but nx + ny + nz + 3 beyond 4093 with more than one layer
    thin = (5000, 2, 2)
    st = mgf.lib().mgf_op_extrapolate_fc_noalloc(ctx.h, mgf.grid(thin), mgf.C.c_void_p(16), mgf.C.c_void_p(16),
                                                 mgf.C.c_void_p(16), 1000, mgf.C.byref(used))
    assert st == mgf.ERR_INVALID
