"""The C-ABI library (no GPU needed): it loads, exports every function that
include/mgf.h declares, and its host-only entry points behave."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mgf


def test_library_exports_header_symbols():
    names = mgf.exported_names()
    assert len(names) >= 30
    L = mgf.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and nothing extra of ours is exported that the header does not declare
    out = subprocess.run(["nm", "-D", "--defined-only", mgf.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    ours = {ln.split()[-1] for ln in out.splitlines() if " T mgf_" in ln}
    assert ours == set(names), sorted(ours ^ set(names))


def test_every_sig_is_in_header():
    assert set(mgf._SIGS) == set(mgf.exported_names())


def test_desc_defaults_are_the_reference_defaults():
    d = mgf.desc_defaults((10, 11, 12), (0.5, -1, 2), (0.25, 0.5, 1.0), 0.01)
    assert (d.nx, d.ny, d.nz) == (10, 11, 12)
    assert list(d.origin) == [0.5, -1.0, 2.0]
    assert d.rk_order == mgf.RK3                            # GridFluidSolver.cpp:404
    assert d.tol_mode == mgf.TOL_ABS_SQRT_NORM and d.tolerance == 1e-3   # CuMatrixFreePCG.h:28
    assert d.max_iterations == 0                            # -> cells (PressureSolver.cpp:82)
    assert d.extrapolation_layers == 1000                   # GridFluidSolver.cpp:176
    assert np.float32(d.gravity[1]) == np.float32(-9.8)     # ExternalForcesSolver.h:25
    assert d.density == 1.0                                 # PressureSolver.h:61
    assert d.precision == mgf.PRECISION_PLAIN
    assert d.preconditioner == mgf.PRECOND_IDENTITY                  # PressureSolver.cpp:12-15
    assert d.mg_sweeps == 3 and np.float32(d.mg_omega) == np.float32(0.9)


def test_status_strings():
    L = mgf.lib()
    assert L.mgf_status_string(0) == b"ok"
    assert L.mgf_status_string(mgf.ERR_NOT_CONVERGED) == b"not converged"


def test_null_arguments_are_rejected_without_a_device():
    L = mgf.lib()
    assert L.mgf_scene_destroy(None) == mgf.ERR_INVALID
    assert L.mgf_step(None, 1, None) == mgf.ERR_INVALID
    assert L.mgf_shutdown(None) == mgf.ERR_INVALID
    assert L.mgf_init(0, None) == mgf.ERR_INVALID
    assert L.mgf_batch_step(None, None, 0, 1, 0, None, None, None) == mgf.ERR_INVALID
    assert L.mgf_server_create(None, 1 << 30, 2, None) == mgf.ERR_INVALID
    assert L.mgf_server_destroy(None) == mgf.ERR_INVALID
    assert L.mgf_server_submit(None, None, 1, None) == mgf.ERR_INVALID
    assert L.mgf_server_poll(None, None, 0, 0, None) == mgf.ERR_INVALID
    assert L.mgf_scene_bytes_for(None, None, None) == mgf.ERR_INVALID
    assert L.mgf_status_string(mgf.ERR_BUDGET) == b"HBM budget full"


def test_ragged_scene_generator():
    """The batched/served workloads' scenes: seeded, per-axis sizes from the
    set, walls negative exactly on the ghost layer, fluid block and spheres
    inside the domain."""
    import scenes
    a, b = scenes.ragged_scene(11), scenes.ragged_scene(11)
    assert a.name == b.name and a.spheres == b.spheres and a.fluid_max == b.fluid_max
    seen = set()
    for seed in range(40):
        sp = scenes.ragged_scene(seed)
        assert all(n in scenes.BATCH_SIZES for n in sp.nxyz)
        assert 0 <= len(sp.spheres) <= 3
        seen.add(sp.nxyz)
        for k in range(3):
            assert 0 < sp.fluid_max[k] <= sp.extent[k] + 1e-12
        vol = np.prod(sp.fluid_max) / np.prod(sp.extent)
        assert 0.2 - 1e-9 <= vol <= 0.5 + 1e-9
    assert len(seen) > 20
    sp = scenes.ragged_scene(3, sizes=(8, 12), h=1.0 / 16)
    w = sp.wall_sdf()
    assert w.shape == sp.res[::-1]
    inner = np.zeros(w.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    assert np.array_equal(w < 0, ~inner)
    for seed in range(20):
        sp = scenes.ragged_scene(seed, log_uniform=(32, 256), max_spheres=4)
        assert all(32 <= n <= 256 for n in sp.nxyz) and len(sp.spheres) <= 4


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is present")
def test_init_without_gpu_reports_device_error():
    h = C.c_void_p()
    rc = mgf.lib().mgf_init(0, C.byref(h))
    assert rc == mgf.ERR_DEVICE
    assert b"device" in mgf.lib().mgf_last_error().lower()
