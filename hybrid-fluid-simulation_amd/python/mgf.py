"""ctypes binding of libmgf.so (include/mgf.h) for Python callers.

The library is the product: HIP kernels for gfx950 behind a C ABI.  This
module only marshals arguments; it raises if the library is missing or a call
fails, and never computes anything itself.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG, "lib", "libmgf.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "mgf.h")

OK, ERR_INVALID, ERR_DEVICE, ERR_NOMEM, ERR_NOT_CONVERGED, ERR_NONFINITE, ERR_UNSUPPORTED, ERR_BUDGET = range(8)
RK1, RK2, RK3 = 1, 2, 3
TOL_ABS_SQRT_NORM, TOL_RELATIVE = 0, 1
PRECISION_PLAIN, PRECISION_MIXED = 0, 1
PRECOND_IDENTITY, PRECOND_JACOBI, PRECOND_MULTIGRID = 0, 1, 2
MG_AUTO, MG_PER_SWEEP, MG_FUSED = 0, 1, 2   # MGF_MG_*
PG = dict(p2gsum=0, g2pnearest=1, linear=2, quadratic=3, cubic=4)   # MGF_PG_*
FIELD = dict(vel_x=0, vel_y=1, vel_z=2, fluid_domain=3, solid_sdf=4, solid_vel_x=5,
             solid_vel_y=6, solid_vel_z=7, markers=8, pressure=9, rhs=10, matrix_a=11,
             divergence=12, color_x=13, color_y=14, color_z=15, fluid_sdf=16, fluid_density=17)


class SceneDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("origin", C.c_float * 3), ("spacing", C.c_float * 3), ("dt", C.c_float),
                ("density", C.c_float), ("gravity", C.c_float * 3),
                ("external_force", C.c_float * 3), ("rk_order", C.c_int32),
                ("tol_mode", C.c_int32), ("tolerance", C.c_double),
                ("max_iterations", C.c_int64), ("precision", C.c_int32),
                ("preconditioner", C.c_int32), ("extrapolation_layers", C.c_int32),
                ("max_refinements", C.c_int32), ("inner_tolerance", C.c_double),
                ("mg_sweeps", C.c_int32), ("mg_omega", C.c_float), ("warm_start", C.c_int32),
                ("mg_kernels", C.c_int32)]


class StepStats(C.Structure):
    _fields_ = [("cg_iterations", C.c_int64), ("refinement_sweeps", C.c_int32),
                ("converged", C.c_int32), ("residual_rel", C.c_double),
                ("residual_recurrence", C.c_double), ("b_norm", C.c_double),
                ("extrapolation_layers", C.c_int32), ("steps", C.c_int32),
                ("ms_advect", C.c_double), ("ms_pressure", C.c_double),
                ("ms_extrapolate", C.c_double), ("ms_total", C.c_double),
                ("fluid_cells", C.c_int64), ("div_residual", C.c_double),
                ("true_converged", C.c_int32), ("reserved0", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelProbe(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("ms", C.c_double),
                ("busy_launches", C.c_int64), ("busy_ms", C.c_double), ("busy_bytes", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("scene", C.c_void_p), ("status", C.c_int32),
                ("steps", C.c_int32), ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_end_ms", C.c_double), ("stats", StepStats)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("origin", C.c_float * 3), ("spacing", C.c_float * 3)]


class MgfError(RuntimeError):
    def __init__(self, status, where, msg):
        super().__init__("%s failed: %s (%s)" % (where, msg, status))
        self.status = status


_lib = None
_F = C.POINTER(C.c_float)
_VP = C.c_void_p

# name -> argtypes (all return mgf_status = int32 unless listed in _RESTYPE)
_SIGS = {
    "mgf_init": [C.c_int32, C.POINTER(C.c_void_p)],
    "mgf_shutdown": [_VP],
    "mgf_status_string": [C.c_int32],
    "mgf_last_error": [],
    "mgf_scene_desc_defaults": [C.POINTER(SceneDesc), C.c_int32, C.c_int32, C.c_int32, _F, _F,
                                C.c_float],
    "mgf_scene_create": [_VP, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)],
    "mgf_scene_destroy": [_VP],
    "mgf_scene_get_desc": [_VP, C.POINTER(SceneDesc)],
    "mgf_scene_set_params": [_VP, C.POINTER(SceneDesc)],
    "mgf_scene_device_bytes": [_VP, C.POINTER(C.c_int64)],
    "mgf_scene_add_solid": [_VP, _F, C.c_int64, _F],
    "mgf_scene_fluid_box": [_VP, _F, _F],
    "mgf_scene_upload": [_VP, C.c_int32, _F, C.c_int64],
    "mgf_scene_download": [_VP, C.c_int32, _F, C.c_int64],
    "mgf_scene_upload_device": [_VP, C.c_int32, _VP, C.c_int64],
    "mgf_scene_download_device": [_VP, C.c_int32, _VP, C.c_int64],
    "mgf_scene_field_count": [_VP, C.c_int32, C.POINTER(C.c_int64)],
    "mgf_scene_update_boundaries": [_VP],
    "mgf_scene_advect": [_VP],
    "mgf_scene_apply_forces": [_VP],
    "mgf_scene_pressure_solve": [_VP, C.POINTER(StepStats)],
    "mgf_scene_extrapolate_velocity": [_VP, C.POINTER(C.c_int32)],
    "mgf_step": [_VP, C.c_int32, C.POINTER(StepStats)],
    "mgf_scene_sync": [_VP],
    "mgf_scene_set_timing": [_VP, C.c_int32],
    "mgf_scene_set_kernel_probes": [_VP, C.c_int32],
    "mgf_scene_kernel_probes": [_VP, C.POINTER(KernelProbe), C.c_int32, C.POINTER(C.c_int32), C.c_int32],
    "mgf_scene_set_divergence_check": [_VP, C.c_int32],
    "mgf_scene_set_block_skip": [_VP, C.c_int32],
    "mgf_batch_step": [_VP, C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32, C.POINTER(StepStats),
                       C.POINTER(C.c_int32), C.POINTER(C.c_double)],
    "mgf_scene_bytes_for": [_VP, C.POINTER(SceneDesc), C.POINTER(C.c_int64)],
    "mgf_server_create": [_VP, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)],
    "mgf_server_destroy": [_VP],
    "mgf_server_open_scene": [_VP, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)],
    "mgf_server_close_scene": [_VP, _VP],
    "mgf_server_submit": [_VP, _VP, C.c_int32, C.POINTER(C.c_int64)],
    "mgf_server_poll": [_VP, C.POINTER(RequestResult), C.c_int32, C.c_int32, C.POINTER(C.c_int32)],
    "mgf_server_state": [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "mgf_device_alloc": [_VP, C.c_int64, C.POINTER(C.c_void_p)],
    "mgf_device_free": [_VP, _VP],
    "mgf_memcpy_h2d": [_VP, _VP, _VP, C.c_int64],
    "mgf_memcpy_d2h": [_VP, _VP, _VP, C.c_int64],
    "mgf_memcpy_d2d": [_VP, _VP, _VP, C.c_int64],
    "mgf_op_backtrace": [_VP, C.POINTER(Grid), _VP, _VP, _VP, C.c_float, C.c_int32, _VP, _VP, _VP,
                         C.c_int64],
    "mgf_op_sample_cc": [_VP, C.POINTER(Grid), _VP, _VP, _VP, _VP, _VP, C.c_int64],
    "mgf_op_advect_fc": [_VP, C.POINTER(Grid), _VP, _VP, _VP, _VP, _VP, _VP, C.c_float, C.c_int32,
                         _VP, _VP, _VP],
    "mgf_op_advect_cc": [_VP, C.POINTER(Grid), _VP, _VP, _VP, _VP, C.c_float, C.c_int32, _VP],
    "mgf_op_pressure_matvec": [_VP, C.POINTER(Grid), _F, _VP, _VP, _VP],
    "mgf_op_apply_forces": [_VP, C.POINTER(Grid), _VP, _VP, _VP, _F, _F, C.c_float],
    "mgf_op_particles_to_grid": [_VP, C.POINTER(Grid), C.c_int32, C.c_int64, _VP, _VP, C.c_int32, C.c_int32,
                                 C.c_int32, _VP, _VP, C.c_int32],
    "mgf_op_grid_to_particles": [_VP, C.POINTER(Grid), C.c_int32, C.c_int64, _VP, _VP, C.c_int32, _VP, C.c_int32,
                                 C.c_int32],
    "mgf_op_advect_particles": [_VP, C.POINTER(Grid), _VP, _VP, _VP, C.c_int64, _VP, C.c_float, C.c_float,
                                C.c_int32, C.c_int32],
    "mgf_op_particle_markers": [_VP, C.POINTER(Grid), _VP, C.c_int64, _VP, _VP],
    "mgf_op_axpy": [_VP, C.c_int64, C.c_float, _VP, _VP],
    "mgf_particles_create": [_VP, C.c_int64, _F, _F, _F, C.POINTER(C.c_void_p)],
    "mgf_particles_generate": [_VP, C.POINTER(Grid), _VP, _VP, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)],
    "mgf_particles_destroy": [_VP],
    "mgf_particles_count": [_VP, C.POINTER(C.c_int64)],
    "mgf_particles_data": [_VP, C.c_int32, C.POINTER(C.c_void_p)],
    "mgf_particles_upload": [_VP, C.c_int32, _F, C.c_int64],
    "mgf_particles_download": [_VP, C.c_int32, _F, C.c_int64],
    "mgf_particles_advect": [_VP, C.POINTER(Grid), _VP, _VP, _VP, C.c_float, C.c_float, C.c_int32, C.c_int32],
    "mgf_psolver_create": [_VP, C.POINTER(Grid), C.c_int32, _VP, C.c_float, C.c_int32, C.c_float,
                           C.POINTER(C.c_void_p)],
    "mgf_psolver_destroy": [_VP],
    "mgf_psolver_advect_velocity": [_VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_float, C.c_int32, C.c_int32, _VP, _VP,
                                    _VP],
    "mgf_psolver_advect_scalar": [_VP, _VP, _VP],
    "mgf_psolver_advect_vector": [_VP, _VP, _VP, _VP, _VP, _VP, _VP],
    "mgf_scene_pressure_solve_particles": [_VP, _VP, C.POINTER(StepStats)],
    "mgf_step_particles": [_VP, _VP, C.c_int32, C.POINTER(StepStats)],
    "mgf_scene_generate_particles": [_VP, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)],
    "mgf_scene_apply_preconditioner": [_VP, _VP, _VP, C.POINTER(C.c_double)],
    "mgf_op_scale": [_VP, C.c_int64, C.c_float, _VP],
    "mgf_op_extrapolate_fc": [_VP, C.POINTER(Grid), _VP, _VP, _VP, C.c_int32, C.POINTER(C.c_int32)],
    "mgf_batch_set_lockstep": [_VP, C.c_int32],
    "mgf_batch_autotune": [_VP, _VP, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)],
    "mgf_op_extrapolate_fc_noalloc": [_VP, C.POINTER(Grid), _VP, _VP, _VP, C.c_int32, C.POINTER(C.c_int32)],
}
_RESTYPE = {"mgf_status_string": C.c_char_p, "mgf_last_error": C.c_char_p}


def exported_names():
    """Function names declared with MGF_API in include/mgf.h."""
    import re
    names = []
    with open(HEADER) as f:
        for line in f:
            m = re.match(r"MGF_API\s+[\w\s\*]+?\b(mgf_\w+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmgf.so is not built (%s): run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int32)
        _lib = L
    return _lib


def check(status, where):
    if status != OK:
        msg = lib().mgf_last_error().decode(errors="replace")
        raise MgfError(status, where, msg)
    return status


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    return a.ctypes.data_as(_F)


def farr3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().mgf_init(device, C.byref(h)), "mgf_init")
        self.h = h

    def close(self):
        if self.h:
            lib().mgf_shutdown(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # device buffers for the operator entry points
    def alloc(self, nbytes):
        p = C.c_void_p()
        check(lib().mgf_device_alloc(self.h, int(nbytes), C.byref(p)), "mgf_device_alloc")
        return p

    def free(self, p):
        check(lib().mgf_device_free(self.h, p), "mgf_device_free")

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        check(lib().mgf_memcpy_h2d(self.h, p, a.ctypes.data_as(C.c_void_p), a.nbytes), "h2d")
        return p

    def to_host(self, p, n, dtype=np.float32):
        out = np.empty(n, dtype)
        check(lib().mgf_memcpy_d2h(self.h, out.ctypes.data_as(C.c_void_p), p, out.nbytes), "d2h")
        return out


def desc_defaults(res, origin=(0, 0, 0), spacing=(1, 1, 1), dt=1.0):
    d = SceneDesc()
    o, s = f32(origin), f32(spacing)
    check(lib().mgf_scene_desc_defaults(C.byref(d), int(res[0]), int(res[1]), int(res[2]), fp(o), fp(s),
                                        C.c_float(dt)), "mgf_scene_desc_defaults")
    return d


def make_desc(res, origin=(0, 0, 0), spacing=(1, 1, 1), dt=1.0, rk=RK3, tol_mode=TOL_ABS_SQRT_NORM,
              tol=1e-3, max_it=0, precision=PRECISION_PLAIN, layers=1000, gravity=(0, -9.8, 0),
              force=(0, 0, 0), density=1.0, max_refinements=8, inner_tol=1e-3,
              precond=PRECOND_IDENTITY, mg_sweeps=3, mg_omega=0.9, warm_start=False, mg_kernels=0):
    d = desc_defaults(res, origin, spacing, dt)
    d.rk_order = rk
    d.tol_mode = tol_mode
    d.tolerance = tol
    d.max_iterations = max_it
    d.precision = precision
    d.extrapolation_layers = layers
    d.gravity = farr3(gravity)
    d.external_force = farr3(force)
    d.density = density
    d.max_refinements = max_refinements
    d.inner_tolerance = inner_tol
    d.preconditioner = precond
    d.mg_sweeps = mg_sweeps
    d.mg_omega = mg_omega
    d.warm_start = 1 if warm_start else 0
    d.mg_kernels = mg_kernels
    return d


class Scene:
    """CGridFluidSolver on one MI355X: fields resident in HBM."""

    def __init__(self, ctx, desc, handle=None):
        self.ctx = ctx
        self.desc = desc
        if handle is not None:          # a scene a Server admitted (the server owns it)
            self.h = handle
            self.owned = False
            return
        h = C.c_void_p()
        check(lib().mgf_scene_create(ctx.h, C.byref(desc), C.byref(h)), "mgf_scene_create")
        self.h = h
        self.owned = True

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                lib().mgf_scene_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def count(self, field):
        n = C.c_int64()
        check(lib().mgf_scene_field_count(self.h, field, C.byref(n)), "mgf_scene_field_count")
        return n.value

    def add_solid(self, sdf, vel=(0, 0, 0)):
        sdf = f32(sdf).ravel()
        check(lib().mgf_scene_add_solid(self.h, fp(sdf), sdf.size, fp(f32(vel))), "mgf_scene_add_solid")

    def fluid_box(self, mn, mx):
        check(lib().mgf_scene_fluid_box(self.h, fp(f32(mn)), fp(f32(mx))), "mgf_scene_fluid_box")

    def upload(self, field, data):
        data = f32(data).ravel()
        check(lib().mgf_scene_upload(self.h, field, fp(data), data.size), "mgf_scene_upload")

    def download(self, field):
        out = np.empty(self.count(field), np.float32)
        check(lib().mgf_scene_download(self.h, field, fp(out), out.size), "mgf_scene_download")
        return out

    def set_params(self, desc):
        check(lib().mgf_scene_set_params(self.h, C.byref(desc)), "mgf_scene_set_params")
        self.desc = desc

    def device_bytes(self):
        n = C.c_int64()
        check(lib().mgf_scene_device_bytes(self.h, C.byref(n)), "mgf_scene_device_bytes")
        return n.value

    def update_boundaries(self):
        check(lib().mgf_scene_update_boundaries(self.h), "mgf_scene_update_boundaries")

    def advect(self):
        check(lib().mgf_scene_advect(self.h), "mgf_scene_advect")

    def apply_forces(self):
        check(lib().mgf_scene_apply_forces(self.h), "mgf_scene_apply_forces")

    def pressure_solve(self, allow_unconverged=False):
        st = StepStats()
        rc = lib().mgf_scene_pressure_solve(self.h, C.byref(st))
        if not (allow_unconverged and rc == ERR_NOT_CONVERGED):
            check(rc, "mgf_scene_pressure_solve")
        return st.as_dict()

    def apply_preconditioner(self, r):
        """z = M^-1 r once (mgf_scene_apply_preconditioner) for a host array r
        of nx*ny*nz floats; returns (z, r . z)."""
        r = f32(r).ravel()
        R = self.ctx.to_device(r)
        Z = self.ctx.alloc(4 * r.size)
        rz = C.c_double()
        try:
            check(lib().mgf_scene_apply_preconditioner(self.h, R, Z, C.byref(rz)), "mgf_scene_apply_preconditioner")
            return self.ctx.to_host(Z, r.size), rz.value
        finally:
            self.ctx.free(R)
            self.ctx.free(Z)

    def extrapolate(self):
        n = C.c_int32()
        check(lib().mgf_scene_extrapolate_velocity(self.h, C.byref(n)), "mgf_scene_extrapolate_velocity")
        return n.value

    def step(self, n=1, allow_unconverged=False):
        st = StepStats()
        rc = lib().mgf_step(self.h, n, C.byref(st))
        if not (allow_unconverged and rc == ERR_NOT_CONVERGED):
            check(rc, "mgf_step")
        return st.as_dict()

    def pressure_solve_particles(self, particles, allow_unconverged=False):
        st = StepStats()
        rc = lib().mgf_scene_pressure_solve_particles(self.h, particles.h, C.byref(st))
        if not (allow_unconverged and rc == ERR_NOT_CONVERGED):
            check(rc, "mgf_scene_pressure_solve_particles")
        return st.as_dict()

    def generate_particles(self, per_cell, seed):
        """Particles in the scene's fluid cells (outside its solids)."""
        h = C.c_void_p()
        check(lib().mgf_scene_generate_particles(self.h, per_cell, seed, C.byref(h)), "mgf_scene_generate_particles")
        return Particles(self.ctx, handle=h)

    def step_particles(self, solver, n=1, allow_unconverged=False):
        st = StepStats()
        rc = lib().mgf_step_particles(self.h, solver.h, n, C.byref(st))
        if not (allow_unconverged and rc == ERR_NOT_CONVERGED):
            check(rc, "mgf_step_particles")
        return st.as_dict()

    def sync(self):
        check(lib().mgf_scene_sync(self.h), "mgf_scene_sync")

    def set_kernel_probes(self, on):
        check(lib().mgf_scene_set_kernel_probes(self.h, 1 if on else 0), "mgf_scene_set_kernel_probes")

    def kernel_probes(self, reset=True):
        """Per kernel launched while probes were on (template instantiations
        under one name, in order of first appearance): every launch and its
        summed ms, as a kernel trace counts them, and the busy launches (those
        that ran at least a quarter of the kernel's longest launch) with their
        ms and algorithmic bytes."""
        cap = 256
        arr = (KernelProbe * cap)()
        n = C.c_int32()
        check(lib().mgf_scene_kernel_probes(self.h, arr, cap, C.byref(n), 1 if reset else 0),
              "mgf_scene_kernel_probes")
        out = {}
        for i in range(min(n.value, cap)):
            e = arr[i]
            out[e.name.decode()] = {"launches": e.launches, "ms": e.ms, "busy_launches": e.busy_launches,
                                    "busy_ms": e.busy_ms, "busy_bytes": e.busy_bytes}
        return out

    def set_divergence_check(self, on):
        check(lib().mgf_scene_set_divergence_check(self.h, 1 if on else 0), "mgf_scene_set_divergence_check")

    def set_block_skip(self, on):
        check(lib().mgf_scene_set_block_skip(self.h, 1 if on else 0), "mgf_scene_set_block_skip")

    def set_timing(self, on):
        check(lib().mgf_scene_set_timing(self.h, 1 if on else 0), "mgf_scene_set_timing")


def grid(res, origin=(0, 0, 0), spacing=(1, 1, 1)):
    g = Grid()
    g.nx, g.ny, g.nz = [int(r) for r in res]
    g.origin = farr3(origin)
    g.spacing = farr3(spacing)
    return g


def batch_autotune(ctx, scenes, trial_steps=3):
    """mgf_batch_autotune over Scene objects: (chosen mg_kernels, [ms AUTO, ms FUSED])."""
    n = len(scenes)
    arr = (C.c_void_p * n)(*[sc.h.value if isinstance(sc.h, C.c_void_p) else sc.h for sc in scenes])
    chosen = C.c_int32()
    ms = (C.c_double * 2)()
    check(lib().mgf_batch_autotune(ctx.h, arr, n, trial_steps, C.byref(chosen), ms), "mgf_batch_autotune")
    return chosen.value, [ms[0], ms[1]]


def batch_step(ctx, scenes, nsteps, workers=0, step_latency=False):
    """mgf_batch_step over Scene objects: returns (statuses, stats, step_ms)
    with step_ms an (n, nsteps) array when step_latency, else None."""
    n = len(scenes)
    arr = (C.c_void_p * n)(*[sc.h.value if isinstance(sc.h, C.c_void_p) else sc.h for sc in scenes])
    stats = (StepStats * n)()
    status = (C.c_int32 * n)()
    ms = np.zeros((n, nsteps), dtype=np.float64) if step_latency else None
    rc = lib().mgf_batch_step(ctx.h, arr, n, nsteps, workers, stats, status,
                              ms.ctypes.data_as(C.POINTER(C.c_double)) if ms is not None else None)
    if rc not in (OK, ERR_NOT_CONVERGED):
        check(rc, "mgf_batch_step")
    return list(status), [st for st in stats], ms


def scene_bytes_for(ctx, desc):
    out = C.c_int64()
    check(lib().mgf_scene_bytes_for(ctx.h, C.byref(desc), C.byref(out)), "mgf_scene_bytes_for")
    return out.value


class Server:
    """The served request loop (mgf_server_*): admitted scenes stay resident;
    step requests run on `slots` workers and overlap on the device."""

    def __init__(self, ctx, budget_bytes, slots=0):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().mgf_server_create(ctx.h, int(budget_bytes), int(slots), C.byref(h)), "mgf_server_create")
        self.h = h

    def open_scene(self, desc):
        """Returns a Scene, or None when the HBM budget refuses admission."""
        h = C.c_void_p()
        rc = lib().mgf_server_open_scene(self.h, C.byref(desc), C.byref(h))
        if rc == ERR_BUDGET:
            return None
        check(rc, "mgf_server_open_scene")
        return Scene(self.ctx, desc, handle=h)

    def close_scene(self, scene):
        check(lib().mgf_server_close_scene(self.h, scene.h), "mgf_server_close_scene")
        scene.h = None

    def submit(self, scene, nsteps=1):
        t = C.c_int64()
        check(lib().mgf_server_submit(self.h, scene.h, int(nsteps), C.byref(t)), "mgf_server_submit")
        return t.value

    def poll(self, cap=256, wait_ms=0):
        out = (RequestResult * cap)()
        n = C.c_int32()
        check(lib().mgf_server_poll(self.h, out, cap, int(wait_ms), C.byref(n)), "mgf_server_poll")
        return [out[i] for i in range(n.value)]

    def state(self):
        b, q, r = C.c_int64(), C.c_int32(), C.c_int32()
        check(lib().mgf_server_state(self.h, C.byref(b), C.byref(q), C.byref(r)), "mgf_server_state")
        return b.value, q.value, r.value

    def close(self):
        if getattr(self, "h", None):
            check(lib().mgf_server_destroy(self.h), "mgf_server_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def field_extent(res, kind):
    """Values of a field on the cells (kind 0) or the x/y/z face grid (1/2/3)."""
    nx, ny, nz = res
    return {0: nx * ny * nz, 1: (nx + 1) * ny * nz, 2: nx * (ny + 1) * nz, 3: nx * ny * (nz + 1)}[kind]


def particles_to_grid(ctx, res, origin, spacing, kind, pos, values, alg, span=1, off=0, normalize=True):
    """mgf_op_particles_to_grid on host arrays (copied to the device and back)."""
    pos = f32(pos).ravel()
    n = pos.size // 3
    m = field_extent(res, kind) if kind in (0, 1, 2, 3) else 1   # the library rejects other kinds
    P, V = ctx.to_device(pos), ctx.to_device(f32(values).ravel())
    F, W = ctx.alloc(4 * m), ctx.alloc(4 * m)
    try:
        check(lib().mgf_op_particles_to_grid(ctx.h, grid(res, origin, spacing), kind, n, P, V, span, off, PG[alg],
                                             F, W, 1 if normalize else 0), "mgf_op_particles_to_grid")
        return ctx.to_host(F, m), ctx.to_host(W, m)
    finally:
        for p in (P, V, F, W):
            ctx.free(p)


def grid_to_particles(ctx, res, origin, spacing, kind, pos, field, alg, span=1, off=0, out=None):
    pos = f32(pos).ravel()
    n = pos.size // 3
    out = np.zeros(span * n, np.float32) if out is None else f32(out).copy()
    P, Fd, O = ctx.to_device(pos), ctx.to_device(f32(field).ravel()), ctx.to_device(out)
    try:
        check(lib().mgf_op_grid_to_particles(ctx.h, grid(res, origin, spacing), kind, n, P, Fd, PG[alg], O, span,
                                             off), "mgf_op_grid_to_particles")
        return ctx.to_host(O, out.size)
    finally:
        for p in (P, Fd, O):
            ctx.free(p)


def advect_particles(ctx, res, origin, spacing, u, v, w, pos, dt, cfl, order, fix=False):
    pos = f32(pos).ravel()
    D = [ctx.to_device(f32(a).ravel()) for a in (u, v, w)]
    P = ctx.to_device(pos)
    try:
        check(lib().mgf_op_advect_particles(ctx.h, grid(res, origin, spacing), D[0], D[1], D[2], pos.size // 3, P,
                                            C.c_float(dt), C.c_float(cfl), order, 1 if fix else 0),
              "mgf_op_advect_particles")
        return ctx.to_host(P, pos.size)
    finally:
        for p in D + [P]:
            ctx.free(p)


def particle_markers(ctx, res, origin, spacing, solid, pos):
    pos = f32(pos).ravel()
    m = int(np.prod(res))
    P, O = ctx.to_device(pos), ctx.alloc(4 * m)
    S = ctx.to_device(f32(solid).ravel()) if solid is not None else None
    try:
        check(lib().mgf_op_particle_markers(ctx.h, grid(res, origin, spacing), S, pos.size // 3, P, O),
              "mgf_op_particle_markers")
        return ctx.to_host(O, m)
    finally:
        for p in (P, O) + ((S,) if S is not None else ()):
            ctx.free(p)


PARTICLE = dict(pos=0, vel=1, color=2, scalar=3, vector=4)   # MGF_PARTICLE_*
PSOLVER = dict(pic=0, flip=1, mix=2)                        # MGF_PSOLVER_*


class Particles:
    """CEulerParticles: a particle set resident in HBM."""

    def __init__(self, ctx, pos=None, vel=None, color=None, n=None, handle=None):
        self.ctx = ctx
        if handle is not None:
            self.h = handle
            return
        arrs = [None if a is None else f32(a).ravel() for a in (pos, vel, color)]
        if n is None:
            n = next(a.size // 3 for a in arrs if a is not None)
        h = C.c_void_p()
        check(lib().mgf_particles_create(ctx.h, n, *[fp(a) if a is not None else None for a in arrs], C.byref(h)),
              "mgf_particles_create")
        self.h = h

    @classmethod
    def generate(cls, ctx, res, origin, spacing, fluid, solid, per_cell, seed):
        """generateParticlesInFluid from host fields (copied to the device)."""
        F = ctx.to_device(f32(fluid).ravel())
        S = ctx.to_device(f32(solid).ravel()) if solid is not None else None
        try:
            h = C.c_void_p()
            check(lib().mgf_particles_generate(ctx.h, grid(res, origin, spacing), F, S, per_cell, seed, C.byref(h)),
                  "mgf_particles_generate")
        finally:
            ctx.free(F)
            if S is not None:
                ctx.free(S)
        return cls(ctx, handle=h)

    def close(self):
        if getattr(self, "h", None):
            lib().mgf_particles_destroy(self.h)
            self.h = None

    def __len__(self):
        n = C.c_int64()
        check(lib().mgf_particles_count(self.h, C.byref(n)), "mgf_particles_count")
        return n.value

    def _count(self, which):
        return len(self) * (1 if which == "scalar" else 3)

    def get(self, which):
        out = np.empty(self._count(which), np.float32)
        check(lib().mgf_particles_download(self.h, PARTICLE[which], fp(out), out.size), "mgf_particles_download")
        return out

    def set(self, which, data):
        data = f32(data).ravel()
        check(lib().mgf_particles_upload(self.h, PARTICLE[which], fp(data), data.size), "mgf_particles_upload")

    def data(self, which):
        p = C.c_void_p()
        check(lib().mgf_particles_data(self.h, PARTICLE[which], C.byref(p)), "mgf_particles_data")
        return p

    def advect(self, res, origin, spacing, u, v, w, dt, cfl, order, fix=False):
        D = [self.ctx.to_device(f32(a).ravel()) for a in (u, v, w)]
        try:
            check(lib().mgf_particles_advect(self.h, grid(res, origin, spacing), D[0], D[1], D[2], C.c_float(dt),
                                             C.c_float(cfl), order, 1 if fix else 0), "mgf_particles_advect")
        finally:
            for d in D:
                self.ctx.free(d)


class ParticleSolver:
    """CParticleInCell / CFluidImplicitParticle / CMixPICAndFLIP on a
    Particles set (borrowed; keep it alive)."""

    def __init__(self, ctx, res, origin, spacing, kind, particles, cfl=1.0, alg="linear", mixing=0.1):
        self.ctx = ctx
        self.res = tuple(int(r) for r in res)
        self.particles = particles
        h = C.c_void_p()
        check(lib().mgf_psolver_create(ctx.h, grid(res, origin, spacing), PSOLVER[kind], particles.h,
                                       C.c_float(cfl), PG[alg], C.c_float(mixing), C.byref(h)), "mgf_psolver_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mgf_psolver_destroy(self.h)
            self.h = None

    def advect_velocity(self, inp, vel, dt, order, fix):
        """Host arrays in and out (copied through device buffers)."""
        ctx = self.ctx
        I = [ctx.to_device(f32(a).ravel()) for a in inp]
        V = [ctx.to_device(f32(a).ravel()) for a in vel]
        n = [field_extent(self.res, k) for k in (1, 2, 3)]
        O = [ctx.alloc(4 * m) for m in n]
        try:
            check(lib().mgf_psolver_advect_velocity(self.h, *I, *V, C.c_float(dt), order, 1 if fix else 0, *O),
                  "mgf_psolver_advect_velocity")
            return [ctx.to_host(o, m) for o, m in zip(O, n)]
        finally:
            for p in I + V + O:
                ctx.free(p)

    def advect_scalar(self, inp):
        ctx = self.ctx
        m = field_extent(self.res, 0)
        I, O = ctx.to_device(f32(inp).ravel()), ctx.alloc(4 * m)
        try:
            check(lib().mgf_psolver_advect_scalar(self.h, I, O), "mgf_psolver_advect_scalar")
            return ctx.to_host(O, m)
        finally:
            ctx.free(I)
            ctx.free(O)

    def advect_vector(self, inp):
        ctx = self.ctx
        m = field_extent(self.res, 0)
        I = [ctx.to_device(f32(a).ravel()) for a in inp]
        O = [ctx.alloc(4 * m) for _ in range(3)]
        try:
            check(lib().mgf_psolver_advect_vector(self.h, *I, *O), "mgf_psolver_advect_vector")
            return [ctx.to_host(o, m) for o in O]
        finally:
            for p in I + O:
                ctx.free(p)
