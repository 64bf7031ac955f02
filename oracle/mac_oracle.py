"""ctypes wrapper of the CPU oracle (libmac_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the CPU baseline,
never as the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmac_oracle.so")

FIELD = dict(vel_x=0, vel_y=1, vel_z=2, fluid_domain=3, solid_sdf=4,
             solid_vel_x=5, solid_vel_y=6, solid_vel_z=7, markers=8,
             pressure=9, rhs=10, matrix_a=11, color_x=13, color_y=14, color_z=15)


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("origin", C.c_double * 3), ("spacing", C.c_double * 3)]


class OrcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("density", C.c_double),
                ("gravity", C.c_double * 3), ("external_force", C.c_double * 3),
                ("rk_order", C.c_int32), ("tol_mode", C.c_int32),
                ("tolerance", C.c_double), ("max_iterations", C.c_int64),
                ("jacobi", C.c_int32), ("extrapolation_layers", C.c_int32)]


class OrcStats(C.Structure):
    _fields_ = [("cg_iterations", C.c_int64), ("converged", C.c_int32),
                ("residual_rel", C.c_double), ("residual_recurrence", C.c_double),
                ("b_norm", C.c_double), ("extrapolation_layers", C.c_int32),
                ("fluid_cells", C.c_int64), ("ms_advect", C.c_double),
                ("ms_pressure", C.c_double), ("ms_extrapolate", C.c_double),
                ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build():
    """Compile libmac_oracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


def _setup(L):
    g = C.POINTER(OrcGrid)
    L.orc_sample_cc_f32.argtypes = [g, _F, _F, _F, _F, _F, C.c_int64]
    L.orc_sample_fc_f32.argtypes = [g, _F, _F, _F, _F, _F, _F, _F, _F, _F, C.c_int64]
    L.orc_backtrace_f32.argtypes = [g, _F, _F, _F, C.c_float, C.c_int32, _F, _F, _F, C.c_int64]
    L.orc_advect_cc_f32.argtypes = [g, _F, _F, _F, _F, C.c_float, C.c_int32, _F]
    L.orc_advect_fc_f32.argtypes = [g, _F, _F, _F, _F, _F, _F, C.c_float, C.c_int32, _F, _F, _F]
    L.orc_external_forces_f32.argtypes = [g, _F, _F, _F, _F, _F, C.c_float]
    L.orc_divergence_f32.argtypes = [g, _F, _F, _F, _F]
    L.orc_matvec_f32.argtypes = [g, _F, _F, _F]
    L.orc_build_matrix_f32.argtypes = [g, _F, _F, _F]
    L.orc_cg_dense_f32.argtypes = [_F, _F, _F, _F, C.c_int64, C.c_int32, C.c_double,
                                   C.c_int64, C.POINTER(C.c_int64)]
    L.orc_extrapolate_grid_f32.argtypes = [_F, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _I]
    L.orc_extrapolate_fc_f32.argtypes = [g, _F, _F, _F, C.c_int32]
    L.orc_fluid_density_f32.argtypes = [g, _F, _F, _F]
    L.orc_p2g_f32.argtypes = [g, C.c_int32, C.c_int64, _F, _F, C.c_int32, C.c_int32, C.c_int32, _F, _F, C.c_int32]
    L.orc_g2p_f32.argtypes = [g, C.c_int32, C.c_int64, _F, _F, C.c_int32, _F, C.c_int32, C.c_int32]
    L.orc_particle_markers_f32.argtypes = [g, _F, C.c_int64, _F, _F, C.c_int32]
    L.orc_advect_particles_f32.argtypes = [g, _F, _F, _F, C.c_int64, _F, C.c_float, C.c_float, C.c_int32,
                                           C.c_int32]
    L.orc_fluid_sdf_f32.argtypes = [g, _F, _F, C.c_int32]
    L.orc_pressure_solve_f32.argtypes = [g, C.c_float, C.c_float, _F, _F, _F, _F, _F, _F, _F,
                                         _F, C.c_int32, C.c_double, C.c_int64, C.c_int32,
                                         _F, _F, _F, _F, C.POINTER(OrcStats)]
    L.orc_pressure_solve_f64.argtypes = [g, C.c_float, C.c_float, _F, _F, _F, _F, _F, _F, _F,
                                         _F, C.c_int32, C.c_double, C.c_int64, C.c_int32,
                                         _D, _D, _D, _D, C.POINTER(OrcStats)]
    L.orc_scene_create.argtypes = [g, C.c_int32]
    L.orc_scene_create.restype = C.c_void_p
    L.orc_scene_destroy.argtypes = [C.c_void_p]
    L.orc_scene_add_solid.argtypes = [C.c_void_p, _F, C.c_int64, _F]
    L.orc_scene_fluid_box.argtypes = [C.c_void_p, _F, _F]
    L.orc_scene_upload.argtypes = [C.c_void_p, C.c_int32, _F, C.c_int64]
    L.orc_scene_download.argtypes = [C.c_void_p, C.c_int32, _F, C.c_int64]
    L.orc_scene_download_f64.argtypes = [C.c_void_p, C.c_int32, _D, C.c_int64]
    P = C.POINTER(OrcParams)
    S = C.POINTER(OrcStats)
    L.orc_scene_update_boundaries.argtypes = [C.c_void_p, P]
    L.orc_scene_advect.argtypes = [C.c_void_p, P]
    L.orc_scene_apply_forces.argtypes = [C.c_void_p, P]
    L.orc_scene_pressure_solve.argtypes = [C.c_void_p, P, S]
    L.orc_scene_extrapolate.argtypes = [C.c_void_p, P]
    L.orc_scene_step.argtypes = [C.c_void_p, P, C.c_int32, S]
    L.orc_scene_solid_state.argtypes = [C.c_void_p, C.c_int32, _F, _F, C.c_int64]
    L.orc_num_threads.restype = C.c_int
    L.orc_splitmix64.argtypes = [C.c_uint64]
    L.orc_splitmix64.restype = C.c_uint64
    L.orc_generate_particles.argtypes = [g, _F, _F, C.c_int32, C.c_uint64, _F, C.c_int64]
    L.orc_generate_particles.restype = C.c_int64
    L.orc_psolver_create.argtypes = [g, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_int64, _F]
    L.orc_psolver_create.restype = C.c_void_p
    L.orc_psolver_destroy.argtypes = [C.c_void_p]
    L.orc_psolver_advect_velocity.argtypes = [C.c_void_p] + [_F] * 6 + [C.c_float, C.c_int32, C.c_int32] + [_F] * 3
    L.orc_psolver_advect_scalar.argtypes = [C.c_void_p, _F, _F]
    L.orc_psolver_advect_vector.argtypes = [C.c_void_p] + [_F] * 6
    L.orc_psolver_get.argtypes = [C.c_void_p, C.c_int32, _F, C.c_int64]
    L.orc_scene_step_particles.argtypes = [C.c_void_p, C.c_void_p, P, C.c_int32, S]
    L.orc_scene_pressure_solve_particles.argtypes = [C.c_void_p, C.c_void_p, P, S]


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    return a.ctypes.data_as(_F)


def dp(a):
    return a.ctypes.data_as(_D)


def grid(res, origin=(0, 0, 0), spacing=(1, 1, 1)):
    g = OrcGrid()
    g.nx, g.ny, g.nz = [int(v) for v in res]
    for i in range(3):
        g.origin[i] = float(np.float32(origin[i]))
        g.spacing[i] = float(np.float32(spacing[i]))
    return g


def face_shape(res, comp):
    nx, ny, nz = res
    return (nz + (comp == 2), ny + (comp == 1), nx + (comp == 0))


def params(dt, rk=3, tol_mode=0, tol=1e-3, max_it=0, jacobi=False, layers=1000,
           gravity=(0.0, -9.8, 0.0), force=(0.0, 0.0, 0.0), density=1.0):
    p = OrcParams()
    p.dt = float(np.float32(dt))
    p.density = float(np.float32(density))
    for i in range(3):
        p.gravity[i] = float(np.float32(gravity[i]))
        p.external_force[i] = float(np.float32(force[i]))
    p.rk_order = rk
    p.tol_mode = tol_mode
    p.tolerance = tol
    p.max_iterations = max_it
    p.jacobi = 1 if jacobi else 0
    p.extrapolation_layers = layers
    return p


# ---- operators ------------------------------------------------------------
def sample_cc(res, origin, spacing, field, px, py, pz):
    g = grid(res, origin, spacing)
    field, px, py, pz = f32(field), f32(px), f32(py), f32(pz)
    out = np.empty_like(px)
    lib().orc_sample_cc_f32(C.byref(g), fp(field), fp(px), fp(py), fp(pz), fp(out), px.size)
    return out


def sample_fc(res, origin, spacing, u, v, w, px, py, pz):
    g = grid(res, origin, spacing)
    u, v, w, px, py, pz = map(f32, (u, v, w, px, py, pz))
    ou, ov, ow = np.empty_like(px), np.empty_like(px), np.empty_like(px)
    lib().orc_sample_fc_f32(C.byref(g), fp(u), fp(v), fp(w), fp(px), fp(py), fp(pz),
                            fp(ou), fp(ov), fp(ow), px.size)
    return ou, ov, ow


def backtrace(res, origin, spacing, u, v, w, dt, order, px, py, pz):
    g = grid(res, origin, spacing)
    u, v, w = f32(u), f32(v), f32(w)
    px, py, pz = f32(px).copy(), f32(py).copy(), f32(pz).copy()
    lib().orc_backtrace_f32(C.byref(g), fp(u), fp(v), fp(w), C.c_float(dt), order,
                            fp(px), fp(py), fp(pz), px.size)
    return px, py, pz


def advect_cc(res, origin, spacing, field, u, v, w, dt, order):
    g = grid(res, origin, spacing)
    field, u, v, w = map(f32, (field, u, v, w))
    out = np.empty_like(field)
    lib().orc_advect_cc_f32(C.byref(g), fp(field), fp(u), fp(v), fp(w), C.c_float(dt),
                            order, fp(out))
    return out


def advect_fc(res, origin, spacing, iu, iv, iw, u, v, w, dt, order):
    g = grid(res, origin, spacing)
    iu, iv, iw, u, v, w = map(f32, (iu, iv, iw, u, v, w))
    ou, ov, ow = np.empty_like(iu), np.empty_like(iv), np.empty_like(iw)
    lib().orc_advect_fc_f32(C.byref(g), fp(iu), fp(iv), fp(iw), fp(u), fp(v), fp(w),
                            C.c_float(dt), order, fp(ou), fp(ov), fp(ow))
    return ou, ov, ow


def external_forces(res, u, v, w, gravity, force, dt):
    g = grid(res)
    u, v, w = f32(u).copy(), f32(v).copy(), f32(w).copy()
    gr, fo = f32(gravity), f32(force)
    lib().orc_external_forces_f32(C.byref(g), fp(u), fp(v), fp(w), fp(gr), fp(fo),
                                  C.c_float(dt))
    return u, v, w


def divergence(res, origin, spacing, u, v, w):
    g = grid(res, origin, spacing)
    u, v, w = f32(u), f32(v), f32(w)
    out = np.empty(tuple(reversed(res)), np.float32)
    lib().orc_divergence_f32(C.byref(g), fp(u), fp(v), fp(w), fp(out))
    return out


def build_matrix(res, markers, scale):
    g = grid(res)
    markers, scale = f32(markers), f32(scale)
    A = np.empty(markers.size * 4, np.float32)
    lib().orc_build_matrix_f32(C.byref(g), fp(markers), fp(scale), fp(A))
    return A


def matvec(res, A, x):
    g = grid(res)
    A, x = f32(A), f32(x)
    y = np.empty_like(x)
    lib().orc_matvec_f32(C.byref(g), fp(A), fp(x), fp(y))
    return y


def cg_dense(A, b, minv=None, tol_mode=0, tol=1e-3, max_it=1000, x0=None):
    A, b = f32(A), f32(b)
    n = b.size
    x = np.zeros(n, np.float32) if x0 is None else f32(x0).copy()
    it = C.c_int64(0)
    mp = fp(f32(minv)) if minv is not None else None
    lib().orc_cg_dense_f32(fp(A), mp, fp(b), fp(x), n, tol_mode, tol, max_it, C.byref(it))
    return x, it.value


def extrapolate_grid(data, layers):
    """data shaped (nz, ny, nx); returns (values, layer ids, layers used)."""
    d = f32(data).copy()
    nz, ny, nx = d.shape
    ids = np.empty(d.shape, np.int32)
    used = lib().orc_extrapolate_grid_f32(fp(d), nx, ny, nz, layers,
                                          ids.ctypes.data_as(_I))
    return d, ids, used


def extrapolate_fc(res, u, v, w, layers):
    g = grid(res)
    u, v, w = f32(u).copy(), f32(v).copy(), f32(w).copy()
    used = lib().orc_extrapolate_fc_f32(C.byref(g), fp(u), fp(v), fp(w), layers)
    return u, v, w, used


PG = dict(p2gsum=0, g2pnearest=1, linear=2, quadratic=3, cubic=4)   # EPGTransferAlgorithm


def _field_extent(res, kind):
    nx, ny, nz = res
    return {0: nx * ny * nz, 1: (nx + 1) * ny * nz, 2: nx * (ny + 1) * nz, 3: nx * ny * (nz + 1)}[kind]


def p2g(res, origin, spacing, kind, pos, val, alg, span=1, off=0, normalize=True):
    """Particles to a cell-centred grid (kind 0) or to the x/y/z face grid
    (kind 1/2/3) of the grid res; returns (field, weight)."""
    g = grid(res, origin, spacing)
    n = _field_extent(res, kind)
    f, wt = np.empty(n, np.float32), np.empty(n, np.float32)
    pos = f32(pos).ravel()
    lib().orc_p2g_f32(C.byref(g), kind, pos.size // 3, fp(pos), fp(f32(val).ravel()), span, off, PG[alg], fp(f),
                      fp(wt), 1 if normalize else 0)
    return f, wt


def g2p(res, origin, spacing, kind, pos, field, alg, span=1, off=0, out=None):
    g = grid(res, origin, spacing)
    pos = f32(pos).ravel()
    np_ = pos.size // 3
    out = np.zeros(span * np_, np.float32) if out is None else f32(out).copy()
    lib().orc_g2p_f32(C.byref(g), kind, np_, fp(pos), fp(f32(field)), PG[alg], fp(out), span, off)
    return out


def particle_markers(res, origin, spacing, solid, pos, counts_only=False):
    g = grid(res, origin, spacing)
    pos = f32(pos).ravel()
    out = np.empty(int(np.prod(res)), np.float32)
    lib().orc_particle_markers_f32(C.byref(g), fp(f32(solid) if solid is not None else out), pos.size // 3,
                                   fp(pos), fp(out), 1 if counts_only else 0)
    return out


def advect_particles(res, origin, spacing, u, v, w, pos, dt, cfl, order, fix=False):
    g = grid(res, origin, spacing)
    pos = f32(pos).ravel().copy()
    lib().orc_advect_particles_f32(C.byref(g), fp(f32(u)), fp(f32(v)), fp(f32(w)), pos.size // 3, fp(pos),
                                   C.c_float(dt), C.c_float(cfl), order, 1 if fix else 0)
    return pos


def generate_particles(res, origin, spacing, fluid, solid, per_cell, seed):
    """CEulerParticles::generateParticlesInFluid (EulerParticles.cpp:163-201)
    with the seeded counter-based jitter; returns (n, 3) positions."""
    g = grid(res, origin, spacing)
    fluid = f32(fluid)
    solid_p = fp(f32(solid)) if solid is not None else None
    n = lib().orc_generate_particles(C.byref(g), fp(fluid), solid_p, per_cell, seed, None, 0)
    out = np.empty(3 * n, np.float32)
    lib().orc_generate_particles(C.byref(g), fp(fluid), solid_p, per_cell, seed, fp(out), n)
    return out.reshape(-1, 3)


PSOLVER = dict(pic=0, flip=1, mix=2)


class ParticleSolver:
    """CParticleInCell / CFluidImplicitParticle / CMixPICAndFLIP restated
    (ParticleInCell.cpp:58-106, FluidImplicitParticle.cpp:60-125,
    MixPICAndFLIP.cpp:64-134) on the CPU, FP32."""

    def __init__(self, res, origin, spacing, kind, pos, cfl=1.0, alg="linear", mixing=0.1):
        self.res = tuple(int(r) for r in res)
        self.g = grid(res, origin, spacing)
        pos = f32(pos).ravel()
        self.n = pos.size // 3
        self.h = lib().orc_psolver_create(C.byref(self.g), PSOLVER[kind], PG[alg], C.c_float(cfl),
                                          C.c_float(mixing), self.n, fp(pos))

    def close(self):
        if self.h:
            lib().orc_psolver_destroy(self.h)
            self.h = None

    __del__ = close

    def advect_velocity(self, inp, vel, dt, order, fix):
        nx, ny, nz = self.res
        out = [np.empty((nx + 1) * ny * nz, np.float32), np.empty(nx * (ny + 1) * nz, np.float32),
               np.empty(nx * ny * (nz + 1), np.float32)]
        a = [f32(x).ravel() for x in list(inp) + list(vel)]
        lib().orc_psolver_advect_velocity(self.h, *[fp(x) for x in a], C.c_float(dt), order, 1 if fix else 0,
                                          *[fp(x) for x in out])
        return out

    def advect_scalar(self, inp):
        out = np.empty(int(np.prod(self.res)), np.float32)
        lib().orc_psolver_advect_scalar(self.h, fp(f32(inp).ravel()), fp(out))
        return out

    def advect_vector(self, inp):
        n = int(np.prod(self.res))
        out = [np.empty(n, np.float32) for _ in range(3)]
        a = [f32(x).ravel() for x in inp]
        lib().orc_psolver_advect_vector(self.h, *[fp(x) for x in a], *[fp(x) for x in out])
        return out

    def get(self, which):
        count = {0: 3, 1: 3, 3: 1, 4: 3}[which] * self.n
        out = np.empty(count, np.float32)
        assert lib().orc_psolver_get(self.h, which, fp(out), count) == 0
        return out


def fluid_density(res, phi, solid):
    """buildFluidDensity (EulerSolverTool.cu:408-448) on host arrays."""
    g = grid(res)
    out = np.empty(int(np.prod(res)), np.float32)
    lib().orc_fluid_density_f32(C.byref(g), fp(f32(phi)), fp(f32(solid)), fp(out))
    return out


def fluid_sdf(res, phi, passes=1000):
    """buildFluidSDFInvoker (EulerSolverTool.cu:324-406, 760-785) on host arrays."""
    g = grid(res)
    out = np.empty(int(np.prod(res)), np.float32)
    lib().orc_fluid_sdf_f32(C.byref(g), fp(f32(phi)), fp(out), passes)
    return out


def pressure_solve(res, origin, spacing, dt, u, v, w, us, vs, ws, solid, fluid,
                   tol_mode=0, tol=1e-3, max_it=0, jacobi=False, density=1.0):
    """FP32 CPressureSolver::solvePressure; returns dict of outputs."""
    g = grid(res, origin, spacing)
    u, v, w = f32(u).copy(), f32(v).copy(), f32(w).copy()
    us, vs, ws, solid, fluid = map(f32, (us, vs, ws, solid, fluid))
    n = solid.size
    markers, A, b, x = (np.empty(n, np.float32), np.empty(4 * n, np.float32),
                        np.empty(n, np.float32), np.empty(n, np.float32))
    st = OrcStats()
    lib().orc_pressure_solve_f32(C.byref(g), C.c_float(dt), C.c_float(density), fp(u), fp(v),
                                 fp(w), fp(us), fp(vs), fp(ws), fp(solid), fp(fluid),
                                 tol_mode, tol, max_it, 1 if jacobi else 0, fp(markers),
                                 fp(A), fp(b), fp(x), C.byref(st))
    return dict(u=u, v=v, w=w, markers=markers, A=A, b=b, x=x, stats=st.as_dict())


def pressure_solve_f64(res, origin, spacing, dt, u, v, w, us, vs, ws, solid, fluid,
                       tol_mode=1, tol=1e-12, max_it=0, jacobi=False, density=1.0):
    g = grid(res, origin, spacing)
    u, v, w, us, vs, ws, solid, fluid = map(f32, (u, v, w, us, vs, ws, solid, fluid))
    uo, vo, wo = (np.empty(u.size, np.float64), np.empty(v.size, np.float64),
                  np.empty(w.size, np.float64))
    x = np.empty(solid.size, np.float64)
    st = OrcStats()
    lib().orc_pressure_solve_f64(C.byref(g), C.c_float(dt), C.c_float(density), fp(u), fp(v),
                                 fp(w), fp(us), fp(vs), fp(ws), fp(solid), fp(fluid),
                                 tol_mode, tol, max_it, 1 if jacobi else 0, dp(uo), dp(vo),
                                 dp(wo), dp(x), C.byref(st))
    return dict(u=uo, v=vo, w=wo, x=x, stats=st.as_dict())


def num_threads():
    return lib().orc_num_threads()


class Scene:
    """CGridFluidSolver restatement on the CPU (FP32, or FP64 with f64=True)."""

    def __init__(self, res, origin, spacing, f64=False):
        self.res = tuple(int(r) for r in res)
        self.g = grid(res, origin, spacing)
        self.h = lib().orc_scene_create(C.byref(self.g), 1 if f64 else 0)

    def close(self):
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    __del__ = close

    def count(self, field):
        nx, ny, nz = self.res
        n = nx * ny * nz
        return {0: (nx + 1) * ny * nz, 1: nx * (ny + 1) * nz, 2: nx * ny * (nz + 1),
                5: (nx + 1) * ny * nz, 6: nx * (ny + 1) * nz, 7: nx * ny * (nz + 1),
                11: 4 * n}.get(field, n)

    def add_solid(self, sdf, vel=(0, 0, 0)):
        sdf = f32(sdf).ravel()
        rc = lib().orc_scene_add_solid(self.h, fp(sdf), sdf.size, fp(f32(vel)))
        assert rc == 0

    def fluid_box(self, mn, mx):
        lib().orc_scene_fluid_box(self.h, fp(f32(mn)), fp(f32(mx)))

    def upload(self, field, data):
        data = f32(data).ravel()
        rc = lib().orc_scene_upload(self.h, field, fp(data), data.size)
        assert rc == 0, (field, data.size)

    def download(self, field, f64=False):
        n = self.count(field)
        if f64:
            out = np.empty(n, np.float64)
            rc = lib().orc_scene_download_f64(self.h, field, dp(out), n)
        else:
            out = np.empty(n, np.float32)
            rc = lib().orc_scene_download(self.h, field, fp(out), n)
        assert rc == 0, field
        return out

    def update_boundaries(self, p):
        lib().orc_scene_update_boundaries(self.h, C.byref(p))

    def advect(self, p):
        lib().orc_scene_advect(self.h, C.byref(p))

    def apply_forces(self, p):
        lib().orc_scene_apply_forces(self.h, C.byref(p))

    def pressure_solve(self, p):
        st = OrcStats()
        lib().orc_scene_pressure_solve(self.h, C.byref(p), C.byref(st))
        return st.as_dict()

    def extrapolate(self, p):
        return lib().orc_scene_extrapolate(self.h, C.byref(p))

    def solid_state(self, idx):
        t = np.empty(3, np.float32)
        n = self.count(3)
        sdf = np.empty(n, np.float32)
        rc = lib().orc_scene_solid_state(self.h, idx, fp(t), fp(sdf), n)
        assert rc == 0
        return t, sdf

    def step(self, p, n=1):
        st = OrcStats()
        lib().orc_scene_step(self.h, C.byref(p), n, C.byref(st))
        return st.as_dict()

    def step_particles(self, ps, p, n=1):
        """update() with a particle advection solver and particle markers."""
        st = OrcStats()
        assert lib().orc_scene_step_particles(self.h, ps.h, C.byref(p), n, C.byref(st)) == 0
        return st.as_dict()

    def pressure_solve_particles(self, ps, p):
        st = OrcStats()
        assert lib().orc_scene_pressure_solve_particles(self.h, ps.h, C.byref(p), C.byref(st)) == 0
        return st.as_dict()
