"""Synthetic scenes for the MAC-grid step (seeded, reproducible).

A scene is built the way CGridFluidSolver::resizeGridFluidSolver(ptree)
builds one (MPhysics/GridFluidSolver.cpp:48-73): n interior cells per axis
over [0,1], spacing 1/n, one ghost layer per side (origin = -h, resolution
n+2).  Solids are cell-centred SDFs (negative inside): the box walls, which
are negative exactly on the ghost layer, and analytic spheres.  The fluid is
an axis-aligned block given to generateFluidDomainFromBBox
(MPhysics/EulerSolverTool.cu:581-608).
"""
import numpy as np


class SceneSpec:
    """n interior cells per axis (an int, or (nx, ny, nz)) of spacing h over
    [0, L_x] x [0, L_y] x [0, L_z]; by default h = 1/n on a cubic [0,1]^3."""

    def __init__(self, n, dt, fluid_min, fluid_max, spheres, name, h=None):
        nxyz = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        self.nxyz = nxyz
        self.n = nxyz[0] if len(set(nxyz)) == 1 else None
        self.res = tuple(v + 2 for v in nxyz)
        if h is None:
            assert self.n is not None, "non-cubic scenes need an explicit spacing"
            self.h = np.float32(1.0) / np.float32(self.n)
            self.extent = (1.0, 1.0, 1.0)
        else:
            self.h = np.float32(h)
            self.extent = tuple(float(v) * float(self.h) for v in nxyz)
        self.origin = (-float(self.h),) * 3
        self.spacing = (float(self.h),) * 3
        self.dt = dt
        self.fluid_min = fluid_min
        self.fluid_max = fluid_max
        self.spheres = spheres   # list of (cx, cy, cz, r)
        self.name = name

    @property
    def interior_cells(self):
        return self.nxyz[0] * self.nxyz[1] * self.nxyz[2]

    @property
    def padded_cells(self):
        return self.res[0] * self.res[1] * self.res[2]

    def _axis_centres(self, k):
        h = np.float32(self.h)
        o = np.float32(self.origin[k])
        idx = np.arange(self.res[k], dtype=np.float32)
        return (o + np.float32(0.5) * h) + h * idx        # cell centres, as the reference forms them

    def _centres(self, z0, z1):
        cx = self._axis_centres(0)[None, None, :]
        cy = self._axis_centres(1)[None, :, None]
        cz = self._axis_centres(2)[z0:z1][:, None, None]
        return cx, cy, cz

    def _shape(self, z0, z1):
        return (z1 - z0, self.res[1], self.res[0])

    def wall_sdf(self):
        """min distance to the walls of the domain: < 0 exactly on the ghost layer."""
        gz = self.res[2]
        lx, ly, lz = self.extent
        out = np.empty((gz, self.res[1], self.res[0]), np.float32)
        for z0 in range(0, gz, 64):
            z1 = min(gz, z0 + 64)
            cx, cy, cz = self._centres(z0, z1)
            cx = cx.astype(np.float64); cy = cy.astype(np.float64); cz = cz.astype(np.float64)
            d = np.minimum(np.minimum(np.minimum(cx, lx - cx), np.minimum(cy, ly - cy)),
                           np.minimum(cz, lz - cz))
            out[z0:z1] = np.broadcast_to(d, self._shape(z0, z1)).astype(np.float32)
        return out

    def sphere_sdf(self, k):
        cx0, cy0, cz0, r = self.spheres[k]
        gz = self.res[2]
        out = np.empty((gz, self.res[1], self.res[0]), np.float32)
        for z0 in range(0, gz, 64):
            z1 = min(gz, z0 + 64)
            cx, cy, cz = self._centres(z0, z1)
            d = np.sqrt((cx.astype(np.float64) - cx0) ** 2 + (cy.astype(np.float64) - cy0) ** 2 +
                        (cz.astype(np.float64) - cz0) ** 2) - r
            out[z0:z1] = np.broadcast_to(d, self._shape(z0, z1)).astype(np.float32)
        return out

    def solids(self):
        yield self.wall_sdf()
        for k in range(len(self.spheres)):
            yield self.sphere_sdf(k)


def dam_break(n=64, dt=2e-3):
    """Plain single scene: fluid block x < 0.4, y < 0.6 (all z), closed walls."""
    return SceneSpec(n, dt, (0.0, 0.0, 0.0), (0.4, 0.6, 1.0), [], "dam_break_%d" % n)


def obstacles(n=256, seed=0, nspheres=3, dt=None):
    """Seeded fluid block of volume fraction f in [0.2,0.5] ([0,a)x[0,b)x[0,1),
    a ~ U[f,1], b = f/a) and `nspheres` spheres with centres in [0.15,0.85]^3
    and radii in [0.05,0.15]."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.2, 0.5)
    a = rng.uniform(f, 1.0)
    b = f / a
    spheres = []
    for _ in range(nspheres):
        c = rng.uniform(0.15, 0.85, 3)
        r = rng.uniform(0.05, 0.15)
        spheres.append((float(c[0]), float(c[1]), float(c[2]), float(r)))
    if dt is None:
        dt = 2e-3 * 64.0 / n
    return SceneSpec(n, dt, (0.0, 0.0, 0.0), (a, b, 1.0), spheres, "obstacles_%d_s%d" % (n, seed))


BATCH_SIZES = (32, 48, 64, 96, 128)


def ragged_scene(seed, sizes=BATCH_SIZES, max_spheres=3, h=1.0 / 64, dt=2e-3, log_uniform=None):
    """One scene of the batched and served workloads: nx, ny, nz drawn
    independently from `sizes` (or log-uniform in [lo, hi] when
    log_uniform=(lo, hi)), spacing h, a seeded fluid block [0,a)x[0,b)x[0,L_z)
    of volume fraction f in [0.2,0.5] of the domain, and 0..max_spheres spheres
    with centres in [0.15,0.85] of each extent and radii in [0.05,0.15] of the
    shortest extent."""
    rng = np.random.default_rng(seed)
    if log_uniform is None:
        nxyz = tuple(int(v) for v in rng.choice(sizes, 3))
    else:
        lo, hi = log_uniform
        nxyz = tuple(int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3))
    ext = [v * float(np.float32(h)) for v in nxyz]
    f = rng.uniform(0.2, 0.5)
    a = rng.uniform(f, 1.0)
    b = f / a
    spheres = []
    for _ in range(int(rng.integers(0, max_spheres + 1))):
        c = rng.uniform(0.15, 0.85, 3)
        r = rng.uniform(0.05, 0.15) * min(ext)
        spheres.append((float(c[0] * ext[0]), float(c[1] * ext[1]), float(c[2] * ext[2]), float(r)))
    return SceneSpec(nxyz, dt, (0.0, 0.0, 0.0), (a * ext[0], b * ext[1], ext[2]), spheres,
                     "ragged_%dx%dx%d_s%d" % (nxyz + (seed,)), h=h)


def build_gpu_scene(mgf, ctx, spec, **desc_kw):
    d = mgf.make_desc(spec.res, spec.origin, spec.spacing, spec.dt, **desc_kw)
    s = mgf.Scene(ctx, d)
    setup_gpu_scene(s, spec)
    return s


def setup_gpu_scene(s, spec):
    """Solids and fluid block of `spec` on an existing mgf.Scene."""
    for sdf in spec.solids():
        s.add_solid(sdf)
    s.fluid_box(spec.fluid_min, spec.fluid_max)


def build_oracle_scene(orc, spec, f64=False):
    s = orc.Scene(spec.res, spec.origin, spec.spacing, f64=f64)
    for sdf in spec.solids():
        s.add_solid(sdf)
    s.fluid_box(spec.fluid_min, spec.fluid_max)
    return s
