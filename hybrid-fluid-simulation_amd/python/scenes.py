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
    def __init__(self, n, dt, fluid_min, fluid_max, spheres, name):
        self.n = n
        self.res = (n + 2, n + 2, n + 2)
        self.h = np.float32(1.0) / np.float32(n)
        self.origin = (-float(self.h),) * 3
        self.spacing = (float(self.h),) * 3
        self.dt = dt
        self.fluid_min = fluid_min
        self.fluid_max = fluid_max
        self.spheres = spheres   # list of (cx, cy, cz, r)
        self.name = name

    @property
    def interior_cells(self):
        return self.n ** 3

    def _centres(self, z0, z1):
        g = self.res[0]
        h = np.float32(self.h)
        o = np.float32(self.origin[0])
        idx = np.arange(g, dtype=np.float32)
        c = (o + np.float32(0.5) * h) + h * idx        # cell centres, as the reference forms them
        cz = c[z0:z1][:, None, None]
        cy = c[None, :, None]
        cx = c[None, None, :]
        return cx, cy, cz

    def wall_sdf(self):
        """min distance to the walls of [0,1]^3: < 0 exactly on the ghost layer."""
        g = self.res[0]
        out = np.empty((g, g, g), np.float32)
        for z0 in range(0, g, 64):
            z1 = min(g, z0 + 64)
            cx, cy, cz = self._centres(z0, z1)
            cx = cx.astype(np.float64); cy = cy.astype(np.float64); cz = cz.astype(np.float64)
            d = np.minimum(np.minimum(np.minimum(cx, 1 - cx), np.minimum(cy, 1 - cy)),
                           np.minimum(cz, 1 - cz))
            out[z0:z1] = np.broadcast_to(d, (z1 - z0, g, g)).astype(np.float32)
        return out

    def sphere_sdf(self, k):
        cx0, cy0, cz0, r = self.spheres[k]
        g = self.res[0]
        out = np.empty((g, g, g), np.float32)
        for z0 in range(0, g, 64):
            z1 = min(g, z0 + 64)
            cx, cy, cz = self._centres(z0, z1)
            d = np.sqrt((cx.astype(np.float64) - cx0) ** 2 + (cy.astype(np.float64) - cy0) ** 2 +
                        (cz.astype(np.float64) - cz0) ** 2) - r
            out[z0:z1] = d.astype(np.float32)
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


def build_gpu_scene(mgf, ctx, spec, **desc_kw):
    d = mgf.make_desc(spec.res, spec.origin, spec.spacing, spec.dt, **desc_kw)
    s = mgf.Scene(ctx, d)
    for sdf in spec.solids():
        s.add_solid(sdf)
    s.fluid_box(spec.fluid_min, spec.fluid_max)
    return s


def build_oracle_scene(orc, spec, f64=False):
    s = orc.Scene(spec.res, spec.origin, spec.spacing, f64=f64)
    for sdf in spec.solids():
        s.add_solid(sdf)
    s.fluid_box(spec.fluid_min, spec.fluid_max)
    return s
