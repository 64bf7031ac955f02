// mac_oracle.hpp — CPU restatement of the Eulerian MAC-grid step of
// junwei-jiang/Hybrid-Fluid-Simulation.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
// smoke check and the CPU baseline leg of the benchmark call it.  The product
// (the HIP library) never includes, links or calls anything here.
//
// Templated on Real: Real=float is the plain FP32 transliteration of the
// reference, Real=double the FP64 oracle (the reference's own DOUBLE_REAL
// switch, MPhysics/MMathExtra.h:35-59).  Build with -ffp-contract=off: every
// expression is evaluated in the order written, one rounding per operation,
// except the BLAS axpy, which is written as an explicit fused multiply-add.
//
// Pinned by: tests/test_oracle_golden.py (reference golden values, closed
// forms, brute-force dense solves, invariants).  Nothing here is "parity
// unpinned" except where a function's comment says so.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace mac_oracle {

template <class R> struct Limits;
template <> struct Limits<float> {
    static constexpr float unknown() { return FLT_MAX; }   // MMathExtra.h:58
    static constexpr float epsilon() { return FLT_EPSILON; } // MMathExtra.h:55
};
template <> struct Limits<double> {
    static constexpr double unknown() { return DBL_MAX; }  // MMathExtra.h:46
    static constexpr double epsilon() { return DBL_EPSILON; } // MMathExtra.h:43
};

// FluidDomainValue (MPhysics/Common.h:245).
constexpr double kFluidDomainValue = 500.0;

template <class R> struct Grid {
    int nx = 0, ny = 0, nz = 0;
    R ox = 0, oy = 0, oz = 0;
    R hx = 1, hy = 1, hz = 1;
    size_t cells() const { return size_t(nx) * ny * nz; }
};

// ---------------------------------------------------------------------------
// Index helpers.
// limitIndex (MPhysics/EulerMathTool.cu:16-27).
inline int limit_index(int i, int n) {
    if (i < 0) return 0;
    if (i > n - 1) return n - 1;
    return i;
}
// transIndex2LinearWithOffset (MPhysics/EulerMathTool.cu:63-73): each axis
// clamped to the grid, x fastest.
inline size_t lin(int x, int y, int z, int nx, int ny, int nz) {
    x = limit_index(x, nx);
    y = limit_index(y, ny);
    z = limit_index(z, nz);
    return (size_t(z) * ny + y) * nx + x;
}
// floorCUDA (MPhysics/EulerMathTool.cu:93-103): truncation toward zero, minus
// one for negative input (so -1.0 maps to -2).  The float->int conversion is
// saturated to +-2^24 and NaN maps to 0 before the branch; the reference's
// conversion is undefined there, and any such point clamps to the grid edge.
template <class R> inline int floor_cuda(R v) {
    int t;
    if (!(v == v)) t = 0;
    else if (v >= R(16777216)) t = 16777216;
    else if (v <= R(-16777216)) t = -16777216;
    else t = int(v);
    if (v >= 0) return t;
    return t - 1;
}

// ---------------------------------------------------------------------------
// Trilinear sample of a cell-centred scalar grid with resolution (nx,ny,nz),
// origin o and spacing h at point p.
// sampleCellCenteredScalarFieldTrilerp (MPhysics/FieldMathTool.cu:52-104):
//   rel = p - o - h/2; i = floorCUDA(rel/h); t = (rel - i*h)/h; 8 clamped
//   corners; the 8-term sum in the order written there.
template <class R>
inline R sample_trilinear(const R* f, int nx, int ny, int nz, R ox, R oy, R oz,
                          R hx, R hy, R hz, R px, R py, R pz) {
    R rx = px - ox - hx / 2;
    R ry = py - oy - hy / 2;
    R rz = pz - oz - hz / 2;
    R tx = rx / hx, ty = ry / hy, tz = rz / hz;
    int ix = floor_cuda(tx), iy = floor_cuda(ty), iz = floor_cuda(tz);
    R fx = (rx - R(ix) * hx) / hx;
    R fy = (ry - R(iy) * hy) / hy;
    R fz = (rz - R(iz) * hz) / hz;
    R ubl = f[lin(ix, iy + 1, iz, nx, ny, nz)];
    R ubr = f[lin(ix + 1, iy + 1, iz, nx, ny, nz)];
    R ufl = f[lin(ix, iy + 1, iz + 1, nx, ny, nz)];
    R ufr = f[lin(ix + 1, iy + 1, iz + 1, nx, ny, nz)];
    R dbl = f[lin(ix, iy, iz, nx, ny, nz)];
    R dbr = f[lin(ix + 1, iy, iz, nx, ny, nz)];
    R dfl = f[lin(ix, iy, iz + 1, nx, ny, nz)];
    R dfr = f[lin(ix + 1, iy, iz + 1, nx, ny, nz)];
    return dbl * (1 - fx) * (1 - fy) * (1 - fz) +
           dbr * fx * (1 - fy) * (1 - fz) +
           ubl * (1 - fx) * fy * (1 - fz) +
           ubr * fx * fy * (1 - fz) +
           dfl * (1 - fx) * (1 - fy) * fz +
           dfr * fx * (1 - fy) * fz +
           ufl * (1 - fx) * fy * fz +
           ufr * fx * fy * fz;
}

template <class R>
inline R sample_cc(const Grid<R>& g, const R* f, R px, R py, R pz) {
    return sample_trilinear(f, g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g.hx, g.hy,
                            g.hz, px, py, pz);
}

// Face-centred sample: each component is a cell-centred grid with one extra
// layer on its axis and the origin shifted by -h/2 on that axis
// (sampleFaceCenteredVectorFieldInvoker, MPhysics/FieldMathTool.cu:783-892).
template <class R>
inline void sample_fc(const Grid<R>& g, const R* u, const R* v, const R* w,
                      R px, R py, R pz, R& ou, R& ov, R& ow) {
    ou = sample_trilinear(u, g.nx + 1, g.ny, g.nz, g.ox - g.hx / 2, g.oy, g.oz,
                          g.hx, g.hy, g.hz, px, py, pz);
    ov = sample_trilinear(v, g.nx, g.ny + 1, g.nz, g.ox, g.oy - g.hy / 2, g.oz,
                          g.hx, g.hy, g.hz, px, py, pz);
    ow = sample_trilinear(w, g.nx, g.ny, g.nz + 1, g.ox, g.oy, g.oz - g.hz / 2,
                          g.hx, g.hy, g.hz, px, py, pz);
}

// BLAS axpy y += a*x with the early return of plusAlphaX when |a| < EPSILON
// (MPhysics/CellCenteredVectorGrid.cpp plusAlphaX, MPhysics/CuDenseVector.cpp:
// 176-191).  The element update is one fused multiply-add; cuBLAS's rounding
// is closed, so this choice is ours and the kernels make the same one.
template <class R> inline bool axpy_skipped(R a) {
    return std::abs(a) < Limits<R>::epsilon();
}
template <class R> inline R fma_r(R a, R x, R y) { return std::fma(a, x, y); }

// CSemiLagrangian::backTrace for one point (MPhysics/SemiLagrangian.cpp:
// 188-347): stage velocities sampled from the face-centred field, positions
// moved with the axpy above.
template <class R>
inline void backtrace_point(const Grid<R>& vg, const R* u, const R* v,
                            const R* w, R dt, int order, R& px, R& py, R& pz) {
    const R x0 = px, y0 = py, z0 = pz;
    R v1x, v1y, v1z;
    sample_fc(vg, u, v, w, x0, y0, z0, v1x, v1y, v1z);
    if (order == 1) {
        R a = -dt;
        if (!axpy_skipped(a)) {
            px = fma_r(a, v1x, x0); py = fma_r(a, v1y, y0); pz = fma_r(a, v1z, z0);
        }
    } else if (order == 2) {
        R a = R(-0.5) * dt;
        R mx = x0, my = y0, mz = z0;
        if (!axpy_skipped(a)) {
            mx = fma_r(a, v1x, x0); my = fma_r(a, v1y, y0); mz = fma_r(a, v1z, z0);
        }
        R v2x, v2y, v2z;
        sample_fc(vg, u, v, w, mx, my, mz, v2x, v2y, v2z);
        R b = -dt;
        if (!axpy_skipped(b)) {
            px = fma_r(b, v2x, x0); py = fma_r(b, v2y, y0); pz = fma_r(b, v2z, z0);
        }
    } else if (order == 3) {
        R a = R(-0.5) * dt;
        R mx = x0, my = y0, mz = z0;
        if (!axpy_skipped(a)) {
            mx = fma_r(a, v1x, x0); my = fma_r(a, v1y, y0); mz = fma_r(a, v1z, z0);
        }
        R v2x, v2y, v2z;
        sample_fc(vg, u, v, w, mx, my, mz, v2x, v2y, v2z);
        R c = R(-0.75) * dt;
        R tx = x0, ty = y0, tz = z0;
        if (!axpy_skipped(c)) {
            tx = fma_r(c, v2x, x0); ty = fma_r(c, v2y, y0); tz = fma_r(c, v2z, z0);
        }
        R v3x, v3y, v3z;
        sample_fc(vg, u, v, w, tx, ty, tz, v3x, v3y, v3z);
        R a1 = -R(2.0 / 9.0) * dt;
        R a2 = -R(3.0 / 9.0) * dt;
        R a3 = -R(4.0 / 9.0) * dt;
        R ox = x0, oy = y0, oz = z0;
        if (!axpy_skipped(a1)) { ox = fma_r(a1, v1x, ox); oy = fma_r(a1, v1y, oy); oz = fma_r(a1, v1z, oz); }
        if (!axpy_skipped(a2)) { ox = fma_r(a2, v2x, ox); oy = fma_r(a2, v2y, oy); oz = fma_r(a2, v2z, oz); }
        if (!axpy_skipped(a3)) { ox = fma_r(a3, v3x, ox); oy = fma_r(a3, v3y, oy); oz = fma_r(a3, v3z, oz); }
        px = ox; py = oy; pz = oz;
    }
}

// Face and cell positions (MPhysics/SemiLagrangian.cpp:47-97):
//   cell centre (k+0.5)*h + o; X-face (k*h + o, (j+0.5)*h + o, (i+0.5)*h + o).
template <class R> inline R cell_pos(int k, R h, R o) { return (k + R(0.5)) * h + o; }
template <class R> inline R face_pos(int k, R h, R o) { return R(k) * h + o; }

// CSemiLagrangian::advect(CCS) (MPhysics/SemiLagrangian.cpp:135-150):
// back-trace every cell centre of `g` through the velocity on `vg`, sample the
// input cell-centred field (grid `g`) there.
template <class R>
void advect_cc(const Grid<R>& g, const Grid<R>& vg, const R* in, const R* u,
               const R* v, const R* w, R dt, int order, R* out) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                R px = cell_pos(x, g.hx, g.ox), py = cell_pos(y, g.hy, g.oy),
                  pz = cell_pos(z, g.hz, g.oz);
                backtrace_point(vg, u, v, w, dt, order, px, py, pz);
                out[(size_t(z) * g.ny + y) * g.nx + x] = sample_cc(g, in, px, py, pz);
            }
}

// CSemiLagrangian::advect(FCV) (MPhysics/SemiLagrangian.cpp:169-186): every
// face of each component is back-traced from its own position, then the input
// face field is sampled at the three position fields.
template <class R>
void advect_fc(const Grid<R>& g, const Grid<R>& vg, const R* iu, const R* iv,
               const R* iw, const R* u, const R* v, const R* w, R dt, int order,
               R* ou, R* ov, R* ow) {
    for (int c = 0; c < 3; ++c) {
        const int ex = g.nx + (c == 0), ey = g.ny + (c == 1), ez = g.nz + (c == 2);
        R* out = c == 0 ? ou : (c == 1 ? ov : ow);
#pragma omp parallel for schedule(static)
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    R px = c == 0 ? face_pos(x, g.hx, g.ox) : cell_pos(x, g.hx, g.ox);
                    R py = c == 1 ? face_pos(y, g.hy, g.oy) : cell_pos(y, g.hy, g.oy);
                    R pz = c == 2 ? face_pos(z, g.hz, g.oz) : cell_pos(z, g.hz, g.oz);
                    backtrace_point(vg, u, v, w, dt, order, px, py, pz);
                    R su, sv, sw;
                    // FaceCenteredVectorField::sampleField with three position
                    // fields samples component c at its own positions.
                    if (c == 0)
                        su = sample_trilinear(iu, g.nx + 1, g.ny, g.nz, g.ox - g.hx / 2, g.oy, g.oz, g.hx, g.hy, g.hz, px, py, pz), sv = sw = 0;
                    else if (c == 1)
                        sv = sample_trilinear(iv, g.nx, g.ny + 1, g.nz, g.ox, g.oy - g.hy / 2, g.oz, g.hx, g.hy, g.hz, px, py, pz), su = sw = 0;
                    else
                        sw = sample_trilinear(iw, g.nx, g.ny, g.nz + 1, g.ox, g.oy, g.oz - g.hz / 2, g.hx, g.hy, g.hz, px, py, pz), su = sv = 0;
                    out[(size_t(z) * ey + y) * ex + x] = c == 0 ? su : (c == 1 ? sv : sw);
                }
    }
}

// CExternalForcesSolver::applyExternalForces (MPhysics/ExternalForcesSolver.
// cpp:50-59) -> CFaceCenteredVectorGrid::plusAlphaX(unit field, Vector3)
// (MPhysics/FaceCenteredVectorGrid.cpp:162-175): skipped when every component
// of the increment is below EPSILON, else every face of every component gets
// u + a*1 (a zero increment still runs, turning -0 into +0).
template <class R>
void add_uniform(R* data, size_t n, R a) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) data[i] = fma_r(a, R(1), data[i]);
}
template <class R>
void apply_external_forces(const Grid<R>& g, R* u, R* v, R* w, const R grav[3],
                           const R force[3], R dt) {
    const size_t nu = size_t(g.nx + 1) * g.ny * g.nz;
    const size_t nv = size_t(g.nx) * (g.ny + 1) * g.nz;
    const size_t nw = size_t(g.nx) * g.ny * (g.nz + 1);
    const R* inc[2] = {grav, force};
    for (int k = 0; k < 2; ++k) {
        R ax = inc[k][0] * dt, ay = inc[k][1] * dt, az = inc[k][2] * dt;
        if (axpy_skipped(ax) && axpy_skipped(ay) && axpy_skipped(az)) continue;
        add_uniform(u, nu, ax);
        add_uniform(v, nv, ay);
        add_uniform(w, nw, az);
    }
}

// ---------------------------------------------------------------------------
// Boundaries (static solids are re-sampled each step exactly like moving ones).
// transformField (MPhysics/BoundaryTool.cu:85-113): cell centre
// (o + 0.5*h) + h*i, plus the solid's translation.
// CBoundarys::__resamplingCurrentBoundarysSDF (Boundary.cpp:129-136):
// trilinear sample of the original SDF there.
template <class R>
void resample_solid(const Grid<R>& g, const R* sdf0, const R t[3], R* out) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                R cx = g.ox + R(0.5) * g.hx + g.hx * R(x);
                R cy = g.oy + R(0.5) * g.hy + g.hy * R(y);
                R cz = g.oz + R(0.5) * g.hz + g.hz * R(z);
                out[(size_t(z) * g.ny + y) * g.nx + x] =
                    sample_cc(g, sdf0, cx + t[0], cy + t[1], cz + t[2]);
            }
}
// buildSolidsMarkerInvoker / buildSolidsMarker (MPhysics/BoundaryTool.cu:3-26,
// 115-144): zero-filled, then solid i (1-based) sets UNKNOWN everywhere when
// i == 1 and writes -i where its SDF < 0 and no earlier solid claimed the cell.
template <class R>
void build_solids_marker(size_t n, const std::vector<const R*>& sdfs, R* marker) {
    for (size_t c = 0; c < n; ++c) marker[c] = 0;
    for (size_t s = 0; s < sdfs.size(); ++s) {
        const int idx = int(s) + 1;
        for (size_t c = 0; c < n; ++c) {
            if (idx == 1) marker[c] = Limits<R>::unknown();
            if (sdfs[s][c] < 0 && marker[c] == Limits<R>::unknown()) marker[c] = R(-idx);
        }
    }
}
// buildSolidsVel (MPhysics/BoundaryTool.cu:28-83, 146-184): zero-filled; each
// cell with marker < 0 writes its solid's velocity on its six faces.  The
// reference writes with racing atomicExch; here cells are visited in linear
// order and the last write wins, i.e. a face takes the velocity of its
// higher-index solid cell, else of its lower-index solid cell.  Faces between
// two different solids are the only ones this choice affects.
template <class R>
void build_solids_vel(const Grid<R>& g, const R* marker, const R* vel, int nsolids,
                      R* us, R* vs, R* ws) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    std::fill(us, us + size_t(nx + 1) * ny * nz, R(0));
    std::fill(vs, vs + size_t(nx) * (ny + 1) * nz, R(0));
    std::fill(ws, ws + size_t(nx) * ny * (nz + 1), R(0));
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                R m = marker[(size_t(z) * ny + y) * nx + x];
                if (!(m < 0)) continue;
                int s = int(-(1 + m));
                if (s >= nsolids) continue;
                us[lin(x, y, z, nx + 1, ny, nz)] = vel[3 * s];
                us[lin(x + 1, y, z, nx + 1, ny, nz)] = vel[3 * s];
                vs[lin(x, y, z, nx, ny + 1, nz)] = vel[3 * s + 1];
                vs[lin(x, y + 1, z, nx, ny + 1, nz)] = vel[3 * s + 1];
                ws[lin(x, y, z, nx, ny, nz + 1)] = vel[3 * s + 2];
                ws[lin(x, y, z + 1, nx, ny, nz + 1)] = vel[3 * s + 2];
            }
}

// generateFluidDomainFromBBox (MPhysics/EulerSolverTool.cu:581-608): cell
// centre (o + 0.5*h) + h*i strictly inside (min,max) -> -500, else +500.
template <class R>
void fluid_domain_box(const Grid<R>& g, const R mn[3], const R mx[3], R* phi) {
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                R cx = g.ox + R(0.5) * g.hx + g.hx * R(x);
                R cy = g.oy + R(0.5) * g.hy + g.hy * R(y);
                R cz = g.oz + R(0.5) * g.hz + g.hz * R(z);
                bool in = cx > mn[0] && cy > mn[1] && cz > mn[2] && cx < mx[0] &&
                          cy < mx[1] && cz < mx[2];
                phi[(size_t(z) * g.ny + y) * g.nx + x] =
                    in ? R(-kFluidDomainValue) : R(kFluidDomainValue);
            }
}

// ---------------------------------------------------------------------------
// Pressure system.
// buildFluidMarkers (MPhysics/EulerSolverTool.cu:30-56): 2 solid, 1 fluid, 0 air.
template <class R>
void build_fluid_markers(size_t n, const R* solid, const R* fluid, R* marker) {
    for (size_t c = 0; c < n; ++c)
        marker[c] = solid[c] < 0 ? R(2) : (fluid[c] < 0 ? R(1) : R(0));
}

// buildPressureMatrixA (MPhysics/EulerSolverTool.cu:58-126): 4 values per cell
// [diag, +x, +y, +z]; scale = dt/(rho*h^2) per axis (PressureSolver.cpp:148-152).
template <class R>
void build_matrix(const Grid<R>& g, const R* m, const R s[3], R* A) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = lin(x, y, z, nx, ny, nz);
                R d = 0, ax = 0, ay = 0, az = 0;
                if (m[c] == 1) {
                    if (x > 0 && m[lin(x - 1, y, z, nx, ny, nz)] != 2) d += s[0];
                    if (x + 1 < nx && m[lin(x + 1, y, z, nx, ny, nz)] != 2) {
                        d += s[0];
                        if (m[lin(x + 1, y, z, nx, ny, nz)] == 1) ax -= s[0];
                    }
                    if (y > 0 && m[lin(x, y - 1, z, nx, ny, nz)] != 2) d += s[1];
                    if (y + 1 < ny && m[lin(x, y + 1, z, nx, ny, nz)] != 2) {
                        d += s[1];
                        if (m[lin(x, y + 1, z, nx, ny, nz)] == 1) ay -= s[1];
                    }
                    if (z > 0 && m[lin(x, y, z - 1, nx, ny, nz)] != 2) d += s[2];
                    if (z + 1 < nz && m[lin(x, y, z + 1, nx, ny, nz)] != 2) {
                        d += s[2];
                        if (m[lin(x, y, z + 1, nx, ny, nz)] == 1) az -= s[2];
                    }
                } else {
                    d = 1;
                }
                A[4 * c] = d; A[4 * c + 1] = ax; A[4 * c + 2] = ay; A[4 * c + 3] = az;
            }
}

// CUDADivergence for a face-centred field (MPhysics/FieldMathTool.cu:370-395):
// (R-L)/hx + (U-D)/hy + (F-B)/hz, no 1/2.
template <class R>
void divergence_fc(const Grid<R>& g, const R* u, const R* v, const R* w, R* div) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                R l = u[lin(x, y, z, nx + 1, ny, nz)], r = u[lin(x + 1, y, z, nx + 1, ny, nz)];
                R d = v[lin(x, y, z, nx, ny + 1, nz)], up = v[lin(x, y + 1, z, nx, ny + 1, nz)];
                R b = w[lin(x, y, z, nx, ny, nz + 1)], f = w[lin(x, y, z + 1, nx, ny, nz + 1)];
                div[lin(x, y, z, nx, ny, nz)] = (r - l) / g.hx + (up - d) / g.hy + (f - b) / g.hz;
            }
}

// buildPressureVectorb (MPhysics/EulerSolverTool.cu:128-197) with
// scale = 1/h (EulerSolverTool.cu:686): fluid cells b = -div, then each
// clamped neighbour that is solid subtracts/adds scale*(u_face - us_face).
template <class R>
void build_rhs(const Grid<R>& g, const R* u, const R* v, const R* w, const R* div,
               const R* m, const R* us, const R* vs, const R* ws, R* b) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const R sx = R(1) / g.hx, sy = R(1) / g.hy, sz = R(1) / g.hz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = lin(x, y, z, nx, ny, nz);
                if (m[c] != 1) { b[c] = 0; continue; }
                R val = -div[c];
                size_t fl = lin(x, y, z, nx + 1, ny, nz), fr = lin(x + 1, y, z, nx + 1, ny, nz);
                size_t fd = lin(x, y, z, nx, ny + 1, nz), fu = lin(x, y + 1, z, nx, ny + 1, nz);
                size_t fb = lin(x, y, z, nx, ny, nz + 1), ff = lin(x, y, z + 1, nx, ny, nz + 1);
                if (m[lin(x - 1, y, z, nx, ny, nz)] == 2) val -= sx * (u[fl] - us[fl]);
                if (m[lin(x + 1, y, z, nx, ny, nz)] == 2) val += sx * (u[fr] - us[fr]);
                if (m[lin(x, y - 1, z, nx, ny, nz)] == 2) val -= sy * (v[fd] - vs[fd]);
                if (m[lin(x, y + 1, z, nx, ny, nz)] == 2) val += sy * (v[fu] - vs[fu]);
                if (m[lin(x, y, z - 1, nx, ny, nz)] == 2) val -= sz * (w[fb] - ws[fb]);
                if (m[lin(x, y, z + 1, nx, ny, nz)] == 2) val += sz * (w[ff] - ws[ff]);
                b[c] = val;
            }
}

// fdmMatrixVectorMul (MPhysics/EulerSolverTool.cu:199-227): each product is
// formed in Real; the ternaries' 0.0 literal makes the sum double, added in
// the order centre, -x, +x, -y, +y, -z, +z, then stored as Real.
template <class R>
void matvec(const Grid<R>& g, const R* A, const R* xv, R* y) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int yy = 0; yy < ny; ++yy)
            for (int x = 0; x < nx; ++x) {
                size_t c = lin(x, yy, z, nx, ny, nz);
                size_t l = lin(x - 1, yy, z, nx, ny, nz), r = lin(x + 1, yy, z, nx, ny, nz);
                size_t d = lin(x, yy - 1, z, nx, ny, nz), u = lin(x, yy + 1, z, nx, ny, nz);
                size_t b = lin(x, yy, z - 1, nx, ny, nz), f = lin(x, yy, z + 1, nx, ny, nz);
                double acc = double(A[4 * c] * xv[c]);
                acc = acc + ((x > 0) ? double(A[4 * l + 1] * xv[l]) : 0.0);
                acc = acc + ((x + 1 < nx) ? double(A[4 * c + 1] * xv[r]) : 0.0);
                acc = acc + ((yy > 0) ? double(A[4 * d + 2] * xv[d]) : 0.0);
                acc = acc + ((yy + 1 < ny) ? double(A[4 * c + 2] * xv[u]) : 0.0);
                acc = acc + ((z > 0) ? double(A[4 * b + 3] * xv[b]) : 0.0);
                acc = acc + ((z + 1 < nz) ? double(A[4 * c + 3] * xv[f]) : 0.0);
                y[c] = R(acc);
            }
}

// applyPressureGradient (MPhysics/EulerSolverTool.cu:229-322) with
// scale = dt/(rho*h) (PressureSolver.cpp:160-164): each cell owns its -x/-y/-z
// face; max-index faces become UNKNOWN.
template <class R>
void apply_pressure_gradient(const Grid<R>& g, const R s[3], const R* m, const R* p,
                             R* u, R* v, R* w, const R* us, const R* vs, const R* ws) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const R unk = Limits<R>::unknown();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = lin(x, y, z, nx, ny, nz);
                size_t l = lin(x - 1, y, z, nx, ny, nz), d = lin(x, y - 1, z, nx, ny, nz),
                       b = lin(x, y, z - 1, nx, ny, nz);
                size_t fu = lin(x, y, z, nx + 1, ny, nz), fv = lin(x, y, z, nx, ny + 1, nz),
                       fw = lin(x, y, z, nx, ny, nz + 1);
                if (m[l] == 1 || m[c] == 1) {
                    if (m[l] == 2 || m[c] == 2) u[fu] = us[fu];
                    else u[fu] -= s[0] * (p[c] - p[l]);
                } else u[fu] = unk;
                if (m[d] == 1 || m[c] == 1) {
                    if (m[d] == 2 || m[c] == 2) v[fv] = vs[fv];
                    else v[fv] -= s[1] * (p[c] - p[d]);
                } else v[fv] = unk;
                if (m[b] == 1 || m[c] == 1) {
                    if (m[b] == 2 || m[c] == 2) w[fw] = ws[fw];
                    else w[fw] -= s[2] * (p[c] - p[b]);
                } else w[fw] = unk;
                if (x == nx - 1) u[lin(x + 1, y, z, nx + 1, ny, nz)] = unk;
                if (y == ny - 1) v[lin(x, y + 1, z, nx, ny + 1, nz)] = unk;
                if (z == nz - 1) w[lin(x, y, z + 1, nx, ny, nz + 1)] = unk;
            }
}

// ---------------------------------------------------------------------------
// BLAS-1 on Real vectors.  Dot products and norms are accumulated in double
// over 256 fixed index ranges combined in order (deterministic for any thread
// count) and rounded to Real: cuBLAS's reduction order is closed, so the
// rounding of the reference's dots is unpinned; these are correctly ordered
// approximations of the same quantities.
template <class R> double dot_d(const R* a, const R* b, size_t n) {
    const int kChunks = 256;
    double part[kChunks];
#pragma omp parallel for schedule(static)
    for (int ch = 0; ch < kChunks; ++ch) {
        size_t lo = n * ch / kChunks, hi = n * (ch + 1) / kChunks;
        double s = 0;
        for (size_t i = lo; i < hi; ++i) s += double(a[i]) * double(b[i]);
        part[ch] = s;
    }
    double s = 0;
    for (int ch = 0; ch < kChunks; ++ch) s += part[ch];
    return s;
}
template <class R> R dot(const R* a, const R* b, size_t n) { return R(dot_d(a, b, n)); }
template <class R> R nrm2(const R* a, size_t n) { return R(std::sqrt(dot_d(a, a, n))); }

struct CgResult {
    long long iterations = 0;
    bool converged = false;
    double r_norm = 0;   // final recurrence ||r||
    double b_norm = 0;
};

// CCuMatrixFreePCG::solvePCGInvDia (MPhysics/CuMatrixFreePCG.cpp:22-65) with
// the identity preconditioner of CPressureSolver (PressureSolver.cpp:12-15)
// or a Jacobi one.  tol_mode 0: sqrt(||r||) < tol (reference); 1: ||r|| <=
// tol*||b||.  x is used as given (the pressure solver zeroes it first,
// PressureSolver.cpp:145).
template <class R>
CgResult cg_solve(const Grid<R>& g, const R* A, const R* b, R* x, int tol_mode,
                  double tol, long long max_it, bool jacobi) {
    const size_t n = g.cells();
    std::vector<R> r(n), z(n), p(n), ap(n);
    CgResult res;
    std::copy(b, b + n, r.begin());                          // m_R = b
    R bn = nrm2(r.data(), n);
    res.b_norm = bn;
    auto stop = [&](R rn) {
        if (tol_mode == 0) return std::sqrt(rn) < R(tol);
        return double(rn) <= tol * double(bn);
    };
    if (tol_mode == 0 ? std::sqrt(bn) < R(tol) : bn == 0) {
        std::fill(x, x + n, R(0));
        res.converged = true;
        return res;
    }
    matvec(g, A, x, r.data());                               // m_R = A x
    for (size_t i = 0; i < n; ++i) r[i] = R(-1.0) * r[i];    // scale(-1)
    for (size_t i = 0; i < n; ++i) r[i] = fma_r(R(1), b[i], r[i]); // += b
    auto precond = [&]() {
        if (!jacobi) { std::copy(r.begin(), r.end(), z.begin()); return; }
        for (size_t i = 0; i < n; ++i) z[i] = r[i] / A[4 * i];
    };
    precond();
    std::copy(z.begin(), z.end(), p.begin());
    for (long long it = 0; it < max_it; ++it) {
        R rz = dot(r.data(), z.data(), n);
        matvec(g, A, p.data(), ap.data());
        res.iterations++;
        R pap = dot(p.data(), ap.data(), n);
        R alpha = rz / pap;
        if (!axpy_skipped(alpha)) {
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)n; ++i) x[i] = fma_r(alpha, p[i], x[i]);
            R na = -alpha;
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)n; ++i) r[i] = fma_r(na, ap[i], r[i]);
        }
        R rn = nrm2(r.data(), n);
        res.r_norm = rn;
        if (stop(rn)) { res.converged = true; break; }
        precond();
        R rz2 = dot(r.data(), z.data(), n);
        R beta = rz2 / rz;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) p[i] = beta * p[i];
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) p[i] = fma_r(R(1), z[i], p[i]);
    }
    return res;
}

// ---------------------------------------------------------------------------
// Extrapolation (MPhysics/EulerSolverTool.cu:450-557, 871-969) on one grid of
// extents (nx,ny,nz): layer 0 = entries != UNKNOWN; layer i (< layers) =
// UNKNOWN entries with a clamped neighbour in layer i-1; then, in layer order,
// each layer-i entry becomes the mean of its clamped neighbours in layers < i
// (sum in the order -x,+x,-y,+y,-z,+z starting from 0, one divide).  Layers
// are found breadth-first, which visits the same sets the reference's full
// sweeps do.  Returns the number of non-empty layers >= 1.
template <class R>
int extrapolate_grid(R* data, int nx, int ny, int nz, int layers,
                     std::vector<int32_t>* layer_out = nullptr) {
    const size_t n = size_t(nx) * ny * nz;
    const int32_t kUnk = std::numeric_limits<int32_t>::max();
    std::vector<int32_t> L(n, kUnk);
    std::vector<size_t> frontier, next;
    for (size_t c = 0; c < n; ++c)
        if (data[c] != Limits<R>::unknown()) L[c] = 0;
    if (layers <= 0) {
        if (layer_out) *layer_out = L;
        return 0;
    }
    for (size_t c = 0; c < n; ++c) if (L[c] == 0) frontier.push_back(c);
    int used = 0;
    for (int i = 1; i < layers && !frontier.empty(); ++i) {
        next.clear();
        for (size_t c : frontier) {
            int x = int(c % nx), y = int((c / nx) % ny), z = int(c / (size_t(nx) * ny));
            size_t nb[6] = {lin(x - 1, y, z, nx, ny, nz), lin(x + 1, y, z, nx, ny, nz),
                            lin(x, y - 1, z, nx, ny, nz), lin(x, y + 1, z, nx, ny, nz),
                            lin(x, y, z - 1, nx, ny, nz), lin(x, y, z + 1, nx, ny, nz)};
            for (size_t q : nb)
                if (L[q] == kUnk) { L[q] = i; next.push_back(q); }
        }
        if (next.empty()) break;
        used = i;
        // Values of layer i in index order (independent of visiting order).
        std::sort(next.begin(), next.end());
        for (size_t c : next) {
            int x = int(c % nx), y = int((c / nx) % ny), z = int(c / (size_t(nx) * ny));
            size_t nb[6] = {lin(x - 1, y, z, nx, ny, nz), lin(x + 1, y, z, nx, ny, nz),
                            lin(x, y - 1, z, nx, ny, nz), lin(x, y + 1, z, nx, ny, nz),
                            lin(x, y, z - 1, nx, ny, nz), lin(x, y, z + 1, nx, ny, nz)};
            R acc = 0, cnt = 0;
            for (size_t q : nb)
                if (L[q] < i) { acc += data[q]; cnt += R(1.0); }
            if (cnt != R(0.0)) acc /= cnt;
            data[c] = acc;
        }
        frontier.swap(next);
    }
    if (layer_out) *layer_out = L;
    return used;
}

template <class R>
int extrapolate_fc(const Grid<R>& g, R* u, R* v, R* w, int layers) {
    int a = extrapolate_grid(u, g.nx + 1, g.ny, g.nz, layers);
    int b = extrapolate_grid(v, g.nx, g.ny + 1, g.nz, layers);
    int c = extrapolate_grid(w, g.nx, g.ny, g.nz + 1, layers);
    return std::max(a, std::max(b, c));
}

// ---------------------------------------------------------------------------
// CPressureSolver::solvePressure (MPhysics/PressureSolver.cpp:85-107).
template <class R> struct PressureOut {
    std::vector<R> markers, A, b, x, div;
    CgResult cg;
};

template <class R>
PressureOut<R> pressure_solve(const Grid<R>& g, R dt, R density, R* u, R* v, R* w,
                              const R* us, const R* vs, const R* ws, const R* solid,
                              const R* fluid, int tol_mode, double tol,
                              long long max_it, bool jacobi, const R* markers_in = nullptr) {
    const size_t n = g.cells();
    PressureOut<R> o;
    o.markers.resize(n); o.A.resize(4 * n); o.b.resize(n); o.x.assign(n, R(0));
    o.div.resize(n);
    // markers_in: markers built elsewhere (from particles, PressureSolver.cpp:109-134)
    if (markers_in) o.markers.assign(markers_in, markers_in + n);
    else build_fluid_markers(n, solid, fluid, o.markers.data());
    R sa[3] = {dt / (density * g.hx * g.hx), dt / (density * g.hy * g.hy),
               dt / (density * g.hz * g.hz)};
    build_matrix(g, o.markers.data(), sa, o.A.data());
    divergence_fc(g, u, v, w, o.div.data());
    build_rhs(g, u, v, w, o.div.data(), o.markers.data(), us, vs, ws, o.b.data());
    if (max_it <= 0) max_it = (long long)n;
    o.cg = cg_solve(g, o.A.data(), o.b.data(), o.x.data(), tol_mode, tol, max_it, jacobi);
    R sg[3] = {dt / (density * g.hx), dt / (density * g.hy), dt / (density * g.hz)};
    apply_pressure_gradient(g, sg, o.markers.data(), o.x.data(), u, v, w, us, vs, ws);
    return o;
}

// Fields the hybrid coupling reads from the grid step (GridFluidSolver.cpp:
// 571-613).  FluidSurfaceDensity = 0.2 (Common.h:244).
//
// buildFluidDensity (MPhysics/EulerSolverTool.cu:408-448): a cell inside the
// fluid (phi < 0, isInsideSDF EulerMathTool.cu:129) and outside every solid
// (solid marker >= 0) is FluidSurfaceDensity when a clamped neighbour is air
// (phi >= 0 and solid >= 0, isInsideAir EulerMathTool.cu:344-348), else 1;
// every other cell 0.
template <class R>
void build_fluid_density(const Grid<R>& g, const R* phi, const R* solid, R* out) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    auto air = [&](size_t i) { return phi[i] >= R(0) && solid[i] >= R(0); };
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t c = lin(x, y, z, nx, ny, nz);
                if (phi[c] < R(0) && !(solid[c] < R(0))) {
                    const bool surf = air(lin(x - 1, y, z, nx, ny, nz)) || air(lin(x + 1, y, z, nx, ny, nz)) ||
                                      air(lin(x, y - 1, z, nx, ny, nz)) || air(lin(x, y + 1, z, nx, ny, nz)) ||
                                      air(lin(x, y, z - 1, nx, ny, nz)) || air(lin(x, y, z + 1, nx, ny, nz));
                    out[c] = surf ? R(0.2) : R(1.0);
                } else {
                    out[c] = R(0);
                }
            }
}

// buildFluidSDFInvoker / buildFluidSDF (MPhysics/EulerSolverTool.cu:324-406,
// 760-785): passes i = 0 .. passes-1 over the grid, in place.  Pass 0: a
// fluid cell (phi < 0) with a non-fluid clamped neighbour gets 0.5 - 0.2, other
// fluid cells -UNKNOWN; a non-fluid cell with a fluid neighbour gets
// 1.5 - 0.2, others UNKNOWN.  Pass i: an UNKNOWN cell with a neighbour equal
// (|a - b| < FLT_MIN, isRealEQ EulerMathTool.cu:88-91) to i + (1.5 - 0.2) - 1
// becomes i + (1.5 - 0.2); a -UNKNOWN cell with a neighbour equal to
// -(i + 0.2 - 1.5) becomes -(i - (0.5 - 0.2)).  The constants are evaluated in
// double and rounded to Real, as the reference's double literals make them.
// A cell written in pass i holds a pass-i value, which no other cell of pass
// i compares against, so the reference's in-place parallel passes equal this
// sequential loop; a pass that writes nothing ends the loop (the later passes
// would find nothing either).
template <class R>
void build_fluid_sdf(const Grid<R>& g, const R* phi, R* out, int passes) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const R unknown = Limits<R>::unknown();
    const R tiny = std::numeric_limits<R>::min();
    auto eq = [&](R a, R b) { return std::abs(a - b) < tiny; };
    for (int i = 0; i < passes; ++i) {
        bool wrote = false;
        const R prev_out = R(double(i) + (1.5 - 0.2) - 1), cur_out = R(double(i) + (1.5 - 0.2));
        const R prev_in = R(-(double(i) + 0.2 - 1.5)), cur_in = R(-(double(i) - (0.5 - 0.2)));
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const size_t c = lin(x, y, z, nx, ny, nz);
                    const size_t nb[6] = {lin(x - 1, y, z, nx, ny, nz), lin(x + 1, y, z, nx, ny, nz),
                                          lin(x, y - 1, z, nx, ny, nz), lin(x, y + 1, z, nx, ny, nz),
                                          lin(x, y, z - 1, nx, ny, nz), lin(x, y, z + 1, nx, ny, nz)};
                    if (i == 0) {
                        const bool in = phi[c] < R(0);
                        bool other = false;
                        for (size_t n : nb) other = other || ((phi[n] < R(0)) != in);
                        out[c] = in ? (other ? R(0.5 - 0.2) : -unknown) : (other ? R(1.5 - 0.2) : unknown);
                        wrote = true;
                    } else if (out[c] == unknown) {
                        bool hit = false;
                        for (size_t n : nb) hit = hit || eq(out[n], prev_out);
                        if (hit) { out[c] = cur_out; wrote = true; }
                    } else if (out[c] == -unknown) {
                        bool hit = false;
                        for (size_t n : nb) hit = hit || eq(out[n], prev_in);
                        if (hit) { out[c] = cur_in; wrote = true; }
                    }
                }
        if (!wrote) break;
    }
}

// ---------------------------------------------------------------------------
// Particles (MPhysics/EulerParticles.cpp, MPhysics/EulerParticlesTool.cu):
// positions and per-particle vectors are interleaved x,y,z (3*i + k).
//
// Transfer kernels (EulerMathTool.cu:268-294).  The reference writes them with
// double literals, so each is evaluated in double from the Real argument and
// rounded to Real once.
template <class R> inline R pg_linear(R v) {
    const R a = std::abs(v);
    return a < R(1) ? R(1.0 - double(a)) : R(0);
}
template <class R> inline R pg_quadratic(R v) {
    const R a = std::abs(v);
    if (a < R(0.5)) return R(0.75 - double(R(a * a)));
    if (a < R(1.5)) return R(0.5 * (1.5 - double(a)) * (1.5 - double(a)));
    return R(0);
}
template <class R> inline R pg_cubic(R v) {
    const R a = std::abs(v);
    if (a < R(1)) return R(0.5 * double(a) * double(a) * double(a) - double(R(a * a)) + 2.0 / 3.0);
    if (a < R(2)) return R(1.0 / 6.0 * (2.0 - double(a)) * (2.0 - double(a)) * (2.0 - double(a)));
    return R(0);
}

// EPGTransferAlgorithm (MPhysics/EnumType.h:19-26).
enum PgAlg { PG_P2GSUM = 0, PG_G2PNEAREST = 1, PG_LINEAR = 2, PG_QUADRATIC = 3, PG_CUBIC = 4 };

// A cell-centred grid of extents n, origin o, spacing h (a face component is
// one with the component's extents and origin shifted by -h/2).
template <class R> struct PGrid {
    int nx, ny, nz;
    R ox, oy, oz, hx, hy, hz;
};

// transPos2Index (EulerMathTool.cu:29-40): truncation of (p - o) / h,
// clamped to the grid.
template <class R> inline size_t pos_cell(const PGrid<R>& g, R px, R py, R pz) {
    const int ix = limit_index(int((px - g.ox) / g.hx), g.nx);
    const int iy = limit_index(int((py - g.oy) / g.hy), g.ny);
    const int iz = limit_index(int((pz - g.oz) / g.hz), g.nz);
    return lin(ix, iy, iz, g.nx, g.ny, g.nz);
}

// The support of a particle: lower corner index and per-axis weights of the
// kernel (EulerParticlesTool.cu:79-198 and 198-290 share this prologue).
template <class R> struct Stencil {
    int i0[3];
    R w[3][4];
    int n;   // 2 (linear) or 4 (quadratic, cubic): corners i0-1 .. i0+2 for n = 4
};
template <class R> inline Stencil<R> pg_stencil(const PGrid<R>& g, R px, R py, R pz, int alg) {
    const R rel[3] = {px - g.ox - g.hx * R(0.5), py - g.oy - g.hy * R(0.5), pz - g.oz - g.hz * R(0.5)};
    const R h[3] = {g.hx, g.hy, g.hz};
    Stencil<R> st{};
    for (int k = 0; k < 3; ++k) {
        st.i0[k] = floor_cuda(rel[k] / h[k]);
        const R f = (rel[k] - R(st.i0[k]) * h[k]) / h[k];
        if (alg == PG_LINEAR) {
            st.n = 2;
            st.w[k][0] = pg_linear(f);
            st.w[k][1] = pg_linear(R(double(f) - 1.0));
        } else {
            st.n = 4;
            for (int q = 0; q < 4; ++q) {
                const R a = R(double(f) + 1.0 - double(q));
                st.w[k][q] = alg == PG_QUADRATIC ? pg_quadratic(a) : pg_cubic(a);
            }
        }
    }
    return st;
}

// accumulateParticles2CCSField (EulerParticlesTool.cu:79-196): each particle
// adds weight * value to the clamped corners of its support and the weight
// to the weight field (P2GSUM: the value to its own cell, no weight).  The
// reference adds with atomics in no fixed order; here particles are added in
// index order, corners z, y, x, so sums agree to rounding.  Then, when
// normalising, value /= weight where weight > 0 (normalizeCCSField :290-308).
template <class R>
void p2g(const PGrid<R>& g, size_t np, const R* pos, const R* val, int span, int off, int alg, R* field,
         R* weight, bool normalize) {
    const size_t n = size_t(g.nx) * g.ny * g.nz;
    std::fill(field, field + n, R(0));
    std::fill(weight, weight + n, R(0));
    for (size_t p = 0; p < np; ++p) {
        const R px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
        const R v = val[size_t(span) * p + size_t(off)];
        if (alg == PG_P2GSUM) {
            field[pos_cell(g, px, py, pz)] += v;
            continue;
        }
        if (alg != PG_LINEAR && alg != PG_QUADRATIC && alg != PG_CUBIC) continue;
        const Stencil<R> st = pg_stencil(g, px, py, pz, alg);
        const int lo = st.n == 2 ? 0 : -1;
        for (int z = 0; z < st.n; ++z)
            for (int y = 0; y < st.n; ++y)
                for (int x = 0; x < st.n; ++x) {
                    const R wt = st.w[0][x] * st.w[1][y] * st.w[2][z];
                    const size_t c = lin(st.i0[0] + x + lo, st.i0[1] + y + lo, st.i0[2] + z + lo, g.nx, g.ny, g.nz);
                    field[c] += wt * v;
                    weight[c] += wt;
                }
    }
    if (normalize)
        for (size_t c = 0; c < n; ++c)
            if (weight[c] > R(0)) field[c] /= weight[c];
}

// accumulateCCSField2Particles (EulerParticlesTool.cu:198-288): the value at
// the particle, summed from 0 in corner order z, y, x (G2PNEAREST: the value
// of its cell).
template <class R>
void g2p(const PGrid<R>& g, size_t np, const R* pos, const R* field, int alg, R* val, int span, int off) {
    for (size_t p = 0; p < np; ++p) {
        const R px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
        R acc = R(0);
        if (alg == PG_G2PNEAREST) {
            acc = field[pos_cell(g, px, py, pz)];
        } else if (alg == PG_LINEAR || alg == PG_QUADRATIC || alg == PG_CUBIC) {
            const Stencil<R> st = pg_stencil(g, px, py, pz, alg);
            const int lo = st.n == 2 ? 0 : -1;
            for (int z = 0; z < st.n; ++z)
                for (int y = 0; y < st.n; ++y)
                    for (int x = 0; x < st.n; ++x) {
                        const R wt = st.w[0][x] * st.w[1][y] * st.w[2][z];
                        acc += wt * field[lin(st.i0[0] + x + lo, st.i0[1] + y + lo, st.i0[2] + z + lo, g.nx, g.ny, g.nz)];
                    }
        }
        val[size_t(span) * p + size_t(off)] = acc;
    }
}

// The three components of a face-centred field as cell-centred grids
// (tranferParticles2FCVFieldInvoker, EulerParticlesTool.cu:601-711).
template <class R> inline PGrid<R> face_grid(const Grid<R>& g, int c) {
    PGrid<R> f{g.nx + (c == 0), g.ny + (c == 1), g.nz + (c == 2), g.ox, g.oy, g.oz, g.hx, g.hy, g.hz};
    if (c == 0) f.ox = g.ox - g.hx / 2;
    if (c == 1) f.oy = g.oy - g.hy / 2;
    if (c == 2) f.oz = g.oz - g.hz / 2;
    return f;
}
template <class R> inline PGrid<R> cell_grid(const Grid<R>& g) {
    return PGrid<R>{g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g.hx, g.hy, g.hz};
}

// statisticalFluidDensity (EulerParticlesTool.cu:27-46): +1 per particle in
// its cell; buildFluidMarkers (:48-77, 366-411): 2 where the solid marker is
// < 0, else 1 where a particle lies, else 0.
template <class R>
void particle_counts(const Grid<R>& g, size_t np, const R* pos, R* out) {
    const PGrid<R> c = cell_grid(g);
    std::fill(out, out + g.cells(), R(0));
    for (size_t p = 0; p < np; ++p) out[pos_cell(c, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2])] += R(1);
}
template <class R>
void particle_markers(const Grid<R>& g, const R* solid, size_t np, const R* pos, R* out) {
    particle_counts(g, np, pos, out);
    for (size_t c = 0; c < g.cells(); ++c) out[c] = solid[c] < R(0) ? R(2) : (out[c] > R(0) ? R(1) : R(0));
}

// CEulerParticles::advectParticlesInVelField (EulerParticles.cpp:231-325):
// substeps of cfl * min(h) / max|v| (v = the velocity components at the
// particles), the last two shortened to end at dt; RK1/2/3 per substep with
// the axpys of the positions (axpyReal: no early return), then the positions
// clamped to the grid box when a boundary SDF of the grid's resolution is
// given (fixParticlesPosWithBoundarys, EulerParticlesTool.cu:310-333, 870-894:
// the reference launches one thread per *cell*, so only the first
// min(cells, particles) particles are clamped; here the same ones).
template <class R>
void advect_particles(const Grid<R>& g, const R* u, const R* v, const R* w, size_t np, R* pos, R dt, R cfl,
                      int order, bool fix) {
    std::vector<R> v1(3 * np), v2(3 * np), v3(3 * np), mid(3 * np), tf(3 * np);
    auto sample = [&](const R* at, R* out) {
        for (size_t p = 0; p < np; ++p)
            sample_fc(g, u, v, w, at[3 * p], at[3 * p + 1], at[3 * p + 2], out[3 * p], out[3 * p + 1], out[3 * p + 2]);
    };
    auto axpy = [&](const R* x, R* y, R a) {
        for (size_t i = 0; i < 3 * np; ++i) y[i] = fma_r(a, x[i], y[i]);
    };
    const R minh = std::min(std::min(g.hx, g.hy), g.hz);
    R t = R(0);
    bool finished = false;
    while (!finished && np > 0) {
        sample(pos, v1.data());
        R vmax = R(0);
        for (size_t i = 0; i < 3 * np; ++i) vmax = std::max(vmax, std::abs(v1[i]));
        R step = vmax != R(0) ? cfl * minh / vmax : dt;
        if (t + step >= dt) {
            step = dt - t;
            finished = true;
        } else if (t + 2 * step >= dt) {
            step = R(0.5 * double(dt - t));
        }
        if (order == 1) {
            axpy(v1.data(), pos, step);
        } else if (order == 2) {
            const R half = R(0.5 * double(step));
            mid.assign(pos, pos + 3 * np);
            axpy(v1.data(), mid.data(), half);
            sample(mid.data(), v2.data());
            axpy(v2.data(), pos, step);
        } else if (order == 3) {
            const R half = R(0.5 * double(step)), tq = R(0.75 * double(step));
            const R c1 = R(2.0 / 9.0 * double(step)), c2 = R(3.0 / 9.0 * double(step)), c3 = R(4.0 / 9.0 * double(step));
            mid.assign(pos, pos + 3 * np);
            tf.assign(pos, pos + 3 * np);
            axpy(v1.data(), mid.data(), half);
            sample(mid.data(), v2.data());
            axpy(v2.data(), tf.data(), tq);
            sample(tf.data(), v3.data());
            axpy(v1.data(), pos, c1);
            axpy(v2.data(), pos, c2);
            axpy(v3.data(), pos, c3);
        }
        t += step;
    }
    if (fix) {
        const R mn[3] = {g.ox, g.oy, g.oz};
        const R mx[3] = {g.ox + R(g.nx) * g.hx, g.oy + R(g.ny) * g.hy, g.oz + R(g.nz) * g.hz};
        const size_t m = std::min(np, g.cells());
        for (size_t p = 0; p < m; ++p)
            for (int k = 0; k < 3; ++k) pos[3 * p + k] = std::max(mn[k], std::min(pos[3 * p + k], mx[k]));
    }
}

// ---------------------------------------------------------------------------
// CEulerParticles::generateParticlesInFluid (EulerParticles.cpp:163-201):
// per_cell particles in every cell with fluid < 0 and not solid < 0
// (findFluidGrid, EulerParticlesTool.cu:5-25), cells in x-fastest order,
// particle i of the k-th such cell at index k * per_cell + i, placed at
// (x + r) * h + o per axis.  The reference draws r = rand() / RAND_MAX with
// the generator seeded by the clock (not reproducible); here r is drawn from
// SplitMix64's output function (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014) applied to (seed, 3 * index +
// axis), top 24 bits, in [0, 1).
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline float particle_jitter(uint64_t seed, uint64_t index, int axis) {
    return float(splitmix64(seed ^ splitmix64(3 * index + uint64_t(axis))) >> 40) * (1.0f / 16777216.0f);
}
template <class R>
std::vector<R> generate_particles(const Grid<R>& g, const R* fluid, const R* solid, int per_cell, uint64_t seed) {
    std::vector<R> pos;
    size_t rank = 0;
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                const size_t c = lin(x, y, z, g.nx, g.ny, g.nz);
                if (!(fluid[c] < R(0)) || (solid && solid[c] < R(0))) continue;
                for (int i = 0; i < per_cell; ++i) {
                    const uint64_t idx = uint64_t(rank) * uint64_t(per_cell) + uint64_t(i);
                    pos.push_back((R(x) + R(particle_jitter(seed, idx, 0))) * g.hx + g.ox);
                    pos.push_back((R(y) + R(particle_jitter(seed, idx, 1))) * g.hy + g.oy);
                    pos.push_back((R(z) + R(particle_jitter(seed, idx, 2))) * g.hz + g.oz);
                }
                ++rank;
            }
    return pos;
}

// The particle CAdvectionSolver plug-ins on one particle set:
// CParticleInCell (ParticleInCell.cpp:58-106), CFluidImplicitParticle
// (FluidImplicitParticle.cpp:60-125), CMixPICAndFLIP (MixPICAndFLIP.cpp:
// 64-134).  The vector transfers are the three component transfers of
// tranferFCVField2ParticlesInvoker / tranferParticles2FCVFieldInvoker
// (EulerParticlesTool.cu:601-711, 820-868) with span 3, offset = component.
enum PSolverKind { PS_PIC = 0, PS_FLIP = 1, PS_MIX = 2 };
template <class R> struct ParticleSolver {
    Grid<R> g;
    int kind = PS_PIC, alg = PG_LINEAR;
    R cfl = R(1), mixing = R(0.1);
    std::vector<R> pos, vel, scalar, vec3, dvel, picvel;
    std::vector<R> delta[3];                 // m_DeltaGridVelField
    bool sflag = false, vflag = false, velflag = false;

    size_t np() const { return pos.size() / 3; }
    size_t faces(int c) const {
        const PGrid<R> f = face_grid(g, c);
        return size_t(f.nx) * f.ny * f.nz;
    }
    void init(size_t n) {
        vel.assign(3 * n, R(0)); scalar.assign(n, R(0)); vec3.assign(3 * n, R(0));
        dvel.assign(3 * n, R(0)); picvel.assign(3 * n, R(0));
        for (int c = 0; c < 3; ++c) delta[c].assign(faces(c), R(0));
    }
    void g2p_fcv(const R* const f[3], R* out) const {
        for (int c = 0; c < 3; ++c) g2p(face_grid(g, c), np(), pos.data(), f[c], alg, out, 3, c);
    }
    void p2g_fcv(const R* val, R* const out[3]) const {
        for (int c = 0; c < 3; ++c) {
            std::vector<R> wt(faces(c));
            p2g(face_grid(g, c), np(), pos.data(), val, 3, c, alg, out[c], wt.data(), true);
        }
    }
    // m_DeltaGridVelField.plusAlphaX(in, -1); *= -1 (the first call copies in).
    void update_delta(const R* const in[3]) {
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < delta[c].size(); ++i)
                delta[c][i] = velflag ? -fma_r(R(-1), in[c][i], delta[c][i]) : in[c][i];
        velflag = true;
    }
    void advect_velocity(const R* const in[3], const R* const V[3], R dt, int order, bool fix, R* const out[3]) {
        if (kind == PS_PIC) {
            g2p_fcv(in, vel.data());
        } else {
            if (kind == PS_MIX) g2p_fcv(in, picvel.data());
            update_delta(in);
            const R* d[3] = {delta[0].data(), delta[1].data(), delta[2].data()};
            g2p_fcv(d, dvel.data());
            for (size_t i = 0; i < vel.size(); ++i) vel[i] = fma_r(R(1), dvel[i], vel[i]);
            if (kind == PS_MIX) {
                const R keep = R(1.0 - double(mixing));
                for (size_t i = 0; i < vel.size(); ++i) vel[i] = keep * vel[i];
                for (size_t i = 0; i < vel.size(); ++i) vel[i] = fma_r(mixing, picvel[i], vel[i]);
            }
        }
        advect_particles(g, V[0], V[1], V[2], np(), pos.data(), dt, cfl, order, fix);
        p2g_fcv(vel.data(), out);
        if (kind != PS_PIC)
            for (int c = 0; c < 3; ++c) delta[c].assign(out[c], out[c] + faces(c));
    }
    void advect_scalar(const R* in, R* out) {
        const PGrid<R> cg = cell_grid(g);
        if (!sflag) g2p(cg, np(), pos.data(), in, alg, scalar.data(), 1, 0);
        sflag = true;
        std::vector<R> wt(g.cells());
        p2g(cg, np(), pos.data(), scalar.data(), 1, 0, alg, out, wt.data(), true);
    }
    void advect_vector(const R* const in[3], R* const out[3]) {
        const PGrid<R> cg = cell_grid(g);
        if (!vflag)
            for (int c = 0; c < 3; ++c) g2p(cg, np(), pos.data(), in[c], alg, vec3.data(), 3, c);
        vflag = true;
        std::vector<R> wt(g.cells());
        for (int c = 0; c < 3; ++c) p2g(cg, np(), pos.data(), vec3.data(), 3, c, alg, out[c], wt.data(), true);
    }
};

// FP64 true residual ||b - A x|| / ||b|| with A from the markers.
template <class R>
double true_residual_rel(const Grid<R>& g, const R* A, const R* b, const R* x) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double num = 0, den = 0;
    for (int z = 0; z < nz; ++z)
        for (int yy = 0; yy < ny; ++yy)
            for (int xx = 0; xx < nx; ++xx) {
                size_t c = lin(xx, yy, z, nx, ny, nz);
                double acc = double(A[4 * c]) * x[c];
                if (xx > 0) acc += double(A[4 * lin(xx - 1, yy, z, nx, ny, nz) + 1]) * x[lin(xx - 1, yy, z, nx, ny, nz)];
                if (xx + 1 < nx) acc += double(A[4 * c + 1]) * x[lin(xx + 1, yy, z, nx, ny, nz)];
                if (yy > 0) acc += double(A[4 * lin(xx, yy - 1, z, nx, ny, nz) + 2]) * x[lin(xx, yy - 1, z, nx, ny, nz)];
                if (yy + 1 < ny) acc += double(A[4 * c + 2]) * x[lin(xx, yy + 1, z, nx, ny, nz)];
                if (z > 0) acc += double(A[4 * lin(xx, yy, z - 1, nx, ny, nz) + 3]) * x[lin(xx, yy, z - 1, nx, ny, nz)];
                if (z + 1 < nz) acc += double(A[4 * c + 3]) * x[lin(xx, yy, z + 1, nx, ny, nz)];
                double r = double(b[c]) - acc;
                num += r * r;
                den += double(b[c]) * double(b[c]);
            }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

}  // namespace mac_oracle
