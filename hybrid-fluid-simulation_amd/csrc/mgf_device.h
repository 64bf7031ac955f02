// mgf_device.h — device-side arithmetic of the MAC-grid step for gfx950.
//
// Every routine here reproduces one reference operation with the same
// operation order and one rounding per written float operation (the
// library is compiled with -ffp-contract=off); the BLAS axpy of the
// back-trace is one explicit fused multiply-add (__builtin_fmaf).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#define MGF_UNKNOWN 3.40282346638528859812e+38f   /* FLT_MAX, MMathExtra.h:58 */
#define MGF_FLT_EPSILON 1.19209290e-07f          /* FLT_EPSILON, MMathExtra.h:55 */

namespace mgf {

struct GridDev {
    int nx, ny, nz;
    float ox, oy, oz;
    float hx, hy, hz;
    // 1/h per axis and whether every spacing is a power of two whose
    // reciprocal is a normal float (set by set_recip)
    float rhx = 0, rhy = 0, rhz = 0;
    int pow2 = 0;
    // 1/d of the index decodes' divisors (cells: nx, nx ny; u faces: nx + 1,
    // (nx + 1) ny; v faces: nx, nx (ny + 1); w faces as cells), set by set_recip
    double r_nx = 0, r_nxy = 0, r_nx1 = 0, r_nx1y = 0, r_nxy1 = 0;
};

// Division by a spacing h = 2^k and multiplication by 2^-k round the same
// real number once, so they give the same float (overflow, underflow, zero
// signs and NaN included) while the multiplication avoids the correctly
// rounded division's instruction sequence.  Other spacings divide.
inline void set_recip(GridDev& g) {
    auto p2 = [](float h, float& r) {
        int e = 0;
        const float m = std::frexp(h, &e);
        r = 1.0f / h;
        return h > 0.0f && m == 0.5f && std::isnormal(h) && std::isnormal(r);
    };
    const bool a = p2(g.hx, g.rhx), b = p2(g.hy, g.rhy), c = p2(g.hz, g.rhz);
    g.pow2 = (a && b && c) ? 1 : 0;
    const double nx = double(g.nx), ny = double(g.ny);
    g.r_nx = 1.0 / nx;
    g.r_nxy = 1.0 / (nx * ny);
    g.r_nx1 = 1.0 / (nx + 1.0);
    g.r_nx1y = 1.0 / ((nx + 1.0) * ny);
    g.r_nxy1 = 1.0 / (nx * (ny + 1.0));
}

// a / d for a < 2^32, d >= 1, from rd = 1/d in FP64: a * rd is within one of
// the quotient and is corrected once, so the quotient is exact (the integer
// divide's instruction sequence avoided).
__device__ __forceinline__ uint32_t udiv_rd(uint32_t a, uint32_t d, double rd) {
    uint32_t q = uint32_t(double(a) * rd);
    const uint32_t r = a - q * d;
    if (int32_t(r) < 0) --q;
    else if (r >= d) ++q;
    return q;
}

// (x, y, z) of index i in an ex x ey x .. array: rex = 1/ex, rexy = 1/(ex ey).
__device__ __forceinline__ void decode3(uint32_t i, uint32_t ex, uint32_t ey, double rex, double rexy, int& x,
                                        int& y, int& z) {
    const uint32_t exy = ex * ey;
    const uint32_t zz = udiv_rd(i, exy, rexy);
    const uint32_t rem = i - zz * exy;
    const uint32_t yy = udiv_rd(rem, ex, rex);
    x = int(rem - yy * ex);
    y = int(yy);
    z = int(zz);
}

__device__ __forceinline__ float div_h(float x, float h, float rh, int pow2) {
    return pow2 ? x * rh : x / h;
}

// limitIndex (MPhysics/EulerMathTool.cu:16-27).
__device__ __forceinline__ int clampi(int i, int n) {
    return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

// transIndex2LinearWithOffset (MPhysics/EulerMathTool.cu:63-73).
__device__ __forceinline__ uint32_t lin3(int x, int y, int z, int nx, int ny, int nz) {
    x = clampi(x, nx);
    y = clampi(y, ny);
    z = clampi(z, nz);
    return (uint32_t(z) * uint32_t(ny) + uint32_t(y)) * uint32_t(nx) + uint32_t(x);
}

// floorCUDA (MPhysics/EulerMathTool.cu:93-103): truncate, minus one when the
// input is negative.  The conversion saturates at +-2^24 and maps NaN to 0
// (points that far out clamp to the grid edge either way).
__device__ __forceinline__ int floor_ref(float v) {
    int t;
    if (!(v == v)) t = 0;
    else if (v >= 16777216.0f) t = 16777216;
    else if (v <= -16777216.0f) t = -16777216;
    else t = int(v);
    return v >= 0.0f ? t : t - 1;
}

// Trilinear gather on a cell-centred grid (extents nx,ny,nz, origin o,
// spacing h): sampleCellCenteredScalarFieldTrilerp
// (MPhysics/FieldMathTool.cu:52-104).
__device__ __forceinline__ float trilerp(const float* __restrict__ f, int nx, int ny, int nz,
                                         float ox, float oy, float oz, const GridDev& g,
                                         float px, float py, float pz) {
    const float hx = g.hx, hy = g.hy, hz = g.hz;
    const float rx = px - ox - hx / 2;
    const float ry = py - oy - hy / 2;
    const float rz = pz - oz - hz / 2;
    const int ix = floor_ref(div_h(rx, hx, g.rhx, g.pow2));
    const int iy = floor_ref(div_h(ry, hy, g.rhy, g.pow2));
    const int iz = floor_ref(div_h(rz, hz, g.rhz, g.pow2));
    const float fx = div_h(rx - float(ix) * hx, hx, g.rhx, g.pow2);
    const float fy = div_h(ry - float(iy) * hy, hy, g.rhy, g.pow2);
    const float fz = div_h(rz - float(iz) * hz, hz, g.rhz, g.pow2);
    const int x0 = clampi(ix, nx), x1 = clampi(ix + 1, nx);
    const int y0 = clampi(iy, ny), y1 = clampi(iy + 1, ny);
    const int z0 = clampi(iz, nz), z1 = clampi(iz + 1, nz);
    const uint32_t sy = uint32_t(nx), sz = uint32_t(nx) * uint32_t(ny);
    const uint32_t r00 = uint32_t(z0) * sz + uint32_t(y0) * sy;
    const uint32_t r01 = uint32_t(z0) * sz + uint32_t(y1) * sy;
    const uint32_t r10 = uint32_t(z1) * sz + uint32_t(y0) * sy;
    const uint32_t r11 = uint32_t(z1) * sz + uint32_t(y1) * sy;
    const float dbl = f[r00 + x0], dbr = f[r00 + x1];
    const float ubl = f[r01 + x0], ubr = f[r01 + x1];
    const float dfl = f[r10 + x0], dfr = f[r10 + x1];
    const float ufl = f[r11 + x0], ufr = f[r11 + x1];
    return dbl * (1 - fx) * (1 - fy) * (1 - fz) +
           dbr * fx * (1 - fy) * (1 - fz) +
           ubl * (1 - fx) * fy * (1 - fz) +
           ubr * fx * fy * (1 - fz) +
           dfl * (1 - fx) * (1 - fy) * fz +
           dfr * fx * (1 - fy) * fz +
           ufl * (1 - fx) * fy * fz +
           ufr * fx * fy * fz;
}

// Component c of a face-centred field: a cell-centred grid one layer longer on
// axis c with the origin moved by -h/2 on that axis
// (MPhysics/FieldMathTool.cu:783-892).
__device__ __forceinline__ float sample_comp(const float* __restrict__ f, int c, const GridDev& g,
                                             float px, float py, float pz) {
    if (c == 0) return trilerp(f, g.nx + 1, g.ny, g.nz, g.ox - g.hx / 2, g.oy, g.oz, g, px, py, pz);
    if (c == 1) return trilerp(f, g.nx, g.ny + 1, g.nz, g.ox, g.oy - g.hy / 2, g.oz, g, px, py, pz);
    return trilerp(f, g.nx, g.ny, g.nz + 1, g.ox, g.oy, g.oz - g.hz / 2, g, px, py, pz);
}

struct Vel {
    const float* __restrict__ u;
    const float* __restrict__ v;
    const float* __restrict__ w;
};

__device__ __forceinline__ void sample_vel(const Vel& V, const GridDev& g, float px, float py,
                                           float pz, float& a, float& b, float& c) {
    a = sample_comp(V.u, 0, g, px, py, pz);
    b = sample_comp(V.v, 1, g, px, py, pz);
    c = sample_comp(V.w, 2, g, px, py, pz);
}

// plusAlphaX's early return: |alpha| < FLT_EPSILON skips the update.
__device__ __forceinline__ bool axpy_on(float a) { return !(fabsf(a) < MGF_FLT_EPSILON); }

// CSemiLagrangian::backTrace for one point (MPhysics/SemiLagrangian.cpp:
// 188-347).  `coef` holds the host-computed axpy factors:
// [0] -dt  [1] -0.5*dt  [2] -0.75*dt  [3] -(2/9)dt  [4] -(3/9)dt  [5] -(4/9)dt.
struct RkCoef {
    float c[6];
};

__device__ __forceinline__ void backtrace(const Vel& V, const GridDev& g, int order,
                                          const RkCoef& k, float& px, float& py, float& pz) {
    const float x0 = px, y0 = py, z0 = pz;
    float v1x, v1y, v1z;
    sample_vel(V, g, x0, y0, z0, v1x, v1y, v1z);
    if (order == 1) {
        if (axpy_on(k.c[0])) {
            px = __builtin_fmaf(k.c[0], v1x, x0);
            py = __builtin_fmaf(k.c[0], v1y, y0);
            pz = __builtin_fmaf(k.c[0], v1z, z0);
        }
        return;
    }
    float mx = x0, my = y0, mz = z0;
    if (axpy_on(k.c[1])) {
        mx = __builtin_fmaf(k.c[1], v1x, x0);
        my = __builtin_fmaf(k.c[1], v1y, y0);
        mz = __builtin_fmaf(k.c[1], v1z, z0);
    }
    float v2x, v2y, v2z;
    sample_vel(V, g, mx, my, mz, v2x, v2y, v2z);
    if (order == 2) {
        if (axpy_on(k.c[0])) {
            px = __builtin_fmaf(k.c[0], v2x, x0);
            py = __builtin_fmaf(k.c[0], v2y, y0);
            pz = __builtin_fmaf(k.c[0], v2z, z0);
        }
        return;
    }
    float tx = x0, ty = y0, tz = z0;
    if (axpy_on(k.c[2])) {
        tx = __builtin_fmaf(k.c[2], v2x, x0);
        ty = __builtin_fmaf(k.c[2], v2y, y0);
        tz = __builtin_fmaf(k.c[2], v2z, z0);
    }
    float v3x, v3y, v3z;
    sample_vel(V, g, tx, ty, tz, v3x, v3y, v3z);
    float ox = x0, oy = y0, oz = z0;
    if (axpy_on(k.c[3])) {
        ox = __builtin_fmaf(k.c[3], v1x, ox);
        oy = __builtin_fmaf(k.c[3], v1y, oy);
        oz = __builtin_fmaf(k.c[3], v1z, oz);
    }
    if (axpy_on(k.c[4])) {
        ox = __builtin_fmaf(k.c[4], v2x, ox);
        oy = __builtin_fmaf(k.c[4], v2y, oy);
        oz = __builtin_fmaf(k.c[4], v2z, oz);
    }
    if (axpy_on(k.c[5])) {
        ox = __builtin_fmaf(k.c[5], v3x, ox);
        oy = __builtin_fmaf(k.c[5], v3y, oy);
        oz = __builtin_fmaf(k.c[5], v3z, oz);
    }
    px = ox;
    py = oy;
    pz = oz;
}

// Face position of component c at (x,y,z) and cell-centre position
// (MPhysics/SemiLagrangian.cpp:47-97).
__device__ __forceinline__ float cpos(int k, float h, float o) { return (k + 0.5f) * h + o; }
__device__ __forceinline__ float fpos(int k, float h, float o) { return float(k) * h + o; }

// ---------------------------------------------------------------------------
// Markers: 0 air, 1 fluid, 2 solid (MPhysics/EulerSolverTool.cu:30-56).
enum : uint8_t { kAir = 0, kFluid = 1, kSolid = 2 };

// One row of the pressure matrix recomputed from the markers
// (buildPressureMatrixA, MPhysics/EulerSolverTool.cu:58-126): diagonal and the
// three "+" couplings; scale s = dt/(rho*h^2).  The diagonal is accumulated in
// the reference's order so it is bitwise the stored value.
struct Row {
    float d, ax, ay, az;
};

template <class M>
__device__ __forceinline__ Row matrix_row(const M& mk, int x, int y, int z, int nx, int ny, int nz,
                                          float sx, float sy, float sz) {
    Row r{0.f, 0.f, 0.f, 0.f};
    if (mk(x, y, z) != kFluid) {
        r.d = 1.0f;
        return r;
    }
    if (x > 0 && mk(x - 1, y, z) != kSolid) r.d += sx;
    if (x + 1 < nx) {
        const uint8_t m = mk(x + 1, y, z);
        if (m != kSolid) {
            r.d += sx;
            if (m == kFluid) r.ax -= sx;
        }
    }
    if (y > 0 && mk(x, y - 1, z) != kSolid) r.d += sy;
    if (y + 1 < ny) {
        const uint8_t m = mk(x, y + 1, z);
        if (m != kSolid) {
            r.d += sy;
            if (m == kFluid) r.ay -= sy;
        }
    }
    if (z > 0 && mk(x, y, z - 1) != kSolid) r.d += sz;
    if (z + 1 < nz) {
        const uint8_t m = mk(x, y, z + 1);
        if (m != kSolid) {
            r.d += sz;
            if (m == kFluid) r.az -= sz;
        }
    }
    return r;
}

// "+d" coupling of cell (x,y,z) toward its +d neighbour, which is what the
// reference reads from the neighbour's slot for the -d term.
__device__ __forceinline__ float plus_coupling(uint8_t mc, uint8_t mn, float s) {
    return (mc == kFluid && mn == kFluid) ? -s : 0.0f;
}

}  // namespace mgf
