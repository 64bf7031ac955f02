// mgf_kernels.hip — HIP kernels of the MAC-grid step for gfx950 (wave64).
//
// Layout: structure-of-arrays, x fastest, face arrays with one extra layer on
// their own axis (the reference's CFaceCenteredVectorField layout).  Markers
// are uint8 (0 air, 1 fluid, 2 solid).  Every kernel is memory-bound; blocks
// are 256 work-items and block IDs are remapped so that each of the 8 XCDs
// walks one contiguous slab of the grid (neighbouring planes then meet in the
// same L2).
#include <cstring>
#include <vector>

#include "mgf_internal.h"
#include "mgf_kcommon.h"

namespace mgf {

// ---------------------------------------------------------------------------
// Back-trace of arbitrary points and cell-centred sampling (operator seams).
__global__ void __launch_bounds__(kBlock) k_backtrace(GridDev g, Vel V, RkCoef k, int order,
                                                      float* px, float* py, float* pz, int64_t n) {
    const int64_t i = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (i >= n) return;
    float x = px[i], y = py[i], z = pz[i];
    backtrace(V, g, order, k, x, y, z);
    px[i] = x; py[i] = y; pz[i] = z;
}

__global__ void __launch_bounds__(kBlock) k_sample_cc(GridDev g, const float* f, const float* px,
                                                      const float* py, const float* pz, float* out,
                                                      int64_t n) {
    const int64_t i = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = trilerp(f, g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g, px[i], py[i], pz[i]);
}

// ---------------------------------------------------------------------------
// Semi-Lagrangian advection of the face-centred velocity
// (CSemiLagrangian::advect(FCV), MPhysics/SemiLagrangian.cpp:169-186) with
// the uniform body-force increments of CExternalForcesSolver fused after the
// sample (MPhysics/ExternalForcesSolver.cpp:50-59).
struct Inc {
    float a[3];
    float b[3];
    int add_a, add_b;
};

__device__ __forceinline__ void k_advect_fc_run(GridDev g, Vel in, Vel V, RkCoef k, int order,
                                                      float* __restrict__ ou, float* __restrict__ ov,
                                                      float* __restrict__ ow, uint32_t nfx,
                                                      uint32_t nfy, uint32_t nfz, Inc inc, uint32_t blk) {
    uint32_t i = blk * kBlock + threadIdx.x;
    int c;
    if (i < nfx) c = 0;
    else if (i < nfx + nfy) { c = 1; i -= nfx; }
    else if (i < nfx + nfy + nfz) { c = 2; i -= nfx + nfy; }
    else return;
    const uint32_t ex = uint32_t(g.nx + (c == 0)), ey = uint32_t(g.ny + (c == 1));
    int x, y, z;
    decode3(i, ex, ey, c == 0 ? g.r_nx1 : g.r_nx, c == 0 ? g.r_nx1y : (c == 1 ? g.r_nxy1 : g.r_nxy), x, y, z);
    float px = c == 0 ? fpos(x, g.hx, g.ox) : cpos(x, g.hx, g.ox);
    float py = c == 1 ? fpos(y, g.hy, g.oy) : cpos(y, g.hy, g.oy);
    float pz = c == 2 ? fpos(z, g.hz, g.oz) : cpos(z, g.hz, g.oz);
    backtrace(V, g, order, k, px, py, pz);
    const float* src = c == 0 ? in.u : (c == 1 ? in.v : in.w);
    float val = sample_comp(src, c, g, px, py, pz);
    // the component's increments by selection (an index at run time would
    // copy the struct to scratch memory)
    const float ia = c == 0 ? inc.a[0] : (c == 1 ? inc.a[1] : inc.a[2]);
    const float ib = c == 0 ? inc.b[0] : (c == 1 ? inc.b[1] : inc.b[2]);
    if (inc.add_a) val = val + ia;
    if (inc.add_b) val = val + ib;
    float* dst = c == 0 ? ou : (c == 1 ? ov : ow);
    dst[i] = val;
}

__global__ void __launch_bounds__(kBlock) k_advect_fc(GridDev g, Vel in, Vel V, RkCoef k, int order,
                                                      float* __restrict__ ou, float* __restrict__ ov,
                                                      float* __restrict__ ow, uint32_t nfx,
                                                      uint32_t nfy, uint32_t nfz, Inc inc) {
    k_advect_fc_run(g, in, V, k, order, ou, ov, ow, nfx, nfy, nfz, inc, xcd_block());
}

struct MAdvectFc {
    static constexpr int kThreads = kBlock;
    GridDev g;
    Vel in;
    Vel V;
    RkCoef k;
    int order;
    float* ou;
    float* ov;
    float* ow;
    uint32_t nfx;
    uint32_t nfy;
    uint32_t nfz;
    Inc inc;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_advect_fc_run(g, in, V, k, order, ou, ov, ow, nfx, nfy, nfz, inc, xcd_of(vb, G)); }
};

// CSemiLagrangian::advect(CCS) (MPhysics/SemiLagrangian.cpp:135-150).
__device__ __forceinline__ void k_advect_cc_run(GridDev g, const float* __restrict__ in, Vel V,
                                                      RkCoef k, int order, float* __restrict__ out,
                                                      uint32_t n, uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t nx = uint32_t(g.nx), ny = uint32_t(g.ny);
    int x, y, z;
    decode3(i, nx, ny, g.r_nx, g.r_nxy, x, y, z);
    float px = cpos(x, g.hx, g.ox), py = cpos(y, g.hy, g.oy), pz = cpos(z, g.hz, g.oz);
    backtrace(V, g, order, k, px, py, pz);
    out[i] = trilerp(in, g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g, px, py, pz);
}

__global__ void __launch_bounds__(kBlock) k_advect_cc(GridDev g, const float* __restrict__ in, Vel V,
                                                      RkCoef k, int order, float* __restrict__ out,
                                                      uint32_t n) {
    k_advect_cc_run(g, in, V, k, order, out, n, xcd_block());
}

struct MAdvectCc {
    static constexpr int kThreads = kBlock;
    GridDev g;
    const float* in;
    Vel V;
    RkCoef k;
    int order;
    float* out;
    uint32_t n;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_advect_cc_run(g, in, V, k, order, out, n, xcd_of(vb, G)); }
};

// CSemiLagrangian::advect(CCS) of the fluid domain and advect(CCV) of the
// colour (MPhysics/SemiLagrangian.cpp:135-167): the four advects of one step
// back-trace the same cell centre through the same velocity, so one back-trace
// serves all four trilinear samples (sampleCellCenteredVectorFieldInvoker,
// FieldMathTool.cu:575-640, samples each component like a scalar field).
struct Cc4 {
    const float* in[4];
    float* out[4];
};
__device__ __forceinline__ void k_advect_cc4_run(GridDev g, Cc4 F, Vel V, RkCoef k, int order, uint32_t n, uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t nx = uint32_t(g.nx), ny = uint32_t(g.ny);
    int x, y, z;
    decode3(i, nx, ny, g.r_nx, g.r_nxy, x, y, z);
    float px = cpos(x, g.hx, g.ox), py = cpos(y, g.hy, g.oy), pz = cpos(z, g.hz, g.oz);
    backtrace(V, g, order, k, px, py, pz);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        F.out[q][i] = trilerp(F.in[q], g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g, px, py, pz);
}

__global__ void __launch_bounds__(kBlock) k_advect_cc4(GridDev g, Cc4 F, Vel V, RkCoef k, int order, uint32_t n) {
    k_advect_cc4_run(g, F, V, k, order, n, xcd_block());
}

struct MAdvectCc4 {
    static constexpr int kThreads = kBlock;
    GridDev g;
    Cc4 F;
    Vel V;
    RkCoef k;
    int order;
    uint32_t n;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_advect_cc4_run(g, F, V, k, order, n, xcd_of(vb, G)); }
};

// plusAlphaX with a unit field: a[i] = inc*1 + a[i].
__global__ void __launch_bounds__(kBlock) k_add_uniform(float* a, size_t n, float inc) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) a[i] = a[i] + inc;
}

// ---------------------------------------------------------------------------
// buildFluidMarkers (MPhysics/EulerSolverTool.cu:30-56).
__device__ __forceinline__ void k_markers_run(uint32_t n, const uint8_t* __restrict__ solid,
                                                    const float* __restrict__ fluid,
                                                    uint8_t* __restrict__ m, uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i >= n) return;
    m[i] = solid[i] ? kSolid : (fluid[i] < 0.0f ? kFluid : kAir);
}

__global__ void __launch_bounds__(kBlock) k_markers(uint32_t n, const uint8_t* __restrict__ solid,
                                                    const float* __restrict__ fluid,
                                                    uint8_t* __restrict__ m) {
    k_markers_run(n, solid, fluid, m, xcd_block());
}

struct MMarkers {
    static constexpr int kThreads = kBlock;
    uint32_t n;
    const uint8_t* solid;
    const float* fluid;
    uint8_t* m;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_markers_run(n, solid, fluid, m, xcd_of(vb, G)); }
};

struct MarkerView {
    const uint8_t* __restrict__ m;
    int nx, ny, nz;
    __device__ __forceinline__ uint8_t operator()(int x, int y, int z) const {
        return m[lin3(x, y, z, nx, ny, nz)];
    }
};

// Divergence + buildPressureVectorb (MPhysics/FieldMathTool.cu:370-395,
// MPhysics/EulerSolverTool.cu:128-197): b = -div on fluid cells, corrected by
// (u - u_solid)/h on each face shared with a solid (clamped) neighbour.
__device__ __forceinline__ void k_rhs_run(GridDev g, Vel U, Vel S, const uint8_t* __restrict__ m,
                                                float sx, float sy, float sz, float* __restrict__ b,
                                                uint32_t n, uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i >= n) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int x, y, z;
    decode3(uint32_t(i), uint32_t(nx), uint32_t(ny), g.r_nx, g.r_nxy, x, y, z);
    if (m[i] != kFluid) { b[i] = 0.0f; return; }
    const uint32_t fl = lin3(x, y, z, nx + 1, ny, nz), fr = lin3(x + 1, y, z, nx + 1, ny, nz);
    const uint32_t fd = lin3(x, y, z, nx, ny + 1, nz), fu = lin3(x, y + 1, z, nx, ny + 1, nz);
    const uint32_t fb = lin3(x, y, z, nx, ny, nz + 1), ff = lin3(x, y, z + 1, nx, ny, nz + 1);
    const float ul = U.u[fl], ur = U.u[fr], vd = U.v[fd], vu = U.v[fu], wb = U.w[fb], wf = U.w[ff];
    const float div = (ur - ul) / g.hx + (vu - vd) / g.hy + (wf - wb) / g.hz;
    float val = -div;
    if (m[lin3(x - 1, y, z, nx, ny, nz)] == kSolid) val -= sx * (ul - S.u[fl]);
    if (m[lin3(x + 1, y, z, nx, ny, nz)] == kSolid) val += sx * (ur - S.u[fr]);
    if (m[lin3(x, y - 1, z, nx, ny, nz)] == kSolid) val -= sy * (vd - S.v[fd]);
    if (m[lin3(x, y + 1, z, nx, ny, nz)] == kSolid) val += sy * (vu - S.v[fu]);
    if (m[lin3(x, y, z - 1, nx, ny, nz)] == kSolid) val -= sz * (wb - S.w[fb]);
    if (m[lin3(x, y, z + 1, nx, ny, nz)] == kSolid) val += sz * (wf - S.w[ff]);
    b[i] = val;
}

__global__ void __launch_bounds__(kBlock) k_rhs(GridDev g, Vel U, Vel S, const uint8_t* __restrict__ m,
                                                float sx, float sy, float sz, float* __restrict__ b,
                                                uint32_t n) {
    k_rhs_run(g, U, S, m, sx, sy, sz, b, n, xcd_block());
}

struct MRhs {
    static constexpr int kThreads = kBlock;
    GridDev g;
    Vel U;
    Vel S;
    const uint8_t* m;
    float sx;
    float sy;
    float sz;
    float* b;
    uint32_t n;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_rhs_run(g, U, S, m, sx, sy, sz, b, n, xcd_of(vb, G)); }
};

// CUDADivergence alone (download of MGF_FIELD_DIVERGENCE).
__global__ void __launch_bounds__(kBlock) k_divergence(GridDev g, Vel U, float* __restrict__ d,
                                                       uint32_t n) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int x, y, z;
    decode3(uint32_t(i), uint32_t(nx), uint32_t(ny), g.r_nx, g.r_nxy, x, y, z);
    const float ul = U.u[lin3(x, y, z, nx + 1, ny, nz)], ur = U.u[lin3(x + 1, y, z, nx + 1, ny, nz)];
    const float vd = U.v[lin3(x, y, z, nx, ny + 1, nz)], vu = U.v[lin3(x, y + 1, z, nx, ny + 1, nz)];
    const float wb = U.w[lin3(x, y, z, nx, ny, nz + 1)], wf = U.w[lin3(x, y, z + 1, nx, ny, nz + 1)];
    d[i] = (ur - ul) / g.hx + (vu - vd) / g.hy + (wf - wb) / g.hz;
}

// The compact matrix [diag,+x,+y,+z] (download of MGF_FIELD_MATRIX_A).
__global__ void __launch_bounds__(kBlock) k_matrix_a(GridDev g, const uint8_t* __restrict__ m,
                                                     float sx, float sy, float sz,
                                                     float* __restrict__ A, uint32_t n) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int x, y, z;
    decode3(uint32_t(i), uint32_t(nx), uint32_t(ny), g.r_nx, g.r_nxy, x, y, z);
    MarkerView mv{m, nx, ny, nz};
    const Row r = matrix_row(mv, x, y, z, nx, ny, nz, sx, sy, sz);
    A[4 * size_t(i)] = r.d;
    A[4 * size_t(i) + 1] = r.ax;
    A[4 * size_t(i) + 2] = r.ay;
    A[4 * size_t(i) + 3] = r.az;
}

// ---------------------------------------------------------------------------
// Stencil of the pressure matrix at cell (x,y,z) given a value accessor.
// Returns the seven (coefficient, value) pairs in the reference's order:
// centre, -x, +x, -y, +y, -z, +z; out-of-grid terms are flagged absent.
struct Stencil {
    float coef[7];
    bool on[7];
    uint32_t idx[7];
};

__device__ __forceinline__ Stencil stencil_at(const uint8_t* __restrict__ m, int x, int y, int z,
                                              int nx, int ny, int nz, float sx, float sy, float sz) {
    Stencil st;
    const uint32_t c = lin3(x, y, z, nx, ny, nz);
    st.idx[0] = c;
    st.idx[1] = lin3(x - 1, y, z, nx, ny, nz);
    st.idx[2] = lin3(x + 1, y, z, nx, ny, nz);
    st.idx[3] = lin3(x, y - 1, z, nx, ny, nz);
    st.idx[4] = lin3(x, y + 1, z, nx, ny, nz);
    st.idx[5] = lin3(x, y, z - 1, nx, ny, nz);
    st.idx[6] = lin3(x, y, z + 1, nx, ny, nz);
    st.on[0] = true;
    st.on[1] = x > 0; st.on[2] = x + 1 < nx;
    st.on[3] = y > 0; st.on[4] = y + 1 < ny;
    st.on[5] = z > 0; st.on[6] = z + 1 < nz;
    const uint8_t mc = m[c];
    uint8_t mn[7];
#pragma unroll
    for (int q = 1; q < 7; ++q) mn[q] = m[st.idx[q]];
    float d = 0.0f;
    if (mc == kFluid) {
        if (st.on[1] && mn[1] != kSolid) d += sx;
        if (st.on[2] && mn[2] != kSolid) d += sx;
        if (st.on[3] && mn[3] != kSolid) d += sy;
        if (st.on[4] && mn[4] != kSolid) d += sy;
        if (st.on[5] && mn[5] != kSolid) d += sz;
        if (st.on[6] && mn[6] != kSolid) d += sz;
    } else {
        d = 1.0f;
    }
    st.coef[0] = d;
    st.coef[1] = plus_coupling(mn[1], mc, sx);
    st.coef[2] = plus_coupling(mc, mn[2], sx);
    st.coef[3] = plus_coupling(mn[3], mc, sy);
    st.coef[4] = plus_coupling(mc, mn[4], sy);
    st.coef[5] = plus_coupling(mn[5], mc, sz);
    st.coef[6] = plus_coupling(mc, mn[6], sz);
    return st;
}

// fdmMatrixVectorMul (MPhysics/EulerSolverTool.cu:199-227): products in FP32,
// summed in FP64 in the reference's order, stored as FP32.
__global__ void __launch_bounds__(kBlock) k_matvec(GridDev g, const uint8_t* __restrict__ m, float sx,
                                                   float sy, float sz, const float* __restrict__ xv,
                                                   float* __restrict__ y, uint32_t n) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int x, yy, z;
    decode3(uint32_t(i), uint32_t(nx), uint32_t(ny), g.r_nx, g.r_nxy, x, yy, z);
    const Stencil st = stencil_at(m, x, yy, z, nx, ny, nz, sx, sy, sz);
    double acc = double(st.coef[0] * xv[st.idx[0]]);
#pragma unroll
    for (int q = 1; q < 7; ++q) acc = acc + (st.on[q] ? double(st.coef[q] * xv[st.idx[q]]) : 0.0);
    y[i] = float(acc);
}

// ---------------------------------------------------------------------------
// Plain BLAS-1 (the reference's cuBLAS calls, CuDenseVector.cpp:120-240).
__global__ void __launch_bounds__(kBlock) k_dot_partials(const float* __restrict__ a,
                                                         const float* __restrict__ b, size_t n,
                                                         double* partials) {
    double s[1] = {0.0};
    const size_t stride = size_t(gridDim.x) * kBlock;
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        s[0] += double(a[i]) * double(b[i]);
    block_sum<1>(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kBlock) k_axpy(float a, const float* __restrict__ x,
                                                 float* __restrict__ y, size_t n) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) y[i] = __builtin_fmaf(a, x[i], y[i]);
}

// m_P *= beta; m_P += z  (scale, then axpy with alpha 1: two roundings)
__global__ void __launch_bounds__(kBlock) k_scal_add(float beta, float* __restrict__ p,
                                                     const float* __restrict__ z, size_t n) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) {
        const float t = beta * p[i];
        p[i] = __builtin_fmaf(1.0f, z[i], t);
    }
}

__global__ void __launch_bounds__(kBlock) k_fill(float* a, size_t n, float v) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_fill64(double* a, size_t n, double v) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_u8_to_f32(const uint8_t* a, float* b, size_t n) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}

__global__ void __launch_bounds__(kBlock) k_f64_to_f32(const double* a, float* b, size_t n) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}


// FP64 ||b - A x||^2 and ||b||^2 partials for an FP32 pressure.
__global__ void __launch_bounds__(kBlock) k_residual32(GridDev g, const uint8_t* __restrict__ m,
                                                       float sx, float sy, float sz,
                                                       const float* __restrict__ xv,
                                                       const float* __restrict__ b,
                                                       double* partials) {
    const uint32_t n = uint32_t(g.nx) * uint32_t(g.ny) * uint32_t(g.nz);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double s[2] = {0.0, 0.0};
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        int x, y, z;
        decode3(uint32_t(i), uint32_t(nx), uint32_t(ny), g.r_nx, g.r_nxy, x, y, z);
        const Stencil st = stencil_at(m, x, y, z, nx, ny, nz, sx, sy, sz);
        double ax = double(st.coef[0]) * double(xv[st.idx[0]]);
#pragma unroll
        for (int q = 1; q < 7; ++q)
            if (st.on[q]) ax += double(st.coef[q]) * double(xv[st.idx[q]]);
        const double bi = double(b[i]);
        const double ri = bi - ax;
        s[0] += ri * ri;
        s[1] += bi * bi;
    }
    block_sum<2>(s);
    if (threadIdx.x == 0) {
        partials[2 * size_t(blockIdx.x)] = s[0];
        partials[2 * size_t(blockIdx.x) + 1] = s[1];
    }
}

// ---------------------------------------------------------------------------
// applyPressureGradient (MPhysics/EulerSolverTool.cu:229-322).  With an FP64
// pressure the update is formed in FP64 and rounded once.
__device__ __forceinline__ void k_project_run(GridDev g, const uint8_t* __restrict__ m,
                                                    float sx, float sy, float sz,
                                                    const float* __restrict__ p32,
                                                    const double* __restrict__ p64,
                                                    float* __restrict__ u, float* __restrict__ v,
                                                    float* __restrict__ w,
                                                    const float* __restrict__ us,
                                                    const float* __restrict__ vs,
                                                    const float* __restrict__ ws, uint32_t n, uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i >= n) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int x, y, z;
    decode3(uint32_t(i), uint32_t(nx), uint32_t(ny), g.r_nx, g.r_nxy, x, y, z);
    const uint32_t l = lin3(x - 1, y, z, nx, ny, nz), d = lin3(x, y - 1, z, nx, ny, nz),
                   bk = lin3(x, y, z - 1, nx, ny, nz);
    const uint8_t mc = m[i], ml = m[l], md = m[d], mb = m[bk];
    const uint32_t fu = lin3(x, y, z, nx + 1, ny, nz), fv = lin3(x, y, z, nx, ny + 1, nz),
                   fw = lin3(x, y, z, nx, ny, nz + 1);
    struct F {
        uint8_t mo;
        uint32_t o;
        float s;
        float* a;
        const float* as;
        uint32_t f;
    } faces[3] = {{ml, l, sx, u, us, fu}, {md, d, sy, v, vs, fv}, {mb, bk, sz, w, ws, fw}};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const F& F_ = faces[q];
        if (F_.mo == kFluid || mc == kFluid) {
            if (F_.mo == kSolid || mc == kSolid) {
                F_.a[F_.f] = F_.as[F_.f];
            } else if (p64) {
                F_.a[F_.f] = float(double(F_.a[F_.f]) - double(F_.s) * (p64[i] - p64[F_.o]));
            } else {
                F_.a[F_.f] = F_.a[F_.f] - F_.s * (p32[i] - p32[F_.o]);
            }
        } else {
            F_.a[F_.f] = MGF_UNKNOWN;
        }
    }
    if (x == nx - 1) u[lin3(x + 1, y, z, nx + 1, ny, nz)] = MGF_UNKNOWN;
    if (y == ny - 1) v[lin3(x, y + 1, z, nx, ny + 1, nz)] = MGF_UNKNOWN;
    if (z == nz - 1) w[lin3(x, y, z + 1, nx, ny, nz + 1)] = MGF_UNKNOWN;
}

__global__ void __launch_bounds__(kBlock) k_project(GridDev g, const uint8_t* __restrict__ m,
                                                    float sx, float sy, float sz,
                                                    const float* __restrict__ p32,
                                                    const double* __restrict__ p64,
                                                    float* __restrict__ u, float* __restrict__ v,
                                                    float* __restrict__ w,
                                                    const float* __restrict__ us,
                                                    const float* __restrict__ vs,
                                                    const float* __restrict__ ws, uint32_t n) {
    k_project_run(g, m, sx, sy, sz, p32, p64, u, v, w, us, vs, ws, n, xcd_block());
}

struct MProject {
    static constexpr int kThreads = kBlock;
    GridDev g;
    const uint8_t* m;
    float sx;
    float sy;
    float sz;
    const float* p32;
    const double* p64;
    float* u;
    float* v;
    float* w;
    const float* us;
    const float* vs;
    const float* ws;
    uint32_t n;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_project_run(g, m, sx, sy, sz, p32, p64, u, v, w, us, vs, ws, n, xcd_of(vb, G)); }
};

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Boundaries (MPhysics/Boundary.cpp:115-150, MPhysics/BoundaryTool.cu:3-113).
__global__ void __launch_bounds__(kBlock) k_resample_solid(GridDev g, const float* __restrict__ sdf0,
                                                           float tx, float ty, float tz,
                                                           float* __restrict__ out, uint32_t n) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t nx = uint32_t(g.nx), ny = uint32_t(g.ny);
    int x, y, z;
    decode3(i, nx, ny, g.r_nx, g.r_nxy, x, y, z);
    const float cx = g.ox + 0.5f * g.hx + g.hx * float(x);
    const float cy = g.oy + 0.5f * g.hy + g.hy * float(y);
    const float cz = g.oz + 0.5f * g.hz + g.hz * float(z);
    out[i] = trilerp(sdf0, g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g, cx + tx, cy + ty,
                     cz + tz);
}

__global__ void __launch_bounds__(kBlock) k_solids_marker(size_t n, const float* __restrict__ sdf,
                                                          int index1, float* __restrict__ marker) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i >= n) return;
    float mk = marker[i];
    if (index1 == 1) mk = MGF_UNKNOWN;
    if (sdf[i] < 0.0f && mk == MGF_UNKNOWN) mk = float(-index1);
    marker[i] = mk;
}

__global__ void __launch_bounds__(kBlock) k_solid_mask(size_t n, const float* __restrict__ total,
                                                       uint8_t* __restrict__ mask) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) mask[i] = total[i] < 0.0f ? 1 : 0;
}

struct Vel3Table {
    float v[3 * 64];
};

// Solid face velocity: a face takes its higher-index adjacent solid cell's
// velocity, else its lower-index one's, else 0 (the order in which the
// reference's racing writes land when cells are visited by index).
__global__ void __launch_bounds__(kBlock) k_solids_vel(GridDev g, const float* __restrict__ total,
                                                       Vel3Table tab, int nsolids,
                                                       float* __restrict__ us,
                                                       float* __restrict__ vs,
                                                       float* __restrict__ ws, uint32_t nfx,
                                                       uint32_t nfy, uint32_t nfz) {
    uint32_t i = xcd_block() * kBlock + threadIdx.x;
    int c;
    if (i < nfx) c = 0;
    else if (i < nfx + nfy) { c = 1; i -= nfx; }
    else if (i < nfx + nfy + nfz) { c = 2; i -= nfx + nfy; }
    else return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const uint32_t ex = uint32_t(nx + (c == 0)), ey = uint32_t(ny + (c == 1));
    int x, y, z;
    decode3(i, ex, ey, c == 0 ? g.r_nx1 : g.r_nx, c == 0 ? g.r_nx1y : (c == 1 ? g.r_nxy1 : g.r_nxy), x, y, z);
    int cell[2][3] = {{x, y, z}, {x, y, z}};  // [0] upper (same index), [1] lower
    cell[1][c] -= 1;
    const int up_lim = (c == 0 ? nx : (c == 1 ? ny : nz));
    float val = 0.0f;
    bool set = false;
    for (int q = 0; q < 2 && !set; ++q) {
        const int a = cell[q][c];
        if (a < 0 || a >= up_lim) continue;
        const float mk = total[(uint32_t(cell[q][2]) * uint32_t(ny) + uint32_t(cell[q][1])) * uint32_t(nx) +
                               uint32_t(cell[q][0])];
        if (!(mk < 0.0f)) continue;
        const int s = int(-(1.0f + mk));
        if (s >= nsolids || s < 0) continue;
        val = tab.v[3 * s + c];
        set = true;
    }
    float* dst = c == 0 ? us : (c == 1 ? vs : ws);
    dst[i] = val;
}

// generateFluidDomainFromBBox (MPhysics/EulerSolverTool.cu:581-608).
__global__ void __launch_bounds__(kBlock) k_fluid_box(GridDev g, float ax, float ay, float az,
                                                      float bx, float by, float bz,
                                                      float* __restrict__ phi, uint32_t n) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t nx = uint32_t(g.nx), ny = uint32_t(g.ny);
    int x, y, z;
    decode3(i, nx, ny, g.r_nx, g.r_nxy, x, y, z);
    const float cx = g.ox + 0.5f * g.hx + g.hx * float(x);
    const float cy = g.oy + 0.5f * g.hy + g.hy * float(y);
    const float cz = g.oz + 0.5f * g.hz + g.hz * float(z);
    const bool in = cx > ax && cy > ay && cz > az && cx < bx && cy < by && cz < bz;
    phi[i] = in ? -500.0f : 500.0f;
}

// ===========================================================================
// Launchers.
RkCoef rk_coef(float dt) {
    RkCoef k;
    k.c[0] = -dt;
    k.c[1] = -0.5f * dt;
    k.c[2] = -0.75f * dt;
    k.c[3] = -float(2.0 / 9.0) * dt;
    k.c[4] = -float(3.0 / 9.0) * dt;
    k.c[5] = -float(4.0 / 9.0) * dt;
    return k;
}



hipError_t launch_backtrace(hipStream_t s, const GridDev& g, Vel V, float dt, int order, float* px,
                            float* py, float* pz, int64_t n) {
    if (n <= 0) return hipSuccess;
    return LAUNCH(k_backtrace, grid_for(size_t(n)), g, V, rk_coef(dt), order, px, py, pz, n);
}

hipError_t launch_sample_cc(hipStream_t s, const GridDev& g, const float* f, const float* px,
                            const float* py, const float* pz, float* out, int64_t n) {
    if (n <= 0) return hipSuccess;
    return LAUNCH(k_sample_cc, grid_for(size_t(n)), g, f, px, py, pz, out, n);
}

hipError_t launch_advect_fc(hipStream_t s, const GridDev& g, Vel in, Vel V, float dt, int order,
                            float* ou, float* ov, float* ow, int add0, const float inc0[3], int add1,
                            const float inc1[3]) {
    const uint32_t nfx = uint32_t(g.nx + 1) * g.ny * g.nz, nfy = uint32_t(g.nx) * (g.ny + 1) * g.nz,
                   nfz = uint32_t(g.nx) * g.ny * (g.nz + 1);
    Inc inc{};
    for (int q = 0; q < 3; ++q) {
        inc.a[q] = inc0 ? inc0[q] : 0.0f;
        inc.b[q] = inc1 ? inc1[q] : 0.0f;
    }
    inc.add_a = add0;
    inc.add_b = add1;
    return LAUNCH(k_advect_fc, grid_for(size_t(nfx) + nfy + nfz), g, in, V, rk_coef(dt), order, ou, ov,
                  ow, nfx, nfy, nfz, inc);
}

hipError_t launch_advect_cc(hipStream_t s, const GridDev& g, const float* in, Vel V, float dt,
                            int order, float* out) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_advect_cc, grid_for(n), g, in, V, rk_coef(dt), order, out, n);
}

hipError_t launch_advect_cc4(hipStream_t s, const GridDev& g, const float* phi, const float* const c[3], Vel V,
                             float dt, int order, float* ophi, float* const oc[3]) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    Cc4 F;
    F.in[0] = phi;
    F.out[0] = ophi;
    for (int q = 0; q < 3; ++q) {
        F.in[q + 1] = c[q];
        F.out[q + 1] = oc[q];
    }
    return LAUNCH(k_advect_cc4, grid_for(n), g, F, V, rk_coef(dt), order, n);
}

hipError_t launch_add_uniform(hipStream_t s, float* a, size_t n, float inc) {
    return LAUNCH(k_add_uniform, grid_for(n), a, n, inc);
}

hipError_t launch_markers(hipStream_t s, const GridDev& g, const uint8_t* solid_mask,
                          const float* fluid, uint8_t* markers) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_markers, grid_for(n), n, solid_mask, fluid, markers);
}

hipError_t launch_rhs(hipStream_t s, const GridDev& g, Vel U, Vel S, const uint8_t* markers,
                      float* b) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_rhs, grid_for(n), g, U, S, markers, 1.0f / g.hx, 1.0f / g.hy, 1.0f / g.hz, b, n);
}

hipError_t launch_matrix_a(hipStream_t s, const GridDev& g, const float scale[3],
                           const uint8_t* markers, float* A) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_matrix_a, grid_for(n), g, markers, scale[0], scale[1], scale[2], A, n);
}

hipError_t launch_divergence(hipStream_t s, const GridDev& g, Vel U, float* div) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_divergence, grid_for(n), g, U, div, n);
}

hipError_t launch_matvec(hipStream_t s, const GridDev& g, const float scale[3],
                         const uint8_t* markers, const float* x, float* y) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_matvec, grid_for(n), g, markers, scale[0], scale[1], scale[2], x, y, n);
}

hipError_t launch_dot_partials(hipStream_t s, const float* a, const float* b, size_t n,
                               double* partials, int blocks) {
    return LAUNCH(k_dot_partials, blocks, a, b, n, partials);
}

hipError_t launch_axpy(hipStream_t s, float alpha, const float* x, float* y, size_t n) {
    return LAUNCH(k_axpy, grid_for(n), alpha, x, y, n);
}

hipError_t launch_scal_add(hipStream_t s, float beta, float* p, const float* z, size_t n) {
    return LAUNCH(k_scal_add, grid_for(n), beta, p, z, n);
}

hipError_t launch_fill(hipStream_t s, float* a, size_t n, float v) {
    if (n == 0) return hipSuccess;
    return LAUNCH(k_fill, grid_for(n), a, n, v);
}

hipError_t launch_fill64(hipStream_t s, double* a, size_t n, double v) {
    if (n == 0) return hipSuccess;
    return LAUNCH(k_fill64, grid_for(n), a, n, v);
}

hipError_t launch_u8_to_f32(hipStream_t s, const uint8_t* a, float* b, size_t n) {
    return LAUNCH(k_u8_to_f32, grid_for(n), a, b, n);
}

hipError_t launch_f64_to_f32(hipStream_t s, const double* a, float* b, size_t n) {
    return LAUNCH(k_f64_to_f32, grid_for(n), a, b, n);
}

hipError_t launch_true_residual32(hipStream_t s, const GridDev& g, const float scale[3],
                                  const uint8_t* markers, const float* x, const float* b,
                                  double* partials, int blocks) {
    return LAUNCH(k_residual32, blocks, g, markers, scale[0], scale[1], scale[2], x, b, partials);
}

hipError_t launch_project(hipStream_t s, const GridDev& g, const float sg[3], const uint8_t* markers,
                          const float* p32, const double* p64, float* u, float* v, float* w,
                          const float* us, const float* vs, const float* ws) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_project, grid_for(n), g, markers, sg[0], sg[1], sg[2], p32, p64, u, v, w, us, vs,
                  ws, n);
}

hipError_t launch_resample_solid(hipStream_t s, const GridDev& g, const float* sdf0, const float t[3],
                                 float* out) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_resample_solid, grid_for(n), g, sdf0, t[0], t[1], t[2], out, n);
}

hipError_t launch_solids_marker(hipStream_t s, size_t n, const float* sdf, int index1, float* marker) {
    return LAUNCH(k_solids_marker, grid_for(n), n, sdf, index1, marker);
}

hipError_t launch_solid_mask(hipStream_t s, size_t n, const float* total, uint8_t* mask) {
    return LAUNCH(k_solid_mask, grid_for(n), n, total, mask);
}

hipError_t launch_solids_vel(hipStream_t s, const GridDev& g, const float* total, const float* vel3,
                             int nsolids, float* us, float* vs, float* ws) {
    Vel3Table tab{};
    const int m = nsolids < 64 ? nsolids : 64;
    for (int q = 0; q < 3 * m; ++q) tab.v[q] = vel3[q];
    const uint32_t nfx = uint32_t(g.nx + 1) * g.ny * g.nz, nfy = uint32_t(g.nx) * (g.ny + 1) * g.nz,
                   nfz = uint32_t(g.nx) * g.ny * (g.nz + 1);
    return LAUNCH(k_solids_vel, grid_for(size_t(nfx) + nfy + nfz), g, total, tab, m, us, vs, ws, nfx,
                  nfy, nfz);
}

hipError_t launch_fluid_box(hipStream_t s, const GridDev& g, const float mn[3], const float mx[3],
                            float* phi) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    return LAUNCH(k_fluid_box, grid_for(n), g, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], phi, n);
}

// ===========================================================================
// Multi-scene records of the launches launch_advect_fc, launch_advect_cc,
// launch_advect_cc4, launch_markers, launch_rhs and launch_project make (the
// same arguments and grids).
template <class A>
static void push_rec_k(std::vector<MRec>& out, int32_t kind, const A& a, uint32_t blocks) {
    static_assert(sizeof(A) <= kMRecBytes, "argument struct fits a record");
    MRec r;
    r.kind = kind;
    r.blocks = blocks;
    r.bytes = uint32_t(sizeof(A));
    std::memcpy(r.args, &a, sizeof(A));
    out.push_back(r);
}

void rec_advect_fc(std::vector<MRec>& out, const GridDev& g, Vel in, Vel V, float dt, int order, float* ou, float* ov,
                   float* ow, int add0, const float inc0[3], int add1, const float inc1[3]) {
    const uint32_t nfx = uint32_t(g.nx + 1) * g.ny * g.nz, nfy = uint32_t(g.nx) * (g.ny + 1) * g.nz,
                   nfz = uint32_t(g.nx) * g.ny * (g.nz + 1);
    Inc inc{};
    for (int q = 0; q < 3; ++q) {
        inc.a[q] = inc0 ? inc0[q] : 0.0f;
        inc.b[q] = inc1 ? inc1[q] : 0.0f;
    }
    inc.add_a = add0;
    inc.add_b = add1;
    const uint32_t G = grid_for(size_t(nfx) + nfy + nfz);
    push_rec_k(out, kMkAdvectFc, MAdvectFc{g, in, V, rk_coef(dt), order, ou, ov, ow, nfx, nfy, nfz, inc, G}, G);
}

void rec_advect_cc(std::vector<MRec>& out, const GridDev& g, const float* in, Vel V, float dt, int order, float* o) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    const uint32_t G = grid_for(n);
    push_rec_k(out, kMkAdvectCc, MAdvectCc{g, in, V, rk_coef(dt), order, o, n, G}, G);
}

void rec_advect_cc4(std::vector<MRec>& out, const GridDev& g, const float* phi, const float* const c[3], Vel V,
                    float dt, int order, float* ophi, float* const oc[3]) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    Cc4 F;
    F.in[0] = phi;
    F.out[0] = ophi;
    for (int q = 0; q < 3; ++q) {
        F.in[q + 1] = c[q];
        F.out[q + 1] = oc[q];
    }
    const uint32_t G = grid_for(n);
    push_rec_k(out, kMkAdvectCc4, MAdvectCc4{g, F, V, rk_coef(dt), order, n, G}, G);
}

void rec_markers(std::vector<MRec>& out, const GridDev& g, const uint8_t* solid_mask, const float* fluid,
                 uint8_t* markers) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    const uint32_t G = grid_for(n);
    push_rec_k(out, kMkMarkers, MMarkers{n, solid_mask, fluid, markers, G}, G);
}

void rec_rhs(std::vector<MRec>& out, const GridDev& g, Vel U, Vel S, const uint8_t* markers, float* b) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    const uint32_t G = grid_for(n);
    push_rec_k(out, kMkRhs, MRhs{g, U, S, markers, 1.0f / g.hx, 1.0f / g.hy, 1.0f / g.hz, b, n, G}, G);
}

void rec_project(std::vector<MRec>& out, const GridDev& g, const float sg[3], const uint8_t* markers, const float* p32,
                 const double* p64, float* u, float* v, float* w, const float* us, const float* vs, const float* ws) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    const uint32_t G = grid_for(n);
    push_rec_k(out, kMkProject, MProject{g, markers, sg[0], sg[1], sg[2], p32, p64, u, v, w, us, vs, ws, n, G}, G);
}

// Copies one 32-bit word per scene into a packed array (one work-item).
struct MCopyWord {
    static constexpr int kThreads = 64;
    const uint32_t* src;
    uint32_t* dst;
    __device__ void run(uint32_t) const {
        if (threadIdx.x == 0) *dst = *src;
    }
};

void rec_copy_word(std::vector<MRec>& out, const uint32_t* src, uint32_t* dst) {
    push_rec_k(out, kMkCopyWord, MCopyWord{src, dst}, 1u);
}

template <class A>
static hipError_t launch_mk(hipStream_t s, const char* name, const void* tab, const uint32_t* map, uint32_t nblocks,
                            const uint8_t* active) {
    return launch_dims(k_multi<A>, name, dim3(nblocks), dim3(A::kThreads), 0u, s, static_cast<const A*>(tab), map,
                       active);
}

hipError_t launch_multi_k(hipStream_t s, int32_t kind, const void* tab, const uint32_t* map, uint32_t nblocks,
                          uint32_t shmem, const uint8_t* active) {
    switch (kind) {
        case kMkAdvectFc: return launch_mk<MAdvectFc>(s, "k_multi_advect_fc", tab, map, nblocks, active);
        case kMkAdvectCc: return launch_mk<MAdvectCc>(s, "k_multi_advect_cc", tab, map, nblocks, active);
        case kMkAdvectCc4: return launch_mk<MAdvectCc4>(s, "k_multi_advect_cc4", tab, map, nblocks, active);
        case kMkMarkers: return launch_mk<MMarkers>(s, "k_multi_markers", tab, map, nblocks, active);
        case kMkRhs: return launch_mk<MRhs>(s, "k_multi_rhs", tab, map, nblocks, active);
        case kMkProject: return launch_mk<MProject>(s, "k_multi_project", tab, map, nblocks, active);
        case kMkCopyWord: return launch_mk<MCopyWord>(s, "k_multi_copy_word", tab, map, nblocks, active);
        default: return launch_multi_ext(s, kind, tab, map, nblocks, shmem, active);
    }
}

}  // namespace mgf
