// mgf_kernels_particles.hip — particle <-> grid transfers and particle
// advection of CEulerParticles (MPhysics/EulerParticles.cpp,
// MPhysics/EulerParticlesTool.cu), the primitives of the PIC / FLIP / mixed
// advection solvers (MPhysics/ParticleInCell.cpp, FluidImplicitParticle.cpp,
// MixPICAndFLIP.cpp).
//
// Particle positions and per-particle vectors are interleaved x, y, z
// (3 * i + k), as the reference stores them.  One work-item per particle.
// Particle-to-grid sums use FP32 atomics as the reference does, so their
// rounding depends on the order the atomics land in; grid-to-particle sums,
// counts and the advection are order-free and bitwise the oracle's.
#include "mgf_internal.h"
#include "mgf_kcommon.h"

namespace mgf {

// A cell-centred grid of extents n, origin o, spacing h: a field of cells, or
// one component of a face-centred field (its extents and origin shifted by
// -h/2 on its axis).
struct PGrid {
    int nx, ny, nz;
    float ox, oy, oz, hx, hy, hz;
};

// Transfer kernels (MPhysics/EulerMathTool.cu:268-294): written with double
// literals in the reference, so evaluated in double from the float argument
// and rounded once.
__device__ __forceinline__ float pg_linear(float v) {
    const float a = fabsf(v);
    return a < 1.0f ? float(1.0 - double(a)) : 0.0f;
}
__device__ __forceinline__ float pg_quadratic(float v) {
    const float a = fabsf(v);
    if (a < 0.5f) return float(0.75 - double(a * a));
    if (a < 1.5f) return float(0.5 * (1.5 - double(a)) * (1.5 - double(a)));
    return 0.0f;
}
__device__ __forceinline__ float pg_cubic(float v) {
    const float a = fabsf(v);
    if (a < 1.0f) return float(0.5 * double(a) * double(a) * double(a) - double(a * a) + 2.0 / 3.0);
    if (a < 2.0f) return float(1.0 / 6.0 * (2.0 - double(a)) * (2.0 - double(a)) * (2.0 - double(a)));
    return 0.0f;
}

// transPos2Index (EulerMathTool.cu:29-40): truncation of (p - o) / h, clamped.
__device__ __forceinline__ uint32_t pos_cell(const PGrid& g, float px, float py, float pz) {
    const int ix = clampi(int((px - g.ox) / g.hx), g.nx);
    const int iy = clampi(int((py - g.oy) / g.hy), g.ny);
    const int iz = clampi(int((pz - g.oz) / g.hz), g.nz);
    return (uint32_t(iz) * uint32_t(g.ny) + uint32_t(iy)) * uint32_t(g.nx) + uint32_t(ix);
}

// The support of a particle (EulerParticlesTool.cu:97-101, 216-220): lower
// corner and per-axis weights; n = 2 (linear) or 4 (corners i0-1 .. i0+2).
struct PStencil {
    int i0[3];
    float w[3][4];
};
__device__ __forceinline__ void pg_stencil(const PGrid& g, float px, float py, float pz, int alg, PStencil& st) {
    const float rel[3] = {px - g.ox - g.hx * 0.5f, py - g.oy - g.hy * 0.5f, pz - g.oz - g.hz * 0.5f};
    const float h[3] = {g.hx, g.hy, g.hz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        st.i0[k] = floor_ref(rel[k] / h[k]);
        const float f = (rel[k] - float(st.i0[k]) * h[k]) / h[k];
        if (alg == MGF_PG_LINEAR) {
            st.w[k][0] = pg_linear(f);
            st.w[k][1] = pg_linear(float(double(f) - 1.0));
            st.w[k][2] = st.w[k][3] = 0.0f;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float a = float(double(f) + 1.0 - double(q));
                st.w[k][q] = alg == MGF_PG_QUADRATIC ? pg_quadratic(a) : pg_cubic(a);
            }
        }
    }
}

__device__ __forceinline__ uint32_t pg_index(const PGrid& g, int x, int y, int z) {
    return lin3(x, y, z, g.nx, g.ny, g.nz);
}

// accumulateParticles2CCSField (EulerParticlesTool.cu:79-196).
__global__ void __launch_bounds__(kBlock) k_p2g(PGrid g, int64_t np, const float* __restrict__ pos,
                                                const float* __restrict__ val, int span, int off, int alg,
                                                float* field, float* weight) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    const float v = val[int64_t(span) * p + off];
    if (alg == MGF_PG_P2GSUM) {
        atomicAdd(&field[pos_cell(g, px, py, pz)], v);
        return;
    }
    PStencil st;
    pg_stencil(g, px, py, pz, alg, st);
    const int n = alg == MGF_PG_LINEAR ? 2 : 4, lo = alg == MGF_PG_LINEAR ? 0 : -1;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const float wt = st.w[0][x] * st.w[1][y] * st.w[2][z];
                const uint32_t c = pg_index(g, st.i0[0] + x + lo, st.i0[1] + y + lo, st.i0[2] + z + lo);
                atomicAdd(&field[c], wt * v);
                atomicAdd(&weight[c], wt);
            }
}

// normalizeCCSField (EulerParticlesTool.cu:290-308).
__global__ void __launch_bounds__(kBlock) k_p2g_normalize(size_t n, float* __restrict__ field,
                                                          const float* __restrict__ weight) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n && weight[i] > 0.0f) field[i] /= weight[i];
}

// accumulateCCSField2Particles (EulerParticlesTool.cu:198-288): summed from
// 0 in corner order z, y, x.
__global__ void __launch_bounds__(kBlock) k_g2p(PGrid g, int64_t np, const float* __restrict__ pos,
                                                const float* __restrict__ field, int alg, float* val, int span,
                                                int off) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    float acc = 0.0f;
    if (alg == MGF_PG_G2PNEAREST) {
        acc = field[pos_cell(g, px, py, pz)];
    } else if (alg == MGF_PG_LINEAR || alg == MGF_PG_QUADRATIC || alg == MGF_PG_CUBIC) {
        PStencil st;
        pg_stencil(g, px, py, pz, alg, st);
        const int n = alg == MGF_PG_LINEAR ? 2 : 4, lo = alg == MGF_PG_LINEAR ? 0 : -1;
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    const float wt = st.w[0][x] * st.w[1][y] * st.w[2][z];
                    acc += wt * field[pg_index(g, st.i0[0] + x + lo, st.i0[1] + y + lo, st.i0[2] + z + lo)];
                }
    }
    val[int64_t(span) * p + off] = acc;
}

// statisticalFluidDensity (EulerParticlesTool.cu:27-46): +1 per particle.
__global__ void __launch_bounds__(kBlock) k_particle_counts(PGrid g, int64_t np, const float* __restrict__ pos,
                                                            float* out) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    atomicAdd(&out[pos_cell(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2])], 1.0f);
}

// buildFluidMarkers (EulerParticlesTool.cu:48-77): 2 solid, 1 with a
// particle, 0 else.
__global__ void __launch_bounds__(kBlock) k_particle_markers(size_t n, const float* __restrict__ solid,
                                                             float* __restrict__ io) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i >= n) return;
    io[i] = solid[i] < 0.0f ? 2.0f : (io[i] > 0.0f ? 1.0f : 0.0f);
}

// The face-centred velocity at each particle (sampleFaceCenteredVectorField
// Invoker on a position array: the trilinear sampler per component).
__global__ void __launch_bounds__(kBlock) k_sample_particles(GridDev g, Vel V, int64_t np,
                                                             const float* __restrict__ pos, float* __restrict__ out) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    float a, b, c;
    sample_vel(V, g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], a, b, c);
    out[3 * p] = a;
    out[3 * p + 1] = b;
    out[3 * p + 2] = c;
}

// max |v| over the sampled components (getMaxValue / getMinValue through
// cublasI[sa]max/min, EulerParticles.cpp:251): the bit pattern of a
// non-negative float orders like its value.
__global__ void __launch_bounds__(kBlock) k_absmax(size_t n, const float* __restrict__ v, uint32_t* out) {
    const size_t i0 = size_t(blockIdx.x) * kBlock + threadIdx.x;
    uint32_t m = 0;
    for (size_t i = i0; i < n; i += size_t(gridDim.x) * kBlock) m = max(m, __float_as_uint(fabsf(v[i])));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), o, 64)));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// axpyReal (ThrustWapper.cu:281-292, cuBLAS axpy): y = a x + y as one fused
// multiply-add; scaleReal (:294-304): x = a x.
__global__ void __launch_bounds__(kBlock) k_axpy_n(size_t n, float a, const float* __restrict__ x, float* y) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) y[i] = __builtin_fmaf(a, x[i], y[i]);
}
__global__ void __launch_bounds__(kBlock) k_scale_n(size_t n, float a, float* x) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) x[i] = a * x[i];
}

// fixParticlesPosWithBoundarys (EulerParticlesTool.cu:310-333, 870-894):
// positions clamped to the grid box; the reference launches one work-item per
// cell, so only the first min(cells, particles) particles are clamped.
__global__ void __launch_bounds__(kBlock) k_clamp_particles(int64_t m, float* pos, float lx, float ly, float lz,
                                                            float hx, float hy, float hz) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= m) return;
    pos[3 * p] = fmaxf(lx, fminf(pos[3 * p], hx));
    pos[3 * p + 1] = fmaxf(ly, fminf(pos[3 * p + 1], hy));
    pos[3 * p + 2] = fmaxf(lz, fminf(pos[3 * p + 2], hz));
}

// ===========================================================================
static PGrid pgrid_of(const GridDev& g, int kind) {
    PGrid p{g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g.hx, g.hy, g.hz};
    if (kind == 1) { p.nx += 1; p.ox = g.ox - g.hx / 2; }
    if (kind == 2) { p.ny += 1; p.oy = g.oy - g.hy / 2; }
    if (kind == 3) { p.nz += 1; p.oz = g.oz - g.hz / 2; }
    return p;
}

size_t pg_field_count(const GridDev& g, int kind) {
    const PGrid p = pgrid_of(g, kind);
    return size_t(p.nx) * p.ny * p.nz;
}

hipError_t launch_p2g(hipStream_t s, const GridDev& g, int kind, int64_t np, const float* pos, const float* val,
                      int span, int off, int alg, float* field, float* weight, int normalize) {
    const PGrid p = pgrid_of(g, kind);
    const size_t n = size_t(p.nx) * p.ny * p.nz;
    hipError_t e = launch_fill(s, field, n, 0.0f);
    if (e == hipSuccess) e = launch_fill(s, weight, n, 0.0f);
    if (e == hipSuccess && np > 0)
        e = LAUNCH(k_p2g, grid_for(size_t(np)), p, np, pos, val, span, off, alg, field, weight);
    if (e == hipSuccess && normalize) e = LAUNCH(k_p2g_normalize, grid_for(n), n, field, weight);
    return e;
}

hipError_t launch_g2p(hipStream_t s, const GridDev& g, int kind, int64_t np, const float* pos, const float* field,
                      int alg, float* val, int span, int off) {
    if (np <= 0) return hipSuccess;
    return LAUNCH(k_g2p, grid_for(size_t(np)), pgrid_of(g, kind), np, pos, field, alg, val, span, off);
}

hipError_t launch_particle_markers(hipStream_t s, const GridDev& g, const float* solid, int64_t np, const float* pos,
                                   float* out) {
    const size_t n = size_t(g.nx) * g.ny * g.nz;
    hipError_t e = launch_fill(s, out, n, 0.0f);
    if (e == hipSuccess && np > 0) e = LAUNCH(k_particle_counts, grid_for(size_t(np)), pgrid_of(g, 0), np, pos, out);
    if (e == hipSuccess && solid) e = LAUNCH(k_particle_markers, grid_for(n), n, solid, out);
    return e;
}

hipError_t launch_sample_particles(hipStream_t s, const GridDev& g, Vel V, int64_t np, const float* pos, float* out) {
    if (np <= 0) return hipSuccess;
    return LAUNCH(k_sample_particles, grid_for(size_t(np)), g, V, np, pos, out);
}

hipError_t launch_absmax(hipStream_t s, size_t n, const float* v, uint32_t* out) {
    hipError_t e = hipMemsetAsync(out, 0, sizeof(uint32_t), s);
    if (e != hipSuccess || n == 0) return e;
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 2048);
    return LAUNCH(k_absmax, blocks, n, v, out);
}

hipError_t launch_axpy_n(hipStream_t s, size_t n, float a, const float* x, float* y) {
    if (n == 0) return hipSuccess;
    return LAUNCH(k_axpy_n, grid_for(n), n, a, x, y);
}

hipError_t launch_scale_n(hipStream_t s, size_t n, float a, float* x) {
    if (n == 0) return hipSuccess;
    return LAUNCH(k_scale_n, grid_for(n), n, a, x);
}

hipError_t launch_clamp_particles(hipStream_t s, const GridDev& g, int64_t np, float* pos) {
    const int64_t m = std::min<int64_t>(np, int64_t(g.nx) * g.ny * g.nz);
    if (m <= 0) return hipSuccess;
    const float hx = g.ox + float(g.nx) * g.hx, hy = g.oy + float(g.ny) * g.hy, hz = g.oz + float(g.nz) * g.hz;
    return LAUNCH(k_clamp_particles, grid_for(size_t(m)), m, pos, g.ox, g.oy, g.oz, hx, hy, hz);
}

}  // namespace mgf
