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

// ---- particle generation (generateParticlesInFluid, EulerParticles.cpp:
// 163-201; findFluidGrid, EulerParticlesTool.cu:5-25) -----------------------
// A tile of kGenTile consecutive cells per work-group: the count pass sums the
// fluid cells of each tile, one work-group scans the tile sums, and the emit
// pass ranks the fluid cells inside its tile (wave ballots, x-fastest order)
// and writes their particles.
constexpr int kGenRows = 8;
constexpr int kGenTile = kBlock * kGenRows;

__device__ __forceinline__ bool fluid_cell(const float* fluid, const float* solid, size_t c) {
    return fluid[c] < 0.0f && !(solid && solid[c] < 0.0f);
}

// Fluid cells of this work-group's tile among rows [0, kGenRows): for each row
// the rank of this work-item's cell among the row's fluid cells and the row
// total.  wsum holds one count per wave.
__device__ __forceinline__ void row_rank(bool f, uint32_t* wsum, uint32_t& before, uint32_t& total) {
    const uint64_t b = __ballot(f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = uint32_t(__popcll(b));
    __syncthreads();
    uint32_t pre = uint32_t(__popcll(b & ((uint64_t(1) << lane) - 1)));
    total = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) {
        if (q < wave) pre += wsum[q];
        total += wsum[q];
    }
    before = pre;
    __syncthreads();
}

__global__ void __launch_bounds__(kBlock) k_gen_count(size_t n, const float* __restrict__ fluid,
                                                      const float* __restrict__ solid, uint32_t* __restrict__ tiles) {
    __shared__ uint32_t wsum[kBlock / 64];
    const size_t base = size_t(blockIdx.x) * kGenTile;
    uint32_t sum = 0;
    for (int j = 0; j < kGenRows; ++j) {
        const size_t c = base + size_t(j) * kBlock + threadIdx.x;
        uint32_t before, total;
        row_rank(c < n && fluid_cell(fluid, solid, c), wsum, before, total);
        sum += total;
    }
    if (threadIdx.x == 0) tiles[blockIdx.x] = sum;
}

// In place: tiles[0..nt) -> exclusive prefix sums, tiles[nt] = the total.
__global__ void __launch_bounds__(1024) k_gen_scan(uint32_t nt, uint32_t* tiles) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x, per = (nt + 1023) / 1024;
    const uint32_t b = min(nt, t * per), e = min(nt, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += tiles[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {   // inclusive Hillis-Steele scan
        const uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = tiles[i];
        tiles[i] = run;
        run += v;
    }
    if (t == 1023) tiles[nt] = part[1023];
}

// SplitMix64's output function (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014).
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float particle_jitter(uint64_t seed, uint64_t index, int axis) {
    return float(splitmix64(seed ^ splitmix64(3 * index + uint64_t(axis))) >> 40) * (1.0f / 16777216.0f);
}

__global__ void __launch_bounds__(kBlock) k_gen_emit(GridDev g, size_t n, const float* __restrict__ fluid,
                                                     const float* __restrict__ solid,
                                                     const uint32_t* __restrict__ tiles, int per_cell, uint64_t seed,
                                                     float* __restrict__ pos) {
    __shared__ uint32_t wsum[kBlock / 64];
    const size_t base = size_t(blockIdx.x) * kGenTile;
    uint32_t rank = tiles[blockIdx.x];
    for (int j = 0; j < kGenRows; ++j) {
        const size_t c = base + size_t(j) * kBlock + threadIdx.x;
        const bool f = c < n && fluid_cell(fluid, solid, c);
        uint32_t before, total;
        row_rank(f, wsum, before, total);
        if (f) {
            const uint32_t x = uint32_t(c % uint32_t(g.nx));
            const uint32_t y = uint32_t((c / uint32_t(g.nx)) % uint32_t(g.ny));
            const uint32_t z = uint32_t(c / (size_t(g.nx) * uint32_t(g.ny)));
            const uint64_t first = uint64_t(rank + before) * uint64_t(per_cell);
            for (int i = 0; i < per_cell; ++i) {
                const uint64_t idx = first + uint64_t(i);
                pos[3 * idx] = (float(x) + particle_jitter(seed, idx, 0)) * g.hx + g.ox;
                pos[3 * idx + 1] = (float(y) + particle_jitter(seed, idx, 1)) * g.hy + g.oy;
                pos[3 * idx + 2] = (float(z) + particle_jitter(seed, idx, 2)) * g.hz + g.oz;
            }
        }
        rank += total;
    }
}

// m_DeltaGridVelField.plusAlphaX(in, -1); m_DeltaGridVelField *= -1
// (FluidImplicitParticle.cpp:105-106): one axpy with alpha -1, then a sign flip.
__global__ void __launch_bounds__(kBlock) k_flip_delta(size_t n, const float* __restrict__ in, float* delta) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) delta[i] = -__builtin_fmaf(-1.0f, in[i], delta[i]);
}

// Pressure markers from particles (buildFluidMarkers, EulerParticlesTool.cu:
// 48-77): 2 in solid cells, then 1 in every other cell holding a particle.
__global__ void __launch_bounds__(kBlock) k_markers_solid(size_t n, const uint8_t* __restrict__ solid_mask,
                                                          uint8_t* __restrict__ markers) {
    const size_t i = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (i < n) markers[i] = solid_mask[i] ? uint8_t(2) : uint8_t(0);
}
__global__ void __launch_bounds__(kBlock) k_markers_particles(PGrid g, int64_t np, const float* __restrict__ pos,
                                                              uint8_t* markers) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    const uint32_t c = pos_cell(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    if (markers[c] != 2) markers[c] = 1;
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

namespace mgf {

uint32_t gen_tiles(size_t ncell) { return uint32_t((ncell + kGenTile - 1) / kGenTile); }

hipError_t launch_gen_count(hipStream_t s, size_t ncell, const float* fluid, const float* solid, uint32_t* tiles) {
    const uint32_t nt = gen_tiles(ncell);
    hipLaunchKernelGGL(k_gen_count, dim3(nt), dim3(kBlock), 0, s, ncell, fluid, solid, tiles);
    hipLaunchKernelGGL(k_gen_scan, dim3(1), dim3(1024), 0, s, nt, tiles);
    return hipGetLastError();
}

hipError_t launch_gen_emit(hipStream_t s, const GridDev& g, const float* fluid, const float* solid,
                           const uint32_t* tiles, int per_cell, uint64_t seed, float* pos) {
    const size_t ncell = size_t(g.nx) * g.ny * g.nz;
    hipLaunchKernelGGL(k_gen_emit, dim3(gen_tiles(ncell)), dim3(kBlock), 0, s, g, ncell, fluid, solid, tiles,
                       per_cell, seed, pos);
    return hipGetLastError();
}

hipError_t launch_flip_delta(hipStream_t s, size_t n, const float* in, float* delta) {
    if (n == 0) return hipSuccess;
    return LAUNCH(k_flip_delta, grid_for(n), n, in, delta);
}

hipError_t launch_particle_markers_u8(hipStream_t s, const GridDev& g, const uint8_t* solid_mask, int64_t np,
                                      const float* pos, uint8_t* markers) {
    const size_t n = size_t(g.nx) * g.ny * g.nz;
    hipError_t e = LAUNCH(k_markers_solid, grid_for(n), n, solid_mask, markers);
    if (e == hipSuccess && np > 0)
        e = LAUNCH(k_markers_particles, grid_for(size_t(np)), pgrid_of(g, 0), np, pos, markers);
    return e;
}

}  // namespace mgf
