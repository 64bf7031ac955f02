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

// The same sums with a work-group's particles first merged in an LDS hash
// table keyed by (component, node): particles of a work-group are neighbours
// in index order (generation writes them cell by cell) and mostly share the
// nodes of their supports, so a table entry gathers several contributions and
// the flush issues one global atomic per distinct node.  Contributions that
// find no slot within kProbe probes go straight to global memory.  NC
// components (1: one field; 3: the three face grids of a velocity with
// values[3 p + c]) share the table; the key's top two bits are the component
// (every field holds fewer than 2^30 values).
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
constexpr int kProbe = 32;

struct P2GOut {
    float* field[3];
    float* weight[3];
};

template <int HB>
__device__ __forceinline__ void lds_accumulate(uint32_t* keys, float* sv, float* sw, uint32_t key, float a, float b,
                                               bool with_w, float* gf, float* gw) {
    constexpr uint32_t T = 1u << HB;
    uint32_t h = (key * 2654435761u) >> (32 - HB);
    for (int probe = 0; probe < kProbe; ++probe) {
        // a slot once claimed keeps its key: read first, claim empty slots only
        uint32_t k = __atomic_load_n(&keys[h], __ATOMIC_RELAXED);
        if (k == kEmptyKey) k = atomicCAS(&keys[h], kEmptyKey, key);
        if (k == kEmptyKey || k == key) {
            atomicAdd(&sv[h], a);
            if (with_w) atomicAdd(&sw[h], b);
            return;
        }
        h = (h + 1) & (T - 1);
    }
    atomicAdd(gf, a);
    if (with_w) atomicAdd(gw, b);
}

template <int NC, int HB>
__global__ void __launch_bounds__(kBlock) k_p2g_lds(PGrid g0, PGrid g1, PGrid g2, int64_t np,
                                                    const float* __restrict__ pos, const float* __restrict__ val,
                                                    int span, int off, int alg, P2GOut out) {
    constexpr uint32_t T = 1u << HB;
    __shared__ uint32_t keys[T];
    __shared__ float sv[T], sw[T];
    for (uint32_t i = threadIdx.x; i < T; i += kBlock) {
        keys[i] = kEmptyKey;
        sv[i] = 0.0f;
        sw[i] = 0.0f;
    }
    __syncthreads();
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    const bool with_w = alg != MGF_PG_P2GSUM;
    if (p < np) {
        const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const PGrid& g = c == 0 ? g0 : (c == 1 ? g1 : g2);
            const float v = val[int64_t(span) * p + (NC == 1 ? off : c)];
            const uint32_t tag = uint32_t(c) << 30;
            if (alg == MGF_PG_P2GSUM) {
                const uint32_t node = pos_cell(g, px, py, pz);
                lds_accumulate<HB>(keys, sv, sw, tag | node, v, 0.0f, false, &out.field[c][node], nullptr);
                continue;
            }
            PStencil st;
            pg_stencil(g, px, py, pz, alg, st);
            const int n = alg == MGF_PG_LINEAR ? 2 : 4, lo = alg == MGF_PG_LINEAR ? 0 : -1;
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y)
                    for (int x = 0; x < n; ++x) {
                        const float wt = st.w[0][x] * st.w[1][y] * st.w[2][z];
                        const uint32_t node = pg_index(g, st.i0[0] + x + lo, st.i0[1] + y + lo, st.i0[2] + z + lo);
                        lds_accumulate<HB>(keys, sv, sw, tag | node, wt * v, wt, true, &out.field[c][node],
                                           &out.weight[c][node]);
                    }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < T; i += kBlock) {
        const uint32_t k = keys[i];
        if (k == kEmptyKey) continue;
        const uint32_t c = k >> 30, node = k & 0x3FFFFFFFu;
        atomicAdd(&out.field[c][node], sv[i]);
        if (with_w) atomicAdd(&out.weight[c][node], sw[i]);
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

// The value of one field at a particle (the body of k_g2p).
__device__ __forceinline__ float g2p_one(const PGrid& g, float px, float py, float pz, const float* __restrict__ f,
                                         int alg) {
    if (alg == MGF_PG_G2PNEAREST) return f[pos_cell(g, px, py, pz)];
    PStencil st;
    pg_stencil(g, px, py, pz, alg, st);
    const int n = alg == MGF_PG_LINEAR ? 2 : 4, lo = alg == MGF_PG_LINEAR ? 0 : -1;
    float acc = 0.0f;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const float wt = st.w[0][x] * st.w[1][y] * st.w[2][z];
                acc += wt * f[pg_index(g, st.i0[0] + x + lo, st.i0[1] + y + lo, st.i0[2] + z + lo)];
            }
    return acc;
}

// The particle-velocity update of the three solvers in one pass, with the
// operations and rounding of the reference's separate calls:
//   PIC  (ParticleInCell.cpp:101):  v = G2P(in)
//   FLIP (FluidImplicitParticle.cpp:109-110): v = fma(1, G2P(delta), v)
//   MIX  (MixPICAndFLIP.cpp:114-129): p = G2P(in); v = fma(1, G2P(delta), v);
//        v = keep * v; v = fma(mixing, p, v)
struct Faces3 {
    const float* f[3];
};
__global__ void __launch_bounds__(kBlock) k_transfer_vel(PGrid g0, PGrid g1, PGrid g2, int64_t np,
                                                         const float* __restrict__ pos, Faces3 in, Faces3 delta,
                                                         int kind, int alg, float keep, float mixing,
                                                         float* __restrict__ vel) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const PGrid& g = c == 0 ? g0 : (c == 1 ? g1 : g2);
        if (kind == MGF_PSOLVER_PIC) {
            vel[3 * p + c] = g2p_one(g, px, py, pz, in.f[c], alg);
            continue;
        }
        const float pic = kind == MGF_PSOLVER_MIX ? g2p_one(g, px, py, pz, in.f[c], alg) : 0.0f;
        float v = __builtin_fmaf(1.0f, g2p_one(g, px, py, pz, delta.f[c], alg), vel[3 * p + c]);
        if (kind == MGF_PSOLVER_MIX) v = __builtin_fmaf(mixing, pic, keep * v);
        vel[3 * p + c] = v;
    }
}

// max over the work-group, then one atomic per work-group.
__device__ __forceinline__ void block_max_atomic(uint32_t m, uint32_t* out) {
    __shared__ uint32_t wmax[kBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), o, 64)));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = wmax[0];
#pragma unroll
        for (int q = 1; q < kBlock / 64; ++q) b = max(b, wmax[q]);
        atomicMax(out, b);
    }
}

// One particle advection substep in one pass (EulerParticles.cpp:273-313:
// the axpys of RK1/2/3 with the samples between them, same order and
// rounding), v1 = the velocity sampled at the positions it starts from.  With
// resample, the velocity at the new positions is written to v1 and its
// largest magnitude to vmax_bits, for the next substep's length.
__global__ void __launch_bounds__(kBlock) k_rk_substep(GridDev g, Vel V, int64_t np, float* __restrict__ pos,
                                                       float* __restrict__ v1, int order, float step, float half,
                                                       float three_q, float c1, float c2, float c3, int resample,
                                                       uint32_t* vmax_bits) {
    uint32_t m = 0;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x; p < np; p += stride) {
        const float x0 = pos[3 * p], y0 = pos[3 * p + 1], z0 = pos[3 * p + 2];
        const float ax = v1[3 * p], ay = v1[3 * p + 1], az = v1[3 * p + 2];
        float x, y, z;
        if (order == 1) {
            x = __builtin_fmaf(step, ax, x0); y = __builtin_fmaf(step, ay, y0); z = __builtin_fmaf(step, az, z0);
        } else if (order == 2) {
            float bx, by, bz;
            sample_vel(V, g, __builtin_fmaf(half, ax, x0), __builtin_fmaf(half, ay, y0), __builtin_fmaf(half, az, z0),
                       bx, by, bz);
            x = __builtin_fmaf(step, bx, x0); y = __builtin_fmaf(step, by, y0); z = __builtin_fmaf(step, bz, z0);
        } else if (order == 3) {
            float bx, by, bz, cx, cy, cz;
            sample_vel(V, g, __builtin_fmaf(half, ax, x0), __builtin_fmaf(half, ay, y0), __builtin_fmaf(half, az, z0),
                       bx, by, bz);
            sample_vel(V, g, __builtin_fmaf(three_q, bx, x0), __builtin_fmaf(three_q, by, y0),
                       __builtin_fmaf(three_q, bz, z0), cx, cy, cz);
            x = __builtin_fmaf(c3, cx, __builtin_fmaf(c2, bx, __builtin_fmaf(c1, ax, x0)));
            y = __builtin_fmaf(c3, cy, __builtin_fmaf(c2, by, __builtin_fmaf(c1, ay, y0)));
            z = __builtin_fmaf(c3, cz, __builtin_fmaf(c2, bz, __builtin_fmaf(c1, az, z0)));
        } else {
            x = x0; y = y0; z = z0;
        }
        pos[3 * p] = x; pos[3 * p + 1] = y; pos[3 * p + 2] = z;
        if (resample) {
            float a, b, c;
            sample_vel(V, g, x, y, z, a, b, c);
            v1[3 * p] = a; v1[3 * p + 1] = b; v1[3 * p + 2] = c;
            m = max(m, max(__float_as_uint(fabsf(a)), max(__float_as_uint(fabsf(b)), __float_as_uint(fabsf(c)))));
        }
    }
    if (resample) block_max_atomic(m, vmax_bits);
}

// The velocity at each particle and its largest magnitude (the first
// sample of a substep loop).
__global__ void __launch_bounds__(kBlock) k_sample_absmax(GridDev g, Vel V, int64_t np, const float* __restrict__ pos,
                                                          float* __restrict__ out, uint32_t* vmax_bits) {
    uint32_t m = 0;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x; p < np; p += stride) {
        float a, b, c;
        sample_vel(V, g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], a, b, c);
        out[3 * p] = a; out[3 * p + 1] = b; out[3 * p + 2] = c;
        m = max(m, max(__float_as_uint(fabsf(a)), max(__float_as_uint(fabsf(b)), __float_as_uint(fabsf(c)))));
    }
    block_max_atomic(m, vmax_bits);
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
    if (e == hipSuccess && np > 0) {
        const P2GOut out{{field, field, field}, {weight, weight, weight}};
        // table of 1024 slots for the 8-node supports, 2048 for the 64-node ones
        if (alg == MGF_PG_LINEAR || alg == MGF_PG_P2GSUM)
            e = LAUNCH((k_p2g_lds<1, 10>), grid_for(size_t(np)), p, p, p, np, pos, val, span, off, alg, out);
        else
            e = LAUNCH((k_p2g_lds<1, 11>), grid_for(size_t(np)), p, p, p, np, pos, val, span, off, alg, out);
    }
    if (e == hipSuccess && normalize) e = LAUNCH(k_p2g_normalize, grid_for(n), n, field, weight);
    return e;
}

hipError_t launch_p2g_velocity(hipStream_t s, const GridDev& g, int64_t np, const float* pos, const float* vel,
                               int alg, float* const field[3], float* const weight[3]) {
    PGrid p[3];
    size_t n[3];
    hipError_t e = hipSuccess;
    for (int c = 0; c < 3; ++c) {
        p[c] = pgrid_of(g, c + 1);
        n[c] = size_t(p[c].nx) * p[c].ny * p[c].nz;
        if (e == hipSuccess) e = launch_fill(s, field[c], n[c], 0.0f);
        if (e == hipSuccess) e = launch_fill(s, weight[c], n[c], 0.0f);
    }
    if (e == hipSuccess && np > 0) {
        const P2GOut out{{field[0], field[1], field[2]}, {weight[0], weight[1], weight[2]}};
        if (alg == MGF_PG_LINEAR || alg == MGF_PG_P2GSUM)
            e = LAUNCH((k_p2g_lds<3, 11>), grid_for(size_t(np)), p[0], p[1], p[2], np, pos, vel, 3, 0, alg, out);
        else
            e = LAUNCH((k_p2g_lds<3, 12>), grid_for(size_t(np)), p[0], p[1], p[2], np, pos, vel, 3, 0, alg, out);
    }
    for (int c = 0; c < 3; ++c)
        if (e == hipSuccess) e = LAUNCH(k_p2g_normalize, grid_for(n[c]), n[c], field[c], weight[c]);
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

namespace mgf {

// Work-groups of the kernels that reduce a maximum: a multiple of the XCD
// count, a few per CU, each looping over its share.
constexpr uint32_t kMaxReduceBlocks = 4096;

hipError_t launch_transfer_vel(hipStream_t s, const GridDev& g, int64_t np, const float* pos,
                               const float* const in[3], const float* const delta[3], int kind, int alg, float keep,
                               float mixing, float* vel) {
    if (np <= 0) return hipSuccess;
    const Faces3 fi{{in[0], in[1], in[2]}};
    const Faces3 fd{{delta ? delta[0] : nullptr, delta ? delta[1] : nullptr, delta ? delta[2] : nullptr}};
    return LAUNCH(k_transfer_vel, grid_for(size_t(np)), pgrid_of(g, 1), pgrid_of(g, 2), pgrid_of(g, 3), np, pos, fi,
                  fd, kind, alg, keep, mixing, vel);
}

hipError_t launch_sample_absmax(hipStream_t s, const GridDev& g, Vel V, int64_t np, const float* pos, float* out,
                                uint32_t* vmax_bits) {
    hipError_t e = hipMemsetAsync(vmax_bits, 0, sizeof(uint32_t), s);
    if (e != hipSuccess || np <= 0) return e;
    return LAUNCH(k_sample_absmax, std::min<uint32_t>(grid_for(size_t(np)), kMaxReduceBlocks), g, V, np, pos, out,
                  vmax_bits);
}

hipError_t launch_rk_substep(hipStream_t s, const GridDev& g, Vel V, int64_t np, float* pos, float* v1, int order,
                             float step, float half, float three_q, float c1, float c2, float c3, int resample,
                             uint32_t* vmax_bits) {
    hipError_t e = hipSuccess;
    if (resample) e = hipMemsetAsync(vmax_bits, 0, sizeof(uint32_t), s);
    if (e != hipSuccess || np <= 0) return e;
    return LAUNCH(k_rk_substep, std::min<uint32_t>(grid_for(size_t(np)), kMaxReduceBlocks), g, V, np, pos, v1, order,
                  step, half, three_q, c1, c2, c3, resample, vmax_bits);
}

}  // namespace mgf
