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
    float rhx, rhy, rhz;   // 1/h, used when every spacing is a power of two (pow2)
    int pow2;
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
    const int ix = clampi(int(div_h(px - g.ox, g.hx, g.rhx, g.pow2)), g.nx);
    const int iy = clampi(int(div_h(py - g.oy, g.hy, g.rhy, g.pow2)), g.ny);
    const int iz = clampi(int(div_h(pz - g.oz, g.hz, g.rhz, g.pow2)), g.nz);
    return (uint32_t(iz) * uint32_t(g.ny) + uint32_t(iy)) * uint32_t(g.nx) + uint32_t(ix);
}

// The support of a particle (EulerParticlesTool.cu:97-101, 216-220): lower
// corner and per-axis weights; n = 2 (linear) or 4 (corners i0-1 .. i0+2).
struct PStencil {
    int i0[3];
    float w[3][4];
};
// One axis of the support: lower corner and the kernel's weights.
__device__ __forceinline__ void pg_axis(float rel, float h, float rh, int pow2, int alg, int& i0, float (&w)[4]) {
    i0 = floor_ref(div_h(rel, h, rh, pow2));
    const float f = div_h(rel - float(i0) * h, h, rh, pow2);
    if (alg == MGF_PG_LINEAR) {
        w[0] = pg_linear(f);
        w[1] = pg_linear(float(double(f) - 1.0));
        w[2] = w[3] = 0.0f;
    } else if (alg == MGF_PG_QUADRATIC) {
        w[0] = pg_quadratic(float(double(f) + 1.0));
        w[1] = pg_quadratic(float(double(f) + 1.0 - 1.0));
        w[2] = pg_quadratic(float(double(f) + 1.0 - 2.0));
        w[3] = pg_quadratic(float(double(f) + 1.0 - 3.0));
    } else {
        w[0] = pg_cubic(float(double(f) + 1.0));
        w[1] = pg_cubic(float(double(f) + 1.0 - 1.0));
        w[2] = pg_cubic(float(double(f) + 1.0 - 2.0));
        w[3] = pg_cubic(float(double(f) + 1.0 - 3.0));
    }
}

__device__ __forceinline__ void pg_stencil(const PGrid& g, float px, float py, float pz, int alg, PStencil& st) {
    pg_axis(px - g.ox - g.hx * 0.5f, g.hx, g.rhx, g.pow2, alg, st.i0[0], st.w[0]);
    pg_axis(py - g.oy - g.hy * 0.5f, g.hy, g.rhy, g.pow2, alg, st.i0[1], st.w[1]);
    pg_axis(pz - g.oz - g.hz * 0.5f, g.hz, g.rhz, g.pow2, alg, st.i0[2], st.w[2]);
}

__device__ __forceinline__ uint32_t pg_index(const PGrid& g, int x, int y, int z) {
    return lin3(x, y, z, g.nx, g.ny, g.nz);
}

// The corners of a support in the reference's order (z, y, x), with their
// weights w_x * w_y * w_z; N = 2 (LINEAR, corners i0 .. i0+1) or 4
// (QUADRATIC / CUBIC, corners i0-1 .. i0+2), unrolled so the weights stay in
// registers.
template <int N, class F>
__device__ __forceinline__ void for_corners(const PGrid& g, const PStencil& st, F&& f) {
    constexpr int lo = N == 2 ? 0 : -1;
#pragma unroll
    for (int z = 0; z < N; ++z)
#pragma unroll
        for (int y = 0; y < N; ++y)
#pragma unroll
            for (int x = 0; x < N; ++x)
                f(pg_index(g, st.i0[0] + x + lo, st.i0[1] + y + lo, st.i0[2] + z + lo),
                  st.w[0][x] * st.w[1][y] * st.w[2][z]);
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

// One particle's value on one grid (component c) into the table.
template <int HB, int N>
__device__ __forceinline__ void p2g_one(uint32_t* keys, float* sv, float* sw, const PGrid& g, int c, float px,
                                        float py, float pz, float v, int alg, const P2GOut& out) {
    const uint32_t tag = uint32_t(c) << 30;
    if (alg == MGF_PG_P2GSUM) {
        const uint32_t node = pos_cell(g, px, py, pz);
        lds_accumulate<HB>(keys, sv, sw, tag | node, v, 0.0f, false, &out.field[c][node], nullptr);
        return;
    }
    PStencil st;
    pg_stencil(g, px, py, pz, alg, st);
    auto add = [&](uint32_t node, float wt) {
        lds_accumulate<HB>(keys, sv, sw, tag | node, wt * v, wt, true, &out.field[c][node], &out.weight[c][node]);
    };
    for_corners<N>(g, st, add);
}

// N = 2: the P2GSUM and LINEAR kernels; N = 4: QUADRATIC and CUBIC.
template <int NC, int HB, int N>
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
        p2g_one<HB, N>(keys, sv, sw, g0, 0, px, py, pz, val[int64_t(span) * p + (NC == 1 ? off : 0)], alg, out);
        if (NC == 3) {
            p2g_one<HB, N>(keys, sv, sw, g1, 1, px, py, pz, val[int64_t(span) * p + 1], alg, out);
            p2g_one<HB, N>(keys, sv, sw, g2, 2, px, py, pz, val[int64_t(span) * p + 2], alg, out);
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

// The value of one field at a particle: summed from 0 in corner order
// z, y, x (G2PNEAREST: the value of its cell).
template <int N>
__device__ __forceinline__ float g2p_one(const PGrid& g, float px, float py, float pz, const float* __restrict__ f,
                                         int alg) {
    if (alg == MGF_PG_G2PNEAREST) return f[pos_cell(g, px, py, pz)];
    PStencil st;
    pg_stencil(g, px, py, pz, alg, st);
    float acc = 0.0f;
    for_corners<N>(g, st, [&](uint32_t node, float wt) { acc += wt * f[node]; });
    return acc;
}

// accumulateCCSField2Particles (EulerParticlesTool.cu:198-288): summed from
// 0 in corner order z, y, x.
template <int N>
__global__ void __launch_bounds__(kBlock) k_g2p(PGrid g, int64_t np, const float* __restrict__ pos,
                                                const float* __restrict__ field, int alg, float* val, int span,
                                                int off) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    val[int64_t(span) * p + off] =
        (alg == MGF_PG_G2PNEAREST || alg == MGF_PG_LINEAR || alg == MGF_PG_QUADRATIC || alg == MGF_PG_CUBIC)
            ? g2p_one<N>(g, px, py, pz, field, alg) : 0.0f;
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
template <int N>
__device__ __forceinline__ void transfer_one(const PGrid& g, float px, float py, float pz, const float* fin,
                                             const float* fdelta, int kind, int alg, float keep, float mixing,
                                             float* v) {
    if (kind == MGF_PSOLVER_PIC) {
        *v = g2p_one<N>(g, px, py, pz, fin, alg);
        return;
    }
    const float pic = kind == MGF_PSOLVER_MIX ? g2p_one<N>(g, px, py, pz, fin, alg) : 0.0f;
    float u = __builtin_fmaf(1.0f, g2p_one<N>(g, px, py, pz, fdelta, alg), *v);
    if (kind == MGF_PSOLVER_MIX) u = __builtin_fmaf(mixing, pic, keep * u);
    *v = u;
}

template <int N>
__global__ void __launch_bounds__(kBlock) k_transfer_vel(PGrid g0, PGrid g1, PGrid g2, int64_t np,
                                                         const float* __restrict__ pos, Faces3 in, Faces3 delta,
                                                         int kind, int alg, float keep, float mixing,
                                                         float* __restrict__ vel) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    transfer_one<N>(g0, px, py, pz, in.f[0], delta.f[0], kind, alg, keep, mixing, &vel[3 * p]);
    transfer_one<N>(g1, px, py, pz, in.f[1], delta.f[1], kind, alg, keep, mixing, &vel[3 * p + 1]);
    transfer_one<N>(g2, px, py, pz, in.f[2], delta.f[2], kind, alg, keep, mixing, &vel[3 * p + 2]);
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

// ---- the particle-to-grid splat as a gather over binned particles ---------
// Particles are counting-sorted by the lower corner s of their support on the
// cell grid (s = floor((p - o)/h - 1/2) per axis, clamped to [-1, n-1]: bins
// (nx+1)(ny+1)(nz+1)).  A node of a target grid (the cells, or the faces of
// one axis) then sums, with no atomics, over the particles of the bins whose
// supports can reach it: on the target grid a node i receives from supports
// with lower corner in [i - lo - n + 1, i - lo] (lo, n: the kernel's corner
// offset and width), and on the axis the target grid is shifted along, that
// corner is the cell grid's or the next one (the same floor taken half a cell
// further), so one more bin is visited there.  Each visited particle's
// support on the target grid is evaluated as in the per-particle splat; the
// corners that clamp onto the node add their weights.
struct BinGrid {
    int nx, ny, nz;   // cells of the grid the bins follow
};

__device__ __forceinline__ uint32_t bin_of(const PGrid& c, float px, float py, float pz) {
    const float rel[3] = {px - c.ox - c.hx * 0.5f, py - c.oy - c.hy * 0.5f, pz - c.oz - c.hz * 0.5f};
    const float h[3] = {c.hx, c.hy, c.hz};
    const float rh[3] = {c.rhx, c.rhy, c.rhz};
    const int n[3] = {c.nx, c.ny, c.nz};
    uint32_t b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int s = floor_ref(div_h(rel[k], h[k], rh[k], c.pow2));
        s = s < -1 ? -1 : (s > n[k] - 1 ? n[k] - 1 : s);
        b[k] = uint32_t(s + 1);
    }
    return (b[2] * uint32_t(c.ny + 1) + b[1]) * uint32_t(c.nx + 1) + b[0];
}

__global__ void __launch_bounds__(kBlock) k_bin_count(PGrid c, int64_t np, const float* __restrict__ pos,
                                                      uint32_t* counts) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    atomicAdd(&counts[bin_of(c, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2])], 1u);
}

// Tile sums of a u32 array (kGenTile entries per work-group).
__global__ void __launch_bounds__(kBlock) k_scan_tiles(size_t n, const uint32_t* __restrict__ a,
                                                       uint32_t* __restrict__ tiles) {
    __shared__ uint32_t wsum[kBlock / 64];
    const size_t base = size_t(blockIdx.x) * kGenTile + size_t(threadIdx.x) * kGenRows;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kGenRows; ++j)
        if (base + j < n) sum += a[base + j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum += uint32_t(__shfl_xor(int(sum), o, 64));
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int q = 0; q < kBlock / 64; ++q) t += wsum[q];
        tiles[blockIdx.x] = t;
    }
}

// start[i] = exclusive prefix of a (tile offsets already scanned), and a copy
// in cursor; start[n] = the total.
__global__ void __launch_bounds__(kBlock) k_scan_apply(size_t n, const uint32_t* __restrict__ a,
                                                       const uint32_t* __restrict__ tiles, uint32_t nt,
                                                       uint32_t* __restrict__ start, uint32_t* __restrict__ cursor) {
    __shared__ uint32_t part[kBlock];
    const size_t base = size_t(blockIdx.x) * kGenTile + size_t(threadIdx.x) * kGenRows;
    uint32_t v[kGenRows], sum = 0;
#pragma unroll
    for (int j = 0; j < kGenRows; ++j) {
        v[j] = base + j < n ? a[base + j] : 0u;
        sum += v[j];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {   // inclusive scan of the work-items' sums
        const uint32_t t = threadIdx.x >= uint32_t(o) ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = tiles[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0u);
#pragma unroll
    for (int j = 0; j < kGenRows; ++j) {
        if (base + j < n) {
            start[base + j] = run;
            cursor[base + j] = run;
        }
        run += v[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[n] = tiles[nt];
}

// Each particle's slot in bin order, with sorted copies of its position and
// (when vel is given) its velocity, so the splat reads them contiguously.
__global__ void __launch_bounds__(kBlock) k_bin_scatter(PGrid c, int64_t np, const float* __restrict__ pos,
                                                        uint32_t* cursor, uint32_t* __restrict__ order,
                                                        float* __restrict__ pos_s, const float* __restrict__ vel,
                                                        float* __restrict__ vel_s) {
    const int64_t p = int64_t(xcd_block()) * kBlock + threadIdx.x;
    if (p >= np) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    const uint32_t slot = atomicAdd(&cursor[bin_of(c, px, py, pz)], 1u);
    order[slot] = uint32_t(p);
    pos_s[3 * size_t(slot)] = px;
    pos_s[3 * size_t(slot) + 1] = py;
    pos_s[3 * size_t(slot) + 2] = pz;
    if (vel) {
        vel_s[3 * size_t(slot)] = vel[3 * p];
        vel_s[3 * size_t(slot) + 1] = vel[3 * p + 1];
        vel_s[3 * size_t(slot) + 2] = vel[3 * p + 2];
    }
}

// Two passes per target grid (the cells, or the faces of one axis; LINEAR
// kernel).  Pass 1, one work-item per bin: every particle of the bin adds
// its weights, and weight times value, to slots indexed by the target node
// relative to the bin's first node clamp(s) (two slots per axis, three on the
// axis the target grid is shifted along: its support starts at the cell
// grid's corner or the next).  A support's corners that clamp onto one node
// share a slot, so an edge bin collects far-outside particles too.  Empty
// bins write nothing.  Pass 2, one work-item per node: the sum, in a fixed bin
// order, of the slots of the non-empty neighbouring bins that point at it;
// then the normalisation.
constexpr int kSlotsMax = 12;

__device__ __forceinline__ int clamp_bin(int s, int cells) { return s < -1 ? -1 : (s > cells - 1 ? cells - 1 : s); }

// pos_s: positions in bin order; val is in bin order too when val_sorted,
// else indexed by particle through order.
template <int SHIFT>
__global__ void __launch_bounds__(kBlock) k_p2g_bins(PGrid g, BinGrid bg, const uint32_t* __restrict__ start,
                                                     const uint32_t* __restrict__ order, const float* __restrict__ pos_s,
                                                     const float* __restrict__ val, int span, int off, int val_sorted,
                                                     float* __restrict__ acc) {
    const size_t nb = size_t(bg.nx + 1) * (bg.ny + 1) * (bg.nz + 1);
    const size_t b = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (b >= nb) return;
    const int s[3] = {int(b % uint32_t(bg.nx + 1)) - 1, int((b / uint32_t(bg.nx + 1)) % uint32_t(bg.ny + 1)) - 1,
                      int(b / (size_t(bg.nx + 1) * uint32_t(bg.ny + 1))) - 1};
    const int len[3] = {g.nx, g.ny, g.nz};
    constexpr int ns[3] = {SHIFT == 0 ? 3 : 2, SHIFT == 1 ? 3 : 2, SHIFT == 2 ? 3 : 2};
    int base[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) base[a] = clampi(s[a], len[a]);
    const uint32_t q0 = start[b], q1 = start[b + 1];
    if (q0 == q1) return;   // an empty bin's slots are never read
    float sv[kSlotsMax], sw[kSlotsMax];
#pragma unroll
    for (int q = 0; q < kSlotsMax; ++q) sv[q] = sw[q] = 0.0f;
    for (uint32_t q = q0; q < q1; ++q) {
        PStencil st;
        pg_stencil(g, pos_s[3 * size_t(q)], pos_s[3 * size_t(q) + 1], pos_s[3 * size_t(q) + 2], MGF_PG_LINEAR, st);
        const float v = val[size_t(span) * (val_sorted ? q : order[q]) + off];
        // per-axis slot weights: the corners that clamp onto slot t
        float W[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int t = 0; t < 3; ++t) W[a][t] = 0.0f;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int t = clampi(st.i0[a] + c, len[a]) - base[a];
#pragma unroll
                for (int u = 0; u < 3; ++u)
                    if (t == u) W[a][u] += st.w[a][c];
            }
        }
#pragma unroll
        for (int tz = 0; tz < ns[2]; ++tz)
#pragma unroll
            for (int ty = 0; ty < ns[1]; ++ty)
#pragma unroll
                for (int tx = 0; tx < ns[0]; ++tx) {
                    const int slot = (tz * ns[1] + ty) * ns[0] + tx;
                    const float wt = W[0][tx] * W[1][ty] * W[2][tz];
                    sv[slot] += wt * v;
                    sw[slot] += wt;
                }
    }
    constexpr int nslot = ns[0] * ns[1] * ns[2];
#pragma unroll
    for (int u = 0; u < kSlotsMax; ++u)
        if (u < nslot) {
            acc[(2 * size_t(u)) * nb + b] = sv[u];
            acc[(2 * size_t(u) + 1) * nb + b] = sw[u];
        }
}

__global__ void __launch_bounds__(kBlock) k_p2g_nodes(PGrid g, int shift, BinGrid bg, const uint32_t* __restrict__ start,
                                                      const float* __restrict__ acc, float* __restrict__ field,
                                                      float* __restrict__ weight, int normalize) {
    const size_t total = size_t(g.nx) * g.ny * g.nz;
    const size_t node = size_t(xcd_block()) * kBlock + threadIdx.x;
    if (node >= total) return;
    const size_t nb = size_t(bg.nx + 1) * (bg.ny + 1) * (bg.nz + 1);
    const int idx[3] = {int(node % uint32_t(g.nx)), int((node / uint32_t(g.nx)) % uint32_t(g.ny)),
                        int(node / (size_t(g.nx) * uint32_t(g.ny)))};
    const int len[3] = {g.nx, g.ny, g.nz}, cells[3] = {bg.nx, bg.ny, bg.nz};
    int ns[3], s0[3], s1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ns[a] = a == shift ? 3 : 2;
        s0[a] = clamp_bin(idx[a] - ns[a], cells[a]);
        s1[a] = clamp_bin(idx[a], cells[a]);
    }
    float sv = 0.0f, sw = 0.0f;
    for (int bz = s0[2]; bz <= s1[2]; ++bz) {
        const int tz = idx[2] - clampi(bz, len[2]);
        if (tz < 0 || tz >= ns[2]) continue;
        for (int by = s0[1]; by <= s1[1]; ++by) {
            const int ty = idx[1] - clampi(by, len[1]);
            if (ty < 0 || ty >= ns[1]) continue;
            for (int bx = s0[0]; bx <= s1[0]; ++bx) {
                const int tx = idx[0] - clampi(bx, len[0]);
                if (tx < 0 || tx >= ns[0]) continue;
                const size_t b = (size_t(bz + 1) * uint32_t(bg.ny + 1) + uint32_t(by + 1)) * uint32_t(bg.nx + 1) +
                                 uint32_t(bx + 1);
                if (start[b] == start[b + 1]) continue;   // empty bin: no slots written
                const size_t slot = size_t((tz * ns[1] + ty) * ns[0] + tx);
                sv += acc[2 * slot * nb + b];
                sw += acc[(2 * slot + 1) * nb + b];
            }
        }
    }
    field[node] = normalize && sw > 0.0f ? sv / sw : sv;
    weight[node] = sw;
}

// ===========================================================================
static PGrid pgrid_of(const GridDev& g, int kind) {
    PGrid p{g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g.hx, g.hy, g.hz, g.rhx, g.rhy, g.rhz, g.pow2};
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
            e = LAUNCH((k_p2g_lds<1, 10, 2>), grid_for(size_t(np)), p, p, p, np, pos, val, span, off, alg, out);
        else
            e = LAUNCH((k_p2g_lds<1, 11, 4>), grid_for(size_t(np)), p, p, p, np, pos, val, span, off, alg, out);
    }
    if (e == hipSuccess && normalize) e = LAUNCH(k_p2g_normalize, grid_for(n), n, field, weight);
    return e;
}

hipError_t launch_g2p(hipStream_t s, const GridDev& g, int kind, int64_t np, const float* pos, const float* field,
                      int alg, float* val, int span, int off) {
    if (np <= 0) return hipSuccess;
    if (alg == MGF_PG_QUADRATIC || alg == MGF_PG_CUBIC)
        return LAUNCH(k_g2p<4>, grid_for(size_t(np)), pgrid_of(g, kind), np, pos, field, alg, val, span, off);
    return LAUNCH(k_g2p<2>, grid_for(size_t(np)), pgrid_of(g, kind), np, pos, field, alg, val, span, off);
}

hipError_t launch_particle_markers(hipStream_t s, const GridDev& g, const float* solid, int64_t np, const float* pos,
                                   float* out) {
    const size_t n = size_t(g.nx) * g.ny * g.nz;
    hipError_t e = launch_fill(s, out, n, 0.0f);
    if (e == hipSuccess && np > 0) e = LAUNCH(k_particle_counts, grid_for(size_t(np)), pgrid_of(g, 0), np, pos, out);
    if (e == hipSuccess && solid) e = LAUNCH(k_particle_markers, grid_for(n), n, solid, out);
    return e;
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
    (void)launch_dims(k_gen_count, "k_gen_count", dim3(nt), dim3(kBlock), 0, s, ncell, fluid, solid, tiles);
    (void)launch_dims(k_gen_scan, "k_gen_scan", dim3(1), dim3(1024), 0, s, nt, tiles);
    return hipGetLastError();
}

hipError_t launch_gen_emit(hipStream_t s, const GridDev& g, const float* fluid, const float* solid,
                           const uint32_t* tiles, int per_cell, uint64_t seed, float* pos) {
    const size_t ncell = size_t(g.nx) * g.ny * g.nz;
    (void)launch_dims(k_gen_emit, "k_gen_emit", dim3(gen_tiles(ncell)), dim3(kBlock), 0, s, g, ncell, fluid, solid, tiles,
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
    if (alg == MGF_PG_QUADRATIC || alg == MGF_PG_CUBIC)
        return LAUNCH(k_transfer_vel<4>, grid_for(size_t(np)), pgrid_of(g, 1), pgrid_of(g, 2), pgrid_of(g, 3), np,
                      pos, fi, fd, kind, alg, keep, mixing, vel);
    return LAUNCH(k_transfer_vel<2>, grid_for(size_t(np)), pgrid_of(g, 1), pgrid_of(g, 2), pgrid_of(g, 3), np, pos,
                  fi, fd, kind, alg, keep, mixing, vel);
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

namespace mgf {

size_t bin_count(const GridDev& g) { return size_t(g.nx + 1) * size_t(g.ny + 1) * size_t(g.nz + 1); }
uint32_t bin_tiles(const GridDev& g) { return gen_tiles(bin_count(g)); }

hipError_t launch_bin_particles(hipStream_t s, const GridDev& g, int64_t np, const float* pos, uint32_t* counts,
                                uint32_t* tiles, uint32_t* start, uint32_t* cursor, uint32_t* order, float* pos_s,
                                const float* vel, float* vel_s) {
    const size_t nb = bin_count(g);
    const uint32_t nt = gen_tiles(nb);
    const PGrid c = pgrid_of(g, 0);
    hipError_t e = hipMemsetAsync(counts, 0, nb * sizeof(uint32_t), s);
    if (e == hipSuccess && np > 0) e = LAUNCH(k_bin_count, grid_for(size_t(np)), c, np, pos, counts);
    if (e != hipSuccess) return e;
    (void)launch_dims(k_scan_tiles, "k_scan_tiles", dim3(nt), dim3(kBlock), 0, s, nb, counts, tiles);
    (void)launch_dims(k_gen_scan, "k_gen_scan", dim3(1), dim3(1024), 0, s, nt, tiles);
    (void)launch_dims(k_scan_apply, "k_scan_apply", dim3(nt), dim3(kBlock), 0, s, nb, counts, tiles, nt, start, cursor);
    e = hipGetLastError();
    if (e == hipSuccess && np > 0)
        e = LAUNCH(k_bin_scatter, grid_for(size_t(np)), c, np, pos, cursor, order, pos_s, vel, vel_s);
    return e;
}

size_t p2g_acc_floats(const GridDev& g) { return 2 * size_t(kSlotsMax) * bin_count(g); }

hipError_t launch_p2g_binned(hipStream_t s, const GridDev& g, int kind, const uint32_t* start, const uint32_t* order,
                             const float* pos_s, const float* val, int span, int off, int val_sorted, float* acc,
                             float* field, float* weight, int normalize) {
    const PGrid p = pgrid_of(g, kind);
    const size_t n = size_t(p.nx) * p.ny * p.nz;
    const BinGrid bg{g.nx, g.ny, g.nz};
    const uint32_t gb = grid_for(bin_count(g));
    hipError_t e;
    switch (kind) {
        case 1: e = LAUNCH(k_p2g_bins<0>, gb, p, bg, start, order, pos_s, val, span, off, val_sorted, acc); break;
        case 2: e = LAUNCH(k_p2g_bins<1>, gb, p, bg, start, order, pos_s, val, span, off, val_sorted, acc); break;
        case 3: e = LAUNCH(k_p2g_bins<2>, gb, p, bg, start, order, pos_s, val, span, off, val_sorted, acc); break;
        default: e = LAUNCH(k_p2g_bins<-1>, gb, p, bg, start, order, pos_s, val, span, off, val_sorted, acc); break;
    }
    if (e == hipSuccess) e = LAUNCH(k_p2g_nodes, grid_for(n), p, kind - 1, bg, start, acc, field, weight, normalize);
    return e;
}

}  // namespace mgf
