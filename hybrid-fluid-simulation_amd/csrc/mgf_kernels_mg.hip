// mgf_kernels_mg.hip — stencil codes, the fused kernels of the mixed-precision
// CG and the geometric multigrid V-cycle used as its preconditioner.
//
// Layout and access pattern: one work-item per cell, cells x-fastest, so every
// load of a wave is one coalesced row.  Every load of a stencil is issued
// unconditionally — a neighbour outside the grid is read at a clamped index
// and its value discarded by a mask taken from the cell's stencil code — so
// all loads of a cell are in flight together.
//
// Arithmetic: (A v)_c = aw_c v_c + sum over fluid neighbours k of s_k (v_c -
// v_k) in FP32, where aw_c = sum of s_k over the air neighbours.  This is the
// matrix row of buildPressureMatrixA (diagonal = sum of s_k over non-solid
// neighbours, -s_k to fluid neighbours) rearranged so that no term of the
// size of the diagonal cancels: for smooth v the differences are exact and
// the result is accurate to FP32 rounding of itself.  Dot products and the
// x / r updates of the CG are formed in FP64.
//
// Reductions: each block writes one FP64 partial sum with a plain store; a
// one-block kernel sums the partials in a fixed order and updates the CG
// scalars.  The kernel boundary makes the partials visible: no device-scope
// fence (which writes back and invalidates the XCD's L2) is executed per block.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mgf_internal.h"
#include "mgf_kcommon.h"

namespace mgf {

// ---------------------------------------------------------------------------
// Stencil code of a cell (uint16), from the markers of the cell and its six
// neighbours, in the order -x, +x, -y, +y, -z, +z:
//   bit 0      fluid cell
//   bits 1-6   the neighbour is in the grid and air (Dirichlet term only)
//   bits 7-12  the neighbour is in the grid and fluid (coupled)
// Non-fluid cells have code 0 (identity rows of the reference matrix).
// (MPhysics/EulerSolverTool.cu:58-126 builds the same rows.)
enum : uint16_t {
    kCodeFluid = 1u << 0,
    kAirShift = 1,
    kCplShift = 7,
};

__device__ __forceinline__ void k_codes_run(int nx, int ny, int nz, const uint8_t* __restrict__ m,
                                                  uint16_t* __restrict__ code, uint32_t n, uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i >= n) return;
    const int x = int(i % uint32_t(nx)), y = int((i / uint32_t(nx)) % uint32_t(ny)),
              z = int(i / (uint32_t(nx) * uint32_t(ny)));
    if (m[i] != kFluid) { code[i] = 0; return; }
    const uint32_t sy = uint32_t(nx), sz = uint32_t(nx) * uint32_t(ny);
    const bool in[6] = {x > 0, x + 1 < nx, y > 0, y + 1 < ny, z > 0, z + 1 < nz};
    const uint32_t j[6] = {i - 1, i + 1, i - sy, i + sy, i - sz, i + sz};
    uint32_t c = kCodeFluid;
    for (int k = 0; k < 6; ++k) {
        if (!in[k]) continue;
        const uint8_t q = m[j[k]];
        if (q == kAir) c |= 1u << (kAirShift + k);
        if (q == kFluid) c |= 1u << (kCplShift + k);
    }
    code[i] = uint16_t(c);
}

__global__ void __launch_bounds__(kBlock) k_codes(int nx, int ny, int nz, const uint8_t* __restrict__ m,
                                                  uint16_t* __restrict__ code, uint32_t n) {
    k_codes_run(nx, ny, nz, m, code, n, xcd_block());
}

struct MCodes {
    static constexpr int kThreads = kBlock;
    int nx;
    int ny;
    int nz;
    const uint8_t* m;
    uint16_t* code;
    uint32_t n;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_codes_run(nx, ny, nz, m, code, n, xcd_of(vb, G)); }
};

// Per-level tables indexed by a 6-bit neighbour mask (-x,+x,-y,+y,-z,+z):
// sum of the scales of the masked directions, added in the reference's order
// (x terms, then y, then z: MPhysics/EulerSolverTool.cu:58-126), so that
// table[air | coupled] is bitwise the reference's diagonal.
__device__ __forceinline__ float mask_scale_sum(uint32_t m, float sx, float sy, float sz) {
    float d = 0.0f;
    if (m & 1u) d += sx;
    if (m & 2u) d += sx;
    if (m & 4u) d += sy;
    if (m & 8u) d += sy;
    if (m & 16u) d += sz;
    if (m & 32u) d += sz;
    return d;
}

// Block-shared tables: aw[mask] (air weights) and, with W, w[mask] = omega / d
// for mask = the non-solid neighbours (0 where d == 0).  All threads call it.
template <bool W>
__device__ __forceinline__ void build_tables(float* aw, float* w, float sx, float sy, float sz, float omega) {
    if (threadIdx.x < 64) {
        const float d = mask_scale_sum(threadIdx.x, sx, sy, sz);
        aw[threadIdx.x] = d;
        if (W) w[threadIdx.x] = d != 0.0f ? omega / d : 0.0f;
    }
    __syncthreads();
}

__device__ __forceinline__ uint32_t air_mask(uint16_t cd) { return (uint32_t(cd) >> kAirShift) & 63u; }
__device__ __forceinline__ uint32_t cpl_mask(uint16_t cd) { return (uint32_t(cd) >> kCplShift) & 63u; }

// v if bit `b` of the code is set, else +0 (bit pattern AND with 0 / ~0).
__device__ __forceinline__ float keep_if(float v, uint16_t cd, int b) {
    const int32_t m = __builtin_amdgcn_sbfe(int32_t(cd), b, 1);
    return __int_as_float(__float_as_int(v) & m);
}

// (A v)_c, FP32, difference form; identity row for a non-fluid cell.
__device__ __forceinline__ float arow(uint16_t cd, float vc, const float (&v)[6], float sx, float sy, float sz,
                                      const float* aw) {
    const float ex = keep_if(vc - v[0], cd, kCplShift + 0) + keep_if(vc - v[1], cd, kCplShift + 1);
    const float ey = keep_if(vc - v[2], cd, kCplShift + 2) + keep_if(vc - v[3], cd, kCplShift + 3);
    const float ez = keep_if(vc - v[4], cd, kCplShift + 4) + keep_if(vc - v[5], cd, kCplShift + 5);
    const float a = aw[air_mask(cd)] * vc + sx * ex + sy * ey + sz * ez;
    return (cd & kCodeFluid) ? a : vc;
}

// omega / d of a cell: 0 for non-fluid cells (code 0, entry 0 of the table)
// and for zero rows.  No test of the code, so no branch holds up the loads.
__device__ __forceinline__ float wdiag(uint16_t cd, const float* w) {
    return w[air_mask(cd) | cpl_mask(cd)];
}

// ---------------------------------------------------------------------------
// Flat geometry: n cells, row stride sy = nx, plane stride sz = nx * ny.
struct Flat {
    uint32_t n, sy, sz;
};

inline Flat make_flat(int nx, int ny, int nz) {
    Flat f;
    f.sy = uint32_t(nx);
    f.sz = uint32_t(nx) * uint32_t(ny);
    f.n = f.sz * uint32_t(nz);
    return f;
}

// The six neighbour indices, clamped into [0, n): where a coupling bit is set
// the neighbour is in the grid and the clamped index is the true one.
__device__ __forceinline__ void nbrs(uint32_t c, const Flat& F, uint32_t (&j)[6]) {
    const uint32_t last = F.n - 1;
    j[0] = c > 0 ? c - 1 : 0;
    j[1] = c < last ? c + 1 : last;
    j[2] = c >= F.sy ? c - F.sy : 0;
    j[3] = c + F.sy <= last ? c + F.sy : last;
    j[4] = c >= F.sz ? c - F.sz : 0;
    j[5] = c + F.sz <= last ? c + F.sz : last;
}

// Block sum of one double, stored to part[slot] by thread 0.
__device__ __forceinline__ void block_partial(double v, double* part, uint32_t slot) {
    double a[1] = {v};
    block_sum<1>(a);
    if (threadIdx.x == 0) part[slot] = a[0];
}

// a / d for a < 2^31, d >= 1 via the FP64 reciprocal rd = 1/d: the estimate
// is within one of the quotient and corrected once.
__device__ __forceinline__ uint32_t udiv(uint32_t a, uint32_t d, double rd) {
    uint32_t q = uint32_t(double(a) * rd);
    const uint32_t r = a - q * d;
    if (int32_t(r) < 0) --q;
    else if (r >= d) ++q;
    return q;
}

// Cells per work-item of the fine-grid kernels (256 apart): fewer block
// partials to finish, the LDS tables amortised, more loads in flight.
constexpr int kCpt = 4;

// ---------------------------------------------------------------------------
// Fused CG.  Each work-item handles kCpt cells 256 apart: first all loads of
// its cells, then the arithmetic and the stores.  Only the last block can
// hold cells past the end; the others run the body without the range test
// (a block-uniform branch), so nothing stands between the loads.
#define MGF_CELLS(CHECK, F, base)                                                  \
    uint32_t cidx[kCpt];                                                           \
    _Pragma("unroll") for (int k = 0; k < kCpt; ++k) {                             \
        const uint32_t c_ = (base) + uint32_t(k) * kBlock;                         \
        cidx[k] = (CHECK && c_ >= (F).n) ? (F).n - 1 : c_;                         \
    }

// K_I: Az = A z (the stencil), then in place p = z + beta p and
// Ap = Az + beta Ap (p = z, Ap = Az on the first iteration); block partials
// of p . Ap.  Only z is read at the neighbours, so p and Ap are updated in
// place without a race.
template <bool CHECK>
__device__ __forceinline__ double cg_pap_body(const Flat& F, uint32_t base, const uint16_t* __restrict__ code,
                                              float sx, float sy, float sz, const float* __restrict__ zv,
                                              float* __restrict__ p, float* __restrict__ ap, float beta,
                                              bool first, const float* aw) {
    MGF_CELLS(CHECK, F, base)
    uint16_t cd[kCpt];
    float zc[kCpt], zn[kCpt][6], pc[kCpt], apc[kCpt];
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        uint32_t j[6];
        nbrs(cidx[k], F, j);
        cd[k] = code[cidx[k]];
        zc[k] = zv[cidx[k]];
#pragma unroll
        for (int m = 0; m < 6; ++m) zn[k][m] = zv[j[m]];
        if (!first) {
            pc[k] = p[cidx[k]];
            apc[k] = ap[cidx[k]];
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const float az = arow(cd[k], zc[k], zn[k], sx, sy, sz, aw);
        const float pn = first ? zc[k] : __builtin_fmaf(beta, pc[k], zc[k]);
        const float an = first ? az : __builtin_fmaf(beta, apc[k], az);
        const uint32_t c = base + uint32_t(k) * kBlock;
        if (!CHECK || c < F.n) {
            p[c] = pn;
            ap[c] = an;
            acc += double(pn) * double(an);
        }
    }
    return acc;
}

// Blocks of the flat CG kernels (kBlock * kCpt consecutive cells each) that
// hold no fluid cell (wet[block] == 0).  On non-fluid cells b = 0, so r and
// x stay +0, z = M^-1 r is 0, and from the first K_I of a sweep on p and Ap
// are +0 too: such a block would store the values its cells hold and add +0
// to the dot products, so K_I (after the first iteration) and K_II return at
// once with a zero partial.
__device__ __forceinline__ void k_flat_wet_run(Flat F, const uint16_t* __restrict__ code,
                                                     uint8_t* __restrict__ wet, uint32_t blk) {
        const uint32_t b0 = blk * (kBlock * kCpt);
    int any = 0;
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const uint32_t c = b0 + threadIdx.x + uint32_t(k) * kBlock;
        if (c < F.n) any |= code[c] & kCodeFluid;
    }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) wet[blk] = uint8_t(any != 0);
}

__global__ void __launch_bounds__(kBlock) k_flat_wet(Flat F, const uint16_t* __restrict__ code,
                                                     uint8_t* __restrict__ wet) {
    k_flat_wet_run(F, code, wet, xcd_block());
}

struct MFlatWet {
    static constexpr int kThreads = kBlock;
    Flat F;
    const uint16_t* code;
    uint8_t* wet;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_flat_wet_run(F, code, wet, xcd_of(vb, G)); }
};

// A kernel body runs block `blk` of its launch (after the XCD remapping) and
// writes its partial sums to slot `slot` (the launch's own block index).
__device__ __forceinline__ void cg_pap_run(Flat F, const uint16_t* __restrict__ code, float sx, float sy, float sz,
                                           const float* __restrict__ zv, float* __restrict__ p,
                                           float* __restrict__ ap, const CgScalars* __restrict__ sc,
                                           double* __restrict__ part, const uint8_t* __restrict__ wet, uint32_t blk,
                                           uint32_t slot) {
    __shared__ float aw[64];
    if (sc->done) return;
    const bool first = sc->iters == 0;
    if (!first && wet && !wet[blk]) {
        if (threadIdx.x == 0) part[slot] = 0.0;
        return;
    }
    build_tables<false>(aw, nullptr, sx, sy, sz, 0.0f);
    const float beta = float(sc->beta);
    const uint32_t b0 = blk * (kBlock * kCpt);
    const uint32_t base = b0 + threadIdx.x;
    const double acc = (b0 + kBlock * kCpt <= F.n)
                           ? cg_pap_body<false>(F, base, code, sx, sy, sz, zv, p, ap, beta, first, aw)
                           : cg_pap_body<true>(F, base, code, sx, sy, sz, zv, p, ap, beta, first, aw);
    block_partial(acc, part, slot);
}

__global__ void __launch_bounds__(kBlock) k_cg_pap(Flat F, const uint16_t* __restrict__ code, float sx, float sy,
                                                   float sz, const float* __restrict__ zv, float* __restrict__ p,
                                                   float* __restrict__ ap, const CgScalars* __restrict__ sc,
                                                   double* __restrict__ part, const uint8_t* __restrict__ wet) {
    cg_pap_run(F, code, sx, sy, sz, zv, p, ap, sc, part, wet, xcd_block(), blockIdx.x);
}

struct MPap {
    static constexpr int kThreads = kBlock;
    Flat F;
    const uint16_t* code;
    float sx, sy, sz;
    const float* zv;
    float *p, *ap;
    const CgScalars* sc;
    double* part;
    const uint8_t* wet;
    uint32_t G;
    __device__ void run(uint32_t b) const { cg_pap_run(F, code, sx, sy, sz, zv, p, ap, sc, part, wet, xcd_of(b, G), b); }
};

// K_II (pointwise): x += alpha p; r -= alpha Ap (each formed in FP64, stored
// FP32); block partials of r . r.
template <bool CHECK>
__device__ __forceinline__ double cg_update_body(const Flat& F, uint32_t base, const float* __restrict__ p,
                                                 const float* __restrict__ ap, float* __restrict__ x,
                                                 float* __restrict__ r, double alpha) {
    MGF_CELLS(CHECK, F, base)
    float pc[kCpt], ac[kCpt], x0[kCpt], r0[kCpt];
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        pc[k] = p[cidx[k]];
        ac[k] = ap[cidx[k]];
        x0[k] = x[cidx[k]];
        r0[k] = r[cidx[k]];
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const float xn = float(__builtin_fma(alpha, double(pc[k]), double(x0[k])));
        const float rn = float(__builtin_fma(-alpha, double(ac[k]), double(r0[k])));
        const uint32_t c = base + uint32_t(k) * kBlock;
        if (!CHECK || c < F.n) {
            x[c] = xn;
            r[c] = rn;
            acc += double(rn) * double(rn);
        }
    }
    return acc;
}

__device__ __forceinline__ void cg_update_run(Flat F, const float* __restrict__ p, const float* __restrict__ ap,
                                              float* __restrict__ x, float* __restrict__ r,
                                              const CgScalars* __restrict__ sc, double* __restrict__ part,
                                              const uint8_t* __restrict__ wet, uint32_t blk, uint32_t slot) {
    if (sc->done) return;
    if (wet && !wet[blk]) {
        if (threadIdx.x == 0) part[slot] = 0.0;
        return;
    }
    const double alpha = sc->alpha;
    const uint32_t b0 = blk * (kBlock * kCpt);
    const uint32_t base = b0 + threadIdx.x;
    const double acc = (b0 + kBlock * kCpt <= F.n) ? cg_update_body<false>(F, base, p, ap, x, r, alpha)
                                                   : cg_update_body<true>(F, base, p, ap, x, r, alpha);
    block_partial(acc, part, slot);
}

__global__ void __launch_bounds__(kBlock) k_cg_update(Flat F, const float* __restrict__ p,
                                                      const float* __restrict__ ap, float* __restrict__ x,
                                                      float* __restrict__ r, const CgScalars* __restrict__ sc,
                                                      double* __restrict__ part, const uint8_t* __restrict__ wet) {
    cg_update_run(F, p, ap, x, r, sc, part, wet, xcd_block(), blockIdx.x);
}

struct MUpdate {
    static constexpr int kThreads = kBlock;
    Flat F;
    const float *p, *ap;
    float *x, *r;
    const CgScalars* sc;
    double* part;
    const uint8_t* wet;
    uint32_t G;
    __device__ void run(uint32_t b) const { cg_update_run(F, p, ap, x, r, sc, part, wet, xcd_of(b, G), b); }
};

// Refinement (outer sweeps of the mixed solver).
// MODE 0: x64 = 0; partials of ||b||^2.
// MODE 1: x64 = xc (the FP32 correction of the first sweep; the neighbours'
// new values come from xc, x64 is only written); MODE 2: x64 read only (the
// correction was added by k_acc64 before).  Then r = b - A x64 formed in FP64
// and stored FP32 when r != null; partials of ||r||^2 and ||b||^2.  (Reading
// x64 at the neighbours while updating it in place would race.)
__device__ __forceinline__ double arow64(uint16_t cd, double vc, const double (&v)[6], double sx, double sy,
                                         double sz, const float* aw) {
    auto keep = [&](double d, int b) { return ((cd >> b) & 1) ? d : 0.0; };
    const double ex = keep(vc - v[0], kCplShift + 0) + keep(vc - v[1], kCplShift + 1);
    const double ey = keep(vc - v[2], kCplShift + 2) + keep(vc - v[3], kCplShift + 3);
    const double ez = keep(vc - v[4], kCplShift + 4) + keep(vc - v[5], kCplShift + 5);
    const double a = double(aw[air_mask(cd)]) * vc + sx * ex + sy * ey + sz * ez;
    return (cd & kCodeFluid) ? a : vc;
}

template <int MODE, bool CHECK>
__device__ __forceinline__ void refine_body(const Flat& F, uint32_t base, const uint16_t* __restrict__ code,
                                            float sx, float sy, float sz, const float* __restrict__ xc,
                                            double* __restrict__ x64, const float* __restrict__ b,
                                            float* __restrict__ r, const float* aw, double& rr, double& bb) {
    MGF_CELLS(CHECK, F, base)
    constexpr int K = kCpt;
    uint16_t cd[K];
    float bc[K], xcc[K], xcn[K][6];
    double xo[K], xon[K][6];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        bc[k] = b[cidx[k]];
        if (MODE != 0) {
            uint32_t j[6];
            nbrs(cidx[k], F, j);
            cd[k] = code[cidx[k]];
            if (MODE == 1) {
                xcc[k] = xc[cidx[k]];
#pragma unroll
                for (int m = 0; m < 6; ++m) xcn[k][m] = xc[j[m]];
            } else {
                xo[k] = x64[cidx[k]];
#pragma unroll
                for (int m = 0; m < 6; ++m) xon[k][m] = x64[j[m]];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t c = base + uint32_t(k) * kBlock;
        if (CHECK && c >= F.n) continue;
        bb += double(bc[k]) * double(bc[k]);
        if (MODE == 0) {
            x64[c] = 0.0;
            continue;
        }
        double vc, v[6];
        if (MODE == 1) {
            vc = double(xcc[k]);
#pragma unroll
            for (int m = 0; m < 6; ++m) v[m] = double(xcn[k][m]);
            x64[c] = vc;
        } else {
            vc = xo[k];
#pragma unroll
            for (int m = 0; m < 6; ++m) v[m] = xon[k][m];
            // a warm start keeps the fluid cells' pressure and zeroes the
            // rest (fluid rows never read a non-fluid neighbour's value)
            if (MODE == 3 && !(cd[k] & kCodeFluid)) {
                vc = 0.0;
                x64[c] = 0.0;
            }
        }
        const double res = double(bc[k]) - arow64(cd[k], vc, v, sx, sy, sz, aw);
        if (r) r[c] = float(res);
        rr += res * res;
    }
}

template <int MODE>
__device__ __forceinline__ void refine_run(Flat F, const uint16_t* __restrict__ code, float sx, float sy, float sz,
                                           const float* __restrict__ xc, double* __restrict__ x64,
                                           const float* __restrict__ b, float* __restrict__ r,
                                           double* __restrict__ part, uint32_t blk, uint32_t slot) {
    __shared__ float aw[64];
    build_tables<false>(aw, nullptr, sx, sy, sz, 0.0f);
    const uint32_t b0 = blk * (kBlock * kCpt);
    const uint32_t base = b0 + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (b0 + kBlock * kCpt <= F.n)
        refine_body<MODE, false>(F, base, code, sx, sy, sz, xc, x64, b, r, aw, v[0], v[1]);
    else
        refine_body<MODE, true>(F, base, code, sx, sy, sz, xc, x64, b, r, aw, v[0], v[1]);
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * size_t(slot)] = v[0];
        part[2 * size_t(slot) + 1] = v[1];
    }
}

template <int MODE>
__global__ void __launch_bounds__(kBlock) k_refine(Flat F, const uint16_t* __restrict__ code, float sx, float sy,
                                                   float sz, const float* __restrict__ xc, double* __restrict__ x64,
                                                   const float* __restrict__ b, float* __restrict__ r,
                                                   double* __restrict__ part) {
    refine_run<MODE>(F, code, sx, sy, sz, xc, x64, b, r, part, xcd_block(), blockIdx.x);
}

template <int MODE>
struct MRefine {
    static constexpr int kThreads = kBlock;
    Flat F;
    const uint16_t* code;
    float sx, sy, sz;
    const float* xc;
    double* x64;
    const float* b;
    float* r;
    double* part;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        refine_run<MODE>(F, code, sx, sy, sz, xc, x64, b, r, part, xcd_of(blk, G), blk);
    }
};

// x64 += xc, pointwise.
__device__ __forceinline__ void acc64_run(uint32_t n, const float* __restrict__ xc, double* __restrict__ x64,
                                          uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i < n) x64[i] = x64[i] + double(xc[i]);
}

__global__ void __launch_bounds__(kBlock) k_acc64(uint32_t n, const float* __restrict__ xc, double* __restrict__ x64) {
    acc64_run(n, xc, x64, xcd_block());
}

struct MAcc64 {
    static constexpr int kThreads = kBlock;
    uint32_t n;
    const float* xc;
    double* x64;
    uint32_t G;
    __device__ void run(uint32_t b) const { acc64_run(n, xc, x64, xcd_of(b, G)); }
};

// Start of an inner sweep: r = rsrc (when distinct), x = 0; partials of r.r.
template <bool CHECK>
__device__ __forceinline__ double start_body(const Flat& F, uint32_t base, const float* __restrict__ rsrc,
                                             float* __restrict__ r, float* __restrict__ x) {
    MGF_CELLS(CHECK, F, base)
    float v[kCpt];
#pragma unroll
    for (int k = 0; k < kCpt; ++k) v[k] = rsrc[cidx[k]];
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const uint32_t c = base + uint32_t(k) * kBlock;
        if (CHECK && c >= F.n) continue;
        if (r != rsrc) r[c] = v[k];
        x[c] = 0.0f;
        acc += double(v[k]) * double(v[k]);
    }
    return acc;
}

__device__ __forceinline__ void cg_start_run(Flat F, const float* __restrict__ rsrc, float* r, float* __restrict__ x,
                                             double* __restrict__ part, uint32_t blk, uint32_t slot) {
    const uint32_t b0 = blk * (kBlock * kCpt);
    const uint32_t base = b0 + threadIdx.x;
    const double acc = (b0 + kBlock * kCpt <= F.n) ? start_body<false>(F, base, rsrc, r, x)
                                                   : start_body<true>(F, base, rsrc, r, x);
    block_partial(acc, part, slot);
}

__global__ void __launch_bounds__(kBlock) k_cg_start(Flat F, const float* __restrict__ rsrc, float* r,
                                                     float* __restrict__ x, double* __restrict__ part) {
    cg_start_run(F, rsrc, r, x, part, xcd_block(), blockIdx.x);
}

struct MStart {
    static constexpr int kThreads = kBlock;
    Flat F;
    const float* rsrc;
    float *r, *x;
    double* part;
    uint32_t G;
    __device__ void run(uint32_t b) const { cg_start_run(F, rsrc, r, x, part, xcd_of(b, G), b); }
};

// Sums `count` pairs of partials (fixed order) into out[0], out[1].
__device__ __forceinline__ void sum2_run(const double* __restrict__ part, uint32_t count, double* __restrict__ out) {
    double s[2] = {0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < count; i += 1024) {
        s[0] += part[2 * size_t(i)];
        s[1] += part[2 * size_t(i) + 1];
    }
    __shared__ double sh[2][16];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double w = wave_sum(s[q]);
        if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            double t = 0.0;
            for (int w = 0; w < 16; ++w) t += sh[q][w];
            out[q] = t;
        }
    }
}

__global__ void __launch_bounds__(1024) k_sum2(const double* __restrict__ part, uint32_t count,
                                               double* __restrict__ out) {
    sum2_run(part, count, out);
}

struct MSum2 {
    static constexpr int kThreads = 1024;
    const double* part;
    uint32_t count;
    double* out;
    __device__ void run(uint32_t) const { sum2_run(part, count, out); }
};

// The one-block finishing kernel: sums `count` partials in a fixed order and
// updates the scalars.
enum FinishKind : int {
    kFinishInit = 4,      // r.r -> rz, rr, sweep_r0, target2, done; iters, alpha, beta = 0
    kFinishAlpha = 0,     // pap -> alpha
    kFinishRrIdent = 1,   // rr -> iters, done; beta = rr / rz, rz = rr (identity)
    kFinishRr = 2,        // rr -> iters, done (preconditioned: beta comes later)
    kFinishRz = 3,        // r . z -> beta = rz / rz_old (0 before the first iteration), rz
};
constexpr int kFinishBlock = 1024;

__device__ __forceinline__ void finish_run(const double* __restrict__ part, uint32_t count, CgScalars* sc, int kind,
                                           double rel_inner, double abs_target2) {
    if (kind != kFinishInit && sc->done) return;
    // four independent chains: the loads of a thread are in flight together
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    uint32_t i = threadIdx.x;
    for (; i + 3 * kFinishBlock < count; i += 4 * kFinishBlock) {
#pragma unroll
        for (int q = 0; q < 4; ++q) s4[q] += part[i + uint32_t(q) * kFinishBlock];
    }
    for (; i < count; i += kFinishBlock) s4[0] += part[i];
    double s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    __shared__ double sh[kFinishBlock / 64];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < kFinishBlock / 64; ++q) t += sh[q];
    switch (kind) {
        case kFinishInit: {
            sc->rz = t;
            sc->rr = t;
            sc->sweep_r0 = t;
            const double rel2 = rel_inner * rel_inner * t;
            sc->target2 = rel2 > abs_target2 ? rel2 : abs_target2;
            sc->done = (t <= sc->target2 || t == 0.0) ? 1 : 0;
            sc->iters = 0;
            sc->beta = 0.0;
            sc->alpha = 0.0;
            break;
        }
        case kFinishAlpha:
            sc->pap = t;
            if (!isfinite(t)) { sc->done = 2; sc->alpha = 0.0; }
            else if (t == 0.0) { sc->done = 1; sc->alpha = 0.0; }
            else sc->alpha = sc->rz / t;
            break;
        case kFinishRrIdent:
        case kFinishRr:
            sc->rr = t;
            sc->iters = sc->iters + 1;
            if (kind == kFinishRrIdent) {
                sc->beta = t / sc->rz;
                sc->rz = t;
            }
            if (t <= sc->target2) sc->done = 1;
            else if (!isfinite(t)) sc->done = 2;
            break;
        default:  // kFinishRz
            sc->beta = sc->iters == 0 ? 0.0 : t / sc->rz;
            sc->rz = t;
            if (!isfinite(t)) sc->done = 2;
            break;
    }
}

__global__ void __launch_bounds__(kFinishBlock) k_finish(const double* __restrict__ part, uint32_t count,
                                                         CgScalars* sc, int kind, double rel_inner,
                                                         double abs_target2) {
    finish_run(part, count, sc, kind, rel_inner, abs_target2);
}

struct MFinish {
    static constexpr int kThreads = kFinishBlock;
    const double* part;
    uint32_t count;
    CgScalars* sc;
    int kind;
    double rel_inner, abs_target2;
    __device__ void run(uint32_t) const { finish_run(part, count, sc, kind, rel_inner, abs_target2); }
};

// ---------------------------------------------------------------------------
// Divergence check of the projected velocity: per block the FP64 sums of
// d^2 and b^2 over the fluid cells, d = (Right - Left)/hx + (Up - Down)/hy +
// (Front - Back)/hz from the stored FP32 faces (CUDADivergence,
// MPhysics/FieldMathTool.cu:370-395, evaluated in FP64).  Cells on the last
// layer of an axis are left out: the projection sets their upper face to
// UNKNOWN (MPhysics/EulerSolverTool.cu:318-321).  Grid-stride loop; the pair
// of partials of block k goes to part[2k], part[2k+1].
__global__ void __launch_bounds__(kBlock) k_div_check(GridDev g, const uint8_t* __restrict__ m,
                                                      const float* __restrict__ u, const float* __restrict__ v,
                                                      const float* __restrict__ w, const float* __restrict__ b,
                                                      double* __restrict__ part, uint32_t n) {
    const uint32_t nx = uint32_t(g.nx), ny = uint32_t(g.ny), nz = uint32_t(g.nz);
    const double hx = g.hx, hy = g.hy, hz = g.hz;
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        if (m[i] != kFluid) continue;
        const uint32_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
        if (x + 1 >= nx || y + 1 >= ny || z + 1 >= nz) continue;
        const uint32_t fu = (z * ny + y) * (nx + 1) + x;
        const uint32_t fv = (z * (ny + 1) + y) * nx + x;
        const uint32_t fw = (z * ny + y) * nx + x;
        const double d = (double(u[fu + 1]) - double(u[fu])) / hx + (double(v[fv + nx]) - double(v[fv])) / hy +
                         (double(w[fw + nx * ny]) - double(w[fw])) / hz;
        const double bc = double(b[i]);
        acc[0] += d * d;
        acc[1] += bc * bc;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[2 * size_t(blockIdx.x)] = acc[0];
        part[2 * size_t(blockIdx.x) + 1] = acc[1];
    }
}

// out[0] = max(out[0], sqrt(sums[0] / sums[1])) (sqrt(sums[0]) when ||b|| = 0);
// a NaN ratio is kept.
__global__ void k_div_fold(const double* __restrict__ sums, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const double r = sums[1] > 0.0 ? sqrt(sums[0] / sums[1]) : sqrt(sums[0]);
    if (!(r <= out[0])) out[0] = r;
}

// ---------------------------------------------------------------------------
// Multigrid (preconditioner).  Level l: cells with code_l, right-hand side
// b_l, iterate x_l; a coarse cell is a 2x2x2 block of fine cells (one layer
// on an axis whose extent is 1).  Damped Jacobi smoothing with the diagonal
// D (zero rows are left at 0), restriction = weight * sum over the fluid
// children's residuals, prolongation = injection into fluid children: with as
// many sweeps after the coarse correction as before it, the V-cycle is a
// symmetric operator.
using MgLevel = MgDims;

// Coarse markers: with fluid_first, fluid if any child is fluid, else air if
// any child is air, else solid; otherwise air wins over fluid.
__device__ __forceinline__ void k_mg_coarsen_run(MgLevel F, MgLevel Cc, const uint8_t* __restrict__ mf,
                                                       uint8_t* __restrict__ mc, int fluid_first, uint32_t n, uint32_t blk) {
    const uint32_t i = blk * kBlock + threadIdx.x;
    if (i >= n) return;
    const int X = int(i % uint32_t(Cc.nx)), Y = int((i / uint32_t(Cc.nx)) % uint32_t(Cc.ny)),
              Z = int(i / (uint32_t(Cc.nx) * uint32_t(Cc.ny)));
    const int fx = F.nx > 1 ? 2 : 1, fy = F.ny > 1 ? 2 : 1, fz = F.nz > 1 ? 2 : 1;
    bool any_f = false, any_a = false;
    for (int dz = 0; dz < fz; ++dz)
        for (int dy = 0; dy < fy; ++dy)
            for (int dx = 0; dx < fx; ++dx) {
                const int x = X * fx + dx, y = Y * fy + dy, z = Z * fz + dz;
                if (x >= F.nx || y >= F.ny || z >= F.nz) continue;
                const uint8_t q = mf[(uint32_t(z) * uint32_t(F.ny) + uint32_t(y)) * uint32_t(F.nx) + uint32_t(x)];
                any_f |= q == kFluid;
                any_a |= q == kAir;
            }
    uint8_t out;
    if (fluid_first) out = any_f ? kFluid : (any_a ? kAir : kSolid);
    else out = any_a ? kAir : (any_f ? kFluid : kSolid);
    mc[i] = out;
}

__global__ void __launch_bounds__(kBlock) k_mg_coarsen(MgLevel F, MgLevel Cc, const uint8_t* __restrict__ mf,
                                                       uint8_t* __restrict__ mc, int fluid_first, uint32_t n) {
    k_mg_coarsen_run(F, Cc, mf, mc, fluid_first, n, xcd_block());
}

struct MCoarsen {
    static constexpr int kThreads = kBlock;
    MgLevel F;
    MgLevel Cc;
    const uint8_t* mf;
    uint8_t* mc;
    int fluid_first;
    uint32_t n;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_mg_coarsen_run(F, Cc, mf, mc, fluid_first, n, xcd_of(vb, G)); }
};

// One damped Jacobi sweep xo = xi + omega D^-1 (b - A xi) on fluid cells (0
// elsewhere, and on zero rows: a fluid cell whose neighbours are all solid
// or outside the grid); FIRST: the sweep from xi = 0, xo = omega D^-1 b.  RZ
// (the level-0 final sweep, b = r there) also writes block partials of r . z.
template <bool FIRST, bool CHECK>
__device__ __forceinline__ double smooth_body(const Flat& F, const MgLevel& L, uint32_t base,
                                              const uint16_t* __restrict__ code, const float* __restrict__ b,
                                              const float* __restrict__ xi, float* __restrict__ xo,
                                              const float* aw, const float* w) {
    MGF_CELLS(CHECK, F, base)
    uint16_t cd[kCpt];
    float bc[kCpt], vc[kCpt], vn[kCpt][6];
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        cd[k] = code[cidx[k]];
        bc[k] = b[cidx[k]];
        if (!FIRST) {
            uint32_t j[6];
            nbrs(cidx[k], F, j);
            vc[k] = xi[cidx[k]];
#pragma unroll
            for (int m = 0; m < 6; ++m) vn[k][m] = xi[j[m]];
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const float wd = wdiag(cd[k], w);
        float out;
        if (FIRST) {
            out = wd * bc[k];
        } else {
            const float ax = arow(cd[k], vc[k], vn[k], L.sx, L.sy, L.sz, aw);
            out = (cd[k] & kCodeFluid) ? vc[k] + wd * (bc[k] - ax) : 0.0f;
        }
        const uint32_t c = base + uint32_t(k) * kBlock;
        if (!CHECK || c < F.n) {
            xo[c] = out;
            acc += double(bc[k]) * double(out);
        }
    }
    return acc;
}

template <bool FIRST, bool RZ>
__device__ __forceinline__ void smooth_run(Flat F, MgLevel L, const uint16_t* __restrict__ code,
                                           const float* __restrict__ b, const float* __restrict__ xi,
                                           float* __restrict__ xo, float omega, const CgScalars* __restrict__ sc,
                                           double* __restrict__ part, uint32_t blk, uint32_t slot) {
    __shared__ float aw[64], w[64];
    if (sc->done) return;
    build_tables<true>(aw, w, L.sx, L.sy, L.sz, omega);
    const uint32_t b0 = blk * (kBlock * kCpt);
    const uint32_t base = b0 + threadIdx.x;
    const double acc = (b0 + kBlock * kCpt <= F.n)
                           ? smooth_body<FIRST, false>(F, L, base, code, b, xi, xo, aw, w)
                           : smooth_body<FIRST, true>(F, L, base, code, b, xi, xo, aw, w);
    if (RZ) block_partial(acc, part, slot);
}

template <bool FIRST, bool RZ>
__global__ void __launch_bounds__(kBlock) k_mg_smooth(Flat F, MgLevel L, const uint16_t* __restrict__ code,
                                                      const float* __restrict__ b, const float* __restrict__ xi,
                                                      float* __restrict__ xo, float omega,
                                                      const CgScalars* __restrict__ sc, double* __restrict__ part) {
    smooth_run<FIRST, RZ>(F, L, code, b, xi, xo, omega, sc, part, xcd_block(), blockIdx.x);
}

template <bool FIRST, bool RZ>
struct MSmooth {
    static constexpr int kThreads = kBlock;
    Flat F;
    MgLevel L;
    const uint16_t* code;
    const float *b, *xi;
    float* xo;
    float omega;
    const CgScalars* sc;
    double* part;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        smooth_run<FIRST, RZ>(F, L, code, b, xi, xo, omega, sc, part, xcd_of(blk, G), blk);
    }
};

// The first two sweeps from x = 0 in one pass: x1 = omega D^-1 b is
// pointwise, so x2 = x1 + omega D^-1 (b - A x1) needs only b and the codes at
// the seven points.
template <bool CHECK>
__device__ __forceinline__ double smooth2_body(const Flat& F, const MgLevel& L, uint32_t base,
                                               const uint16_t* __restrict__ code, const float* __restrict__ b,
                                               float* __restrict__ xo, const float* aw, const float* w) {
    MGF_CELLS(CHECK, F, base)
    uint16_t cd[kCpt], cn[kCpt][6];
    float bc[kCpt], bn[kCpt][6];
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        uint32_t j[6];
        nbrs(cidx[k], F, j);
        cd[k] = code[cidx[k]];
        bc[k] = b[cidx[k]];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            cn[k][m] = code[j[m]];
            bn[k][m] = b[j[m]];
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const float wd = wdiag(cd[k], w);
        const float x1c = wd * bc[k];
        float x1[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) x1[m] = wdiag(cn[k][m], w) * bn[k][m];
        const float ax = arow(cd[k], x1c, x1, L.sx, L.sy, L.sz, aw);
        const float out = (cd[k] & kCodeFluid) ? x1c + wd * (bc[k] - ax) : 0.0f;
        const uint32_t c = base + uint32_t(k) * kBlock;
        if (!CHECK || c < F.n) {
            xo[c] = out;
            acc += double(bc[k]) * double(out);
        }
    }
    return acc;
}

template <bool RZ>
__device__ __forceinline__ void smooth2_run(Flat F, MgLevel L, const uint16_t* __restrict__ code,
                                            const float* __restrict__ b, float* __restrict__ xo, float omega,
                                            const CgScalars* __restrict__ sc, double* __restrict__ part,
                                            uint32_t blk, uint32_t slot) {
    __shared__ float aw[64], w[64];
    if (sc->done) return;
    build_tables<true>(aw, w, L.sx, L.sy, L.sz, omega);
    const uint32_t b0 = blk * (kBlock * kCpt);
    const uint32_t base = b0 + threadIdx.x;
    const double acc = (b0 + kBlock * kCpt <= F.n) ? smooth2_body<false>(F, L, base, code, b, xo, aw, w)
                                                   : smooth2_body<true>(F, L, base, code, b, xo, aw, w);
    if (RZ) block_partial(acc, part, slot);
}

template <bool RZ>
__global__ void __launch_bounds__(kBlock) k_mg_smooth2_first(Flat F, MgLevel L, const uint16_t* __restrict__ code,
                                                             const float* __restrict__ b, float* __restrict__ xo,
                                                             float omega, const CgScalars* __restrict__ sc,
                                                             double* __restrict__ part) {
    smooth2_run<RZ>(F, L, code, b, xo, omega, sc, part, xcd_block(), blockIdx.x);
}

template <bool RZ>
struct MSmooth2 {
    static constexpr int kThreads = kBlock;
    Flat F;
    MgLevel L;
    const uint16_t* code;
    const float* b;
    float* xo;
    float omega;
    const CgScalars* sc;
    double* part;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        smooth2_run<RZ>(F, L, code, b, xo, omega, sc, part, xcd_of(blk, G), blk);
    }
};

// The first sweep after the coarse correction, with the prolongation by
// injection applied on the fly: x_in = x + xc[parent] at the seven points
// (only coupled, hence fluid, neighbours are used), xo = x_in + omega D^-1
// (b - A x_in); x_in is not stored.  h[xyz] = 1 where the axis was halved.
struct Prolong {
    Flat FC;
    int hx, hy, hz;
    double rnx, rplane;
};

template <bool CHECK>
__device__ __forceinline__ double smooth_prolong_body(const Flat& F, const MgLevel& L, const Prolong& P,
                                                      uint32_t base, const uint16_t* __restrict__ code,
                                                      const float* __restrict__ b, const float* __restrict__ xi,
                                                      const float* __restrict__ xc, float* __restrict__ xo,
                                                      const float* aw, const float* w) {
    MGF_CELLS(CHECK, F, base)
    uint16_t cd[kCpt];
    float bc[kCpt], vc[kCpt], vn[kCpt][6];
    const uint32_t nx = uint32_t(L.nx), ny = uint32_t(L.ny), nz = uint32_t(L.nz);
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const uint32_t cc = cidx[k];
        const uint32_t z = udiv(cc, F.sz, P.rplane);
        const uint32_t rem = cc - z * F.sz;
        const uint32_t y = udiv(rem, F.sy, P.rnx);
        const uint32_t x = rem - y * F.sy;
        const uint32_t X = x >> P.hx, Y = y >> P.hy, Z = z >> P.hz;
        const uint32_t Xm = (x > 0 ? x - 1 : 0) >> P.hx, Xp = (x + 1 < nx ? x + 1 : x) >> P.hx;
        const uint32_t Ym = (y > 0 ? y - 1 : 0) >> P.hy, Yp = (y + 1 < ny ? y + 1 : y) >> P.hy;
        const uint32_t Zm = (z > 0 ? z - 1 : 0) >> P.hz, Zp = (z + 1 < nz ? z + 1 : z) >> P.hz;
        const uint32_t pz = Z * P.FC.sz, py = Y * P.FC.sy;
        const uint32_t pp[6] = {pz + py + Xm, pz + py + Xp, pz + Ym * P.FC.sy + X,
                                pz + Yp * P.FC.sy + X, Zm * P.FC.sz + py + X, Zp * P.FC.sz + py + X};
        uint32_t j[6];
        nbrs(cc, F, j);
        cd[k] = code[cc];
        bc[k] = b[cc];
        vc[k] = xi[cc] + xc[pz + py + X];
#pragma unroll
        for (int m = 0; m < 6; ++m) vn[k][m] = xi[j[m]] + xc[pp[m]];
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
        const float wd = wdiag(cd[k], w);
        const float ax = arow(cd[k], vc[k], vn[k], L.sx, L.sy, L.sz, aw);
        const float out = (cd[k] & kCodeFluid) ? vc[k] + wd * (bc[k] - ax) : 0.0f;
        const uint32_t c = base + uint32_t(k) * kBlock;
        if (!CHECK || c < F.n) {
            xo[c] = out;
            acc += double(bc[k]) * double(out);
        }
    }
    return acc;
}

template <bool RZ>
__device__ __forceinline__ void smooth_prolong_run(Flat F, MgLevel L, Prolong P, const uint16_t* __restrict__ code,
                                                   const float* __restrict__ b, const float* __restrict__ xi,
                                                   const float* __restrict__ xc, float* __restrict__ xo, float omega,
                                                   const CgScalars* __restrict__ sc, double* __restrict__ part,
                                                   uint32_t blk, uint32_t slot) {
    __shared__ float aw[64], w[64];
    if (sc->done) return;
    build_tables<true>(aw, w, L.sx, L.sy, L.sz, omega);
    const uint32_t b0 = blk * (kBlock * kCpt);
    const uint32_t base = b0 + threadIdx.x;
    const double acc = (b0 + kBlock * kCpt <= F.n)
                           ? smooth_prolong_body<false>(F, L, P, base, code, b, xi, xc, xo, aw, w)
                           : smooth_prolong_body<true>(F, L, P, base, code, b, xi, xc, xo, aw, w);
    if (RZ) block_partial(acc, part, slot);
}

template <bool RZ>
__global__ void __launch_bounds__(kBlock) k_mg_smooth_prolong(Flat F, MgLevel L, Prolong P,
                                                              const uint16_t* __restrict__ code,
                                                              const float* __restrict__ b,
                                                              const float* __restrict__ xi,
                                                              const float* __restrict__ xc, float* __restrict__ xo,
                                                              float omega, const CgScalars* __restrict__ sc,
                                                              double* __restrict__ part) {
    smooth_prolong_run<RZ>(F, L, P, code, b, xi, xc, xo, omega, sc, part, xcd_block(), blockIdx.x);
}

template <bool RZ>
struct MProlong {
    static constexpr int kThreads = kBlock;
    Flat F;
    MgLevel L;
    Prolong P;
    const uint16_t* code;
    const float *b, *xi, *xc;
    float* xo;
    float omega;
    const CgScalars* sc;
    double* part;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        smooth_prolong_run<RZ>(F, L, P, code, b, xi, xc, xo, omega, sc, part, xcd_of(blk, G), blk);
    }
};

// Restriction of the residual, one work-item per coarse cell:
// bc = weight * sum over fluid children of (b - A x).
__device__ __forceinline__ void restrict_run(Flat FC, MgLevel Cd, Flat FF, MgLevel Fd,
                                             const uint16_t* __restrict__ codef, const float* __restrict__ bf,
                                             const float* __restrict__ xf, const uint16_t* __restrict__ codec,
                                             float* __restrict__ bc, float weight, double rnx, double rplane,
                                             const CgScalars* __restrict__ sc, uint32_t blk) {
    __shared__ float aw[64];
    if (sc->done) return;
    build_tables<false>(aw, nullptr, Fd.sx, Fd.sy, Fd.sz, 0.0f);
    const uint32_t c = blk * kBlock + threadIdx.x;
    const bool ok = c < FC.n;
    const uint32_t cc = ok ? c : FC.n - 1;
    const uint32_t Z = udiv(cc, FC.sz, rplane);
    const uint32_t rem = cc - Z * FC.sz;
    const uint32_t Y = udiv(rem, FC.sy, rnx);
    const uint32_t X = rem - Y * FC.sy;
    const int fx = Fd.nx > 1 ? 2 : 1, fy = Fd.ny > 1 ? 2 : 1, fz = Fd.nz > 1 ? 2 : 1;
    const uint32_t x0 = X * uint32_t(fx), y0 = Y * uint32_t(fy), z0 = Z * uint32_t(fz);
    const uint32_t f0 = z0 * FF.sz + y0 * FF.sy + x0;
    float sum = 0.0f;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const bool in = dx < fx && dy < fy && dz < fz && x0 + uint32_t(dx) < uint32_t(Fd.nx) &&
                                y0 + uint32_t(dy) < uint32_t(Fd.ny) && z0 + uint32_t(dz) < uint32_t(Fd.nz);
                const uint32_t f = in ? f0 + uint32_t(dz) * FF.sz + uint32_t(dy) * FF.sy + uint32_t(dx) : f0;
                uint32_t j[6];
                nbrs(f, FF, j);
                const uint16_t cd = codef[f];
                const float bv = bf[f];
                const float vc = xf[f];
                float v[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) v[m] = xf[j[m]];
                const float res = bv - arow(cd, vc, v, Fd.sx, Fd.sy, Fd.sz, aw);
                sum += (in && (cd & kCodeFluid)) ? res : 0.0f;
            }
    if (ok) bc[c] = (codec[c] & kCodeFluid) ? sum * weight : 0.0f;
}

__global__ void __launch_bounds__(kBlock) k_mg_restrict(Flat FC, MgLevel Cd, Flat FF, MgLevel Fd,
                                                        const uint16_t* __restrict__ codef,
                                                        const float* __restrict__ bf, const float* __restrict__ xf,
                                                        const uint16_t* __restrict__ codec, float* __restrict__ bc,
                                                        float weight, double rnx, double rplane,
                                                        const CgScalars* __restrict__ sc) {
    restrict_run(FC, Cd, FF, Fd, codef, bf, xf, codec, bc, weight, rnx, rplane, sc, xcd_block());
}

struct MRestrict {
    static constexpr int kThreads = kBlock;
    Flat FC;
    MgLevel Cd;
    Flat FF;
    MgLevel Fd;
    const uint16_t* codef;
    const float *bf, *xf;
    const uint16_t* codec;
    float* bc;
    float weight;
    double rnx, rplane;
    const CgScalars* sc;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        restrict_run(FC, Cd, FF, Fd, codef, bf, xf, codec, bc, weight, rnx, rplane, sc, xcd_of(blk, G));
    }
};

// Restriction when every axis is halved: one work-item per fine cell.  A
// block covers 64 fine cells along x in each of the 2 x 2 rows (y, z) of one
// coarse row: each wave computes its row's residuals, sums x-pairs with a
// lane shuffle, and 32 work-items add the four rows' pair sums (a fixed
// order) into the 32 coarse cells.
__device__ __forceinline__ void restrict2_run(Flat FF, MgLevel Fd, Flat FC, MgLevel Cd, uint32_t gx,
                                              const uint16_t* __restrict__ codef, const float* __restrict__ bf,
                                              const float* __restrict__ xf, const uint16_t* __restrict__ codec,
                                              float* __restrict__ bc, float weight, const CgScalars* __restrict__ sc,
                                              uint32_t blk) {
    __shared__ float aw[64];
    __shared__ float sh[4][32];
    if (sc->done) return;
    build_tables<false>(aw, nullptr, Fd.sx, Fd.sy, Fd.sz, 0.0f);
    const uint32_t xcb = blk % gx;                  // 64-cell chunk along x
    const uint32_t YZ = blk / gx;
    const uint32_t Y = YZ % uint32_t(Cd.ny), Z = YZ / uint32_t(Cd.ny);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t x = xcb * 64 + uint32_t(lane);
    const uint32_t y = 2 * Y + uint32_t(wv & 1), z = 2 * Z + uint32_t(wv >> 1);
    const bool in = Z < uint32_t(Cd.nz) && x < uint32_t(Fd.nx) && y < uint32_t(Fd.ny) && z < uint32_t(Fd.nz);
    const uint32_t f = in ? (z * FF.sz + y * FF.sy + x) : 0u;
    uint32_t j[6];
    nbrs(f, FF, j);
    const uint16_t cd = codef[f];
    const float bv = bf[f];
    const float vc = xf[f];
    float v[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) v[m] = xf[j[m]];
    float res = (in && (cd & kCodeFluid)) ? bv - arow(cd, vc, v, Fd.sx, Fd.sy, Fd.sz, aw) : 0.0f;
    res += __shfl_xor(res, 1, 64);
    if ((lane & 1) == 0) sh[wv][lane >> 1] = res;
    __syncthreads();
    if (threadIdx.x < 32) {
        const uint32_t X = xcb * 32 + threadIdx.x;
        if (Z < uint32_t(Cd.nz) && X < uint32_t(Cd.nx)) {
            const float sum = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
            const uint32_t c = Z * FC.sz + Y * FC.sy + X;
            bc[c] = (codec[c] & kCodeFluid) ? sum * weight : 0.0f;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_mg_restrict2(Flat FF, MgLevel Fd, Flat FC, MgLevel Cd, uint32_t gx,
                                                         const uint16_t* __restrict__ codef,
                                                         const float* __restrict__ bf, const float* __restrict__ xf,
                                                         const uint16_t* __restrict__ codec, float* __restrict__ bc,
                                                         float weight, const CgScalars* __restrict__ sc) {
    restrict2_run(FF, Fd, FC, Cd, gx, codef, bf, xf, codec, bc, weight, sc, xcd_block());
}

struct MRestrict2 {
    static constexpr int kThreads = kBlock;
    Flat FF;
    MgLevel Fd;
    Flat FC;
    MgLevel Cd;
    uint32_t gx;
    const uint16_t* codef;
    const float *bf, *xf;
    const uint16_t* codec;
    float* bc;
    float weight;
    const CgScalars* sc;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        restrict2_run(FF, Fd, FC, Cd, gx, codef, bf, xf, codec, bc, weight, sc, xcd_of(blk, G));
    }
};

// ---------------------------------------------------------------------------
// Tiled restriction.  A block owns a 64 x 4 x 4 tile of fine cells (work-item
// (lx, ly) the four cells of its column lz = 0..3) and first stages the
// 66 x 6 x 6 box of x around the tile in LDS with coalesced loads (0 outside
// the grid); the residual stencils then read LDS.  (For the smoothers the
// same staging measured slower than the flat kernels: the halo of a 4 x 4
// cross-section costs 2.3 loads per cell and the staging barrier removes the
// overlap of loads and arithmetic.)
constexpr int kTX = 64, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kBox = kHX * kHY * kHZ;
constexpr int kBoxIters = (kBox + kBlock - 1) / kBlock;

struct Tiles {
    uint32_t tx, ty, tz;       // tiles along x, y, z
    int nx, ny, nz;            // level extents
    uint32_t sy, sz;           // strides
};

inline Tiles make_tiles(const MgDims& d) {
    Tiles T;
    T.nx = d.nx; T.ny = d.ny; T.nz = d.nz;
    T.tx = uint32_t((d.nx + kTX - 1) / kTX);
    T.ty = uint32_t((d.ny + kTY - 1) / kTY);
    T.tz = uint32_t((d.nz + kTZ - 1) / kTZ);
    T.sy = uint32_t(d.nx);
    T.sz = uint32_t(d.nx) * uint32_t(d.ny);
    return T;
}
inline uint32_t tiles_grid(const Tiles& T) {
    const size_t g = size_t(T.tx) * T.ty * T.tz;
    return uint32_t((g + kXcds - 1) / kXcds * kXcds);
}

// Tile origin of this block (the extra blocks of the rounded grid get a z
// beyond the grid: every cell of theirs is outside).
__device__ __forceinline__ void tile_origin(const Tiles& T, uint32_t blk, int& x0, int& y0, int& z0) {
    const uint32_t tx = blk % T.tx, r = blk / T.tx;
    const uint32_t ty = r % T.ty, tz = r / T.ty;
    x0 = int(tx) * kTX;
    y0 = int(ty) * kTY;
    z0 = int(tz) * kTZ;
}

// Global index of box element e (box origin = tile origin - 1), or -1.
__device__ __forceinline__ int64_t box_global(const Tiles& T, int x0, int y0, int z0, int e) {
    const int hz = e / (kHX * kHY), rr = e - hz * (kHX * kHY), hy = rr / kHX, hx = rr - hy * kHX;
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
    const bool in = e < kBox && unsigned(gx) < unsigned(T.nx) && unsigned(gy) < unsigned(T.ny) &&
                    unsigned(gz) < unsigned(T.nz);
    return in ? int64_t(uint32_t(gz) * T.sz + uint32_t(gy) * T.sy + uint32_t(gx)) : int64_t(-1);
}

// LDS index of the work-item's cell lz (the box interior starts at 1,1,1).
__device__ __forceinline__ int box_cell(int lz) {
    return ((lz + 1) * kHY + (int(threadIdx.x) >> 6) + 1) * kHX + (int(threadIdx.x) & 63) + 1;
}

template <class T>
__device__ __forceinline__ void load_box(T* dst, const T* __restrict__ src, const Tiles& TL, int x0, int y0,
                                         int z0) {
    T v[kBoxIters];
#pragma unroll
    for (int q = 0; q < kBoxIters; ++q) {
        const int64_t g = box_global(TL, x0, y0, z0, int(threadIdx.x) + q * kBlock);
        v[q] = src[g >= 0 ? uint32_t(g) : 0u];
        if (g < 0) v[q] = T(0);
    }
#pragma unroll
    for (int q = 0; q < kBoxIters; ++q) {
        const int e = int(threadIdx.x) + q * kBlock;
        if (e < kBox) dst[e] = v[q];
    }
}

__device__ __forceinline__ void box_nbrs(const float* box, int e, float (&v)[6]) {
    v[0] = box[e - 1];
    v[1] = box[e + 1];
    v[2] = box[e - kHX];
    v[3] = box[e + kHX];
    v[4] = box[e - kHX * kHY];
    v[5] = box[e + kHX * kHY];
}

// Restriction on a tile (every axis halved): the residuals of the tile's
// cells go to LDS, then 128 work-items sum each 2 x 2 x 2 block in a fixed
// order into the 32 x 2 x 2 coarse cells of the tile.
__device__ __forceinline__ void restrict_tile_run(Tiles TL, MgLevel Fd, Tiles TC, const uint16_t* __restrict__ codef,
                                                  const float* __restrict__ bf, const float* __restrict__ xf,
                                                  const uint16_t* __restrict__ codec, float* __restrict__ bc,
                                                  const CgScalars* __restrict__ sc, uint32_t blk) {
    __shared__ float aw[64];
    __shared__ float xb[kBox];
    __shared__ float res[kTZ][kTY][kTX];
    if (sc->done) return;
    build_tables<false>(aw, nullptr, Fd.sx, Fd.sy, Fd.sz, 0.0f);
    int x0, y0, z0;
    tile_origin(TL, blk, x0, y0, z0);
    const int lx = int(threadIdx.x) & 63, ly = int(threadIdx.x) >> 6;
    const int gx = x0 + lx, gy = y0 + ly;
    uint16_t cd[kTZ];
    float bv[kTZ];
    bool in[kTZ];
#pragma unroll
    for (int lz = 0; lz < kTZ; ++lz) {
        const int gz = z0 + lz;
        in[lz] = gx < TL.nx && gy < TL.ny && gz < TL.nz;
        const uint32_t gi = in[lz] ? uint32_t(gz) * TL.sz + uint32_t(gy) * TL.sy + uint32_t(gx) : 0u;
        cd[lz] = codef[gi];
        bv[lz] = bf[gi];
    }
    load_box(xb, xf, TL, x0, y0, z0);
    __syncthreads();
#pragma unroll
    for (int lz = 0; lz < kTZ; ++lz) {
        const int e = box_cell(lz);
        float v[6];
        box_nbrs(xb, e, v);
        const float r = bv[lz] - arow(cd[lz], xb[e], v, Fd.sx, Fd.sy, Fd.sz, aw);
        res[lz][ly][lx] = (in[lz] && (cd[lz] & kCodeFluid)) ? r : 0.0f;
    }
    __syncthreads();
    if (threadIdx.x < 128) {
        const int cx = int(threadIdx.x) & 31, cy = (int(threadIdx.x) >> 5) & 1, cz = int(threadIdx.x) >> 6;
        const int X = x0 / 2 + cx, Y = y0 / 2 + cy, Z = z0 / 2 + cz;
        if (X < TC.nx && Y < TC.ny && Z < TC.nz) {
            const int fx = 2 * cx, fy = 2 * cy, fz = 2 * cz;
            const float s = ((res[fz][fy][fx] + res[fz][fy][fx + 1]) + (res[fz][fy + 1][fx] + res[fz][fy + 1][fx + 1])) +
                            ((res[fz + 1][fy][fx] + res[fz + 1][fy][fx + 1]) +
                             (res[fz + 1][fy + 1][fx] + res[fz + 1][fy + 1][fx + 1]));
            const uint32_t c = uint32_t(Z) * TC.sz + uint32_t(Y) * TC.sy + uint32_t(X);
            bc[c] = (codec[c] & kCodeFluid) ? s * 0.125f : 0.0f;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_mg_restrict_tile(Tiles TL, MgLevel Fd, Tiles TC,
                                                             const uint16_t* __restrict__ codef,
                                                             const float* __restrict__ bf,
                                                             const float* __restrict__ xf,
                                                             const uint16_t* __restrict__ codec,
                                                             float* __restrict__ bc,
                                                             const CgScalars* __restrict__ sc) {
    restrict_tile_run(TL, Fd, TC, codef, bf, xf, codec, bc, sc, xcd_block());
}

struct MRestrictTile {
    static constexpr int kThreads = kBlock;
    Tiles TL;
    MgLevel Fd;
    Tiles TC;
    const uint16_t* codef;
    const float *bf, *xf;
    const uint16_t* codec;
    float* bc;
    const CgScalars* sc;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        restrict_tile_run(TL, Fd, TC, codef, bf, xf, codec, bc, sc, xcd_of(blk, G));
    }
};

// ---------------------------------------------------------------------------
// The V-cycle on the smallest levels in one work-group, iterates in LDS.
// Level 0 of the argument list takes its right-hand side from global memory
// (written by the restriction of the level above) and writes its result back
// (read by the prolongation of the level above); deeper levels keep b, x and
// the Jacobi partner in LDS.  The arithmetic is the multi-kernel cycle's.
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseMaxLevels = 8;

struct CoarseLevel {
    int nx, ny, nz;
    float sx, sy, sz;
    const uint16_t* code;
};

struct CoarseArgs {
    int nlev, nu, last_sweeps;
    float omega;
    const float* b0;
    float* x0;
    CoarseLevel lv[kCoarseMaxLevels];
};

__device__ __forceinline__ void lds_sweep(const CoarseLevel& L, const uint16_t* code, const float* aw,
                                          const float* w, const float* b, const float* xi, float* xo, bool first) {
    const uint32_t n = uint32_t(L.nx) * uint32_t(L.ny) * uint32_t(L.nz);
    const Flat F{n, uint32_t(L.nx), uint32_t(L.nx) * uint32_t(L.ny)};
    for (uint32_t c = threadIdx.x; c < n; c += kCoarseThreads) {
        const uint16_t cd = code[c];
        const float wd = wdiag(cd, w);
        const float bc = b[c];
        float out;
        if (first) {
            out = wd * bc;
        } else {
            uint32_t j[6];
            nbrs(c, F, j);
            float v[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) v[m] = xi[j[m]];
            const float vc = xi[c];
            out = (cd & kCodeFluid) ? vc + wd * (bc - arow(cd, vc, v, L.sx, L.sy, L.sz, aw)) : 0.0f;
        }
        xo[c] = out;
    }
    __syncthreads();
}

// LDS layout (floats): per level the tables aw[64], w[64]; then per level the
// codes (uint16, each level padded to an even count); then level 0: x, t;
// deeper levels: b, x, t.
__device__ __forceinline__ void coarse_run(const CoarseArgs& A, const CgScalars* __restrict__ sc) {
    if (sc->done) return;
    extern __shared__ float lds[];
    // Offsets (floats from lds) of each level's codes, right-hand side,
    // iterate and scratch iterate, in LDS rather than in per-work-item arrays
    // (which a runtime level index would put in scratch memory); par bit l
    // swaps level l's iterate and scratch iterate.
    __shared__ uint32_t s_code[kCoarseMaxLevels], s_b[kCoarseMaxLevels], s_x[kCoarseMaxLevels],
        s_t[kCoarseMaxLevels], s_n[kCoarseMaxLevels];
    float* tab = lds;
    if (threadIdx.x == 0) {
        uint32_t q = 128u * uint32_t(A.nlev);
        for (int l = 0; l < A.nlev; ++l) {
            s_n[l] = uint32_t(A.lv[l].nx) * uint32_t(A.lv[l].ny) * uint32_t(A.lv[l].nz);
            s_code[l] = q;
            q += (s_n[l] + 1) / 2;
        }
        for (int l = 0; l < A.nlev; ++l) {
            s_b[l] = q;
            if (l > 0) q += s_n[l];
            s_x[l] = q; q += s_n[l];
            s_t[l] = q; q += s_n[l];
        }
    }
    __syncthreads();
    uint32_t par = 0;
    auto codes = [&](int l) { return reinterpret_cast<uint16_t*>(lds + s_code[l]); };
    auto bs = [&](int l) { return l == 0 ? A.b0 : static_cast<const float*>(lds + s_b[l]); };
    auto xs = [&](int l) { return lds + (((par >> l) & 1u) ? s_t[l] : s_x[l]); };
    auto ts = [&](int l) { return lds + (((par >> l) & 1u) ? s_x[l] : s_t[l]); };
    auto ns = [&](int l) { return s_n[l]; };
    // tables and codes of every level
    for (int l = 0; l < A.nlev; ++l) {
        if (threadIdx.x < 64) {
            const float d = mask_scale_sum(threadIdx.x, A.lv[l].sx, A.lv[l].sy, A.lv[l].sz);
            tab[128 * l + threadIdx.x] = d;
            tab[128 * l + 64 + threadIdx.x] = d != 0.0f ? A.omega / d : 0.0f;
        }
        for (uint32_t c = threadIdx.x; c < ns(l); c += kCoarseThreads) codes(l)[c] = A.lv[l].code[c];
    }
    __syncthreads();
    // descent
    for (int l = 0; l < A.nlev; ++l) {
        const CoarseLevel& L = A.lv[l];
        const float* aw = tab + 128 * l;
        const float* w = aw + 64;
        const int sweeps = l == A.nlev - 1 ? A.last_sweeps : A.nu;
        lds_sweep(L, codes(l), aw, w, bs(l), xs(l), xs(l), true);
        for (int k = 1; k < sweeps; ++k) {
            lds_sweep(L, codes(l), aw, w, bs(l), xs(l), ts(l), false);
            par ^= 1u << l;
        }
        if (l == A.nlev - 1) break;
        // restriction of the residual into level l+1
        const CoarseLevel& C = A.lv[l + 1];
        const uint32_t nc = ns(l + 1);
        const int fx = L.nx > 1 ? 2 : 1, fy = L.ny > 1 ? 2 : 1, fz = L.nz > 1 ? 2 : 1;
        const float weight = 1.0f / float(fx * fy * fz);
        const Flat FF{ns(l), uint32_t(L.nx), uint32_t(L.nx) * uint32_t(L.ny)};
        float* bc = const_cast<float*>(bs(l + 1));
        const float* xf = xs(l);
        const float* bf = bs(l);
        for (uint32_t c = threadIdx.x; c < nc; c += kCoarseThreads) {
            const uint32_t X = c % uint32_t(C.nx), Y = (c / uint32_t(C.nx)) % uint32_t(C.ny),
                           Z = c / (uint32_t(C.nx) * uint32_t(C.ny));
            float sum = 0.0f;
            for (int dz = 0; dz < fz; ++dz)
                for (int dy = 0; dy < fy; ++dy)
                    for (int dx = 0; dx < fx; ++dx) {
                        const uint32_t x = X * fx + dx, y = Y * fy + dy, z = Z * fz + dz;
                        if (x >= uint32_t(L.nx) || y >= uint32_t(L.ny) || z >= uint32_t(L.nz)) continue;
                        const uint32_t f = z * FF.sz + y * FF.sy + x;
                        const uint16_t cd = codes(l)[f];
                        if (!(cd & kCodeFluid)) continue;
                        uint32_t j[6];
                        nbrs(f, FF, j);
                        float v[6];
#pragma unroll
                        for (int m = 0; m < 6; ++m) v[m] = xf[j[m]];
                        sum += bf[f] - arow(cd, xf[f], v, L.sx, L.sy, L.sz, aw);
                    }
            bc[c] = (codes(l + 1)[c] & kCodeFluid) ? sum * weight : 0.0f;
        }
        __syncthreads();
    }
    // ascent
    for (int l = A.nlev - 2; l >= 0; --l) {
        const CoarseLevel& L = A.lv[l];
        const CoarseLevel& C = A.lv[l + 1];
        const float* aw = tab + 128 * l;
        const float* w = aw + 64;
        const int hx = L.nx > 1 ? 1 : 0, hy = L.ny > 1 ? 1 : 0, hz = L.nz > 1 ? 1 : 0;
        const float* xc = xs(l + 1);
        float* xf = xs(l);
        for (uint32_t c = threadIdx.x; c < ns(l); c += kCoarseThreads) {
            const uint32_t x = c % uint32_t(L.nx), y = (c / uint32_t(L.nx)) % uint32_t(L.ny),
                           z = c / (uint32_t(L.nx) * uint32_t(L.ny));
            if (codes(l)[c] & kCodeFluid)
                xf[c] += xc[((z >> hz) * uint32_t(C.ny) + (y >> hy)) * uint32_t(C.nx) + (x >> hx)];
        }
        __syncthreads();
        for (int k = 0; k < A.nu; ++k) {
            lds_sweep(L, codes(l), aw, w, bs(l), xs(l), ts(l), false);
            par ^= 1u << l;
        }
    }
    for (uint32_t c = threadIdx.x; c < ns(0); c += kCoarseThreads) A.x0[c] = xs(0)[c];
}

__global__ void __launch_bounds__(kCoarseThreads) k_mg_coarse_cycle(CoarseArgs A, const CgScalars* __restrict__ sc) {
    coarse_run(A, sc);
}

struct MCoarse {
    static constexpr int kThreads = kCoarseThreads;
    static constexpr bool kByRef = true;
    CoarseArgs A;
    const CgScalars* sc;
    __device__ void run(uint32_t) const { coarse_run(A, sc); }
};

// ===========================================================================
// Fused descent and ascent of one level (every axis halved, three sweeps),
// marching along z.
//
// A block owns a tile of kMTX x kMTY columns and the planes [za, za + chunk)
// of the level.  Its 64 x 16 work-items are the tile's columns and a ring of
// kMH = 3 columns around it (lane = x, wave = y).  The block walks up the
// planes, one per step: the step that loads plane q forms stage 0 (pointwise)
// at q, stage 1 at q - 1, stage 2 at q - 2 and stage 3 at q - 3.  A stage reads
// the previous stage at the work-item's own column from registers (the
// planes below, and the plane above, formed earlier in the same step) and at
// the four xy neighbours from LDS, where it was written one step earlier.
// Each stage has two LDS slots that alternate with the step's parity, so one
// barrier per step orders every write before its reads and every read before
// the slot is written again.  Stage s is exact on the columns at least s
// cells inside the ring's edge, so stage 3 is exact on the tile; what the
// ring's outer columns hold is never stored.  Cells outside the level read as
// code 0 and b = x = 0, so every stage is 0 there.
//
// Descent (PRE): stage 0 = x1 = omega D^-1 b, stages 1 and 2 = the sweeps x2
// and x3 (x3 is stored), stage 3 = the residual b - A x3 of the fluid cells,
// summed over each 2 x 2 x 2 block of children into the coarse right-hand
// side (pairs along x, then y, then z; weight 1/8).  Every value is formed
// by the same operations in the same order as in k_mg_smooth2_first,
// k_mg_smooth and the restriction kernels, so the results are bitwise
// theirs.
// Ascent (POST): stage 0 = x3 + xc[parent] (the prolongation by injection),
// stages 1-3 = three sweeps, stage 3 stored; with RZ (level 0) also the block
// partials of b . x.  Bitwise k_mg_smooth_prolong's and k_mg_smooth's values.
// 32 x 32 work-items (a wave = two rows of 32): 26 x 26 tiles, less ring and
// grid-edge waste than 64 x 16 (58 x 10 tiles), measured 106 / 104 against
// 122 / 114 us per descent / ascent at 256^3.
constexpr int kMThreads = 1024;
constexpr int kMX = 32, kMY = kMThreads / kMX, kMH = 3;
constexpr int kMRows = 64 / kMX;                             // rows per wave
constexpr int kMTX = kMX - 2 * kMH, kMTY = kMY - 2 * kMH;   // the tile

struct March {
    int nx, ny, nz;          // level extents
    uint32_t sy, sz;         // level strides
    int cnz;                 // coarse extent along z
    uint32_t csy, csz;       // coarse strides
    uint32_t tx, ty;         // tiles along x and y
    uint32_t blocks;         // tiles x chunks; blocks past it idle
    int chunk;               // planes per block (even)
    uint32_t tiles;          // tx * ty
    uint32_t gmax;           // blocks of a launch (any schedule's units fit)
    uint32_t wunits;         // units of the fluid-range pass (tile x kWetChunk planes)
    int cmin;                // shortest chunk a schedule may choose (a multiple of kRzSlab)
    int skip;                // units without fluid store their zeros and skip the march
    uint32_t parts;          // partials of r . z: tiles x slabs of kRzSlab planes
};

// The schedule of a fused level, made once per set of codes on the device
// (k_mg_march_wet, then k_mg_march_plan), in one array of 32-bit words:
//   [0] C, the planes per unit (even); [1] U, the units (tiles x chunks of C
//   planes); [2] H, the units whose tile holds fluid in their planes;
//   [kSchedHead + b], b < U: the unit of block b, bit 31 set when it holds
//   fluid;
//   [kSchedHead + gmax + w], w < wunits: planes lo | hi << 16 holding fluid
//   in the tile's columns of the pass unit w (lo > hi: none); the plan
//   overwrites the first `tiles` entries with each tile's range.
// Unit u is tile u % tiles (x fastest) and chunk u / tiles, as the static
// numbering of blocks.  The units with fluid come first in launch order and
// in equal runs per XCD (block b runs on XCD b % 8), so that the heavy
// blocks start at once and no XCD holds more of them than another.  Which
// block runs a unit changes no stored value; the partial sums of r . z are
// kept per unit, so the reduction's order does not depend on it either.
constexpr int kSchedHead = 4;
constexpr int kWetChunk = 32;
constexpr uint32_t kSchedHeavy = 0x80000000u;
// Every chunk is a multiple of kRzSlab planes, and the ascent of level 0
// writes one partial of r . z per tile and slab of kRzSlab planes (each
// work-item's planes in order, then a fixed order over the work-items): the
// sum, and so the CG iterates, do not depend on the schedule.
constexpr int kRzSlab = 2;

// Sum over a wave: xor butterflies within each half of 32 lanes by
// ds_swizzle (its lane pattern is an immediate, so no address registers),
// then the two halves through readlane.  Every lane of a half holds the same
// value after each step (a + b = b + a), so the result is one fixed sum.
template <int M>
__device__ __forceinline__ double swz_xor_add(double v) {
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), 0x1F | (M << 10));
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), 0x1F | (M << 10));
    return v + __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_swz(double v) {
    v = swz_xor_add<1>(v);
    v = swz_xor_add<2>(v);
    v = swz_xor_add<4>(v);
    v = swz_xor_add<8>(v);
    v = swz_xor_add<16>(v);
    const double a = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 0),
                                      __builtin_amdgcn_readlane(__double2loint(v), 0));
    const double b = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 32),
                                      __builtin_amdgcn_readlane(__double2loint(v), 32));
    return a + b;
}

// Per-work-item constants of a march.
struct MarchCol {
    bool in;                 // the column lies in the level
    bool tile;               // ... and in the tile (stage 3 exact)
    bool owner;              // tile column at an even x and y of the tile (PRE)
    uint32_t col, pcol;      // column offset in the level and of its parent
    int za, zs, zr;          // first plane, end of the stored planes, end of the pairs
    int me;                  // LDS index of the column
    int wy;                  // the wave (wave-uniform); it holds rows kMRows wy ..
    uint32_t tix;            // the tile's index (x fastest)
};

// The rolling state of a work-item: codes and b of the planes q-1 .. q-3 and
// each stage at the two planes below the one it forms next.
struct MarchRing {
    uint32_t c1, c2, c3;
    float b1, b2, b3;
    float w1, w2, w3;        // omega / d of the planes q-1 .. q-3
    float a1, a2, a3;        // air weight aw[air mask] of the planes q-1 .. q-3
    float v0a, v0b;          // stage 0 at q-1, q-2
    float v1a, v1b;          // stage 1 at q-2, q-3
    float v2a, v2b;          // stage 2 at q-3, q-4
    float pe;                // PRE: pair sum of the last even plane
    double acc;              // POST: sum of b . x over the stored cells
};

// One damped Jacobi sweep of a cell (the flat kernels' formula).
__device__ __forceinline__ float jacobi(uint16_t cd, float bc, float vc, const float (&v)[6], const MgLevel& L,
                                        const float* aw, const float* w) {
    const float wd = wdiag(cd, w);
    const float ax = arow(cd, vc, v, L.sx, L.sy, L.sz, aw);
    return (cd & kCodeFluid) ? vc + wd * (bc - ax) : 0.0f;
}

// A fluid cell whose six neighbours are all fluid: arow with every mask all
// ones and no air (aw[0]), the same operations in the same order.
constexpr uint16_t kCodeFull = uint16_t(kCodeFluid | (63u << kCplShift));
__device__ __forceinline__ float arow_full(float vc, const float (&v)[6], const MgLevel& L, float aw0) {
    const float ex = (vc - v[0]) + (vc - v[1]);
    const float ey = (vc - v[2]) + (vc - v[3]);
    const float ez = (vc - v[4]) + (vc - v[5]);
    return aw0 * vc + L.sx * ex + L.sy * ey + L.sz * ez;
}

// keep_if with the mask formed by one bit-field extract (the compiler would
// otherwise test the bit and select, three instructions instead of two).
__device__ __forceinline__ float keep_bit(float v, uint32_t cd, int b) {
    int32_t m;
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"(cd), "i"(b));
    return __int_as_float(__float_as_int(v) & m);
}

// arow with the air weight awc = aw[air_mask(cd)] of the cell given.
__device__ __forceinline__ float arow_w(uint32_t cd, float awc, float vc, const float (&v)[6], const MgLevel& L) {
    const float ex = keep_bit(vc - v[0], cd, kCplShift + 0) + keep_bit(vc - v[1], cd, kCplShift + 1);
    const float ey = keep_bit(vc - v[2], cd, kCplShift + 2) + keep_bit(vc - v[3], cd, kCplShift + 3);
    const float ez = keep_bit(vc - v[4], cd, kCplShift + 4) + keep_bit(vc - v[5], cd, kCplShift + 5);
    const float a = awc * vc + L.sx * ex + L.sy * ey + L.sz * ez;
    return (cd & kCodeFluid) ? a : vc;
}

// Wave-uniform shortcuts of a stage: a wave of non-fluid cells gives 0
// without reading its neighbours; a wave of full cells skips the masks.
// Otherwise jacobi() / the restriction's residual with the cell's omega / d
// (wd) and air weight (awc) formed once per plane.  Bitwise the same values.
enum { kWaveMixed = 0, kWaveDry = 1, kWaveFull = 2 };
__device__ __forceinline__ int wave_kind(uint32_t cd) {
    if (__all((cd & kCodeFluid) == 0)) return kWaveDry;
    if (__all(cd == kCodeFull)) return kWaveFull;
    return kWaveMixed;
}

// The four xy neighbours of the work-item in an LDS slot.
__device__ __forceinline__ void xy_nbrs(const float* S, const MarchCol& K, float (&n)[4]) {
    n[0] = S[K.me - 1];
    n[1] = S[K.me + 1];
    n[2] = S[K.me - kMX];
    n[3] = S[K.me + kMX];
}

__device__ __forceinline__ float march_jacobi(uint32_t cd, float wd, float awc, float bc, float vc, const float* S,
                                              const MarchCol& K, float vzm, float vzp, const MgLevel& L) {
    const int kind = wave_kind(cd);
    if (kind == kWaveDry) return 0.0f;
    float xy[4];
    xy_nbrs(S, K, xy);
    const float n[6] = {xy[0], xy[1], xy[2], xy[3], vzm, vzp};
    if (kind == kWaveFull) return vc + wd * (bc - arow_full(vc, n, L, awc));
    return (cd & kCodeFluid) ? vc + wd * (bc - arow_w(cd, awc, vc, n, L)) : 0.0f;
}

__device__ __forceinline__ float march_residual(uint32_t cd, float awc, float bc, float vc, const float* S,
                                                const MarchCol& K, float vzm, float vzp, const MgLevel& L) {
    const int kind = wave_kind(cd);
    if (kind == kWaveDry) return 0.0f;
    float xy[4];
    xy_nbrs(S, K, xy);
    const float n[6] = {xy[0], xy[1], xy[2], xy[3], vzm, vzp};
    if (kind == kWaveFull) return bc - arow_full(vc, n, L, awc);
    return (cd & kCodeFluid) ? bc - arow_w(cd, awc, vc, n, L) : 0.0f;
}

// Loads of plane q of the work-item's column, at the plane clamped into the
// level (the values of a plane or column outside the level are replaced
// where they are used).  POST: also x and the parent's coarse x.  PRE: the
// code of the coarse cell that the pairs of plane q - 4 complete (in ev's
// bits), so that the step issuing the restriction waits on no load.
template <bool POST>
__device__ __forceinline__ void march_load(const March& M, const MarchCol& K, int q,
                                           const uint16_t* __restrict__ code, const float* __restrict__ b,
                                           const float* __restrict__ xi, const float* __restrict__ xc,
                                           const uint16_t* __restrict__ codec, uint16_t& cd, float& bv,
                                           float& xv, float& ev) {
    const int qc = q < 0 ? 0 : (q >= M.nz ? M.nz - 1 : q);
    const uint32_t i = uint32_t(qc) * M.sz + K.col;
    cd = code[i];
    bv = b[i];
    if constexpr (POST) {
        xv = xi[i];
        ev = xc[uint32_t(qc >> 1) * M.csz + K.pcol];
    } else {
        const int Z = (q - 4) >> 1;
        const int Zc = Z < 0 ? 0 : (Z >= M.cnz ? M.cnz - 1 : Z);
        ev = __uint_as_float(uint32_t(codec[uint32_t(Zc) * M.csz + K.pcol]));
    }
}

// PRE: pair sum of the residuals of plane p (in slot RS); an odd plane
// completes its coarse cell, whose code (prefetched with plane p + 4) is cc.
__device__ __forceinline__ void march_pairs(const March& M, const MarchCol& K, int p, const float* RS,
                                            uint32_t cc, float* __restrict__ bco, MarchRing& R) {
    if (!K.owner || p < K.za || p >= K.zr) return;
    const int e = K.me;
    const float P = (RS[e] + RS[e + 1]) + (RS[e + kMX] + RS[e + kMX + 1]);
    if (((p - K.za) & 1) == 0) {
        R.pe = P;
        return;
    }
    const int Z = p >> 1;
    if (Z >= M.cnz) return;
    const uint32_t c = uint32_t(Z) * M.csz + K.pcol;
    const float s = R.pe + P;
    bco[c] = (cc & kCodeFluid) ? s * 0.125f : 0.0f;
}

// Whether the wave wy holds a row at least s rows from the block's edge.
__device__ __forceinline__ bool rows_need(int wy, int s) {
    return wy * kMRows + kMRows - 1 >= s && wy * kMRows < kMY - s;
}

// LDS of a march (floats): the tables aw and w, then two slots of kMThreads
// per stage 0, 1, 2 and (PRE) for the residuals.  One array, so that every
// access is a constant offset from the work-item's index.
// Each slot has kSlotPad floats before and after it (shared with the next
// slot), so that the four xy neighbours of every work-item are constant
// offsets of its own index: the first and last row read the pads and the
// first and last column the row beside, neither ever written in the step
// that reads it.  Those are the outermost ring columns, whose values are
// never needed (stage s is exact s columns inside the ring).
constexpr int kLdsAw = 0, kLdsW = 64, kLdsSlots = 128;
constexpr int kSlotPad = kMX, kSlotStride = kMThreads + kSlotPad;
__device__ __forceinline__ constexpr int slot(int stage, int h) {
    return kLdsSlots + kSlotPad + (2 * stage + h) * kSlotStride;
}

// Planes loaded ahead: the loads of plane q are issued kMPre steps before
// its step (kMPre register sets in rotation), so that about kMPre planes of
// every block are in flight.
constexpr int kMPre = 2;

// One step (the plane q = za - 3 + t; H = t & 1 selects the LDS slots
// written now).  cdN .. eN hold plane q on entry and plane q + kMPre on exit.
template <bool POST, bool RZ, int H>
__device__ __forceinline__ void march_step(const March& M, const MgLevel& L, const MarchCol& K, int t,
                                           uint16_t& cdN, float& bN, float& xN, float& eN, MarchRing& R,
                                           const uint16_t* __restrict__ code, const float* __restrict__ b,
                                           const float* __restrict__ xi, const float* __restrict__ xc,
                                           float* __restrict__ xo, const uint16_t* __restrict__ codec,
                                           float* __restrict__ bco, float* lds, double* red,
                                           double* __restrict__ part) {
    const float* aw = lds + kLdsAw;
    const float* w = lds + kLdsW;
    const int q = K.za - 3 + t;
    const bool in0 = K.in && unsigned(q) < unsigned(M.nz);
    const uint32_t cd0 = in0 ? uint32_t(cdN) : 0u;
    const float b0 = in0 ? bN : 0.0f;
    const float wd0 = wdiag(uint16_t(cd0), w);
    const float aw0 = aw[air_mask(uint16_t(cd0))];
    // stage 0 at q
    const float v0 = POST ? (in0 ? xN + eN : 0.0f) : wd0 * b0;
    const uint32_t ccode = POST ? 0u : __float_as_uint(eN);   // PRE: the coarse code for the pairs of q - 4
    march_load<POST>(M, K, q + kMPre, code, b, xi, xc, codec, cdN, bN, xN, eN);
    lds[slot(0, H) + K.me] = v0;
    // Stage s is needed on the rows at least s from the block's edge only:
    // waves outside skip it (what their registers and slots then hold is
    // never read by a needed value).
    float v1 = 0.0f, v2 = 0.0f;
    // stage 1 at q - 1
    if (rows_need(K.wy, 1)) {
        v1 = march_jacobi(R.c1, R.w1, R.a1, R.b1, R.v0a, lds + slot(0, H ^ 1), K, R.v0b, v0, L);
        lds[slot(1, H) + K.me] = v1;
    }
    // stage 2 at q - 2
    if (rows_need(K.wy, 2)) {
        v2 = march_jacobi(R.c2, R.w2, R.a2, R.b2, R.v1a, lds + slot(1, H ^ 1), K, R.v1b, v1, L);
        lds[slot(2, H) + K.me] = v2;
    }
    // stage 3 at q - 3
    if (rows_need(K.wy, 3)) {
        const float* S = lds + slot(2, H ^ 1);
        if constexpr (POST) {
            const float v3 = march_jacobi(R.c3, R.w3, R.a3, R.b3, R.v2a, S, K, R.v2b, v2, L);
            const int p = q - 3;
            if (K.tile && p >= K.za && p < K.zs) {
                xo[uint32_t(p) * M.sz + K.col] = v3;
                if (RZ) R.acc += double(R.b3) * double(v3);
            }
        } else {
            const int p = q - 2;
            if (K.tile && p >= K.za && p < K.zs) xo[uint32_t(p) * M.sz + K.col] = v2;
            lds[slot(3, H) + K.me] = march_residual(R.c3, R.a3, R.b3, R.v2a, S, K, R.v2b, v2, L);
            march_pairs(M, K, q - 4, lds + slot(3, H ^ 1), ccode, bco, R);
        }
    }
    R.c3 = R.c2; R.c2 = R.c1; R.c1 = cd0;
    R.b3 = R.b2; R.b2 = R.b1; R.b1 = b0;
    R.w3 = R.w2; R.w2 = R.w1; R.w1 = wd0;
    R.a3 = R.a2; R.a2 = R.a1; R.a1 = aw0;
    R.v0b = R.v0a; R.v0a = v0;
    R.v1b = R.v1a; R.v1a = v1;
    R.v2b = R.v2a; R.v2a = v2;
    // RZ: the last stored plane of a slab (the same for the whole block)
    // closes the slab's partial.  Every work-item leaves its sum in red;
    // after the barrier, lane l of wave 0 adds entries l, l + 64, .. in
    // order and wave_sum_swz adds the lanes.  The cross-lane work is in wave
    // 0's branch, where the stages' values are dead.
    int slab = -1;
    if constexpr (RZ) {
        const int p = q - 3;
        if (p >= K.za && p < K.zs && ((p + 1) % kRzSlab == 0 || p + 1 == K.zs)) {
            red[threadIdx.x] = R.acc;
            R.acc = 0.0;
            slab = p / kRzSlab;
        }
    }
    __syncthreads();
    if constexpr (RZ) {
        if (slab >= 0 && threadIdx.x < 64) {
            double t = 0.0;
#pragma unroll 4
            for (int k = 0; k < kMThreads / 64; ++k) t += red[threadIdx.x + 64 * k];
            t = wave_sum_swz(t);
            if (threadIdx.x == 0) part[K.tix + M.tiles * uint32_t(slab)] = t;
        }
    }
}

// The tile of block blk: columns [x0, x1) x [y0, y1), planes [za, zs) and the
// coarse planes [za / 2, zc) its pairs complete.
struct MarchTile {
    int x0, x1, y0, y1, za, zs, zc;
};
__device__ __forceinline__ MarchTile march_tile(const March& M, uint32_t blk) {
    const uint32_t tx = blk % M.tx, rest = blk / M.tx;
    const uint32_t ty = rest % M.ty, tc = rest / M.ty;
    MarchTile T;
    T.x0 = int(tx) * kMTX;
    T.x1 = min(T.x0 + kMTX, M.nx);
    T.y0 = int(ty) * kMTY;
    T.y1 = min(T.y0 + kMTY, M.ny);
    T.za = int(tc) * M.chunk;
    T.zs = min(T.za + M.chunk, M.nz);
    T.zc = min((T.za + M.chunk) / 2, M.cnz);
    return T;
}

// A unit whose tile holds no fluid cell in its planes: every value it
// stores is +0 (the sweeps give 0 on non-fluid cells, and so does the
// residual, hence each restricted pair sum), so it stores the zeros and
// skips the march.
template <bool POST>
__device__ __forceinline__ void march_dry(const March& M, uint32_t blk, float* __restrict__ xo,
                                          float* __restrict__ bco) {
    const MarchTile T = march_tile(M, blk);
    const int w = T.x1 - T.x0, h = T.y1 - T.y0;
    for (int e = int(threadIdx.x); e < w * h; e += kMThreads) {
        const int y = T.y0 + e / w, x = T.x0 + e % w;
        for (int z = T.za; z < T.zs; ++z) xo[uint32_t(z) * M.sz + uint32_t(y) * M.sy + uint32_t(x)] = 0.0f;
    }
    if constexpr (!POST) {
        const int cw = (w + 1) / 2, ch = (h + 1) / 2;
        for (int e = int(threadIdx.x); e < cw * ch; e += kMThreads) {
            const int Y = T.y0 / 2 + e / cw, X = T.x0 / 2 + e % cw;
            for (int Z = T.za / 2; Z < T.zc; ++Z)
                bco[uint32_t(Z) * M.csz + uint32_t(Y) * M.csy + uint32_t(X)] = 0.0f;
        }
    }
}

// Pass unit w (tile w % tiles, planes [kWetChunk k, kWetChunk (k + 1)) with
// k = w / tiles): the first and last plane holding a fluid cell in the
// tile's columns (the codes do not change during a pressure solve: one pass
// per solve).
__device__ __forceinline__ void k_mg_march_wet_run(March M, const uint16_t* __restrict__ code,
                                                   uint32_t* __restrict__ sched, uint32_t blk) {
    __shared__ int lo_s, hi_s;
    if (blk >= M.wunits) return;
    if (threadIdx.x == 0) {
        lo_s = 0xFFFF;
        hi_s = 0;
    }
    __syncthreads();
    const uint32_t t = blk % M.tiles, k = blk / M.tiles;
    const int x0 = int(t % M.tx) * kMTX, y0 = int(t / M.tx) * kMTY;
    const int w = min(kMTX, M.nx - x0), h = min(kMTY, M.ny - y0);
    const int z0 = int(k) * kWetChunk, z1 = min(z0 + kWetChunk, M.nz);
    int lo = 0xFFFF, hi = 0;
    for (int e = int(threadIdx.x); e < w * h; e += kMThreads) {
        const uint32_t col = uint32_t(y0 + e / w) * M.sy + uint32_t(x0 + e % w);
        for (int z = z0; z < z1; ++z)
            if (code[uint32_t(z) * M.sz + col] & kCodeFluid) {
                lo = min(lo, z);
                hi = max(hi, z);
            }
    }
    if (lo <= hi) {
        atomicMin(&lo_s, lo);
        atomicMax(&hi_s, hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) sched[kSchedHead + M.gmax + blk] = uint32_t(lo_s) | (uint32_t(hi_s) << 16);
}

__global__ void __launch_bounds__(kMThreads) k_mg_march_wet(March M, const uint16_t* __restrict__ code,
                                                            uint32_t* __restrict__ sched) {
    k_mg_march_wet_run(M, code, sched, blockIdx.x);
}

struct MMarchWet {
    static constexpr int kThreads = kMThreads;
    March M;
    const uint16_t* code;
    uint32_t* sched;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_mg_march_wet_run(M, code, sched, vb); }
};

// Units of tile range r (lo | hi << 16) at chunk length C: the chunks its
// fluid planes meet.
__device__ __forceinline__ uint32_t range_units(uint32_t r, uint32_t C) {
    const uint32_t lo = r & 0xFFFFu, hi = r >> 16;
    return lo > hi ? 0u : hi / C - lo / C + 1u;
}

// The schedule (one work-group): each tile's fluid range from its pass
// units; the chunk length C (a multiple of kRzSlab, at least M.cmin) with the least
// ceil(H / slots) x (C + 7), the units with fluid H filling `slots` resident
// blocks (two per CU) per round and a unit taking about C + 7 plane steps
// (ties: the longer chunk); then the launch order.
constexpr int kPlanThreads = 1024;

__device__ __forceinline__ void k_mg_march_plan_run(const March& M, uint32_t* __restrict__ sched,
                                                    uint32_t slots, uint32_t force_c) {
    __shared__ unsigned long long best;
    __shared__ uint32_t part[kPlanThreads];
    uint32_t* rng = sched + kSchedHead + M.gmax;
    const uint32_t wch = M.wunits / M.tiles;
    if (threadIdx.x == 0) best = ~0ull;
    for (uint32_t t = threadIdx.x; t < M.tiles; t += kPlanThreads) {
        uint32_t lo = 0xFFFFu, hi = 0u;
        for (uint32_t k = 0; k < wch; ++k) {
            const uint32_t r = rng[t + k * M.tiles];
            lo = min(lo, r & 0xFFFFu);
            hi = max(hi, r >> 16);
        }
        rng[t] = lo <= hi ? (lo | (hi << 16)) : 0xFFFFu;
    }
    __syncthreads();
    const uint32_t cmax = (uint32_t(M.nz) + kRzSlab - 1) / kRzSlab * kRzSlab;
    for (uint32_t C = uint32_t(M.cmin) + kRzSlab * threadIdx.x; C <= max(cmax, uint32_t(M.cmin)) && !force_c;
         C += kRzSlab * kPlanThreads) {
        uint32_t H = 0;
        for (uint32_t t = 0; t < M.tiles; ++t) H += range_units(rng[t], C);
        const unsigned long long cost = (unsigned long long)((H + slots - 1) / slots) * (C + 7u);
        atomicMin(&best, (cost << 20) | (0xFFFFFu - C));
    }
    __syncthreads();
    const uint32_t C = force_c ? force_c : 0xFFFFFu - uint32_t(best & 0xFFFFFull);
    const uint32_t nch = (uint32_t(M.nz) + C - 1) / C, U = M.tiles * nch;
    // heavy units counted per contiguous segment of units, then scanned
    const uint32_t seg = (U + kPlanThreads - 1) / kPlanThreads;
    const uint32_t u0 = min(U, threadIdx.x * seg), u1 = min(U, u0 + seg);
    auto heavy = [&](uint32_t u) {
        const uint32_t r = rng[u % M.tiles], k = u / M.tiles;
        const uint32_t lo = r & 0xFFFFu, hi = r >> 16;
        return lo <= hi && k >= lo / C && k <= hi / C;
    };
    uint32_t n = 0;
    for (uint32_t u = u0; u < u1; ++u) n += heavy(u) ? 1u : 0u;
    part[threadIdx.x] = n;
    __syncthreads();
    for (int off = 1; off < kPlanThreads; off <<= 1) {
        const uint32_t v = threadIdx.x >= uint32_t(off) ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    const uint32_t H = part[kPlanThreads - 1];
    uint32_t hk = part[threadIdx.x] - n;     // heavy units before u0
    uint32_t dk = u0 - hk;                   // units without fluid before u0
    // Blocks [0, H) run the heavy units, [H, U) the others; XCD j runs the
    // blocks b = j (mod 8), each a contiguous run of units.
    auto place = [&](uint32_t k, uint32_t first, uint32_t count) {
        // the k-th of `count` units given to blocks [first, first + count)
        uint32_t before = 0;
        for (uint32_t j = 0; j < kXcds; ++j) {
            const uint32_t a = (first + kXcds - 1 - j) / kXcds;            // first m with j + 8 m >= first
            const uint32_t e = (first + count + kXcds - 1 - j) / kXcds;    // end m
            const uint32_t cj = e > a ? e - a : 0u;
            if (k < before + cj) return j + kXcds * (a + (k - before));
            before += cj;
        }
        return 0u;
    };
    for (uint32_t u = u0; u < u1; ++u) {
        if (heavy(u)) {
            sched[kSchedHead + place(hk, 0u, H)] = u | kSchedHeavy;
            ++hk;
        } else {
            sched[kSchedHead + place(dk, H, U - H)] = u;
            ++dk;
        }
    }
    if (threadIdx.x == 0) {
        sched[0] = C;
        sched[1] = U;
        sched[2] = H;
        sched[3] = 0u;
    }
}

__global__ void __launch_bounds__(kPlanThreads) k_mg_march_plan(March M, uint32_t* __restrict__ sched,
                                                               uint32_t slots, uint32_t force_c) {
    k_mg_march_plan_run(M, sched, slots, force_c);
}

struct MMarchPlan {
    static constexpr int kThreads = kPlanThreads;
    March M;
    uint32_t* sched;
    uint32_t slots;
    uint32_t force_c;
    __device__ void run(uint32_t) const { k_mg_march_plan_run(M, sched, slots, force_c); }
};

template <bool POST, bool RZ>
__device__ __forceinline__ void march_unit(const March& M, const MgLevel& L, const uint16_t* __restrict__ code,
                                           const float* __restrict__ b, const float* __restrict__ xi,
                                           const float* __restrict__ xc, float* __restrict__ xo,
                                           const uint16_t* __restrict__ codec, float* __restrict__ bco, float omega,
                                           double* __restrict__ part, uint32_t blk, bool heavy, float* lds,
                                           double* red) {
    if (M.skip && !heavy) {
        march_dry<POST>(M, blk, xo, bco);
        if (RZ) {
            // the unit's slabs sum to +0
            const MarchTile T = march_tile(M, blk);
            const uint32_t t = blk % M.tiles;
            for (int k = T.za / kRzSlab + int(threadIdx.x); k * kRzSlab < T.zs; k += kMThreads)
                part[t + M.tiles * uint32_t(k)] = 0.0;
        }
        return;
    }
    build_tables<true>(lds + kLdsAw, lds + kLdsW, L.sx, L.sy, L.sz, omega);
    MarchCol K;
    {
        const uint32_t tx = blk % M.tx, rest = blk / M.tx;
        const uint32_t ty = rest % M.ty, tc = rest / M.ty;
        const int cx = int(threadIdx.x) & (kMX - 1), cy = int(threadIdx.x) / kMX;
        const int gx = int(tx) * kMTX - kMH + cx, gy = int(ty) * kMTY - kMH + cy;
        K.in = unsigned(gx) < unsigned(M.nx) && unsigned(gy) < unsigned(M.ny);
        K.tile = K.in && cx >= kMH && cx < kMH + kMTX && cy >= kMH && cy < kMH + kMTY;
        K.owner = !POST && K.tile && ((cx - kMH) & 1) == 0 && ((cy - kMH) & 1) == 0;
        K.col = K.in ? uint32_t(gy) * M.sy + uint32_t(gx) : 0u;
        K.pcol = K.in ? uint32_t(gy >> 1) * M.csy + uint32_t(gx >> 1) : 0u;
        K.za = int(tc) * M.chunk;
        K.zr = K.za + M.chunk;
        K.zs = K.zr < M.nz ? K.zr : M.nz;
        K.me = int(threadIdx.x);
        K.wy = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64);
        K.tix = blk % M.tiles;
    }
    MarchRing R;
    R.c1 = R.c2 = R.c3 = 0u;
    R.b1 = R.b2 = R.b3 = 0.0f;
    R.w1 = R.w2 = R.w3 = 0.0f;
    R.a1 = R.a2 = R.a3 = 0.0f;
    R.v0a = R.v0b = R.v1a = R.v1b = R.v2a = R.v2b = 0.0f;
    R.pe = 0.0f;
    R.acc = 0.0;
    uint16_t cd[kMPre];
    float bv[kMPre], xv[kMPre], ev[kMPre];
#pragma unroll
    for (int k = 0; k < kMPre; ++k) {
        xv[k] = ev[k] = 0.0f;
        march_load<POST>(M, K, K.za - 3 + k, code, b, xi, xc, codec, cd[k], bv[k], xv[k], ev[k]);
    }
    // Stage 3 reaches the last plane za + chunk - 1 at step chunk + 5, and
    // (PRE) its pair sums one step later; the steps are rounded up to whole
    // rotations (the extra planes store nothing).
    const int steps = (M.chunk + (POST ? 6 : 7) + kMPre - 1) / kMPre * kMPre;
    for (int t = 0; t < steps; t += kMPre) {
#pragma unroll
        for (int k = 0; k < kMPre; k += 2) {
            march_step<POST, RZ, 0>(M, L, K, t + k, cd[k], bv[k], xv[k], ev[k], R, code, b, xi, xc, xo, codec,
                                    bco, lds, red, part);
            march_step<POST, RZ, 1>(M, L, K, t + k + 1, cd[k + 1], bv[k + 1], xv[k + 1], ev[k + 1], R, code, b, xi,
                                    xc, xo, codec, bco, lds, red, part);
        }
    }
}


// The unit of block `slot`: from the schedule (in launch order), else the
// static unit vblk (XCD-remapped), every unit marched.  A launch has a block
// for every unit of the schedule it follows: each solve makes its levels'
// schedules before its first V-cycle (a scene alone: any chunk, gmax
// blocks; many scenes' records: the static chunk, the static blocks).
template <bool POST, bool RZ>
__device__ __forceinline__ void march_run(const March& M0, const MgLevel& L, const uint16_t* __restrict__ code,
                                          const float* __restrict__ b, const float* __restrict__ xi,
                                          const float* __restrict__ xc, float* __restrict__ xo,
                                          const uint16_t* __restrict__ codec, float* __restrict__ bco, float omega,
                                          const CgScalars* __restrict__ sc, double* __restrict__ part,
                                          const uint32_t* __restrict__ sched, uint32_t vblk, uint32_t slot) {
    __shared__ float lds[kLdsSlots + kSlotPad + (POST ? 6 : 8) * kSlotStride];
    __shared__ double red[RZ ? kMThreads : 1];
    if (sc->done) return;
    if (!sched) {
        if (vblk < M0.blocks) march_unit<POST, RZ>(M0, L, code, b, xi, xc, xo, codec, bco, omega, part, vblk, true, lds, red);
        return;
    }
    const uint32_t U = sched[1];
    if (slot >= U) return;
    March M = M0;
    M.chunk = int(sched[0]);
    M.blocks = U;
    const uint32_t e = sched[kSchedHead + slot];
    march_unit<POST, RZ>(M, L, code, b, xi, xc, xo, codec, bco, omega, part, e & ~kSchedHeavy,
                         (e & kSchedHeavy) != 0u, lds, red);
}

template <bool POST, bool RZ>
__global__ void __launch_bounds__(kMThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_mg_march(March M, MgLevel L, const uint16_t* __restrict__ code, const float* __restrict__ b,
           const float* __restrict__ xi, const float* __restrict__ xc, float* __restrict__ xo,
           const uint16_t* __restrict__ codec, float* __restrict__ bco, float omega,
           const CgScalars* __restrict__ sc, double* __restrict__ part, const uint32_t* __restrict__ sched) {
    march_run<POST, RZ>(M, L, code, b, xi, xc, xo, codec, bco, omega, sc, part, sched, xcd_block(), blockIdx.x);
}

template <bool POST, bool RZ>
struct MMarch {
    static constexpr int kThreads = kMThreads;
    March M;
    MgLevel L;
    const uint16_t* code;
    const float *b, *xi, *xc;
    float* xo;
    const uint16_t* codec;
    float* bco;
    float omega;
    const CgScalars* sc;
    double* part;
    const uint32_t* sched;
    uint32_t G;
    __device__ void run(uint32_t blk) const {
        march_run<POST, RZ>(M, L, code, b, xi, xc, xo, codec, bco, omega, sc, part, sched, xcd_of(blk, G), blk);
    }
};

// The multi-scene wrapper of the fused march, with its occupancy attribute.
template <class A>
__global__ void __launch_bounds__(kMThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_multi_march(const A* __restrict__ tab, const uint32_t* __restrict__ map, const uint8_t* __restrict__ active) {
    const uint32_t e = map[blockIdx.x];
    const uint32_t s = e >> kMapShift;
    if (active && !active[s]) return;
    const A a = tab[s];
    a.run(e & kMapLocal);
}

// ===========================================================================
// Launchers.
static uint32_t flat_grid(uint32_t n) {
    size_t g = (size_t(n) + size_t(kBlock) * kCpt - 1) / (size_t(kBlock) * kCpt);
    g = (g + kXcds - 1) / kXcds * kXcds;
    if (g == 0) g = kXcds;
    return uint32_t(g);
}

uint32_t col_blocks(const GridDev& g) {
    return flat_grid(make_flat(g.nx, g.ny, g.nz).n);
}

hipError_t launch_codes(hipStream_t s, int nx, int ny, int nz, const uint8_t* markers, uint16_t* codes) {
    const uint32_t n = uint32_t(nx) * ny * nz;
    return LAUNCH(k_codes, grid_for(n), nx, ny, nz, markers, codes, n);
}

static hipError_t finish(hipStream_t s, const double* part, uint32_t count, CgScalars* sc, int kind,
                         double rel_inner = 0.0, double abs_target2 = 0.0) {
    (void)launch_dims(k_finish, "k_finish", dim3(1), dim3(kFinishBlock), 0, s, part, count, sc, kind, rel_inner, abs_target2);
    return hipGetLastError();
}

hipError_t launch_cg_start(hipStream_t s, const GridDev& g, const float* rsrc, float* r, float* x, CgScalars* sc,
                           ReduceBuf red, double rel_inner, double abs_target2) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    hipError_t e = LAUNCH(k_cg_start, G, F, rsrc, r, x, red.partials);
    if (e != hipSuccess) return e;
    return finish(s, red.partials, G, sc, kFinishInit, rel_inner, abs_target2);
}

hipError_t launch_refine_true(hipStream_t s, const GridDev& g, const float scale[3], const uint16_t* codes,
                              const float* xcorr, double* x64, const float* b, float* r, int mode, ReduceBuf red,
                              double* out2) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    hipError_t e;
    if (mode == 0)
        e = LAUNCH(k_refine<0>, G, F, codes, scale[0], scale[1], scale[2], xcorr, x64, b, r, red.partials);
    else if (mode == 3)
        e = LAUNCH(k_refine<3>, G, F, codes, scale[0], scale[1], scale[2], xcorr, x64, b, r, red.partials);
    else if (mode == 1)
        e = LAUNCH(k_refine<1>, G, F, codes, scale[0], scale[1], scale[2], xcorr, x64, b, r, red.partials);
    else {
        e = LAUNCH(k_acc64, grid_for(F.n), F.n, xcorr, x64);
        if (e != hipSuccess) return e;
        e = LAUNCH(k_refine<2>, G, F, codes, scale[0], scale[1], scale[2], xcorr, x64, b, r, red.partials);
    }
    if (e != hipSuccess) return e;
    (void)launch_dims(k_sum2, "k_sum2", dim3(1), dim3(1024), 0, s, red.partials, G, out2);
    return hipGetLastError();
}

hipError_t launch_div_check(hipStream_t s, const GridDev& g, const uint8_t* markers, const float* u,
                            const float* v, const float* w, const float* b, double* partials, int blocks,
                            double* out) {
    const uint32_t n = uint32_t(g.nx) * g.ny * g.nz;
    probe_next_bytes(15.0 * double(n));   // markers 1, three faces 12 (each shared by two cells), b 4 on fluid cells
    hipError_t e = LAUNCH(k_div_check, uint32_t(blocks), g, markers, u, v, w, b, partials, n);
    if (e != hipSuccess) return e;
    e = launch_dims(k_sum2, "k_sum2", dim3(1), dim3(1024), 0u, s, (const double*)partials, uint32_t(blocks), out + 1);
    if (e != hipSuccess) return e;
    return launch_dims(k_div_fold, "k_div_fold", dim3(1), dim3(64), 0u, s, (const double*)(out + 1), out);
}

double flat_wet_cells(const GridDev& g, const uint8_t* wet_host) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t per = uint32_t(kBlock * kCpt), nb = (F.n + per - 1) / per;
    double cells = 0.0;
    for (uint32_t b = 0; b < nb; ++b)
        if (wet_host[b]) cells += double(std::min(per, F.n - b * per));
    return cells;
}

hipError_t launch_flat_wet(hipStream_t s, const GridDev& g, const uint16_t* codes, uint8_t* wet) {
    if (!wet || !codes) return hipErrorInvalidValue;
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    return LAUNCH(k_flat_wet, flat_grid(F.n), F, codes, wet);
}

hipError_t launch_col_p_ap(hipStream_t s, const GridDev& g, const float scale[3], const uint16_t* codes,
                           const float* zv, float* p, float* ap, CgScalars* sc, ReduceBuf red, const uint8_t* wet) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    hipError_t e = LAUNCH(k_cg_pap, G, F, codes, scale[0], scale[1], scale[2], zv, p, ap, sc, red.partials, wet);
    if (e != hipSuccess) return e;
    return finish(s, red.partials, G, sc, kFinishAlpha);
}

hipError_t launch_col_update(hipStream_t s, const GridDev& g, const float* p, const float* ap, float* x, float* r,
                             CgScalars* sc, ReduceBuf red, int identity, const uint8_t* wet) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    hipError_t e = LAUNCH(k_cg_update, G, F, p, ap, x, r, sc, red.partials, wet);
    if (e != hipSuccess) return e;
    return finish(s, red.partials, G, sc, identity ? kFinishRrIdent : kFinishRr);
}

hipError_t launch_mg_coarsen(hipStream_t s, const MgDims& f, const MgDims& c, const uint8_t* mf, uint8_t* mc,
                             int fluid_first) {
    const uint32_t n = uint32_t(c.nx) * c.ny * c.nz;
    return LAUNCH(k_mg_coarsen, grid_for(n), f, c, mf, mc, fluid_first, n);
}

// Levels at least this large restrict with the tiled kernel.
constexpr size_t kTileMinCells = 32768;
static bool use_tiles(const MgDims& l) { return size_t(l.nx) * l.ny * l.nz >= kTileMinCells; }

hipError_t launch_mg_smooth(hipStream_t s, const MgDims& l, const uint16_t* code, const float* b, const float* xi,
                            float* xo, float omega, CgScalars* sc, ReduceBuf red, int first, int with_rz) {
    const Flat F = make_flat(l.nx, l.ny, l.nz);
    const uint32_t G = flat_grid(F.n);
    hipError_t e;
    if (first && with_rz)
        e = LAUNCH((k_mg_smooth<true, true>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    else if (first)
        e = LAUNCH((k_mg_smooth<true, false>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    else if (with_rz)
        e = LAUNCH((k_mg_smooth<false, true>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    else
        e = LAUNCH((k_mg_smooth<false, false>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    if (e != hipSuccess || !with_rz) return e;
    return finish(s, red.partials, G, sc, kFinishRz);
}

hipError_t launch_mg_smooth2_first(hipStream_t s, const MgDims& l, const uint16_t* code, const float* b, float* xo,
                                   float omega, CgScalars* sc, ReduceBuf red, int with_rz) {
    const Flat F = make_flat(l.nx, l.ny, l.nz);
    const uint32_t G = flat_grid(F.n);
    hipError_t e;
    if (with_rz) e = LAUNCH(k_mg_smooth2_first<true>, G, F, l, code, b, xo, omega, sc, red.partials);
    else e = LAUNCH(k_mg_smooth2_first<false>, G, F, l, code, b, xo, omega, sc, red.partials);
    if (e != hipSuccess || !with_rz) return e;
    return finish(s, red.partials, G, sc, kFinishRz);
}

hipError_t launch_mg_smooth_prolong(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code,
                                    const float* b, const float* xi, const float* xc, float* xo, float omega,
                                    CgScalars* sc, ReduceBuf red, int with_rz) {

    const Flat F = make_flat(f.nx, f.ny, f.nz);
    const uint32_t G = flat_grid(F.n);
    Prolong P;
    P.FC = make_flat(c.nx, c.ny, c.nz);
    P.hx = f.nx > 1 ? 1 : 0;
    P.hy = f.ny > 1 ? 1 : 0;
    P.hz = f.nz > 1 ? 1 : 0;
    P.rnx = 1.0 / double(F.sy);
    P.rplane = 1.0 / double(F.sz);
    hipError_t e;
    if (with_rz)
        e = LAUNCH(k_mg_smooth_prolong<true>, G, F, f, P, code, b, xi, xc, xo, omega, sc, red.partials);
    else
        e = LAUNCH(k_mg_smooth_prolong<false>, G, F, f, P, code, b, xi, xc, xo, omega, sc, red.partials);
    if (e != hipSuccess || !with_rz) return e;
    return finish(s, red.partials, G, sc, kFinishRz);
}

// Fused descent / ascent.  The static split (no schedule): chunks of about
// kMarchChunk planes; more chunks (down to 8 planes) until the grid holds
// about kMarchBlocks blocks, two per CU.  A schedule's chunk is at least
// cmin, the shortest that keeps its units within kMarchMaxUnits.
constexpr int kMarchChunk = 32;
constexpr uint32_t kMarchBlocks = 512;
constexpr size_t kMarchMaxUnits = 2048;

bool mg_march_ok(const MgDims& f) { return f.nx > 1 && f.ny > 1 && f.nz > 1 && f.nz < 0xFFFF; }

static March make_march(const MgDims& f, const MgDims& c, int skip = 1) {
    March M;
    M.nx = f.nx; M.ny = f.ny; M.nz = f.nz;
    M.sy = uint32_t(f.nx);
    M.sz = uint32_t(f.nx) * uint32_t(f.ny);
    M.cnz = c.nz;
    M.csy = uint32_t(c.nx);
    M.csz = uint32_t(c.nx) * uint32_t(c.ny);
    M.tx = uint32_t((f.nx + kMTX - 1) / kMTX);
    M.ty = uint32_t((f.ny + kMTY - 1) / kMTY);
    const uint32_t xy = M.tx * M.ty;
    int chunks = (f.nz + kMarchChunk - 1) / kMarchChunk;
    const int want = int((kMarchBlocks + xy - 1) / xy);
    chunks = std::max(chunks, std::min(want, (f.nz + 7) / 8));
    int C = (f.nz + chunks - 1) / chunks;
    C = (C + kRzSlab - 1) / kRzSlab * kRzSlab;
    M.chunk = C;
    M.blocks = xy * uint32_t((f.nz + C - 1) / C);
    M.tiles = xy;
    M.wunits = xy * uint32_t((f.nz + kWetChunk - 1) / kWetChunk);
    int cm = 4;
    while (cm < f.nz && size_t(xy) * size_t((f.nz + cm - 1) / cm) > std::max<size_t>(kMarchMaxUnits, xy))
        cm += kRzSlab;
    M.cmin = cm;
    M.parts = xy * uint32_t((f.nz + kRzSlab - 1) / kRzSlab);
    const uint32_t units = std::max(xy * uint32_t((f.nz + cm - 1) / cm), M.blocks);
    M.gmax = (units + kXcds - 1) / kXcds * kXcds;
    M.skip = skip;
    return M;
}

uint32_t mg_march_grid(const MgDims& f, const MgDims& c) { return make_march(f, c).gmax; }

uint32_t mg_march_parts(const MgDims& f, const MgDims& c) { return make_march(f, c).parts; }

size_t mg_march_sched_words(const MgDims& f, const MgDims& c) {
    const March M = make_march(f, c);
    return size_t(kSchedHead) + M.gmax + M.wunits;
}

static uint32_t march_wet_grid(const March& M) { return (M.wunits + kXcds - 1) / kXcds * kXcds; }

double mg_march_sched_cells(const MgDims& f, const MgDims& c, const uint32_t* sched_host) {
    const March M = make_march(f, c);
    const int C = int(sched_host[0]);
    const uint32_t U = sched_host[1];
    if (C <= 0 || U > M.gmax) return -1.0;
    double cells = 0.0;
    for (uint32_t b = 0; b < U; ++b) {
        const uint32_t e = sched_host[kSchedHead + b];
        if (!(e & kSchedHeavy)) continue;
        const uint32_t u = e & ~kSchedHeavy;
        const uint32_t tx = u % M.tx, rest = u / M.tx, ty = rest % M.ty, tc = rest / M.ty;
        const int w = std::min(kMTX, M.nx - int(tx) * kMTX), h = std::min(kMTY, M.ny - int(ty) * kMTY);
        const int d = std::min(C, M.nz - int(tc) * C);
        cells += double(w) * h * d;
    }
    return cells;
}

hipError_t launch_mg_march_wet(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code,
                               uint32_t* sched) {
    if (!mg_march_ok(f) || !sched || !code) return hipErrorInvalidValue;
    const March M = make_march(f, c);
    return launch_dims(k_mg_march_wet, "k_mg_march_wet", dim3(march_wet_grid(M)), dim3(kMThreads), 0u, s, M, code,
                       sched);
}

hipError_t launch_mg_march_plan(hipStream_t s, const MgDims& f, const MgDims& c, uint32_t* sched, uint32_t slots) {
    if (!mg_march_ok(f) || !sched || slots == 0) return hipErrorInvalidValue;
    const March M = make_march(f, c);
    return launch_dims(k_mg_march_plan, "k_mg_march_plan", dim3(1), dim3(kPlanThreads), 0u, s, M, sched, slots, 0u);
}

hipError_t launch_mg_descend(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                             float* x, const uint16_t* codec, float* bc, float omega, const CgScalars* sc,
                             const uint32_t* sched, int skip) {
    if (!mg_march_ok(f)) return hipErrorInvalidValue;
    const March M = make_march(f, c, skip);
    const uint32_t G = sched ? M.gmax : (M.blocks + kXcds - 1) / kXcds * kXcds;
    return launch_dims(k_mg_march<false, false>, "k_mg_march", dim3(G), dim3(kMThreads), 0u, s, M, f, code, b,
                       static_cast<const float*>(nullptr), static_cast<const float*>(nullptr), x, codec, bc, omega,
                       sc, static_cast<double*>(nullptr), sched);
}

hipError_t launch_mg_ascend(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                            const float* xi, const float* xc, float* xo, float omega, CgScalars* sc, ReduceBuf red,
                            int with_rz, const uint32_t* sched, int skip) {
    if (!mg_march_ok(f)) return hipErrorInvalidValue;
    const March M = make_march(f, c, skip);
    const uint32_t G = sched ? M.gmax : (M.blocks + kXcds - 1) / kXcds * kXcds;
    const CgScalars* csc = sc;
    hipError_t e;
    if (with_rz)
        e = launch_dims(k_mg_march<true, true>, "k_mg_march", dim3(G), dim3(kMThreads), 0u, s, M, f, code, b, xi, xc, xo,
                        static_cast<const uint16_t*>(nullptr), static_cast<float*>(nullptr), omega, csc,
                        red.partials, sched);
    else
        e = launch_dims(k_mg_march<true, false>, "k_mg_march", dim3(G), dim3(kMThreads), 0u, s, M, f, code, b, xi, xc, xo,
                        static_cast<const uint16_t*>(nullptr), static_cast<float*>(nullptr), omega, csc,
                        static_cast<double*>(nullptr), sched);
    if (e != hipSuccess || !with_rz) return e;
    return finish(s, red.partials, M.parts, sc, kFinishRz);
}

size_t mg_coarse_lds_bytes(const MgDims* levels, int nlev) {
    size_t floats = 128 * size_t(nlev);
    for (int l = 0; l < nlev; ++l) {
        const size_t n = size_t(levels[l].nx) * levels[l].ny * levels[l].nz;
        floats += (n + 1) / 2 + (l == 0 ? 2 : 3) * n;
    }
    return floats * sizeof(float);
}

hipError_t launch_mg_coarse_cycle(hipStream_t s, const MgDims* levels, const uint16_t* const* codes, int nlev,
                                  const float* b0, float* x0, int nu, int last_sweeps, float omega,
                                  const CgScalars* sc) {
    if (nlev < 1 || nlev > kCoarseMaxLevels) return hipErrorInvalidValue;
    CoarseArgs A;
    A.nlev = nlev;
    A.nu = nu;
    A.last_sweeps = last_sweeps;
    A.omega = omega;
    A.b0 = b0;
    A.x0 = x0;
    for (int l = 0; l < nlev; ++l) {
        A.lv[l].nx = levels[l].nx; A.lv[l].ny = levels[l].ny; A.lv[l].nz = levels[l].nz;
        A.lv[l].sx = levels[l].sx; A.lv[l].sy = levels[l].sy; A.lv[l].sz = levels[l].sz;
        A.lv[l].code = codes[l];
    }
    (void)launch_dims(k_mg_coarse_cycle, "k_mg_coarse_cycle", dim3(1), dim3(kCoarseThreads), mg_coarse_lds_bytes(levels, nlev), s, A,
                       sc);
    return hipGetLastError();
}

hipError_t launch_mg_restrict(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* codef,
                              const float* bf, const float* xf, const uint16_t* codec, float* bc,
                              const CgScalars* sc) {
    const Flat FC = make_flat(c.nx, c.ny, c.nz), FF = make_flat(f.nx, f.ny, f.nz);
    const int slots = (f.nx > 1 ? 2 : 1) * (f.ny > 1 ? 2 : 1) * (f.nz > 1 ? 2 : 1);
    if (slots == 8 && use_tiles(f)) {
        const Tiles T = make_tiles(f), TC = make_tiles(c);
        return LAUNCH(k_mg_restrict_tile, tiles_grid(T), T, f, TC, codef, bf, xf, codec, bc, sc);
    }
    if (slots == 8) {
        const uint32_t gx = (uint32_t(f.nx) + 63) / 64;
        return LAUNCH(k_mg_restrict2, size_t(gx) * c.ny * c.nz + (8 - (size_t(gx) * c.ny * c.nz) % 8) % 8, FF, f,
                      FC, c, gx, codef, bf, xf, codec, bc, 0.125f, sc);
    }
    return LAUNCH(k_mg_restrict, grid_for(FC.n), FC, c, FF, f, codef, bf, xf, codec, bc, 1.0f / float(slots),
                  1.0 / double(FC.sy), 1.0 / double(FC.sz), sc);
}

// ===========================================================================
// Multi-scene records: the launches launch_cg_start, launch_col_p_ap,
// launch_col_update, launch_refine_true, launch_mg_smooth,
// launch_mg_smooth2_first, launch_mg_smooth_prolong, launch_mg_restrict,
// launch_mg_coarse_cycle, launch_mg_descend and launch_mg_ascend make, as
// argument structs for k_multi (the same arguments, grids and order).
template <class A>
static void push_rec(std::vector<MRec>& out, int32_t kind, const A& a, uint32_t blocks, uint32_t shmem = 0) {
    static_assert(sizeof(A) <= kMRecBytes, "argument struct fits a record");
    MRec r;
    r.kind = kind;
    r.blocks = blocks;
    r.shmem = shmem;
    r.bytes = uint32_t(sizeof(A));
    std::memcpy(r.args, &a, sizeof(A));
    out.push_back(r);
}

static void rec_finish(std::vector<MRec>& out, const double* part, uint32_t count, CgScalars* sc, int kind,
                       double rel_inner = 0.0, double abs_target2 = 0.0) {
    push_rec(out, kMkFinish, MFinish{part, count, sc, kind, rel_inner, abs_target2}, 1u);
}

size_t finish_arg_bytes() { return sizeof(MFinish); }
void finish_set_target(unsigned char* args, double abs_target2) {
    MFinish f;
    std::memcpy(&f, args, sizeof(MFinish));
    f.abs_target2 = abs_target2;
    std::memcpy(args, &f, sizeof(MFinish));
}

void rec_cg_start(std::vector<MRec>& out, const GridDev& g, const float* rsrc, float* r, float* x, CgScalars* sc,
                  ReduceBuf red, double rel_inner, double abs_target2) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    push_rec(out, kMkStart, MStart{F, rsrc, r, x, red.partials, G}, G);
    rec_finish(out, red.partials, G, sc, kFinishInit, rel_inner, abs_target2);
}

void rec_col_p_ap(std::vector<MRec>& out, const GridDev& g, const float scale[3], const uint16_t* codes,
                  const float* zv, float* p, float* ap, CgScalars* sc, ReduceBuf red, const uint8_t* wet) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    push_rec(out, kMkPap, MPap{F, codes, scale[0], scale[1], scale[2], zv, p, ap, sc, red.partials, wet, G}, G);
    rec_finish(out, red.partials, G, sc, kFinishAlpha);
}

void rec_col_update(std::vector<MRec>& out, const GridDev& g, const float* p, const float* ap, float* x, float* r,
                    CgScalars* sc, ReduceBuf red, int identity, const uint8_t* wet) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    push_rec(out, kMkUpdate, MUpdate{F, p, ap, x, r, sc, red.partials, wet, G}, G);
    rec_finish(out, red.partials, G, sc, identity ? kFinishRrIdent : kFinishRr);
}

void rec_refine_true(std::vector<MRec>& out, const GridDev& g, const float scale[3], const uint16_t* codes,
                     const float* xcorr, double* x64, const float* b, float* r, int mode, ReduceBuf red,
                     double* out2) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    if (mode == 0) {
        push_rec(out, kMkRefine0, MRefine<0>{F, codes, scale[0], scale[1], scale[2], xcorr, x64, b, r, red.partials, G},
                 G);
    } else if (mode == 1) {
        push_rec(out, kMkRefine1, MRefine<1>{F, codes, scale[0], scale[1], scale[2], xcorr, x64, b, r, red.partials, G},
                 G);
    } else {
        const uint32_t ga = grid_for(F.n);
        push_rec(out, kMkAcc64, MAcc64{F.n, xcorr, x64, ga}, ga);
        push_rec(out, kMkRefine2, MRefine<2>{F, codes, scale[0], scale[1], scale[2], xcorr, x64, b, r, red.partials, G},
                 G);
    }
    push_rec(out, kMkSum2, MSum2{red.partials, G, out2}, 1u);
}

void rec_codes(std::vector<MRec>& out, int nx, int ny, int nz, const uint8_t* markers, uint16_t* codes) {
    const uint32_t n = uint32_t(nx) * ny * nz;
    const uint32_t G = grid_for(n);
    push_rec(out, kMkCodes, MCodes{nx, ny, nz, markers, codes, n, G}, G);
}

void rec_flat_wet(std::vector<MRec>& out, const GridDev& g, const uint16_t* codes, uint8_t* wet) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n);
    push_rec(out, kMkFlatWet, MFlatWet{F, codes, wet, G}, G);
}

void rec_mg_coarsen(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint8_t* mf, uint8_t* mc,
                    int fluid_first) {
    const uint32_t n = uint32_t(c.nx) * c.ny * c.nz;
    const uint32_t G = grid_for(n);
    push_rec(out, kMkCoarsen, MCoarsen{f, c, mf, mc, fluid_first, n, G}, G);
}

void rec_mg_march_wet(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code,
                      uint32_t* sched) {
    const March M = make_march(f, c);
    const uint32_t G = march_wet_grid(M);
    push_rec(out, kMkMarchWet, MMarchWet{M, code, sched, G}, G);
}

// A record's schedule has the static chunk, and without one the units are
// the static blocks: its launches need only those.
static uint32_t rec_march_grid(const March& M, const uint32_t*) { return (M.blocks + kXcds - 1) / kXcds * kXcds; }

void rec_mg_march_plan(std::vector<MRec>& out, const MgDims& f, const MgDims& c, uint32_t* sched) {
    // many scenes share the device: the static chunk (short blocks that pack
    // the launch), ordered like a scene's own schedule
    const March M = make_march(f, c);
    push_rec(out, kMkMarchPlan, MMarchPlan{M, sched, 8u, uint32_t(M.chunk)}, 1u);
}

void rec_mg_smooth(std::vector<MRec>& out, const MgDims& l, const uint16_t* code, const float* b, const float* xi,
                   float* xo, float omega, CgScalars* sc, ReduceBuf red, int first, int with_rz) {
    const Flat F = make_flat(l.nx, l.ny, l.nz);
    const uint32_t G = flat_grid(F.n);
    if (first && with_rz) push_rec(out, kMkSmoothTT, MSmooth<true, true>{F, l, code, b, xi, xo, omega, sc, red.partials, G}, G);
    else if (first) push_rec(out, kMkSmoothTF, MSmooth<true, false>{F, l, code, b, xi, xo, omega, sc, red.partials, G}, G);
    else if (with_rz) push_rec(out, kMkSmoothFT, MSmooth<false, true>{F, l, code, b, xi, xo, omega, sc, red.partials, G}, G);
    else push_rec(out, kMkSmoothFF, MSmooth<false, false>{F, l, code, b, xi, xo, omega, sc, red.partials, G}, G);
    if (with_rz) rec_finish(out, red.partials, G, sc, kFinishRz);
}

void rec_mg_smooth2_first(std::vector<MRec>& out, const MgDims& l, const uint16_t* code, const float* b, float* xo,
                          float omega, CgScalars* sc, ReduceBuf red, int with_rz) {
    const Flat F = make_flat(l.nx, l.ny, l.nz);
    const uint32_t G = flat_grid(F.n);
    if (with_rz) push_rec(out, kMkSmooth2T, MSmooth2<true>{F, l, code, b, xo, omega, sc, red.partials, G}, G);
    else push_rec(out, kMkSmooth2F, MSmooth2<false>{F, l, code, b, xo, omega, sc, red.partials, G}, G);
    if (with_rz) rec_finish(out, red.partials, G, sc, kFinishRz);
}

void rec_mg_smooth_prolong(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code,
                           const float* b, const float* xi, const float* xc, float* xo, float omega, CgScalars* sc,
                           ReduceBuf red, int with_rz) {
    const Flat F = make_flat(f.nx, f.ny, f.nz);
    const uint32_t G = flat_grid(F.n);
    Prolong P;
    P.FC = make_flat(c.nx, c.ny, c.nz);
    P.hx = f.nx > 1 ? 1 : 0;
    P.hy = f.ny > 1 ? 1 : 0;
    P.hz = f.nz > 1 ? 1 : 0;
    P.rnx = 1.0 / double(F.sy);
    P.rplane = 1.0 / double(F.sz);
    if (with_rz) push_rec(out, kMkProlongT, MProlong<true>{F, f, P, code, b, xi, xc, xo, omega, sc, red.partials, G}, G);
    else push_rec(out, kMkProlongF, MProlong<false>{F, f, P, code, b, xi, xc, xo, omega, sc, red.partials, G}, G);
    if (with_rz) rec_finish(out, red.partials, G, sc, kFinishRz);
}

void rec_mg_restrict(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* codef, const float* bf,
                     const float* xf, const uint16_t* codec, float* bc, const CgScalars* sc) {
    const Flat FC = make_flat(c.nx, c.ny, c.nz), FF = make_flat(f.nx, f.ny, f.nz);
    const int slots = (f.nx > 1 ? 2 : 1) * (f.ny > 1 ? 2 : 1) * (f.nz > 1 ? 2 : 1);
    if (slots == 8 && use_tiles(f)) {
        const Tiles T = make_tiles(f), TC = make_tiles(c);
        const uint32_t G = tiles_grid(T);
        push_rec(out, kMkRestrictTile, MRestrictTile{T, f, TC, codef, bf, xf, codec, bc, sc, G}, G);
        return;
    }
    if (slots == 8) {
        const uint32_t gx = (uint32_t(f.nx) + 63) / 64;
        const uint32_t G = uint32_t(size_t(gx) * c.ny * c.nz + (8 - (size_t(gx) * c.ny * c.nz) % 8) % 8);
        push_rec(out, kMkRestrict2, MRestrict2{FF, f, FC, c, gx, codef, bf, xf, codec, bc, 0.125f, sc, G}, G);
        return;
    }
    const uint32_t G = grid_for(FC.n);
    push_rec(out, kMkRestrict,
             MRestrict{FC, c, FF, f, codef, bf, xf, codec, bc, 1.0f / float(slots), 1.0 / double(FC.sy),
                       1.0 / double(FC.sz), sc, G},
             G);
}

bool rec_mg_coarse_cycle(std::vector<MRec>& out, const MgDims* levels, const uint16_t* const* codes, int nlev,
                         const float* b0, float* x0, int nu, int last_sweeps, float omega, const CgScalars* sc) {
    if (nlev < 1 || nlev > kCoarseMaxLevels) return false;
    MCoarse m;
    m.A.nlev = nlev;
    m.A.nu = nu;
    m.A.last_sweeps = last_sweeps;
    m.A.omega = omega;
    m.A.b0 = b0;
    m.A.x0 = x0;
    for (int l = 0; l < nlev; ++l) {
        m.A.lv[l].nx = levels[l].nx; m.A.lv[l].ny = levels[l].ny; m.A.lv[l].nz = levels[l].nz;
        m.A.lv[l].sx = levels[l].sx; m.A.lv[l].sy = levels[l].sy; m.A.lv[l].sz = levels[l].sz;
        m.A.lv[l].code = codes[l];
    }
    m.sc = sc;
    push_rec(out, kMkCoarse, m, 1u, uint32_t(mg_coarse_lds_bytes(levels, nlev)));
    return true;
}

void rec_mg_descend(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                    float* x, const uint16_t* codec, float* bc, float omega, const CgScalars* sc,
                    const uint32_t* sched, int skip) {
    const March M = make_march(f, c, skip);
    const uint32_t G = rec_march_grid(M, sched);
    push_rec(out, kMkMarchDesc,
             MMarch<false, false>{M, f, code, b, nullptr, nullptr, x, codec, bc, omega, sc, nullptr, sched, G}, G);
}

void rec_mg_ascend(std::vector<MRec>& out, const MgDims& f, const MgDims& c, const uint16_t* code, const float* b,
                   const float* xi, const float* xc, float* xo, float omega, CgScalars* sc, ReduceBuf red,
                   int with_rz, const uint32_t* sched, int skip) {
    const March M = make_march(f, c, skip);
    const uint32_t G = rec_march_grid(M, sched);
    if (with_rz) {
        push_rec(out, kMkMarchAscRz,
                 MMarch<true, true>{M, f, code, b, xi, xc, xo, nullptr, nullptr, omega, sc, red.partials, sched, G}, G);
        rec_finish(out, red.partials, M.parts, sc, kFinishRz);
    } else {
        push_rec(out, kMkMarchAsc,
                 MMarch<true, false>{M, f, code, b, xi, xc, xo, nullptr, nullptr, omega, sc, nullptr, sched, G}, G);
    }
}

// Copies each scene's CG scalars and two refinement sums into packed arrays
// (one work-item per scene).
struct MGather {
    static constexpr int kThreads = 64;
    const CgScalars* sc;
    const double* out2;
    CgScalars* dst_sc;
    double* dst2;
    __device__ void run(uint32_t) const {
        if (threadIdx.x == 0) {
            *dst_sc = *sc;
            dst2[0] = out2[0];
            dst2[1] = out2[1];
        }
    }
};

void rec_gather(std::vector<MRec>& out, const CgScalars* sc, const double* out2, CgScalars* dst_sc, double* dst2) {
    push_rec(out, kMkGather, MGather{sc, out2, dst_sc, dst2}, 1u);
}

template <class A>
static hipError_t launch_m(hipStream_t s, const char* name, const void* tab, const uint32_t* map, uint32_t nblocks,
                           uint32_t shmem, const uint8_t* active) {
    return launch_dims(k_multi<A>, name, dim3(nblocks), dim3(A::kThreads), shmem, s, static_cast<const A*>(tab), map,
                       active);
}

template <class A>
static hipError_t launch_m_march(hipStream_t s, const void* tab, const uint32_t* map, uint32_t nblocks,
                                 const uint8_t* active) {
    return launch_dims(k_multi_march<A>, "k_multi_march", dim3(nblocks), dim3(A::kThreads), 0u, s,
                       static_cast<const A*>(tab), map, active);
}

hipError_t launch_multi(hipStream_t s, int32_t kind, const void* tab, const uint32_t* map, uint32_t nblocks,
                        uint32_t shmem, const uint8_t* active) {
    if (nblocks == 0) return hipSuccess;
    switch (kind) {
        case kMkStart: return launch_m<MStart>(s, "k_multi_cg_start", tab, map, nblocks, 0, active);
        case kMkFinish: return launch_m<MFinish>(s, "k_multi_finish", tab, map, nblocks, 0, active);
        case kMkPap: return launch_m<MPap>(s, "k_multi_cg_pap", tab, map, nblocks, 0, active);
        case kMkUpdate: return launch_m<MUpdate>(s, "k_multi_cg_update", tab, map, nblocks, 0, active);
        case kMkRefine1: return launch_m<MRefine<1>>(s, "k_multi_refine", tab, map, nblocks, 0, active);
        case kMkRefine2: return launch_m<MRefine<2>>(s, "k_multi_refine", tab, map, nblocks, 0, active);
        case kMkAcc64: return launch_m<MAcc64>(s, "k_multi_acc64", tab, map, nblocks, 0, active);
        case kMkSum2: return launch_m<MSum2>(s, "k_multi_sum2", tab, map, nblocks, 0, active);
        case kMkSmoothTT: return launch_m<MSmooth<true, true>>(s, "k_multi_mg_smooth", tab, map, nblocks, 0, active);
        case kMkSmoothTF: return launch_m<MSmooth<true, false>>(s, "k_multi_mg_smooth", tab, map, nblocks, 0, active);
        case kMkSmoothFT: return launch_m<MSmooth<false, true>>(s, "k_multi_mg_smooth", tab, map, nblocks, 0, active);
        case kMkSmoothFF: return launch_m<MSmooth<false, false>>(s, "k_multi_mg_smooth", tab, map, nblocks, 0, active);
        case kMkSmooth2T: return launch_m<MSmooth2<true>>(s, "k_multi_mg_smooth2_first", tab, map, nblocks, 0, active);
        case kMkSmooth2F: return launch_m<MSmooth2<false>>(s, "k_multi_mg_smooth2_first", tab, map, nblocks, 0, active);
        case kMkProlongT: return launch_m<MProlong<true>>(s, "k_multi_mg_smooth_prolong", tab, map, nblocks, 0, active);
        case kMkProlongF: return launch_m<MProlong<false>>(s, "k_multi_mg_smooth_prolong", tab, map, nblocks, 0, active);
        case kMkRestrict: return launch_m<MRestrict>(s, "k_multi_mg_restrict", tab, map, nblocks, 0, active);
        case kMkRestrict2: return launch_m<MRestrict2>(s, "k_multi_mg_restrict", tab, map, nblocks, 0, active);
        case kMkRestrictTile: return launch_m<MRestrictTile>(s, "k_multi_mg_restrict_tile", tab, map, nblocks, 0, active);
        case kMkCoarse: return launch_m<MCoarse>(s, "k_multi_mg_coarse_cycle", tab, map, nblocks, shmem, active);
        case kMkMarchDesc: return launch_m_march<MMarch<false, false>>(s, tab, map, nblocks, active);
        case kMkMarchAsc: return launch_m_march<MMarch<true, false>>(s, tab, map, nblocks, active);
        case kMkMarchAscRz: return launch_m_march<MMarch<true, true>>(s, tab, map, nblocks, active);
        case kMkGather: return launch_m<MGather>(s, "k_multi_gather", tab, map, nblocks, 0, active);
        case kMkRefine0: return launch_m<MRefine<0>>(s, "k_multi_refine", tab, map, nblocks, 0, active);
        case kMkCodes: return launch_m<MCodes>(s, "k_multi_codes", tab, map, nblocks, 0, active);
        case kMkFlatWet: return launch_m<MFlatWet>(s, "k_multi_flat_wet", tab, map, nblocks, 0, active);
        case kMkCoarsen: return launch_m<MCoarsen>(s, "k_multi_mg_coarsen", tab, map, nblocks, 0, active);
        case kMkMarchWet: return launch_m<MMarchWet>(s, "k_multi_mg_march_wet", tab, map, nblocks, 0, active);
        case kMkMarchPlan: return launch_m<MMarchPlan>(s, "k_multi_mg_march_plan", tab, map, nblocks, 0, active);
        default: return launch_multi_k(s, kind, tab, map, nblocks, shmem, active);
    }
}

}  // namespace mgf
