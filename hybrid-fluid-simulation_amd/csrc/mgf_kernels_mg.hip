// mgf_kernels_mg.hip — stencil codes, the fused kernels of the mixed-precision
// CG and the geometric multigrid V-cycle used as its preconditioner.
//
// Layout and access pattern: one work-item per cell (CPT cells per work-item,
// 256 apart, so every load of a wave is one coalesced row), cells x-fastest.
// Every load of a stencil is issued unconditionally — a neighbour outside the
// grid is read at a clamped index and its value discarded by a select on the
// cell's stencil code — so all loads of a cell are in flight together instead
// of one guarded load after another.
//
// Reductions: each block writes one FP64 partial sum with a plain store; a
// one-block kernel sums the partials in a fixed order and updates the CG
// scalars.  The kernel boundary makes the partials visible: no device-scope
// fence (which writes back and invalidates the XCD's L2) is executed per block.
#include "mgf_internal.h"
#include "mgf_kcommon.h"

namespace mgf {

// ---------------------------------------------------------------------------
// Stencil code of a cell (uint16): the matrix row of buildPressureMatrixA
// (MPhysics/EulerSolverTool.cu:58-126) in a form that needs no neighbour
// marker at apply time.  bit 0: fluid; bits 1-2, 3-4, 5-6: number of
// non-solid in-grid neighbours along x, y, z (the diagonal terms); bits 7-12:
// coupling to the -x, +x, -y, +y, -z, +z neighbour (both cells fluid).
enum : uint16_t {
    kCodeFluid = 1u << 0,
    kCplMX = 1u << 7, kCplPX = 1u << 8,
    kCplMY = 1u << 9, kCplPY = 1u << 10,
    kCplMZ = 1u << 11, kCplPZ = 1u << 12,
};
__device__ __forceinline__ int code_cx(uint16_t c) { return (c >> 1) & 3; }
__device__ __forceinline__ int code_cy(uint16_t c) { return (c >> 3) & 3; }
__device__ __forceinline__ int code_cz(uint16_t c) { return (c >> 5) & 3; }

// The diagonal in the reference's accumulation order: the x terms, then y,
// then z — bitwise the value buildPressureMatrixA stores.  Non-fluid rows are
// identity rows.
__device__ __forceinline__ float code_diag(uint16_t c, float sx, float sy, float sz) {
    const int cx = code_cx(c), cy = code_cy(c), cz = code_cz(c);
    float d = 0.0f;
    d = cx > 0 ? d + sx : d;
    d = cx > 1 ? d + sx : d;
    d = cy > 0 ? d + sy : d;
    d = cy > 1 ? d + sy : d;
    d = cz > 0 ? d + sz : d;
    d = cz > 1 ? d + sz : d;
    return (c & kCodeFluid) ? d : 1.0f;
}

__global__ void __launch_bounds__(kBlock) k_codes(int nx, int ny, int nz, const uint8_t* __restrict__ m,
                                                  uint16_t* __restrict__ code, uint32_t n) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const int x = int(i % uint32_t(nx)), y = int((i / uint32_t(nx)) % uint32_t(ny)),
              z = int(i / (uint32_t(nx) * uint32_t(ny)));
    if (m[i] != kFluid) { code[i] = 0; return; }
    const uint32_t sy = uint32_t(nx), sz = uint32_t(nx) * uint32_t(ny);
    uint16_t c = kCodeFluid;
    int cx = 0, cy = 0, cz = 0;
    if (x > 0) { const uint8_t q = m[i - 1]; if (q != kSolid) ++cx; if (q == kFluid) c |= kCplMX; }
    if (x + 1 < nx) { const uint8_t q = m[i + 1]; if (q != kSolid) ++cx; if (q == kFluid) c |= kCplPX; }
    if (y > 0) { const uint8_t q = m[i - sy]; if (q != kSolid) ++cy; if (q == kFluid) c |= kCplMY; }
    if (y + 1 < ny) { const uint8_t q = m[i + sy]; if (q != kSolid) ++cy; if (q == kFluid) c |= kCplPY; }
    if (z > 0) { const uint8_t q = m[i - sz]; if (q != kSolid) ++cz; if (q == kFluid) c |= kCplMZ; }
    if (z + 1 < nz) { const uint8_t q = m[i + sz]; if (q != kSolid) ++cz; if (q == kFluid) c |= kCplPZ; }
    c |= uint16_t(cx << 1) | uint16_t(cy << 3) | uint16_t(cz << 5);
    code[i] = c;
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

// (A v)_c from the code, the centre value and the six neighbour values
// (FP64 accumulation; the order -x, +x, -y, +y, -z, +z).  Branch-free, so the
// neighbour loads are not sunk below a test of the code.
__device__ __forceinline__ double arow(uint16_t cd, double vc, const double (&v)[6], float sx, float sy,
                                       float sz) {
    double a = double(code_diag(cd, sx, sy, sz)) * vc;
    a -= (cd & kCplMX) ? double(sx) * v[0] : 0.0;
    a -= (cd & kCplPX) ? double(sx) * v[1] : 0.0;
    a -= (cd & kCplMY) ? double(sy) * v[2] : 0.0;
    a -= (cd & kCplPY) ? double(sy) * v[3] : 0.0;
    a -= (cd & kCplMZ) ? double(sz) * v[4] : 0.0;
    a -= (cd & kCplPZ) ? double(sz) * v[5] : 0.0;
    return (cd & kCodeFluid) ? a : vc;
}

// FP32 version for the multigrid smoother and residual.
__device__ __forceinline__ float arow_f(uint16_t cd, float vc, const float (&v)[6], float sx, float sy,
                                        float sz) {
    float a = code_diag(cd, sx, sy, sz) * vc;
    a -= (cd & kCplMX) ? sx * v[0] : 0.0f;
    a -= (cd & kCplPX) ? sx * v[1] : 0.0f;
    a -= (cd & kCplMY) ? sy * v[2] : 0.0f;
    a -= (cd & kCplPY) ? sy * v[3] : 0.0f;
    a -= (cd & kCplMZ) ? sz * v[4] : 0.0f;
    a -= (cd & kCplPZ) ? sz * v[5] : 0.0f;
    return (cd & kCodeFluid) ? a : vc;
}

// Block sum of one double, stored to part[blockIdx.x] by thread 0.
__device__ __forceinline__ void block_partial(double v, double* part) {
    double a[1] = {v};
    block_sum<1>(a);
    if (threadIdx.x == 0) part[blockIdx.x] = a[0];
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

// ---------------------------------------------------------------------------
// Fused CG.
// K_I: p_new = z + beta p_old (p_new = z on the first iteration), with the
// neighbours' p_new recomputed from z and p_old; block partials of p . A p.
template <int CPT>
__global__ void __launch_bounds__(kBlock) k_cg_pap(Flat F, const uint16_t* __restrict__ code, float sx, float sy,
                                                   float sz, const float* __restrict__ zv,
                                                   const float* __restrict__ p_old, float* __restrict__ p_new,
                                                   const CgScalars* __restrict__ sc, double* __restrict__ part) {
    if (sc->done) return;
    const double beta = sc->beta;
    const bool first = sc->iters == 0;
    const uint32_t base = xcd_block() * (kBlock * CPT) + threadIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const uint32_t c = base + uint32_t(k) * kBlock;
        const bool ok = c < F.n;
        const uint32_t cc = ok ? c : F.n - 1;
        uint32_t j[6];
        nbrs(cc, F, j);
        const uint16_t cd = code[cc];
        double vc, v[6];
        if (first) {
            vc = double(zv[cc]);
#pragma unroll
            for (int m = 0; m < 6; ++m) v[m] = double(zv[j[m]]);
        } else {
            vc = double(float(__builtin_fma(beta, double(p_old[cc]), double(zv[cc]))));
#pragma unroll
            for (int m = 0; m < 6; ++m)
                v[m] = double(float(__builtin_fma(beta, double(p_old[j[m]]), double(zv[j[m]]))));
        }
        const double a = arow(cd, vc, v, sx, sy, sz);
        if (ok) p_new[c] = float(vc);
        acc += ok ? vc * a : 0.0;
    }
    block_partial(acc, part);
}

// K_II: A p recomputed from the stored p; x += alpha p; r -= alpha A p (each
// formed in FP64, stored FP32); block partials of r . r.
template <int CPT>
__global__ void __launch_bounds__(kBlock) k_cg_update(Flat F, const uint16_t* __restrict__ code, float sx,
                                                      float sy, float sz, const float* __restrict__ p,
                                                      float* __restrict__ x, float* __restrict__ r,
                                                      const CgScalars* __restrict__ sc, double* __restrict__ part) {
    if (sc->done) return;
    const double alpha = sc->alpha;
    const uint32_t base = xcd_block() * (kBlock * CPT) + threadIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const uint32_t c = base + uint32_t(k) * kBlock;
        const bool ok = c < F.n;
        const uint32_t cc = ok ? c : F.n - 1;
        uint32_t j[6];
        nbrs(cc, F, j);
        const uint16_t cd = code[cc];
        const double vc = double(p[cc]);
        double v[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) v[m] = double(p[j[m]]);
        const float x0 = x[cc], r0 = r[cc];
        const double a = arow(cd, vc, v, sx, sy, sz);
        const float xn = float(__builtin_fma(alpha, vc, double(x0)));
        const float rn = float(__builtin_fma(-alpha, a, double(r0)));
        if (ok) {
            x[c] = xn;
            r[c] = rn;
        }
        acc += ok ? double(rn) * double(rn) : 0.0;
    }
    block_partial(acc, part);
}

// The one-block finishing kernel: sums `count` partials in a fixed order and
// updates the scalars.
enum FinishKind : int {
    kFinishAlpha = 0,     // pap -> alpha
    kFinishRrIdent = 1,   // rr -> iters, done; beta = rr / rz, rz = rr (identity)
    kFinishRr = 2,        // rr -> iters, done (preconditioned: beta comes later)
    kFinishRz = 3,        // r . z -> beta = rz / rz_old (0 before the first iteration), rz
};
constexpr int kFinishBlock = 1024;

__global__ void __launch_bounds__(kFinishBlock) k_finish(const double* __restrict__ part, uint32_t count,
                                                         CgScalars* sc, int kind) {
    if (sc->done) return;
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += kFinishBlock) s += part[i];
    __shared__ double sh[kFinishBlock / 64];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < kFinishBlock / 64; ++q) t += sh[q];
    switch (kind) {
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

// ---------------------------------------------------------------------------
// Multigrid (preconditioner).  Level l: cells with code_l, right-hand side
// b_l, iterate x_l; a coarse cell is a 2x2x2 block of fine cells (one layer
// on an axis whose extent is 1).  Damped Jacobi smoothing, restriction =
// weight * sum over the fluid children's residuals, prolongation = injection
// into fluid children: with as many sweeps after the coarse correction as
// before it, the V-cycle is a symmetric operator.
using MgLevel = MgDims;

// Coarse markers: with fluid_first, fluid if any child is fluid, else air if
// any child is air, else solid; otherwise air wins over fluid.
__global__ void __launch_bounds__(kBlock) k_mg_coarsen(MgLevel F, MgLevel Cc, const uint8_t* __restrict__ mf,
                                                       uint8_t* __restrict__ mc, int fluid_first, uint32_t n) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
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

// One damped Jacobi sweep xo = xi + omega (b - A xi) / d on fluid cells (0
// elsewhere, and on a fluid cell whose diagonal is 0: every neighbour solid or
// outside the grid, an all-zero row of the singular system); FIRST: the sweep
// from xi = 0, xo = omega b / d.  RZ (the level-0 final sweep, b = r there)
// also writes block partials of r . z.
template <bool FIRST, bool RZ, int CPT>
__global__ void __launch_bounds__(kBlock) k_mg_smooth(Flat F, MgLevel L, const uint16_t* __restrict__ code,
                                                      const float* __restrict__ b, const float* __restrict__ xi,
                                                      float* __restrict__ xo, float omega,
                                                      const CgScalars* __restrict__ sc, double* __restrict__ part) {
    if (sc->done) return;
    const uint32_t base = xcd_block() * (kBlock * CPT) + threadIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const uint32_t c = base + uint32_t(k) * kBlock;
        const bool ok = c < F.n;
        const uint32_t cc = ok ? c : F.n - 1;
        const uint16_t cd = code[cc];
        const float bc = b[cc];
        const float d = code_diag(cd, L.sx, L.sy, L.sz);
        const bool live = (cd & kCodeFluid) && d != 0.0f;
        // the hardware reciprocal: a smoother needs no correctly rounded divide
        const float wd = omega * __builtin_amdgcn_rcpf(d);
        float out;
        if (FIRST) {
            out = live ? wd * bc : 0.0f;
        } else {
            uint32_t j[6];
            nbrs(cc, F, j);
            const float vc = xi[cc];
            float v[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) v[m] = xi[j[m]];
            const float ax = arow_f(cd, vc, v, L.sx, L.sy, L.sz);
            out = live ? vc + wd * (bc - ax) : 0.0f;
        }
        if (ok) xo[c] = out;
        if (RZ) acc += ok ? double(bc) * double(out) : 0.0;
    }
    if (RZ) block_partial(acc, part);
}

// The first two sweeps from x = 0 in one pass: x1 = omega D^-1 b is
// pointwise, so x2 = x1 + omega D^-1 (b - A x1) needs only b and the codes at
// the seven points.
template <bool RZ>
__global__ void __launch_bounds__(kBlock) k_mg_smooth2_first(Flat F, MgLevel L, const uint16_t* __restrict__ code,
                                                             const float* __restrict__ b, float* __restrict__ xo,
                                                             float omega, const CgScalars* __restrict__ sc,
                                                             double* __restrict__ part) {
    if (sc->done) return;
    const uint32_t c = xcd_block() * kBlock + threadIdx.x;
    const bool ok = c < F.n;
    const uint32_t cc = ok ? c : F.n - 1;
    uint32_t j[6];
    nbrs(cc, F, j);
    const uint16_t cd = code[cc];
    const float bc = b[cc];
    uint16_t cn[6];
    float bn[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        cn[m] = code[j[m]];
        bn[m] = b[j[m]];
    }
    const float d = code_diag(cd, L.sx, L.sy, L.sz);
    const bool live = (cd & kCodeFluid) && d != 0.0f;
    const float wd = omega * __builtin_amdgcn_rcpf(d);
    const float x1c = live ? wd * bc : 0.0f;
    float x1[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const float dn = code_diag(cn[m], L.sx, L.sy, L.sz);
        const bool ln = (cn[m] & kCodeFluid) && dn != 0.0f;
        x1[m] = ln ? omega * __builtin_amdgcn_rcpf(dn) * bn[m] : 0.0f;
    }
    const float ax = arow_f(cd, x1c, x1, L.sx, L.sy, L.sz);
    const float out = live ? x1c + wd * (bc - ax) : 0.0f;
    if (ok) xo[c] = out;
    if (RZ) block_partial(ok ? double(bc) * double(out) : 0.0, part);
}

// The first sweep after the coarse correction, with the prolongation by
// injection applied on the fly: x_in = x + xc[parent] at the seven points
// (only coupled, hence fluid, neighbours are used), xo = x_in + omega D^-1
// (b - A x_in); x_in is not stored.  h[xyz] = 1 where the axis was halved.
template <bool RZ>
__global__ void __launch_bounds__(kBlock) k_mg_smooth_prolong(Flat F, MgLevel L, Flat FC, int hx, int hy, int hz,
                                                              double rnx, double rplane,
                                                              const uint16_t* __restrict__ code,
                                                              const float* __restrict__ b,
                                                              const float* __restrict__ xi,
                                                              const float* __restrict__ xc, float* __restrict__ xo,
                                                              float omega, const CgScalars* __restrict__ sc,
                                                              double* __restrict__ part) {
    if (sc->done) return;
    const uint32_t c = xcd_block() * kBlock + threadIdx.x;
    const bool ok = c < F.n;
    const uint32_t cc = ok ? c : F.n - 1;
    const uint32_t z = udiv(cc, F.sz, rplane);
    const uint32_t rem = cc - z * F.sz;
    const uint32_t y = udiv(rem, F.sy, rnx);
    const uint32_t x = rem - y * F.sy;
    const uint32_t nx = uint32_t(L.nx), ny = uint32_t(L.ny), nz = uint32_t(L.nz);
    const uint32_t X = x >> hx, Y = y >> hy, Z = z >> hz;
    const uint32_t Xm = (x > 0 ? x - 1 : 0) >> hx, Xp = (x + 1 < nx ? x + 1 : x) >> hx;
    const uint32_t Ym = (y > 0 ? y - 1 : 0) >> hy, Yp = (y + 1 < ny ? y + 1 : y) >> hy;
    const uint32_t Zm = (z > 0 ? z - 1 : 0) >> hz, Zp = (z + 1 < nz ? z + 1 : z) >> hz;
    const uint32_t pz = Z * FC.sz, py = Y * FC.sy;
    const uint32_t pc = pz + py + X;
    const uint32_t p[6] = {pz + py + Xm, pz + py + Xp, pz + Ym * FC.sy + X,
                           pz + Yp * FC.sy + X, Zm * FC.sz + py + X, Zp * FC.sz + py + X};
    uint32_t j[6];
    nbrs(cc, F, j);
    const uint16_t cd = code[cc];
    const float bc = b[cc];
    const float vc = xi[cc] + xc[pc];
    float v[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) v[m] = xi[j[m]] + xc[p[m]];
    const float d = code_diag(cd, L.sx, L.sy, L.sz);
    const bool live = (cd & kCodeFluid) && d != 0.0f;
    const float wd = omega * __builtin_amdgcn_rcpf(d);
    const float ax = arow_f(cd, vc, v, L.sx, L.sy, L.sz);
    const float out = live ? vc + wd * (bc - ax) : 0.0f;
    if (ok) xo[c] = out;
    if (RZ) block_partial(ok ? double(bc) * double(out) : 0.0, part);
}

// Restriction of the residual, one work-item per coarse cell:
// bc = weight * sum over fluid children of (b - A x).
__global__ void __launch_bounds__(kBlock) k_mg_restrict(Flat FC, MgLevel Cd, Flat FF, MgLevel Fd,
                                                        const uint16_t* __restrict__ codef,
                                                        const float* __restrict__ bf, const float* __restrict__ xf,
                                                        const uint16_t* __restrict__ codec, float* __restrict__ bc,
                                                        float weight, double rnx, double rplane,
                                                        const CgScalars* __restrict__ sc) {
    if (sc->done) return;
    const uint32_t c = xcd_block() * kBlock + threadIdx.x;
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
                const float res = bv - arow_f(cd, vc, v, Fd.sx, Fd.sy, Fd.sz);
                sum += (in && (cd & kCodeFluid)) ? res : 0.0f;
            }
    if (ok) bc[c] = (codec[c] & kCodeFluid) ? sum * weight : 0.0f;
}

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

__device__ __forceinline__ void lds_sweep(const CoarseLevel& L, const float* b, const float* xi, float* xo,
                                          float omega, bool first) {
    const uint32_t n = uint32_t(L.nx) * uint32_t(L.ny) * uint32_t(L.nz);
    const Flat F{n, uint32_t(L.nx), uint32_t(L.nx) * uint32_t(L.ny)};
    for (uint32_t c = threadIdx.x; c < n; c += kCoarseThreads) {
        const uint16_t cd = L.code[c];
        const float d = code_diag(cd, L.sx, L.sy, L.sz);
        const bool live = (cd & kCodeFluid) && d != 0.0f;
        const float wd = omega * __builtin_amdgcn_rcpf(d);
        const float bc = b[c];
        float out;
        if (first) {
            out = live ? wd * bc : 0.0f;
        } else {
            uint32_t j[6];
            nbrs(c, F, j);
            float v[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) v[m] = xi[j[m]];
            const float vc = xi[c];
            out = live ? vc + wd * (bc - arow_f(cd, vc, v, L.sx, L.sy, L.sz)) : 0.0f;
        }
        xo[c] = out;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kCoarseThreads) k_mg_coarse_cycle(CoarseArgs A, const CgScalars* __restrict__ sc) {
    if (sc->done) return;
    extern __shared__ float lds[];
    float* xs[kCoarseMaxLevels];
    float* ts[kCoarseMaxLevels];
    const float* bs[kCoarseMaxLevels];
    // carve LDS: level 0: x, t; deeper levels: b, x, t
    float* q = lds;
    for (int l = 0; l < A.nlev; ++l) {
        const uint32_t n = uint32_t(A.lv[l].nx) * uint32_t(A.lv[l].ny) * uint32_t(A.lv[l].nz);
        if (l == 0) {
            bs[0] = A.b0;
        } else {
            bs[l] = q;
            q += n;
        }
        xs[l] = q; q += n;
        ts[l] = q; q += n;
    }
    // descent
    for (int l = 0; l < A.nlev; ++l) {
        const CoarseLevel& L = A.lv[l];
        const int sweeps = l == A.nlev - 1 ? A.last_sweeps : A.nu;
        lds_sweep(L, bs[l], xs[l], xs[l], A.omega, true);
        for (int k = 1; k < sweeps; ++k) {
            lds_sweep(L, bs[l], xs[l], ts[l], A.omega, false);
            float* t = xs[l]; xs[l] = ts[l]; ts[l] = t;
        }
        if (l == A.nlev - 1) break;
        // restriction of the residual into level l+1
        const CoarseLevel& C = A.lv[l + 1];
        const uint32_t nc = uint32_t(C.nx) * uint32_t(C.ny) * uint32_t(C.nz);
        const int fx = L.nx > 1 ? 2 : 1, fy = L.ny > 1 ? 2 : 1, fz = L.nz > 1 ? 2 : 1;
        const float weight = 1.0f / float(fx * fy * fz);
        const Flat FF{uint32_t(L.nx) * uint32_t(L.ny) * uint32_t(L.nz), uint32_t(L.nx), uint32_t(L.nx) * uint32_t(L.ny)};
        float* bc = const_cast<float*>(bs[l + 1]);
        const float* xf = xs[l];
        const float* bf = bs[l];
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
                        const uint16_t cd = L.code[f];
                        if (!(cd & kCodeFluid)) continue;
                        uint32_t j[6];
                        nbrs(f, FF, j);
                        float v[6];
#pragma unroll
                        for (int m = 0; m < 6; ++m) v[m] = xf[j[m]];
                        sum += bf[f] - arow_f(cd, xf[f], v, L.sx, L.sy, L.sz);
                    }
            bc[c] = (C.code[c] & kCodeFluid) ? sum * weight : 0.0f;
        }
        __syncthreads();
    }
    // ascent
    for (int l = A.nlev - 2; l >= 0; --l) {
        const CoarseLevel& L = A.lv[l];
        const CoarseLevel& C = A.lv[l + 1];
        const uint32_t n = uint32_t(L.nx) * uint32_t(L.ny) * uint32_t(L.nz);
        const int hx = L.nx > 1 ? 1 : 0, hy = L.ny > 1 ? 1 : 0, hz = L.nz > 1 ? 1 : 0;
        const float* xc = xs[l + 1];
        float* xf = xs[l];
        for (uint32_t c = threadIdx.x; c < n; c += kCoarseThreads) {
            const uint32_t x = c % uint32_t(L.nx), y = (c / uint32_t(L.nx)) % uint32_t(L.ny),
                           z = c / (uint32_t(L.nx) * uint32_t(L.ny));
            if (L.code[c] & kCodeFluid)
                xf[c] += xc[((z >> hz) * uint32_t(C.ny) + (y >> hy)) * uint32_t(C.nx) + (x >> hx)];
        }
        __syncthreads();
        for (int k = 0; k < A.nu; ++k) {
            lds_sweep(L, bs[l], xs[l], ts[l], A.omega, false);
            float* t = xs[l]; xs[l] = ts[l]; ts[l] = t;
        }
    }
    const CoarseLevel& L0 = A.lv[0];
    const uint32_t n0 = uint32_t(L0.nx) * uint32_t(L0.ny) * uint32_t(L0.nz);
    for (uint32_t c = threadIdx.x; c < n0; c += kCoarseThreads) A.x0[c] = xs[0][c];
}

// ===========================================================================
// Launchers.
constexpr int kCpt = 1;   // cells per work-item of the fine-grid kernels

static uint32_t flat_grid(uint32_t n, int cpt) {
    size_t g = (size_t(n) + size_t(kBlock) * cpt - 1) / (size_t(kBlock) * cpt);
    g = (g + kXcds - 1) / kXcds * kXcds;
    if (g == 0) g = kXcds;
    return uint32_t(g);
}

uint32_t col_blocks(const GridDev& g) {
    return flat_grid(make_flat(g.nx, g.ny, g.nz).n, 1);
}

hipError_t launch_codes(hipStream_t s, int nx, int ny, int nz, const uint8_t* markers, uint16_t* codes) {
    const uint32_t n = uint32_t(nx) * ny * nz;
    return LAUNCH(k_codes, grid_for(n), nx, ny, nz, markers, codes, n);
}

static hipError_t finish(hipStream_t s, const double* part, uint32_t count, CgScalars* sc, int kind) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kFinishBlock), 0, s, part, count, sc, kind);
    return hipGetLastError();
}

hipError_t launch_col_p_ap(hipStream_t s, const GridDev& g, const float scale[3], const uint16_t* codes,
                           const float* zv, const float* p_old, float* p_new, CgScalars* sc, ReduceBuf red) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n, kCpt);
    hipError_t e = LAUNCH(k_cg_pap<kCpt>, G, F, codes, scale[0], scale[1], scale[2], zv, p_old, p_new, sc,
                          red.partials);
    if (e != hipSuccess) return e;
    return finish(s, red.partials, G, sc, kFinishAlpha);
}

hipError_t launch_col_update(hipStream_t s, const GridDev& g, const float scale[3], const uint16_t* codes,
                             const float* p, float* x, float* r, CgScalars* sc, ReduceBuf red, int identity) {
    const Flat F = make_flat(g.nx, g.ny, g.nz);
    const uint32_t G = flat_grid(F.n, kCpt);
    hipError_t e = LAUNCH(k_cg_update<kCpt>, G, F, codes, scale[0], scale[1], scale[2], p, x, r, sc,
                          red.partials);
    if (e != hipSuccess) return e;
    return finish(s, red.partials, G, sc, identity ? kFinishRrIdent : kFinishRr);
}

hipError_t launch_mg_coarsen(hipStream_t s, const MgDims& f, const MgDims& c, const uint8_t* mf, uint8_t* mc,
                             int fluid_first) {
    const uint32_t n = uint32_t(c.nx) * c.ny * c.nz;
    return LAUNCH(k_mg_coarsen, grid_for(n), f, c, mf, mc, fluid_first, n);
}

hipError_t launch_mg_smooth(hipStream_t s, const MgDims& l, const uint16_t* code, const float* b, const float* xi,
                            float* xo, float omega, CgScalars* sc, ReduceBuf red, int first, int with_rz) {
    const Flat F = make_flat(l.nx, l.ny, l.nz);
    const uint32_t G = flat_grid(F.n, kCpt);
    hipError_t e;
    if (first && with_rz)
        e = LAUNCH((k_mg_smooth<true, true, kCpt>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    else if (first)
        e = LAUNCH((k_mg_smooth<true, false, kCpt>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    else if (with_rz)
        e = LAUNCH((k_mg_smooth<false, true, kCpt>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    else
        e = LAUNCH((k_mg_smooth<false, false, kCpt>), G, F, l, code, b, xi, xo, omega, sc, red.partials);
    if (e != hipSuccess || !with_rz) return e;
    return finish(s, red.partials, G, sc, kFinishRz);
}

hipError_t launch_mg_smooth2_first(hipStream_t s, const MgDims& l, const uint16_t* code, const float* b, float* xo,
                                   float omega, CgScalars* sc, ReduceBuf red, int with_rz) {
    const Flat F = make_flat(l.nx, l.ny, l.nz);
    const uint32_t G = flat_grid(F.n, 1);
    hipError_t e;
    if (with_rz) e = LAUNCH(k_mg_smooth2_first<true>, G, F, l, code, b, xo, omega, sc, red.partials);
    else e = LAUNCH(k_mg_smooth2_first<false>, G, F, l, code, b, xo, omega, sc, red.partials);
    if (e != hipSuccess || !with_rz) return e;
    return finish(s, red.partials, G, sc, kFinishRz);
}

hipError_t launch_mg_smooth_prolong(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* code,
                                    const float* b, const float* xi, const float* xc, float* xo, float omega,
                                    CgScalars* sc, ReduceBuf red, int with_rz) {
    const Flat F = make_flat(f.nx, f.ny, f.nz), FC = make_flat(c.nx, c.ny, c.nz);
    const uint32_t G = flat_grid(F.n, 1);
    const int hx = f.nx > 1 ? 1 : 0, hy = f.ny > 1 ? 1 : 0, hz = f.nz > 1 ? 1 : 0;
    const double rnx = 1.0 / double(F.sy), rplane = 1.0 / double(F.sz);
    hipError_t e;
    if (with_rz)
        e = LAUNCH(k_mg_smooth_prolong<true>, G, F, f, FC, hx, hy, hz, rnx, rplane, code, b, xi, xc, xo, omega, sc,
                   red.partials);
    else
        e = LAUNCH(k_mg_smooth_prolong<false>, G, F, f, FC, hx, hy, hz, rnx, rplane, code, b, xi, xc, xo, omega, sc,
                   red.partials);
    if (e != hipSuccess || !with_rz) return e;
    return finish(s, red.partials, G, sc, kFinishRz);
}

size_t mg_coarse_lds_bytes(const MgDims* levels, int nlev) {
    size_t floats = 0;
    for (int l = 0; l < nlev; ++l) {
        const size_t n = size_t(levels[l].nx) * levels[l].ny * levels[l].nz;
        floats += (l == 0 ? 2 : 3) * n;
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
    const size_t lds = mg_coarse_lds_bytes(levels, nlev);
    hipLaunchKernelGGL(k_mg_coarse_cycle, dim3(1), dim3(kCoarseThreads), lds, s, A, sc);
    return hipGetLastError();
}

hipError_t launch_mg_restrict(hipStream_t s, const MgDims& f, const MgDims& c, const uint16_t* codef,
                              const float* bf, const float* xf, const uint16_t* codec, float* bc,
                              const CgScalars* sc) {
    const Flat FC = make_flat(c.nx, c.ny, c.nz), FF = make_flat(f.nx, f.ny, f.nz);
    const int slots = (f.nx > 1 ? 2 : 1) * (f.ny > 1 ? 2 : 1) * (f.nz > 1 ? 2 : 1);
    return LAUNCH(k_mg_restrict, grid_for(FC.n), FC, c, FF, f, codef, bf, xf, codec, bc, 1.0f / float(slots),
                  1.0 / double(FC.sy), 1.0 / double(FC.sz), sc);
}


}  // namespace mgf
