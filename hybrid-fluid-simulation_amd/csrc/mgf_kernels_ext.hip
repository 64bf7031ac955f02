// mgf_kernels_ext.hip — velocity extrapolation of extrapolatingDataInvoker
// (MPhysics/EulerSolverTool.cu:450-557, 904-969).
//
// The reference assigns every face of a component a distance marker by
// breadth-first layers (buildExtrapolationMarkers: layer 0 = known faces,
// layer L = unknown faces with a layer L-1 neighbour, for L < N) and then, for
// L = 1 .. N-1 in order, sets each layer-L face to the mean of its neighbours
// whose marker is below L (extrapolatingData).  The breadth-first search runs
// on the full face grid with no obstacles, so the marker of a face is its L1
// distance to the nearest known face (unknown beyond N-1).  Here:
//
//   1. distance transform, exact and separable: along x a wave-wide
//      min-plus scan per row, along y and z a forward and a backward sweep per
//      line (d = min(d, neighbour + 1));
//   2. counting sort of the faces with 1 <= d <= N-1 by d (LDS histogram per
//      block, one global atomic per block and distance into one of 32
//      private copies);
//   3. one gather kernel per pair of layers, over those layers' faces only,
//      in layer order: the neighbours in layers below L are final when layer
//      L runs, and a layer-L+1 face recomputes its layer-L neighbours' values
//      (the sort also stores each face's mask of previous-layer neighbours,
//      per list entry and per face).
//
// The mean is summed in the reference's neighbour order (-x, +x, -y, +y, -z,
// +z) in FP32 and divided once, so the values are bitwise the reference's.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mgf_internal.h"
#include "mgf_kcommon.h"

namespace mgf {

constexpr int32_t kDistInf = 1 << 29;   // unreached (cell distances of the fluid SDF)
// Face distances are stored as uint16 (65535 = unreached): every reachable
// distance is at most nx + ny + nz + 3 < 65535 (the extrapolation refuses
// larger grids), so the transforms compute in int32 and clamp on store.
constexpr int32_t kFaceInf = 0xFFFF;

struct Faces {
    float* f[3];
    uint32_t ex[3], ey[3], ez[3];
    uint32_t off[4];      // first global face index of each component; off[3] = total
    uint32_t rows[4];     // prefix counts of x-rows (ey * ez per component)
    uint32_t ylines[4];   // prefix counts of y-lines (ex * ez)
    uint32_t zlines[4];   // prefix counts of z-lines (ex * ey)
    double rex[3], rexy[3];   // 1/ex, 1/(ex*ey)
};

__device__ __forceinline__ int comp_of(const uint32_t (&pre)[4], uint32_t t) {
    return t < pre[1] ? 0 : (t < pre[2] ? 1 : 2);
}

// Per-component values picked with selects: an index into the argument
// arrays with a per-lane component would load them from memory.
template <class T>
__device__ __forceinline__ T pick(const T (&a)[3], int c) {
    return c == 0 ? a[0] : (c == 1 ? a[1] : a[2]);
}
template <class T>
__device__ __forceinline__ T pick4(const T (&a)[4], int c) {
    return c == 0 ? a[0] : (c == 1 ? a[1] : a[2]);
}

__device__ __forceinline__ uint32_t udiv_r(uint32_t a, uint32_t d, double rd) {
    uint32_t q = uint32_t(double(a) * rd);
    const uint32_t r = a - q * d;
    if (int32_t(r) < 0) --q;
    else if (r >= d) ++q;
    return q;
}

// Min-scans over the 64 lanes of a wave: inclusive prefix (lane l gets the
// minimum over lanes 0..l) and suffix (over lanes l..63).  Within each row of
// 16 lanes four DPP row shifts (a lane whose source lies outside its row keeps
// the identity); across the four rows, each row's end lane read with
// v_readlane.  Integer minima, so any order gives the same result.
constexpr int32_t kMinId = 0x7FFFFFFF;
__device__ __forceinline__ int32_t wave_min_prefix(int32_t g) {
    g = min(g, __builtin_amdgcn_update_dpp(kMinId, g, 0x111, 0xF, 0xF, false));   // row_shr:1
    g = min(g, __builtin_amdgcn_update_dpp(kMinId, g, 0x112, 0xF, 0xF, false));   // row_shr:2
    g = min(g, __builtin_amdgcn_update_dpp(kMinId, g, 0x114, 0xF, 0xF, false));   // row_shr:4
    g = min(g, __builtin_amdgcn_update_dpp(kMinId, g, 0x118, 0xF, 0xF, false));   // row_shr:8
    const int32_t r0 = __builtin_amdgcn_readlane(g, 15);
    const int32_t r1 = min(r0, __builtin_amdgcn_readlane(g, 31));
    const int32_t r2 = min(r1, __builtin_amdgcn_readlane(g, 47));
    const uint32_t row = (threadIdx.x & 63) >> 4;
    return min(g, row == 0 ? kMinId : (row == 1 ? r0 : (row == 2 ? r1 : r2)));
}
__device__ __forceinline__ int32_t wave_min_suffix(int32_t h) {
    h = min(h, __builtin_amdgcn_update_dpp(kMinId, h, 0x101, 0xF, 0xF, false));   // row_shl:1
    h = min(h, __builtin_amdgcn_update_dpp(kMinId, h, 0x102, 0xF, 0xF, false));   // row_shl:2
    h = min(h, __builtin_amdgcn_update_dpp(kMinId, h, 0x104, 0xF, 0xF, false));   // row_shl:4
    h = min(h, __builtin_amdgcn_update_dpp(kMinId, h, 0x108, 0xF, 0xF, false));   // row_shl:8
    const int32_t r3 = __builtin_amdgcn_readlane(h, 48);
    const int32_t r2 = min(r3, __builtin_amdgcn_readlane(h, 32));
    const int32_t r1 = min(r2, __builtin_amdgcn_readlane(h, 16));
    const uint32_t row = (threadIdx.x & 63) >> 4;
    return min(h, row == 3 ? kMinId : (row == 2 ? r3 : (row == 1 ? r2 : r1)));
}

// ---------------------------------------------------------------------------
// 1a. Rows along x: d = 0 on known faces, INF elsewhere; then the 1-D L1
// distance: forward d(x) = x + min_{k<=x} (d(k) - k), backward
// d(x) = min_{k>=x} (d(k) + k) - x.  One wave per row.
__device__ __forceinline__ void k_ext_rows_run(Faces F, uint16_t* __restrict__ dist, uint32_t blk, uint32_t slot) {
    (void)blk;
    (void)slot;
    const uint32_t wave = slot * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= F.rows[3]) return;
    const int c = comp_of(F.rows, wave);
    const uint32_t ex = pick(F.ex, c);
    const uint32_t local = (wave - pick4(F.rows, c)) * ex;     // first face of the row in component c
    const float* f = pick(F.f, c);
    uint16_t* d = dist + pick4(F.off, c) + local;
    // forward
    int32_t carry = kFaceInf;
    for (uint32_t x0 = 0; x0 < ex; x0 += 64) {
        const uint32_t x = x0 + uint32_t(lane);
        const bool ok = x < ex;
        const int32_t d0 = ok ? (f[local + x] != MGF_UNKNOWN ? 0 : kFaceInf) : kFaceInf;
        const int32_t g = min(wave_min_prefix(d0 - int32_t(x)), carry);
        carry = __builtin_amdgcn_readlane(g, 63);
        if (ok) d[x] = uint16_t(min(g + int32_t(x), kFaceInf));
    }
    // backward (each lane re-reads what it wrote)
    carry = 2 * kFaceInf;
    const uint32_t nch = (ex + 63) / 64;
    for (uint32_t ch = nch; ch-- > 0;) {
        const uint32_t x = ch * 64 + uint32_t(lane);
        const bool ok = x < ex;
        const int32_t h = min(wave_min_suffix(ok ? int32_t(d[x]) + int32_t(x) : 2 * kFaceInf), carry);
        carry = __builtin_amdgcn_readlane(h, 0);
        if (ok) d[x] = uint16_t(min(h - int32_t(x), kFaceInf));
    }
}

__global__ void __launch_bounds__(kBlock) k_ext_rows(Faces F, uint16_t* __restrict__ dist) {
    k_ext_rows_run(F, dist, xcd_block(), blockIdx.x);
}

struct MExtRows {
    static constexpr int kThreads = kBlock;
    Faces F;
    uint16_t* dist;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_rows_run(F, dist, xcd_of(vb, G), vb); }
};

// 1b. Lines along y (axis 1) or z (axis 2): d = min(d, d(prev) + 1) forward
// then backward; one work-item per line, lines x-fastest (coalesced), 32
// loads in flight per chunk (a line is a chain of dependent minima, so the
// loads of a chunk are issued before its first minimum).
template <int AXIS>
__device__ __forceinline__ void k_ext_lines_run(Faces F, uint16_t* __restrict__ dist, uint32_t blk, uint32_t slot) {
    (void)blk;
    (void)slot;
    const uint32_t t = slot * kBlock + threadIdx.x;
    const uint32_t(&pre)[4] = AXIS == 1 ? F.ylines : F.zlines;
    if (t >= pre[3]) return;
    const int c = comp_of(pre, t);
    const uint32_t l = t - pick4(pre, c);
    const uint32_t ex = pick(F.ex, c), ey = pick(F.ey, c);
    const uint32_t w = udiv_r(l, ex, pick(F.rex, c));
    const uint32_t x = l - w * ex;
    uint32_t start, stride, len;
    if (AXIS == 1) { start = w * ex * ey + x; stride = ex; len = ey; }
    else { start = w * ex + x; stride = ex * ey; len = pick(F.ez, c); }
    uint16_t* d = dist + pick4(F.off, c) + start;
    constexpr int K = 32;
    int32_t prev = kFaceInf;
    for (uint32_t k0 = 0; k0 < len; k0 += K) {
        int32_t v[K];
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = (k0 + q < len) ? int32_t(d[size_t(k0 + q) * stride]) : 0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            if (k0 + q < len) {
                prev = min(v[q], prev + 1);
                d[size_t(k0 + q) * stride] = uint16_t(prev);
            }
        }
    }
    prev = kFaceInf;
    for (uint32_t k1 = len; k1 > 0;) {
        const uint32_t k0 = k1 >= K ? k1 - K : 0;
        const uint32_t m = k1 - k0;
        int32_t v[K];
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = (uint32_t(q) < m) ? int32_t(d[size_t(k1 - 1 - q) * stride]) : 0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            if (uint32_t(q) < m) {
                prev = min(v[q], prev + 1);
                d[size_t(k1 - 1 - q) * stride] = uint16_t(min(prev, kFaceInf));
            }
        }
        k1 = k0;
    }
}

template <int AXIS>
__global__ void __launch_bounds__(kBlock) k_ext_lines(Faces F, uint16_t* __restrict__ dist) {
    k_ext_lines_run<AXIS>(F, dist, xcd_block(), blockIdx.x);
}

template <int AXIS>
struct MExtLines {
    static constexpr int kThreads = kBlock;
    Faces F;
    uint16_t* dist;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_lines_run<AXIS>(F, dist, xcd_of(vb, G), vb); }
};

// 1b'. The same line transform in one read and one write of the distances
// (k_ext_lines re-reads and re-writes them in its backward sweep).  A block of
// kLineSegWaves waves takes 64 consecutive lines (one per lane); wave k holds
// the k-th segment of every line in registers (at most SEG faces).  With
// P_k = min (d(j) - j) and S_k = min (d(j) + j) over segment k, kept in LDS:
//   forward   F(i) = i + min(min_{m<k} P_m, min_{a_k<=j<=i} (d(j) - j))
//   backward  d'(i) = min(min_{m>k} S_m, min_{i<=j<b_k} (F(j) + j)) - i,
// which is min_j (d(j) + |i - j|), the result of the two full-line sweeps
// (integer minima: the same distances).
constexpr int kLineSegWaves = 16;
constexpr int kLineSegThreads = 64 * kLineSegWaves;
constexpr int kLineSegSizes[4] = {12, 20, 36, 64};   // SEG classes: lines of at most 16 SEG faces
constexpr int32_t kSegNone = 1 << 20;                 // past the line's end: above any d(j) +- j

// Stripes of 64 lines in component c (the last one partial).
template <int AXIS>
__device__ __forceinline__ uint32_t seg_stripes(const Faces& F, int c) {
    const uint32_t(&pre)[4] = AXIS == 1 ? F.ylines : F.zlines;
    const uint32_t n = c == 0 ? pre[1] - pre[0] : (c == 1 ? pre[2] - pre[1] : pre[3] - pre[2]);
    return (n + 63) >> 6;
}

template <int AXIS, int SEG, int WAVES = kLineSegWaves>
__device__ __forceinline__ void k_ext_lines_seg_run(Faces F, uint16_t* __restrict__ dist, uint32_t blk, uint32_t slot) {
    (void)slot;
    __shared__ int32_t sp[WAVES][64], ss[WAVES][64];
    const uint32_t(&pre)[4] = AXIS == 1 ? F.ylines : F.zlines;
    const uint32_t n0 = seg_stripes<AXIS>(F, 0), n1 = seg_stripes<AXIS>(F, 1), n2 = seg_stripes<AXIS>(F, 2);
    if (blk >= n0 + n1 + n2) return;
    const int c = blk < n0 ? 0 : (blk < n0 + n1 ? 1 : 2);
    const uint32_t sb = blk - (c == 0 ? 0u : (c == 1 ? n0 : n0 + n1));
    const uint32_t lines = c == 0 ? pre[1] - pre[0] : (c == 1 ? pre[2] - pre[1] : pre[3] - pre[2]);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t l = sb * 64 + lane;
    const bool ok = l < lines;
    const uint32_t ex = pick(F.ex, c), ey = pick(F.ey, c);
    const uint32_t w = udiv_r(ok ? l : 0u, ex, pick(F.rex, c));
    const uint32_t x = (ok ? l : 0u) - w * ex;
    uint32_t start, stride, len;
    if (AXIS == 1) { start = w * ex * ey + x; stride = ex; len = ey; }
    else { start = w * ex + x; stride = ex * ey; len = pick(F.ez, c); }
    const uint32_t lb = 2u * (pick4(F.off, c) + start);   // this lane's line start, in bytes (< 2^32)
    const uint32_t seg = (len + WAVES - 1) / WAVES;
    const uint32_t a = wv * seg;
    const uint32_t n = a < len ? min(seg, len - a) : 0u;   // faces of this wave's segment
    // Buffer accesses: the lane's line start in the VGPR offset, the row's
    // wave-uniform byte offset in the SGPR offset (no 64-bit address per row).
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(dist, int16_t(0), int32_t(2u * F.off[3]), kBufWord3);
    const uint32_t rb = 2u * stride;
    int32_t v[SEG];
#pragma unroll
    for (int q = 0; q < SEG; ++q) {
        v[q] = (ok && uint32_t(q) < n) ? int32_t(__builtin_amdgcn_raw_buffer_load_b16(rs, lb, (a + q) * rb, 0))
                                       : kSegNone;
    }
    int32_t pm = kSegNone, sm = kSegNone;
#pragma unroll
    for (int q = 0; q < SEG; ++q) {
        pm = min(pm, v[q] - int32_t(a + q));
        sm = min(sm, v[q] + int32_t(a + q));
    }
    sp[wv][lane] = pm;
    ss[wv][lane] = sm;
    __syncthreads();
    int32_t f = kSegNone, g = kSegNone;
#pragma unroll
    for (int m = 0; m < WAVES; ++m) {
        if (uint32_t(m) < wv) f = min(f, sp[m][lane]);
        if (uint32_t(m) > wv) g = min(g, ss[m][lane]);
    }
#pragma unroll
    for (int q = 0; q < SEG; ++q) {
        f = min(f, v[q] - int32_t(a + q));
        v[q] = f + int32_t(a + q);
    }
#pragma unroll
    for (int q = SEG - 1; q >= 0; --q) {
        g = min(g, v[q] + int32_t(a + q));
        v[q] = g - int32_t(a + q);
    }
#pragma unroll
    for (int q = 0; q < SEG; ++q) {
        if (ok && uint32_t(q) < n)
            __builtin_amdgcn_raw_buffer_store_b16(uint16_t(min(v[q], kFaceInf)), rs, lb, (a + q) * rb, 0);
    }
}

template <int AXIS, int SEG, int WAVES = kLineSegWaves>
__global__ void __launch_bounds__(64 * WAVES) k_ext_lines_seg(Faces F, uint16_t* __restrict__ dist) {
    k_ext_lines_seg_run<AXIS, SEG, WAVES>(F, dist, xcd_block(), blockIdx.x);
}

template <int AXIS, int SEG, int WAVES = kLineSegWaves>
struct MExtLinesSeg {
    static constexpr int kThreads = 64 * WAVES;
    Faces F;
    uint16_t* dist;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_lines_seg_run<AXIS, SEG, WAVES>(F, dist, xcd_of(vb, G), vb); }
};

// Short lines (at most kLineShortWaves * SEG faces) in work-groups of
// kLineShortWaves waves: with 16 waves, a line of 33 faces gives each wave
// 3 faces behind the same loads, exchange and barrier.
constexpr int kLineShortWaves = 4;
constexpr int kLineShortSizes[2] = {12, 36};

// 1a'. The row transform from a ballot of the known faces: a row's
// distances start as 0 (known) or INF, and the two min-plus sweeps of
// k_ext_rows leave at x  min(x - p, n - x, INF), p and n the nearest known
// faces at or before and at or after x (INF where neither exists).  One wave a
// row of at most 64 NCH faces, held in registers; per chunk of 64 faces the
// wave takes the ballot of its known lanes, a lane finds p and n in it with a
// leading- and a trailing-zero count, and the nearest known faces in the
// chunks before and after come from scalar scans of the ballots.  The same
// integers as the sweeps, from one read of the faces and one write of the
// distances.  The faces and distances are raw buffers (the lane's face in a
// 32-bit offset).  At 256^3 this takes 104 us against 125 us for wave-wide
// min-plus scans of the same registers; two, four or eight rows a wave
// measured slower.
template <int NCH>
__device__ __forceinline__ void k_ext_rows_reg_run(Faces F, uint16_t* __restrict__ dist, uint32_t blk, uint32_t slot) {
    (void)blk;
    constexpr int32_t kFar = 1 << 20;   // no known face on that side
    // wave-uniform (readfirstlane), so that the row's component, buffer and
    // bounds live in scalar registers
    const uint32_t row = slot * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (row >= F.rows[3]) return;
    const int c = comp_of(F.rows, row);
    const uint32_t ex = pick(F.ex, c);
    const uint32_t local = (row - pick4(F.rows, c)) * ex;
    const uint32_t fb = pick4(F.off, c) + local;   // the row's first face (global index)
    const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc(
        pick(F.f, c), int16_t(0), int32_t(4u * (c == 0 ? F.off[1] - F.off[0] : (c == 1 ? F.off[2] - F.off[1] : F.off[3] - F.off[2]))),
        kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dist, int16_t(0), int32_t(2u * F.off[3]), kBufWord3);
    float fv[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        const uint32_t x = uint32_t(ch) * 64 + lane;
        fv[ch] = x < ex ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rf, 4u * (local + x), 0, 0)) : MGF_UNKNOWN;
    }
    uint64_t m[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) m[ch] = __ballot(fv[ch] != MGF_UNKNOWN);
    // nearest known faces in the chunks before and after each chunk (scalar)
    int32_t pc[NCH], nc[NCH];
    int32_t run = -kFar;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        pc[ch] = run;
        if (m[ch]) run = ch * 64 + 63 - __clzll(m[ch]);
    }
    run = kFar;
#pragma unroll
    for (int ch = NCH - 1; ch >= 0; --ch) {
        nc[ch] = run;
        if (m[ch]) run = ch * 64 + __ffsll((unsigned long long)m[ch]) - 1;
    }
    const uint64_t below = (2ull << lane) - 1ull;   // lanes 0 .. lane
    const uint64_t above = ~0ull << lane;            // lanes lane .. 63
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        const int32_t x = ch * 64 + int32_t(lane);
        if (uint32_t(x) < ex) {
            const uint64_t lo = m[ch] & below, hi = m[ch] & above;
            const int32_t p = lo ? ch * 64 + 63 - __clzll(lo) : pc[ch];
            const int32_t n = hi ? ch * 64 + __ffsll((unsigned long long)hi) - 1 : nc[ch];
            const int32_t d = min(min(x - p, n - x), kFaceInf);
            __builtin_amdgcn_raw_buffer_store_b16(uint16_t(d), rd, 2u * (fb + uint32_t(x)), 0, 0);
        }
    }
}

// Short rows (at most NCH * 64 faces, NCH 1 or 2): the launch is bound by its
// wave count, one wave a row, so a wave takes kRowsPerWave consecutive rows,
// all their loads issued before the first row's distances; each row's
// arithmetic is the one above.
constexpr int kRowsPerWave = 4;

template <int NCH>
__device__ __forceinline__ void k_ext_rows_short_run(Faces F, uint16_t* __restrict__ dist, uint32_t blk,
                                                     uint32_t slot) {
    (void)blk;
    constexpr int32_t kFar = 1 << 20;
    const uint32_t row0 = (slot * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * kRowsPerWave;
    const uint32_t lane = threadIdx.x & 63;
    if (row0 >= F.rows[3]) return;
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dist, int16_t(0), int32_t(2u * F.off[3]), kBufWord3);
    float fv[kRowsPerWave][NCH];
    uint32_t exr[kRowsPerWave], fbr[kRowsPerWave];
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        const uint32_t row = row0 + uint32_t(r);
        const bool ok = row < F.rows[3];
        const int c = comp_of(F.rows, ok ? row : row0);
        const uint32_t ex = ok ? pick(F.ex, c) : 0u;   // a row past the end loads and stores nothing
        const uint32_t local = ok ? (row - pick4(F.rows, c)) * ex : 0u;
        exr[r] = ex;
        fbr[r] = pick4(F.off, c) + local;
        const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc(
            pick(F.f, c), int16_t(0),
            int32_t(4u * (c == 0 ? F.off[1] - F.off[0] : (c == 1 ? F.off[2] - F.off[1] : F.off[3] - F.off[2]))), kBufWord3);
        // loaded unconditionally (past the component's end the buffer reads
        // 0), then the faces past the row replaced
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const uint32_t x = uint32_t(ch) * 64 + lane;
            const float v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rf, 4u * (local + x), 0, 0));
            fv[r][ch] = x < ex ? v : MGF_UNKNOWN;
        }
    }
    // every row's values in registers before the first row's ballots
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r)
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) asm volatile("" : "+v"(fv[r][ch]));
    const uint64_t below = (2ull << lane) - 1ull;   // lanes 0 .. lane
    const uint64_t above = ~0ull << lane;            // lanes lane .. 63
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        uint64_t m[NCH];
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) m[ch] = __ballot(fv[r][ch] != MGF_UNKNOWN);
        int32_t pc[NCH], nc[NCH];
        int32_t run = -kFar;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            pc[ch] = run;
            if (m[ch]) run = ch * 64 + 63 - __clzll(m[ch]);
        }
        run = kFar;
#pragma unroll
        for (int ch = NCH - 1; ch >= 0; --ch) {
            nc[ch] = run;
            if (m[ch]) run = ch * 64 + __ffsll((unsigned long long)m[ch]) - 1;
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int32_t x = ch * 64 + int32_t(lane);
            if (uint32_t(x) < exr[r]) {
                const uint64_t lo = m[ch] & below, hi = m[ch] & above;
                const int32_t p = lo ? ch * 64 + 63 - __clzll(lo) : pc[ch];
                const int32_t n = hi ? ch * 64 + __ffsll((unsigned long long)hi) - 1 : nc[ch];
                const int32_t d = min(min(x - p, n - x), kFaceInf);
                __builtin_amdgcn_raw_buffer_store_b16(uint16_t(d), rd, 2u * (fbr[r] + uint32_t(x)), 0, 0);
            }
        }
    }
}

template <int NCH>
__global__ void __launch_bounds__(kBlock) k_ext_rows_short(Faces F, uint16_t* __restrict__ dist) {
    k_ext_rows_short_run<NCH>(F, dist, xcd_block(), blockIdx.x);
}

template <int NCH>
struct MExtRowsShort {
    static constexpr int kThreads = kBlock;
    Faces F;
    uint16_t* dist;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_rows_short_run<NCH>(F, dist, xcd_of(vb, G), vb); }
};

template <int NCH>
__global__ void __launch_bounds__(kBlock) k_ext_rows_reg(Faces F, uint16_t* __restrict__ dist) {
    k_ext_rows_reg_run<NCH>(F, dist, xcd_block(), blockIdx.x);
}

template <int NCH>
struct MExtRowsReg {
    static constexpr int kThreads = kBlock;
    Faces F;
    uint16_t* dist;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_rows_reg_run<NCH>(F, dist, xcd_of(vb, G), vb); }
};

// ---------------------------------------------------------------------------
// 2. Counting sort by distance.  Bins 1..cap.  A block counts its faces in
// LDS (kLdsBins bins; larger distances go straight to the global counters),
// then adds its non-zero bins to one of kCopies private copies of the global
// histogram (copy = block % kCopies): the atomics of different blocks on one
// distance are spread over kCopies addresses instead of serialising on one.
constexpr int kLdsBins = 2048;
constexpr int kSortCpt = 8;   // faces per work-item
constexpr int kCopies = 32;

// Layout of the counter buffer for bins 0..K+1 (K2 = K + 2 entries a row):
// hist[kCopies][K2], cursor[kCopies][K2], poff[kCopies][K2], offs[K2], maxd.
struct SortBufs {
    uint32_t *hist, *cursor, *poff, *offs, *maxd;
};
static SortBufs sort_bufs(uint32_t* counts, int32_t K) {
    const size_t K2 = size_t(K) + 2;
    SortBufs b;
    b.hist = counts;
    b.cursor = b.hist + kCopies * K2;
    b.poff = b.cursor + kCopies * K2;
    b.offs = b.poff + kCopies * K2;
    b.maxd = b.offs + K2;
    return b;
}

// Counts the active lanes of a wave by key into LDS counters and returns, per
// active lane, a slot in [old counter, old counter + lanes with its key).
// Where every active lane holds the same key (the common case: a row of faces
// at one distance) one lane adds them all with one atomic; otherwise each lane
// adds itself (distances along a row differ, so the addresses mostly differ).
__device__ __forceinline__ uint32_t lds_count(uint32_t* ctr_of_key, int32_t key, bool active) {
    const uint64_t act = __ballot(active);
    if (act == 0) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)act) - 1;
    const int32_t k0 = __builtin_amdgcn_readlane(key, leader);
    const uint64_t same = __ballot(active && key == k0);
    if (same == act) {
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&ctr_of_key[k0], uint32_t(__popcll(act)));
        base = __builtin_amdgcn_readlane(base, leader);
        return base + uint32_t(__popcll(act & ((1ull << lane) - 1ull)));
    }
    return active ? atomicAdd(&ctr_of_key[key], 1u) : 0u;
}

__device__ __forceinline__ void k_ext_zero_run(uint32_t* __restrict__ a, uint32_t n, uint32_t blk, uint32_t slot) {
    (void)blk;
    (void)slot;
    const uint32_t i = slot * kBlock + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void __launch_bounds__(kBlock) k_ext_zero(uint32_t* __restrict__ a, uint32_t n) {
    k_ext_zero_run(a, n, xcd_block(), blockIdx.x);
}

struct MExtZero {
    static constexpr int kThreads = kBlock;
    uint32_t* a;
    uint32_t n;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_zero_run(a, n, xcd_of(vb, G), vb); }
};

__device__ __forceinline__ void k_ext_count_run(const uint16_t* __restrict__ dist, uint32_t total, int32_t cap,
                                                      uint32_t* __restrict__ hist, uint32_t K2, uint32_t blk, uint32_t slot) {
    (void)blk;
    (void)slot;
    __shared__ uint32_t h[kLdsBins];
    const int32_t nb = min(cap + 1, kLdsBins);
    for (int b = threadIdx.x; b < nb; b += kBlock) h[b] = 0;
    __syncthreads();
    uint32_t* hp = hist + size_t(slot % kCopies) * K2;
    const uint32_t base = blk * (kBlock * kSortCpt) + threadIdx.x;
    // all loads first (clamped index, no branch), then the counting
    int32_t dv[kSortCpt];
#pragma unroll
    for (int k = 0; k < kSortCpt; ++k) {
        const uint32_t i = base + uint32_t(k) * kBlock;
        dv[k] = dist[i < total ? i : total - 1];
        if (i >= total) dv[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < kSortCpt; ++k) {
        const bool in = dv[k] >= 1 && dv[k] <= cap;
        (void)lds_count(h, dv[k] < kLdsBins ? dv[k] : 0, in && dv[k] < kLdsBins);
        if (in && dv[k] >= kLdsBins) atomicAdd(&hp[dv[k]], 1u);
    }
    __syncthreads();
    for (int b = 1 + threadIdx.x; b < nb; b += kBlock)
        if (h[b]) atomicAdd(&hp[b], h[b]);
}

__global__ void __launch_bounds__(kBlock) k_ext_count(const uint16_t* __restrict__ dist, uint32_t total, int32_t cap,
                                                      uint32_t* __restrict__ hist, uint32_t K2) {
    k_ext_count_run(dist, total, cap, hist, K2, xcd_block(), blockIdx.x);
}

struct MExtCount {
    static constexpr int kThreads = kBlock;
    const uint16_t* dist;
    uint32_t total;
    int32_t cap;
    uint32_t* hist;
    uint32_t K2;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_count_run(dist, total, cap, hist, K2, xcd_of(vb, G), vb); }
};

// Per distance b: T = sum over copies; offs[b] = exclusive scan of T over
// b = 1..cap (offs[cap+1] = total); poff[p][b] = offs[b] + counts of the
// copies before p; cursors cleared; *maxd = the largest non-empty distance.
// One block.
__device__ __forceinline__ void k_ext_scan_run(const uint32_t* __restrict__ hist, int32_t cap, uint32_t K2,
                                                   uint32_t* __restrict__ offs, uint32_t* __restrict__ poff,
                                                   uint32_t* __restrict__ cursor, uint32_t* __restrict__ maxd, uint32_t blk, uint32_t slot) {
    (void)blk;
    (void)slot;
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    __shared__ int32_t top;
    if (threadIdx.x == 0) { carry = 0; top = 0; }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int32_t b0 = 1; b0 <= cap; b0 += 1024) {
        const int32_t b = b0 + int32_t(threadIdx.x);
        uint32_t v = 0;
        if (b <= cap) {
#pragma unroll 8
            for (int p = 0; p < kCopies; ++p) v += hist[size_t(p) * K2 + b];
        }
        uint32_t incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        uint32_t before = carry;
        for (int q = 0; q < wv; ++q) before += wsum[q];
        if (b <= cap) {
            const uint32_t o = before + incl - v;
            offs[b] = o;
            uint32_t run = o;
            for (int p = 0; p < kCopies; ++p) {
                poff[size_t(p) * K2 + b] = run;
                cursor[size_t(p) * K2 + b] = 0;
                run += hist[size_t(p) * K2 + b];
            }
            if (v) atomicMax(&top, b);
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        offs[cap + 1] = carry;
        *maxd = uint32_t(top);
    }
}

__global__ void __launch_bounds__(1024) k_ext_scan(const uint32_t* __restrict__ hist, int32_t cap, uint32_t K2,
                                                   uint32_t* __restrict__ offs, uint32_t* __restrict__ poff,
                                                   uint32_t* __restrict__ cursor, uint32_t* __restrict__ maxd) {
    k_ext_scan_run(hist, cap, K2, offs, poff, cursor, maxd, xcd_block(), blockIdx.x);
}

struct MExtScan {
    static constexpr int kThreads = 1024;
    const uint32_t* hist;
    int32_t cap;
    uint32_t K2;
    uint32_t* offs;
    uint32_t* poff;
    uint32_t* cursor;
    uint32_t* maxd;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_scan_run(hist, cap, K2, offs, poff, cursor, maxd, xcd_of(vb, G), vb); }
};

// Bit q of the lower-neighbour mask of a face: its in-grid neighbour in
// direction q (-x, +x, -y, +y, -z, +z) has a smaller distance, i.e. lies in
// the previous layer.
__device__ __forceinline__ uint32_t lower_mask(const Faces& F, const uint16_t* __restrict__ dist, uint32_t gi,
                                               int32_t dv) {
    const int c = comp_of(F.off, gi);
    const uint32_t o = pick4(F.off, c);
    const uint32_t i = gi - o;
    const uint32_t ex = pick(F.ex, c), ey = pick(F.ey, c), ez = pick(F.ez, c), sz = ex * ey;
    const uint32_t z = udiv_r(i, sz, pick(F.rexy, c));
    const uint32_t rem = i - z * sz;
    const uint32_t y = udiv_r(rem, ex, pick(F.rex, c));
    const uint32_t x = rem - y * ex;
    const bool in[6] = {x > 0, x + 1 < ex, y > 0, y + 1 < ey, z > 0, z + 1 < ez};
    const uint32_t j[6] = {in[0] ? gi - 1 : gi, in[1] ? gi + 1 : gi, in[2] ? gi - ex : gi,
                           in[3] ? gi + ex : gi, in[4] ? gi - sz : gi, in[5] ? gi + sz : gi};
    int32_t dn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) dn[q] = dist[j[q]];
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) m |= (in[q] && dn[q] < dv) ? (1u << q) : 0u;
    return m;
}

__device__ __forceinline__ void k_ext_scatter_run(Faces F, const uint16_t* __restrict__ dist, uint32_t total,
                                                        int32_t cap, uint32_t K2, const uint32_t* __restrict__ poff,
                                                        uint32_t* cursor, uint32_t* __restrict__ list,
                                                        uint8_t* __restrict__ lmask, uint8_t* __restrict__ fmask, uint32_t blk, uint32_t slot) {
    (void)blk;
    (void)slot;
    __shared__ uint32_t h[kLdsBins];
    const int32_t nb = min(cap + 1, kLdsBins);
    for (int b = threadIdx.x; b < nb; b += kBlock) h[b] = 0;
    __syncthreads();
    const size_t copy = size_t(slot % kCopies) * K2;
    const uint32_t* po = poff + copy;
    uint32_t* cu = cursor + copy;
    const uint32_t base = blk * (kBlock * kSortCpt) + threadIdx.x;
    int32_t dv[kSortCpt];
    uint32_t rank[kSortCpt];
#pragma unroll
    for (int k = 0; k < kSortCpt; ++k) {
        const uint32_t i = base + uint32_t(k) * kBlock;
        dv[k] = dist[i < total ? i : total - 1];
        if (i >= total) dv[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < kSortCpt; ++k) {
        const bool in = dv[k] >= 1 && dv[k] <= cap;
        rank[k] = lds_count(h, dv[k] < kLdsBins ? dv[k] : 0, in && dv[k] < kLdsBins);
        if (in && dv[k] >= kLdsBins) rank[k] = atomicAdd(&cu[dv[k]], 1u);
    }
    __syncthreads();
    // reserve each bin's range in this block's copy: the bin's slot now holds its base
    for (int b = 1 + threadIdx.x; b < nb; b += kBlock)
        if (h[b]) h[b] = atomicAdd(&cu[b], h[b]);
    __syncthreads();
    // Only faces that enter the list need their mask: known faces (0) and
    // those past the cap skip the six neighbour loads.
    uint32_t msk[kSortCpt];
#pragma unroll
    for (int k = 0; k < kSortCpt; ++k) {
        const uint32_t i = base + uint32_t(k) * kBlock;
        msk[k] = (dv[k] >= 1 && dv[k] <= cap) ? lower_mask(F, dist, i < total ? i : total - 1, dv[k]) : 0u;
    }
#pragma unroll
    for (int k = 0; k < kSortCpt; ++k) {
        const uint32_t i = base + uint32_t(k) * kBlock;
        if (dv[k] >= 1 && dv[k] <= cap) {
            const uint32_t pos = po[dv[k]] + (dv[k] < kLdsBins ? h[dv[k]] : 0u) + rank[k];
            list[pos] = i;
            lmask[pos] = uint8_t(msk[k]);
            fmask[i] = uint8_t(msk[k]);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_ext_scatter(Faces F, const uint16_t* __restrict__ dist, uint32_t total,
                                                        int32_t cap, uint32_t K2, const uint32_t* __restrict__ poff,
                                                        uint32_t* cursor, uint32_t* __restrict__ list,
                                                        uint8_t* __restrict__ lmask, uint8_t* __restrict__ fmask) {
    k_ext_scatter_run(F, dist, total, cap, K2, poff, cursor, list, lmask, fmask, xcd_block(), blockIdx.x);
}

struct MExtScatter {
    static constexpr int kThreads = kBlock;
    Faces F;
    const uint16_t* dist;
    uint32_t total;
    int32_t cap;
    uint32_t K2;
    const uint32_t* poff;
    uint32_t* cursor;
    uint32_t* list;
    uint8_t* lmask;
    uint8_t* fmask;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_scatter_run(F, dist, total, cap, K2, poff, cursor, list, lmask, fmask, xcd_of(vb, G), vb); }
};

// ---------------------------------------------------------------------------
// 3. Layer L: each listed face takes the mean of its neighbours in the
// previous layer (lower-neighbour mask from the sort); only those are loaded,
// all at once, and no distance is read.
//
// The mean of face i (component-local) over the neighbours flagged in m, in
// the reference's order (-x, +x, -y, +y, -z, +z) with one divide; the
// conditional sums as selects (an unused neighbour leaves the sum untouched,
// a -0.0 included).
// The flagged neighbours' values (0 for the others, never summed): only the
// flagged ones are loaded.
__device__ __forceinline__ void layer_loads(const float* f, uint32_t i, uint32_t m, uint32_t ex, uint32_t sz,
                                            float (&fv)[6]) {
    const uint32_t j[6] = {(m & 1u) ? i - 1 : i, (m & 2u) ? i + 1 : i, (m & 4u) ? i - ex : i,
                           (m & 8u) ? i + ex : i, (m & 16u) ? i - sz : i, (m & 32u) ? i + sz : i};
#pragma unroll
    for (int q = 0; q < 6; ++q) fv[q] = ((m >> q) & 1u) ? f[j[q]] : 0.0f;
}

// Every loaded value in a register here, before any is used: without this
// the compiler moves the first sum (0 + the first value) into that load's
// branch, whose wait then holds back the other loads by a memory round trip.
template <int N>
__device__ __forceinline__ void loads_done(float (&v)[N]) {
#pragma unroll
    for (int q = 0; q < N; ++q) asm volatile("" : "+v"(v[q]));
}

__device__ __forceinline__ float layer_sum(uint32_t m, const float (&fv)[6]) {
    float acc = 0.0f, cnt = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const bool use = (m >> q) & 1u;
        acc = use ? acc + fv[q] : acc;
        cnt = use ? cnt + 1.0f : cnt;
    }
    if (cnt != 0.0f) acc /= cnt;
    return acc;
}

__device__ __forceinline__ float layer_mean(const float* f, uint32_t i, uint32_t m, uint32_t ex, uint32_t sz) {
    float fv[6];
    layer_loads(f, i, m, ex, sz, fv);
    loads_done(fv);
    return layer_sum(m, fv);
}

__global__ void __launch_bounds__(kBlock) k_ext_layer(Faces F, const uint32_t* __restrict__ list,
                                                      const uint8_t* __restrict__ lmask, uint32_t begin,
                                                      uint32_t count) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const uint32_t gi = list[begin + t];
    const uint32_t m = lmask[begin + t];
    const int c = comp_of(F.off, gi);
    const uint32_t i = gi - pick4(F.off, c);
    const uint32_t ex = pick(F.ex, c), sz = ex * pick(F.ey, c);
    float* f = pick(F.f, c);
    f[i] = layer_mean(f, i, m, ex, sz);
}

// Layers L and L+1 in one launch (their list segments are adjacent: n0 faces
// of layer L from `begin`, then n1 of layer L+1).  A layer-L+1 face needs the
// new values of its layer-L neighbours; it recomputes each of them with the
// same operations from the final layer-(L-1) values (the neighbour's own
// lower mask is in the per-face mask array fmask) instead of reading what the
// layer-L work-items write in this launch, so no work-item reads a value
// written in the launch and the results are bitwise those of two launches.
__global__ void __launch_bounds__(kBlock) k_ext_layer2(Faces F, const uint32_t* __restrict__ list,
                                                       const uint8_t* __restrict__ lmask,
                                                       const uint8_t* __restrict__ fmask, uint32_t begin,
                                                       uint32_t n0, uint32_t n1) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n0 + n1) return;
    const uint32_t gi = list[begin + t];
    const uint32_t m = lmask[begin + t];
    const int c = comp_of(F.off, gi);
    const uint32_t o = pick4(F.off, c);
    const uint32_t i = gi - o;
    const uint32_t ex = pick(F.ex, c), sz = ex * pick(F.ey, c);
    float* f = pick(F.f, c);
    if (t < n0) {
        f[i] = layer_mean(f, i, m, ex, sz);
        return;
    }
    const uint32_t j[6] = {(m & 1u) ? i - 1 : i, (m & 2u) ? i + 1 : i, (m & 4u) ? i - ex : i,
                           (m & 8u) ? i + ex : i, (m & 16u) ? i - sz : i, (m & 32u) ? i + sz : i};
    uint32_t mn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) mn[q] = ((m >> q) & 1u) ? uint32_t(fmask[o + j[q]]) : 0u;
    // the layer-L neighbours' means: all their loads, then the sums
    float nv[6][6];
#pragma unroll
    for (int q = 0; q < 6; ++q) layer_loads(f, j[q], mn[q], ex, sz, nv[q]);
#pragma unroll
    for (int q = 0; q < 6; ++q) loads_done(nv[q]);
    float fv[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) fv[q] = layer_sum(mn[q], nv[q]);
    float acc = 0.0f, cnt = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const bool use = (m >> q) & 1u;
        acc = use ? acc + fv[q] : acc;
        cnt = use ? cnt + 1.0f : cnt;
    }
    if (cnt != 0.0f) acc /= cnt;
    f[i] = acc;
}

// Layers first..last in one work-group: layer by layer, a barrier between
// layers ordering each layer's stores before the next layer's loads (the
// work-group's work-items share one CU), kTailCpt entries per work-item per
// round.  A round is latency-bound (list entry, then the neighbours, then
// the store), so every load of a round is issued before any is used: the
// six neighbours are loaded unconditionally (an unflagged one at the face's
// own index) and only the flagged ones are summed.  For the tail of small
// layers, where a launch per layer costs more than the layer's work, and
// for small scenes' whole extrapolation.
constexpr int kTailThreads = 1024;
constexpr int kTailCpt = 8;

__device__ __forceinline__ void k_ext_tail_run(Faces F, const uint32_t* __restrict__ list,
                                               const uint8_t* __restrict__ lmask, const uint32_t* __restrict__ offs,
                                               int32_t first, int32_t last, const uint32_t* __restrict__ last_dev,
                                               uint32_t blk, uint32_t slot) {
    (void)blk;
    (void)slot;
    if (last_dev) last = int32_t(*last_dev);
    for (int32_t L = first; L <= last; ++L) {
        const uint32_t b = offs[L], e = offs[L + 1];
        for (uint32_t r0 = b; r0 < e; r0 += uint32_t(kTailThreads * kTailCpt)) {
            uint32_t gi[kTailCpt], m[kTailCpt];
#pragma unroll
            for (int k = 0; k < kTailCpt; ++k) {
                const uint32_t t = r0 + uint32_t(k * kTailThreads) + threadIdx.x;
                const uint32_t tc = t < e ? t : e - 1;
                gi[k] = list[tc];
                m[k] = t < e ? uint32_t(lmask[tc]) : 0u;
            }
            float fv[kTailCpt][6];
#pragma unroll
            for (int k = 0; k < kTailCpt; ++k) {
                const int c = comp_of(F.off, gi[k]);
                const uint32_t i = gi[k] - pick4(F.off, c);
                const uint32_t ex = pick(F.ex, c), sz = ex * pick(F.ey, c);
                const float* f = pick(F.f, c);
                const uint32_t mk = m[k];
                const uint32_t j[6] = {(mk & 1u) ? i - 1 : i, (mk & 2u) ? i + 1 : i, (mk & 4u) ? i - ex : i,
                                       (mk & 8u) ? i + ex : i, (mk & 16u) ? i - sz : i, (mk & 32u) ? i + sz : i};
#pragma unroll
                for (int q = 0; q < 6; ++q) fv[k][q] = f[j[q]];
            }
#pragma unroll
            for (int k = 0; k < kTailCpt; ++k) {
                const uint32_t t = r0 + uint32_t(k * kTailThreads) + threadIdx.x;
                float acc = 0.0f, cnt = 0.0f;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const bool use = (m[k] >> q) & 1u;
                    acc = use ? acc + fv[k][q] : acc;
                    cnt = use ? cnt + 1.0f : cnt;
                }
                if (cnt != 0.0f) acc /= cnt;
                if (t < e) {
                    const int c = comp_of(F.off, gi[k]);
                    pick(F.f, c)[gi[k] - pick4(F.off, c)] = acc;
                }
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kTailThreads) k_ext_tail(Faces F, const uint32_t* __restrict__ list,
                                                           const uint8_t* __restrict__ lmask,
                                                           const uint32_t* __restrict__ offs, int32_t first,
                                                           int32_t last, const uint32_t* __restrict__ last_dev) {
    k_ext_tail_run(F, list, lmask, offs, first, last, last_dev, xcd_block(), blockIdx.x);
}

struct MExtTail {
    static constexpr int kThreads = kTailThreads;
    Faces F;
    const uint32_t* list;
    const uint8_t* lmask;
    const uint32_t* offs;
    int32_t first;
    int32_t last;
    const uint32_t* last_dev;
    uint32_t G;
    __device__ void run(uint32_t vb) const { k_ext_tail_run(F, list, lmask, offs, first, last, last_dev, xcd_of(vb, G), vb); }
};

// ===========================================================================
// Fields the hybrid coupling reads (GridFluidSolver.cpp:571-613) on the
// cell-centred grid; FluidSurfaceDensity = 0.2 (Common.h:244).

struct Cells {
    uint32_t nx, ny, nz, n;
};

__device__ __forceinline__ void cell_nbrs(const Cells& C, uint32_t i, uint32_t (&j)[6]) {
    const uint32_t sxy = C.nx * C.ny;
    const uint32_t z = i / sxy, rem = i - z * sxy, y = rem / C.nx, x = rem - y * C.nx;
    j[0] = x > 0 ? i - 1 : i;
    j[1] = x + 1 < C.nx ? i + 1 : i;
    j[2] = y > 0 ? i - C.nx : i;
    j[3] = y + 1 < C.ny ? i + C.nx : i;
    j[4] = z > 0 ? i - sxy : i;
    j[5] = z + 1 < C.nz ? i + sxy : i;
}

// buildFluidDensity (MPhysics/EulerSolverTool.cu:408-448): fluid (phi < 0)
// outside the solids (marker >= 0) is 0.2 next to an air cell (phi >= 0 and
// marker >= 0, isInsideAir EulerMathTool.cu:344-348), else 1; the rest 0.
__global__ void __launch_bounds__(kBlock) k_fluid_density(Cells C, const float* __restrict__ phi,
                                                          const float* __restrict__ solid, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C.n) return;
    uint32_t j[6];
    cell_nbrs(C, i, j);
    bool air = false;
#pragma unroll
    for (int q = 0; q < 6; ++q) air = air || (phi[j[q]] >= 0.0f && solid[j[q]] >= 0.0f);
    const bool fluid = phi[i] < 0.0f && !(solid[i] < 0.0f);
    out[i] = fluid ? (air ? 0.2f : 1.0f) : 0.0f;
}

// buildFluidSDF (MPhysics/EulerSolverTool.cu:324-406) by distance: pass i of
// the reference assigns the cells at breadth-first distance i from the
// boundary layer of their own region (fluid cells with a non-fluid clamped
// neighbour, non-fluid cells with a fluid neighbour).  The nearest boundary
// cell of either region lies in a cell's own region (a path into the other
// region crosses an own boundary cell first), so one L1 distance transform
// from all boundary cells gives both distances.
__global__ void __launch_bounds__(kBlock) k_sdf_seed(Cells C, const float* __restrict__ phi,
                                                     int32_t* __restrict__ d) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C.n) return;
    uint32_t j[6];
    cell_nbrs(C, i, j);
    const bool in = phi[i] < 0.0f;
    bool other = false;
#pragma unroll
    for (int q = 0; q < 6; ++q) other = other || ((phi[j[q]] < 0.0f) != in);
    d[i] = other ? 0 : kDistInf;
}

// 1-D L1 distance along x rows in place (one wave per row): forward and
// backward min-plus scans as in k_ext_rows.
__global__ void __launch_bounds__(kBlock) k_cdt_rows(Cells C, int32_t* __restrict__ dist) {
    const uint32_t row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= C.ny * C.nz) return;
    int32_t* d = dist + size_t(row) * C.nx;
    const uint32_t ex = C.nx;
    int32_t carry = kDistInf;
    for (uint32_t x0 = 0; x0 < ex; x0 += 64) {
        const uint32_t x = x0 + uint32_t(lane);
        const bool ok = x < ex;
        const int32_t g = min(wave_min_prefix((ok ? d[x] : kDistInf) - int32_t(x)), carry);
        carry = __builtin_amdgcn_readlane(g, 63);
        if (ok) d[x] = min(g + int32_t(x), kDistInf);
    }
    carry = 2 * kDistInf;
    for (uint32_t ch = (ex + 63) / 64; ch-- > 0;) {
        const uint32_t x = ch * 64 + uint32_t(lane);
        const bool ok = x < ex;
        const int32_t h = min(wave_min_suffix(ok ? d[x] + int32_t(x) : 2 * kDistInf), carry);
        carry = __builtin_amdgcn_readlane(h, 0);
        if (ok) d[x] = min(h - int32_t(x), kDistInf);
    }
}

// Lines along y (AXIS 1) or z (AXIS 2), one work-item per line, x fastest.
template <int AXIS>
__global__ void __launch_bounds__(kBlock) k_cdt_lines(Cells C, int32_t* __restrict__ dist) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lines = AXIS == 1 ? C.nx * C.nz : C.nx * C.ny;
    if (t >= lines) return;
    const uint32_t w = t / C.nx, x = t - w * C.nx;
    uint32_t start, stride, len;
    if (AXIS == 1) { start = w * C.nx * C.ny + x; stride = C.nx; len = C.ny; }
    else { start = w * C.nx + x; stride = C.nx * C.ny; len = C.nz; }
    int32_t* d = dist + start;
    constexpr int K = 32;
    int32_t prev = kDistInf;
    for (uint32_t k0 = 0; k0 < len; k0 += K) {
        int32_t v[K];
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = (k0 + q < len) ? d[size_t(k0 + q) * stride] : 0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            if (k0 + q < len) {
                prev = min(v[q], prev + 1);
                d[size_t(k0 + q) * stride] = prev;
            }
        }
    }
    prev = kDistInf;
    for (uint32_t k1 = len; k1 > 0;) {
        const uint32_t k0 = k1 >= K ? k1 - K : 0;
        const uint32_t m = k1 - k0;
        int32_t v[K];
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = (uint32_t(q) < m) ? d[size_t(k1 - 1 - q) * stride] : 0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            if (uint32_t(q) < m) {
                prev = min(v[q], prev + 1);
                d[size_t(k1 - 1 - q) * stride] = min(prev, kDistInf);
            }
        }
        k1 = k0;
    }
}

// Values of buildFluidSDF from the distance d and the region: fluid d = 0 ->
// 0.5-0.2, fluid d -> -(d - (0.5-0.2)), non-fluid d -> d + (1.5-0.2), each
// formed in double and rounded once as the reference's double literals make
// them; d >= passes keeps -UNKNOWN (fluid) or UNKNOWN.
__global__ void __launch_bounds__(kBlock) k_sdf_value(Cells C, const float* __restrict__ phi,
                                                      const int32_t* __restrict__ d, int32_t passes,
                                                      float* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C.n) return;
    const bool in = phi[i] < 0.0f;
    const int32_t k = d[i];
    float v;
    if (k >= passes) v = in ? -MGF_UNKNOWN : MGF_UNKNOWN;
    else if (in) v = k == 0 ? float(0.5 - 0.2) : float(-(double(k) - (0.5 - 0.2)));
    else v = float(double(k) + (1.5 - 0.2));
    out[i] = v;
}

hipError_t launch_fluid_density(hipStream_t s, const GridDev& g, const float* phi, const float* solid, float* out) {
    const Cells C{uint32_t(g.nx), uint32_t(g.ny), uint32_t(g.nz), uint32_t(g.nx) * g.ny * g.nz};
    (void)launch_dims(k_fluid_density, "k_fluid_density", dim3((C.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, phi, solid, out);
    return hipGetLastError();
}

hipError_t launch_fluid_sdf(hipStream_t s, const GridDev& g, const float* phi, int32_t passes, int32_t* scratch,
                            float* out) {
    const Cells C{uint32_t(g.nx), uint32_t(g.ny), uint32_t(g.nz), uint32_t(g.nx) * g.ny * g.nz};
    const uint32_t G = (C.n + kBlock - 1) / kBlock;
    (void)launch_dims(k_sdf_seed, "k_sdf_seed", dim3(G), dim3(kBlock), 0, s, C, phi, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    (void)launch_dims(k_cdt_rows, "k_cdt_rows", dim3((C.ny * C.nz + 3) / 4), dim3(kBlock), 0, s, C, scratch);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    (void)launch_dims(k_cdt_lines<1>, "k_cdt_lines", dim3((C.nx * C.nz + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, scratch);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    (void)launch_dims(k_cdt_lines<2>, "k_cdt_lines", dim3((C.nx * C.ny + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, scratch);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    (void)launch_dims(k_sdf_value, "k_sdf_value", dim3(G), dim3(kBlock), 0, s, C, phi, scratch, passes, out);
    return hipGetLastError();
}

// ===========================================================================
static Faces faces_of(const GridDev& g, float* u, float* v, float* w) {
    Faces F;
    F.f[0] = u; F.f[1] = v; F.f[2] = w;
    F.off[0] = F.rows[0] = F.ylines[0] = F.zlines[0] = 0;
    for (int c = 0; c < 3; ++c) {
        F.ex[c] = uint32_t(g.nx + (c == 0));
        F.ey[c] = uint32_t(g.ny + (c == 1));
        F.ez[c] = uint32_t(g.nz + (c == 2));
        F.off[c + 1] = F.off[c] + F.ex[c] * F.ey[c] * F.ez[c];
        F.rows[c + 1] = F.rows[c] + F.ey[c] * F.ez[c];
        F.ylines[c + 1] = F.ylines[c] + F.ex[c] * F.ez[c];
        F.zlines[c + 1] = F.zlines[c] + F.ex[c] * F.ey[c];
        F.rex[c] = 1.0 / double(F.ex[c]);
        F.rexy[c] = 1.0 / (double(F.ex[c]) * double(F.ey[c]));
    }
    return F;
}

int32_t ext_max_distance(const GridDev& g) { return g.nx + g.ny + g.nz + 3; }

// Work-groups of k_ext_rows_reg: a row a wave, 4 waves a group.
static uint32_t rows_grid(const Faces& F, int nch) {
    (void)nch;
    return (F.rows[3] + 3) / 4;
}

// Work-groups of k_ext_rows_short: kRowsPerWave rows a wave, 4 waves a group.
static uint32_t rows_short_grid(const Faces& F) {
    return (F.rows[3] + 4 * kRowsPerWave - 1) / (4 * kRowsPerWave);
}

// The smallest SEG class of k_ext_lines_seg that holds the longest line along
// `axis` (1 = y, 2 = z) in kLineSegWaves segments, or -1 (k_ext_lines then;
// also for 2^31 faces or more, whose byte offsets exceed 32 bits).
static int line_seg_class(const Faces& F, int axis) {
    if (F.off[3] >= (1u << 31)) return -1;
    uint32_t len = 0;
    for (int c = 0; c < 3; ++c) len = std::max(len, axis == 1 ? F.ey[c] : F.ez[c]);
    for (int k = 0; k < 4; ++k)
        if (len <= uint32_t(kLineSegWaves * kLineSegSizes[k])) return k;
    return -1;
}

// The class of k_ext_lines_seg with kLineShortWaves waves for the longest
// line along `axis`, or -1.
static int line_short_class(const Faces& F, int axis) {
    if (F.off[3] >= (1u << 31)) return -1;
    uint32_t len = 0;
    for (int c = 0; c < 3; ++c) len = std::max(len, axis == 1 ? F.ey[c] : F.ez[c]);
    for (int k = 0; k < 2; ++k)
        if (len <= uint32_t(kLineShortWaves * kLineShortSizes[k])) return k;
    return -1;
}

static uint32_t line_seg_grid(const Faces& F, int axis) {
    const uint32_t* pre = axis == 1 ? F.ylines : F.zlines;
    uint32_t n = 0;
    for (int c = 0; c < 3; ++c) n += (pre[c + 1] - pre[c] + 63) / 64;
    return (n + kXcds - 1) / kXcds * kXcds;
}

hipError_t launch_ext_distance(hipStream_t s, const GridDev& g, const float* u, const float* v, const float* w,
                               uint16_t* dist) {
    const Faces F = faces_of(g, const_cast<float*>(u), const_cast<float*>(v), const_cast<float*>(w));
    // rows: held in registers when they fit 16 chunks of 64 faces
    const uint32_t exm = F.ex[0];   // the x-faces have the longest rows
    const double nf = double(F.off[3]);
    probe_next_bytes(6.0 * nf);   // face value 4 read, distance 2 written
    const dim3 b(kBlock);
    if (exm <= 64) (void)launch_dims(k_ext_rows_short<1>, "k_ext_rows_short", dim3(rows_short_grid(F)), b, 0, s, F, dist);
    else if (exm <= 128) (void)launch_dims(k_ext_rows_short<2>, "k_ext_rows_short", dim3(rows_short_grid(F)), b, 0, s, F, dist);
    else if (exm <= 256) (void)launch_dims(k_ext_rows_reg<4>, "k_ext_rows_reg", dim3(rows_grid(F, 4)), b, 0, s, F, dist);
    else if (exm <= 320) (void)launch_dims(k_ext_rows_reg<5>, "k_ext_rows_reg", dim3(rows_grid(F, 5)), b, 0, s, F, dist);
    else if (exm <= 512) (void)launch_dims(k_ext_rows_reg<8>, "k_ext_rows_reg", dim3(rows_grid(F, 8)), b, 0, s, F, dist);
    else if (exm <= 1024) (void)launch_dims(k_ext_rows_reg<16>, "k_ext_rows_reg", dim3(rows_grid(F, 16)), b, 0, s, F, dist);
    else (void)launch_dims(k_ext_rows, "k_ext_rows", dim3((F.rows[3] + 3) / 4), b, 0, s, F, dist);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int axis = 1; axis <= 2; ++axis) {
        probe_next_bytes(4.0 * nf);   // distance 2 read and written
        const int sc = line_short_class(F, axis);
        if (sc >= 0) {
            const dim3 gr(line_seg_grid(F, axis)), bl(64 * kLineShortWaves);
            switch ((axis - 1) * 2 + sc) {
                case 0: (void)launch_dims(k_ext_lines_seg<1, kLineShortSizes[0], kLineShortWaves>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 1: (void)launch_dims(k_ext_lines_seg<1, kLineShortSizes[1], kLineShortWaves>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 2: (void)launch_dims(k_ext_lines_seg<2, kLineShortSizes[0], kLineShortWaves>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                default: (void)launch_dims(k_ext_lines_seg<2, kLineShortSizes[1], kLineShortWaves>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
            }
            const hipError_t e1 = hipGetLastError();
            if (e1 != hipSuccess) return e1;
            continue;
        }
        const int cls = line_seg_class(F, axis);
        if (cls < 0) {
            const uint32_t n = axis == 1 ? F.ylines[3] : F.zlines[3];
            if (axis == 1) (void)launch_dims(k_ext_lines<1>, "k_ext_lines", dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, F, dist);
            else (void)launch_dims(k_ext_lines<2>, "k_ext_lines", dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, F, dist);
        } else {
            const dim3 gr(line_seg_grid(F, axis)), bl(kLineSegThreads);
            switch ((axis - 1) * 4 + cls) {
                case 0: (void)launch_dims(k_ext_lines_seg<1, kLineSegSizes[0]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 1: (void)launch_dims(k_ext_lines_seg<1, kLineSegSizes[1]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 2: (void)launch_dims(k_ext_lines_seg<1, kLineSegSizes[2]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 3: (void)launch_dims(k_ext_lines_seg<1, kLineSegSizes[3]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 4: (void)launch_dims(k_ext_lines_seg<2, kLineSegSizes[0]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 5: (void)launch_dims(k_ext_lines_seg<2, kLineSegSizes[1]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                case 6: (void)launch_dims(k_ext_lines_seg<2, kLineSegSizes[2]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
                default: (void)launch_dims(k_ext_lines_seg<2, kLineSegSizes[3]>, "k_ext_lines_seg", gr, bl, 0, s, F, dist); break;
            }
        }
        const hipError_t e2 = hipGetLastError();
        if (e2 != hipSuccess) return e2;
    }
    return hipSuccess;
}

size_t ext_counts_size(int32_t K) { return (3 * size_t(kCopies) + 1) * (size_t(K) + 2) + 1; }

hipError_t launch_ext_sort(hipStream_t s, const GridDev& g, const uint16_t* dist, int32_t cap, int32_t K,
                           uint32_t* counts, uint32_t* list, uint8_t* lmask, uint8_t* fmask, const uint32_t** offs,
                           const uint32_t** maxd) {
    const Faces F = faces_of(g, nullptr, nullptr, nullptr);
    const uint32_t total = F.off[3];
    const uint32_t K2 = uint32_t(K) + 2;
    const SortBufs B = sort_bufs(counts, K);
    *offs = B.offs;
    *maxd = B.maxd;
    const uint32_t nh = kCopies * K2;
    (void)launch_dims(k_ext_zero, "k_ext_zero", dim3((nh + kBlock - 1) / kBlock), dim3(kBlock), 0, s, B.hist, nh);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t G = (size_t(total) + size_t(kBlock) * kSortCpt - 1) / (size_t(kBlock) * kSortCpt);
    G = (G + kXcds - 1) / kXcds * kXcds;
    probe_next_bytes(2.0 * double(total));
    (void)launch_dims(k_ext_count, "k_ext_count", dim3(uint32_t(G)), dim3(kBlock), 0, s, dist, total, cap, B.hist, K2);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    (void)launch_dims(k_ext_scan, "k_ext_scan", dim3(1), dim3(1024), 0, s, B.hist, cap, K2, B.offs, B.poff, B.cursor, B.maxd);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    probe_next_bytes(4.0 * double(total));   // distance 2 read, fmask 1 and (listed faces) list 4 + lmask 1 written
    (void)launch_dims(k_ext_scatter, "k_ext_scatter", dim3(uint32_t(G)), dim3(kBlock), 0, s, F, dist, total, cap, K2, B.poff,
                       B.cursor, list, lmask, fmask);
    return hipGetLastError();
}

hipError_t launch_ext_layer(hipStream_t s, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                            const uint8_t* lmask, uint32_t begin, uint32_t count) {
    if (count == 0) return hipSuccess;
    const Faces F = faces_of(g, u, v, w);
    return launch_dims(k_ext_layer, "k_ext_layer", dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0u, s, F, list, lmask, begin,
                       count);
}

hipError_t launch_ext_tail(hipStream_t s, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                           const uint8_t* lmask, const uint32_t* offs, int32_t first, int32_t last,
                           const uint32_t* last_dev) {
    if (!last_dev && last < first) return hipSuccess;
    const Faces F = faces_of(g, u, v, w);
    return launch_dims(k_ext_tail, "k_ext_tail", dim3(1), dim3(kTailThreads), 0u, s, F, list, lmask, offs, first, last,
                       last_dev);
}

hipError_t launch_ext_layer2(hipStream_t s, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                             const uint8_t* lmask, const uint8_t* fmask, uint32_t begin, uint32_t n0, uint32_t n1) {
    if (n0 + n1 == 0) return hipSuccess;
    const Faces F = faces_of(g, u, v, w);
    return launch_dims(k_ext_layer2, "k_ext_layer2", dim3((n0 + n1 + kBlock - 1) / kBlock), dim3(kBlock), 0u, s, F, list, lmask,
                       fmask, begin, n0, n1);
}

// ===========================================================================
// Multi-scene records of the distance transform, the counting sort and the
// one-work-group layers (launch_ext_distance, launch_ext_sort and
// launch_ext_tail with the last layer read on the device): the same
// arguments and grids.
template <class A>
static void push_rec_e(std::vector<MRec>& out, int32_t kind, const A& a, uint32_t blocks) {
    static_assert(sizeof(A) <= kMRecBytes, "argument struct fits a record");
    MRec r;
    r.kind = kind;
    r.blocks = blocks;
    r.bytes = uint32_t(sizeof(A));
    std::memcpy(r.args, &a, sizeof(A));
    out.push_back(r);
}

void rec_ext_device(std::vector<MRec>& out, const GridDev& g, float* u, float* v, float* w, uint16_t* dist,
                    int32_t cap, int32_t K, uint32_t* counts, uint32_t* list, uint8_t* lmask, uint8_t* fmask,
                    const uint32_t** maxd, bool with_layers) {
    const Faces F = faces_of(g, u, v, w);
    const uint32_t exm = F.ex[0];
    const uint32_t rg = (F.rows[3] + 3) / 4;
    if (exm <= 64) push_rec_e(out, kMkExtRowsShort1, MExtRowsShort<1>{F, dist, rows_short_grid(F)}, rows_short_grid(F));
    else if (exm <= 128) push_rec_e(out, kMkExtRowsShort2, MExtRowsShort<2>{F, dist, rows_short_grid(F)}, rows_short_grid(F));
    else if (exm <= 256) push_rec_e(out, kMkExtRowsReg4, MExtRowsReg<4>{F, dist, rows_grid(F, 4)}, rows_grid(F, 4));
    else if (exm <= 320) push_rec_e(out, kMkExtRowsReg5, MExtRowsReg<5>{F, dist, rows_grid(F, 5)}, rows_grid(F, 5));
    else if (exm <= 512) push_rec_e(out, kMkExtRowsReg8, MExtRowsReg<8>{F, dist, rows_grid(F, 8)}, rows_grid(F, 8));
    else if (exm <= 1024) push_rec_e(out, kMkExtRowsReg16, MExtRowsReg<16>{F, dist, rows_grid(F, 16)}, rows_grid(F, 16));
    else push_rec_e(out, kMkExtRows, MExtRows{F, dist, rg}, rg);
    for (int axis = 1; axis <= 2; ++axis) {
        const int sc = line_short_class(F, axis);
        if (sc >= 0) {
            const uint32_t gl = line_seg_grid(F, axis);
            const int32_t kind = kMkExtLinesShort + (axis - 1) * 2 + sc;
            switch ((axis - 1) * 2 + sc) {
                case 0: push_rec_e(out, kind, MExtLinesSeg<1, kLineShortSizes[0], kLineShortWaves>{F, dist, gl}, gl); break;
                case 1: push_rec_e(out, kind, MExtLinesSeg<1, kLineShortSizes[1], kLineShortWaves>{F, dist, gl}, gl); break;
                case 2: push_rec_e(out, kind, MExtLinesSeg<2, kLineShortSizes[0], kLineShortWaves>{F, dist, gl}, gl); break;
                default: push_rec_e(out, kind, MExtLinesSeg<2, kLineShortSizes[1], kLineShortWaves>{F, dist, gl}, gl); break;
            }
            continue;
        }
        const int cls = line_seg_class(F, axis);
        if (cls < 0) {
            const uint32_t gl = ((axis == 1 ? F.ylines[3] : F.zlines[3]) + kBlock - 1) / kBlock;
            if (axis == 1) push_rec_e(out, kMkExtLines1, MExtLines<1>{F, dist, gl}, gl);
            else push_rec_e(out, kMkExtLines2, MExtLines<2>{F, dist, gl}, gl);
            continue;
        }
        const uint32_t gl = line_seg_grid(F, axis);
        const int32_t kind = kMkExtLinesSeg + (axis - 1) * 4 + cls;
        switch ((axis - 1) * 4 + cls) {
            case 0: push_rec_e(out, kind, MExtLinesSeg<1, kLineSegSizes[0]>{F, dist, gl}, gl); break;
            case 1: push_rec_e(out, kind, MExtLinesSeg<1, kLineSegSizes[1]>{F, dist, gl}, gl); break;
            case 2: push_rec_e(out, kind, MExtLinesSeg<1, kLineSegSizes[2]>{F, dist, gl}, gl); break;
            case 3: push_rec_e(out, kind, MExtLinesSeg<1, kLineSegSizes[3]>{F, dist, gl}, gl); break;
            case 4: push_rec_e(out, kind, MExtLinesSeg<2, kLineSegSizes[0]>{F, dist, gl}, gl); break;
            case 5: push_rec_e(out, kind, MExtLinesSeg<2, kLineSegSizes[1]>{F, dist, gl}, gl); break;
            case 6: push_rec_e(out, kind, MExtLinesSeg<2, kLineSegSizes[2]>{F, dist, gl}, gl); break;
            default: push_rec_e(out, kind, MExtLinesSeg<2, kLineSegSizes[3]>{F, dist, gl}, gl); break;
        }
    }
    const uint32_t total = F.off[3];
    const uint32_t K2 = uint32_t(K) + 2;
    const SortBufs B = sort_bufs(counts, K);
    *maxd = B.maxd;
    const uint32_t nh = kCopies * K2;
    const uint32_t gz = (nh + kBlock - 1) / kBlock;
    push_rec_e(out, kMkExtZero, MExtZero{B.hist, nh, gz}, gz);
    size_t G = (size_t(total) + size_t(kBlock) * kSortCpt - 1) / (size_t(kBlock) * kSortCpt);
    G = (G + kXcds - 1) / kXcds * kXcds;
    push_rec_e(out, kMkExtCount, MExtCount{dist, total, cap, B.hist, K2, uint32_t(G)}, uint32_t(G));
    push_rec_e(out, kMkExtScan, MExtScan{B.hist, cap, K2, B.offs, B.poff, B.cursor, B.maxd, 1u}, 1u);
    push_rec_e(out, kMkExtScatter,
               MExtScatter{F, dist, total, cap, K2, B.poff, B.cursor, list, lmask, fmask, uint32_t(G)},
               uint32_t(G));
    if (with_layers) push_rec_e(out, kMkExtTail, MExtTail{F, list, lmask, B.offs, 1, 0, B.maxd, 1u}, 1u);
}

// Layer L of many scenes in one launch: block b of a scene takes entries
// offs[L] + b * kBlock * kLayerCpt ... of its list (the bounds read on the
// device), each the mean of its flagged previous-layer neighbours (as
// k_ext_layer).  map entries: (scene << kMapShift) | block.
constexpr int kLayerCpt = 4;

struct MLayer {
    Faces F;
    const uint32_t* list;
    const uint8_t* lmask;
    const uint32_t* offs;
};

__global__ void __launch_bounds__(kBlock) k_multi_layer(const MLayer* __restrict__ tab,
                                                        const uint32_t* __restrict__ map, int32_t L) {
    const uint32_t e = map[blockIdx.x];
    const MLayer a = tab[e >> kMapShift];
    const uint32_t b = a.offs[L], end = a.offs[L + 1];
    const uint32_t base = b + (e & kMapLocal) * uint32_t(kBlock * kLayerCpt) + threadIdx.x;
    uint32_t gi[kLayerCpt], m[kLayerCpt];
#pragma unroll
    for (int k = 0; k < kLayerCpt; ++k) {
        const uint32_t t = base + uint32_t(k * kBlock);
        const uint32_t tc = t < end ? t : end - 1;
        gi[k] = a.list[tc];
        m[k] = t < end ? uint32_t(a.lmask[tc]) : 0u;
    }
    // every entry's loads, then the sums and stores
    float fv[kLayerCpt][6];
#pragma unroll
    for (int k = 0; k < kLayerCpt; ++k) {
        const int c = comp_of(a.F.off, gi[k]);
        const uint32_t i = gi[k] - pick4(a.F.off, c);
        const uint32_t ex = pick(a.F.ex, c), sz = ex * pick(a.F.ey, c);
        layer_loads(pick(a.F.f, c), i, m[k], ex, sz, fv[k]);   // m = 0 past the end: nothing loaded
    }
#pragma unroll
    for (int k = 0; k < kLayerCpt; ++k) loads_done(fv[k]);
#pragma unroll
    for (int k = 0; k < kLayerCpt; ++k) {
        const uint32_t t = base + uint32_t(k * kBlock);
        if (t >= end) continue;
        const int c = comp_of(a.F.off, gi[k]);
        pick(a.F.f, c)[gi[k] - pick4(a.F.off, c)] = layer_sum(m[k], fv[k]);
    }
}

uint32_t layer_entries_per_block() { return uint32_t(kBlock * kLayerCpt); }

size_t layer_arg_bytes() { return sizeof(MLayer); }

void layer_args(unsigned char* dst, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
                const uint8_t* lmask, uint32_t* counts, int32_t K) {
    const MLayer a{faces_of(g, u, v, w), list, lmask, sort_bufs(counts, K).offs};
    std::memcpy(dst, &a, sizeof(MLayer));
}

hipError_t launch_multi_layer(hipStream_t s, const void* tab, const uint32_t* map, uint32_t nblocks, int32_t L) {
    if (nblocks == 0) return hipSuccess;
    return launch_dims(k_multi_layer, "k_multi_layer", dim3(nblocks), dim3(kBlock), 0u, s,
                       static_cast<const MLayer*>(tab), map, L);
}

// Each scene's layer offsets 0..cap+1 and its largest layer, packed for one
// copy to the host (one block per scene).
struct MCopyOffs {
    static constexpr int kThreads = kBlock;
    const uint32_t* offs;
    const uint32_t* maxd;
    uint32_t n;
    uint32_t* dst;
    __device__ void run(uint32_t) const {
        for (uint32_t i = threadIdx.x; i < n; i += kBlock) dst[i] = offs[i];
        if (threadIdx.x == 0) dst[n] = *maxd;
    }
};

void rec_copy_offs(std::vector<MRec>& out, uint32_t* counts, int32_t K, int32_t cap, uint32_t* dst) {
    const SortBufs B = sort_bufs(counts, K);
    push_rec_e(out, kMkCopyOffs, MCopyOffs{B.offs, B.maxd, uint32_t(cap) + 2u, dst}, 1u);
}

// The sort's layers from `first` to the largest present in one work-group
// (the layers a scene's own launches do not take).
size_t tail_arg_bytes() { return sizeof(MExtTail); }

void tail_args(unsigned char* dst, const GridDev& g, float* u, float* v, float* w, const uint32_t* list,
               const uint8_t* lmask, uint32_t* counts, int32_t K, int32_t first, int32_t last) {
    const SortBufs B = sort_bufs(counts, K);
    const MExtTail a{faces_of(g, u, v, w), list, lmask, B.offs, first, last, nullptr, 1u};
    std::memcpy(dst, &a, sizeof(MExtTail));
}

template <class A>
static hipError_t launch_me(hipStream_t s, const char* name, const void* tab, const uint32_t* map, uint32_t nblocks,
                            const uint8_t* active) {
    return launch_dims(k_multi<A>, name, dim3(nblocks), dim3(A::kThreads), 0u, s, static_cast<const A*>(tab), map,
                       active);
}

hipError_t launch_multi_ext(hipStream_t s, int32_t kind, const void* tab, const uint32_t* map, uint32_t nblocks,
                            uint32_t, const uint8_t* active) {
    switch (kind) {
        case kMkExtRows: return launch_me<MExtRows>(s, "k_multi_ext_rows", tab, map, nblocks, active);
        case kMkExtLinesShort + 0: return launch_me<MExtLinesSeg<1, kLineShortSizes[0], kLineShortWaves>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesShort + 1: return launch_me<MExtLinesSeg<1, kLineShortSizes[1], kLineShortWaves>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesShort + 2: return launch_me<MExtLinesSeg<2, kLineShortSizes[0], kLineShortWaves>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesShort + 3: return launch_me<MExtLinesSeg<2, kLineShortSizes[1], kLineShortWaves>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtRowsShort1: return launch_me<MExtRowsShort<1>>(s, "k_multi_ext_rows_short", tab, map, nblocks, active);
        case kMkExtRowsShort2: return launch_me<MExtRowsShort<2>>(s, "k_multi_ext_rows_short", tab, map, nblocks, active);
        case kMkExtRowsReg4: return launch_me<MExtRowsReg<4>>(s, "k_multi_ext_rows_reg", tab, map, nblocks, active);
        case kMkExtRowsReg5: return launch_me<MExtRowsReg<5>>(s, "k_multi_ext_rows_reg", tab, map, nblocks, active);
        case kMkExtRowsReg8: return launch_me<MExtRowsReg<8>>(s, "k_multi_ext_rows_reg", tab, map, nblocks, active);
        case kMkExtRowsReg16: return launch_me<MExtRowsReg<16>>(s, "k_multi_ext_rows_reg", tab, map, nblocks, active);
        case kMkExtLines1: return launch_me<MExtLines<1>>(s, "k_multi_ext_lines", tab, map, nblocks, active);
        case kMkExtLines2: return launch_me<MExtLines<2>>(s, "k_multi_ext_lines", tab, map, nblocks, active);
        case kMkExtLinesSeg + 0: return launch_me<MExtLinesSeg<1, kLineSegSizes[0]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesSeg + 1: return launch_me<MExtLinesSeg<1, kLineSegSizes[1]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesSeg + 2: return launch_me<MExtLinesSeg<1, kLineSegSizes[2]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesSeg + 3: return launch_me<MExtLinesSeg<1, kLineSegSizes[3]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesSeg + 4: return launch_me<MExtLinesSeg<2, kLineSegSizes[0]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesSeg + 5: return launch_me<MExtLinesSeg<2, kLineSegSizes[1]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesSeg + 6: return launch_me<MExtLinesSeg<2, kLineSegSizes[2]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtLinesSeg + 7: return launch_me<MExtLinesSeg<2, kLineSegSizes[3]>>(s, "k_multi_ext_lines_seg", tab, map, nblocks, active);
        case kMkExtZero: return launch_me<MExtZero>(s, "k_multi_ext_zero", tab, map, nblocks, active);
        case kMkExtCount: return launch_me<MExtCount>(s, "k_multi_ext_count", tab, map, nblocks, active);
        case kMkExtScan: return launch_me<MExtScan>(s, "k_multi_ext_scan", tab, map, nblocks, active);
        case kMkExtScatter: return launch_me<MExtScatter>(s, "k_multi_ext_scatter", tab, map, nblocks, active);
        case kMkExtTail: return launch_me<MExtTail>(s, "k_multi_ext_tail", tab, map, nblocks, active);
        case kMkCopyOffs: return launch_me<MCopyOffs>(s, "k_multi_copy_offs", tab, map, nblocks, active);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mgf
