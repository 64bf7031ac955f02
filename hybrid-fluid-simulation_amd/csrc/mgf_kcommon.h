// mgf_kcommon.h — helpers shared by the kernel translation units: XCD-aware
// block numbering, FP64 wave/block reductions, launch wrapper.
#pragma once

#include <hip/hip_ext.h>

#include "mgf_internal.h"

namespace mgf {

constexpr uint32_t kXcds = 8;

// Workgroups are dispatched round-robin over the XCDs: physical block b runs on
// XCD b % 8.  Map it to logical block (b % 8) * (G / 8) + b / 8 so that XCD k
// owns logical blocks [k G/8, (k+1) G/8).  Requires G % 8 == 0.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t b = blockIdx.x, G = gridDim.x;
    return (b % kXcds) * (G / kXcds) + b / kXcds;
}

// The same remapping for block b of a launch of G blocks (multi-scene
// launches apply it to a block's index within its scene's own launch).
__device__ __forceinline__ uint32_t xcd_of(uint32_t b, uint32_t G) {
    return (b % kXcds) * (G / kXcds) + b / kXcds;
}

// Word 3 of a raw buffer resource (__builtin_amdgcn_make_buffer_rsrc) on
// gfx9: 32-bit data format, no swizzle.
constexpr int32_t kBufWord3 = 0x00020000;

inline uint32_t grid_for(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    g = (g + kXcds - 1) / kXcds * kXcds;
    if (g == 0) g = kXcds;
    return uint32_t(g);
}

// ---------------------------------------------------------------------------
// Wave64 / block reductions in FP64.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums `n` doubles of a block (blockDim.x == kBlock); result valid in thread 0.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
    __shared__ double sh[N][kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < kBlock / 64; ++q) s += sh[k][q];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Every launch goes through launch_dims: with a probe sink installed for this
// host thread (ProbeSink, mgf_internal.h) the kernel records its start/stop
// events under `name`.
template <class K, class... A>
inline hipError_t launch_dims(K kern, const char* name, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s,
                              A... args) {
    const double bytes = probe_take_bytes();
    if (ProbeSink* ps = probe_sink()) {
        hipEvent_t a = nullptr, b = nullptr;
        if (ps->next_pair(a, b)) {
            hipExtLaunchKernelGGL(kern, grid, block, shmem, s, a, b, 0u, args...);
            ps->name.push_back(name);
            ps->bytes.push_back(bytes);
            ps->used += 1;
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL(kern, grid, block, shmem, s, args...);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Multi-scene launches.  One launch covers the blocks of many scenes' own
// launches of one kernel: block b reads map[b] = (scene << kMapShift) | its
// index within that scene's launch, skips scenes whose active byte is 0, and
// runs the kernel body with that scene's arguments from a device table
// (A::run).  The body sees the block index, and applies the XCD remapping,
// of the scene's own launch, so each scene computes bitwise what it computes
// alone, partial sums included.

// An argument struct that indexes its own arrays at run time declares
// kByRef: k_multi then runs it in place in the table (a copy would sit in
// scratch memory).
template <class A>
constexpr auto multi_by_ref(int) -> decltype(A::kByRef, bool()) { return A::kByRef; }
template <class A>
constexpr bool multi_by_ref(long) { return false; }

template <class A>
__global__ void __launch_bounds__(A::kThreads) k_multi(const A* __restrict__ tab, const uint32_t* __restrict__ map,
                                                       const uint8_t* __restrict__ active) {
    const uint32_t e = map[blockIdx.x];
    const uint32_t s = e >> kMapShift;
    if (active && !active[s]) return;
    if constexpr (multi_by_ref<A>(0)) {
        tab[s].run(e & kMapLocal);   // arrays indexed at run time stay in the table
    } else {
        const A a = tab[s];
        a.run(e & kMapLocal);
    }
}

template <class K, class... A>
inline hipError_t launch_k(K kern, const char* name, uint32_t grid, hipStream_t s, A... args) {
    return launch_dims(kern, name, dim3(grid), dim3(kBlock), 0u, s, args...);
}
#define LAUNCH(kern, grid, ...) launch_k(kern, #kern, uint32_t(grid), s, __VA_ARGS__)
// with an explicit kernel name (templates: one name for every instantiation)
#define LAUNCHN(kern, name, grid, ...) launch_k(kern, name, uint32_t(grid), s, __VA_ARGS__)

}  // namespace mgf
