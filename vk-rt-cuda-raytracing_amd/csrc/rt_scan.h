// rt_scan.h — wave64 / workgroup scan primitives (ballot + lane shuffles),
// the MI355X counterpart of the reference's warp32 ballot/popc compaction
// (warp_compact.cuh:23-80).
#pragma once

#include <hip/hip_runtime.h>

namespace rts {

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive prefix sum over a workgroup of BLOCK threads (BLOCK a multiple of
// 64, at most 1024). lds must hold BLOCK/64 + 1 ints. Returns the thread's
// exclusive prefix; *total receives the workgroup sum.
template <int BLOCK>
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int* total) {
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    if (w == 0) {
        int s = lane < NW ? lds[lane] : 0;
        const int si = wave_inclusive_scan(s);
        if (lane < NW) lds[lane] = si - s;
        if (lane == NW - 1) lds[NW] = si;
    }
    __syncthreads();
    const int r = lds[w] + inc - v;
    *total = lds[NW];
    __syncthreads();
    return r;
}

}  // namespace rts
