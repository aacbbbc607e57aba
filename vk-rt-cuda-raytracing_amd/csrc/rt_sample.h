// rt_sample.h — random numbers and hemisphere sampling shared by the bounce
// generator and the path tracer. Only IEEE +,-,*,/,sqrt and integer ops, so
// a host restatement reproduces every sample bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtd {

// PCG32 (O'Neill 2014, pcg32_random_r / pcg32_srandom_r).
struct Pcg32 {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    __device__ __forceinline__ void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    // 24 random bits -> [0, 1)
    __device__ __forceinline__ float uniform() { return (float)(next() >> 8) * 0x1p-24f; }
};

// Unit normal (nx, ny, nz) -> cosine-weighted direction about it: Malley's
// method (a uniform point of the unit disk by bounded rejection, lifted to
// the hemisphere) in the branch-free orthonormal basis of Duff et al. 2017,
// renormalised.
__device__ __forceinline__ void cosine_dir(float nx, float ny, float nz, Pcg32& rng, float& dx, float& dy,
                                           float& dz) {
    float x = 0.0f, y = 0.0f, s = 0.0f;
    bool ok = false;
    for (int k = 0; k < 64 && !ok; k++) {
        x = 2.0f * rng.uniform() - 1.0f;
        y = 2.0f * rng.uniform() - 1.0f;
        s = x * x + y * y;
        ok = s < 1.0f;
    }
    if (!ok) { x = 0.0f; y = 0.0f; s = 0.0f; }
    const float z = sqrtf(1.0f - s);
    const float sg = copysignf(1.0f, nz);
    const float ca = -1.0f / (sg + nz);
    const float cb = nx * ny * ca;
    const float t1x = 1.0f + sg * nx * nx * ca, t1y = sg * cb, t1z = -sg * nx;
    const float t2x = cb, t2y = sg + ny * ny * ca, t2z = -ny;
    dx = (x * t1x + y * t2x) + z * nx;
    dy = (x * t1y + y * t2y) + z * ny;
    dz = (x * t1z + y * t2z) + z * nz;
    const float dl = 1.0f / sqrtf((dx * dx + dy * dy) + dz * dz);
    dx = dx * dl; dy = dy * dl; dz = dz * dl;
}

}  // namespace rtd
