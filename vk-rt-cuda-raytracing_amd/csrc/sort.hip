// sort.hip — ray sorting for incoherent traces (cuda_rt_v29.cu:84-139,
// cuda_rt_v37.cu:101-128): key = direction octant << 29 | 29 bits of the
// Morton code of the origin (normalised by the scene box), a stable LSD radix
// sort of (key, ray index) pairs on the GPU, gather of the rays into sorted
// order and scatter of the hits back. The reference sorts on the host
// (backend :1387-1433) or with CUB; this is an own 4 x 8-bit LSD sort:
//
//  * histogram pass: per 2048-key tile, a 256-bin LDS histogram, written
//    digit-major so that one exclusive scan gives every (digit, tile) its
//    global offset;
//  * scatter pass: keys are ranked in tile order, 256 per round; within a
//    wave the lanes holding the same digit find each other with eight
//    ballots (one per digit bit), and per-digit counts in LDS order the four
//    waves and the rounds, so equal keys keep their input order (stable).
#include <hip/hip_runtime.h>

#include "rt_internal.h"
#include "rt_device.h"
#include "trace_launch.h"

namespace rts2 {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kRadix = 256;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Morton code of a point normalised to [0,1]^3 (backend :265-273).
__device__ __forceinline__ uint32_t morton3d(float x, float y, float z) {
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    return spread10((uint32_t)x) * 4u + spread10((uint32_t)y) * 2u + spread10((uint32_t)z);
}

__global__ void k_keys(CudaRayPacket r, int n, float lx, float ly, float lz, float ix, float iy, float iz,
                       uint32_t* keys, int* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float nx = (r.ox[i] - lx) * ix, ny = (r.oy[i] - ly) * iy, nz = (r.oz[i] - lz) * iz;
    const uint32_t oct = (r.dx[i] < 0.f ? 4u : 0u) | (r.dy[i] < 0.f ? 2u : 0u) | (r.dz[i] < 0.f ? 1u : 0u);
    keys[i] = (oct << 29) | (morton3d(nx, ny, nz) & 0x1FFFFFFFu);
    vals[i] = i;
}

__global__ __launch_bounds__(kThreads) void k_hist(const uint32_t* keys, int n, int shift, int tiles, int* hist) {
    __shared__ int h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int k = 0; k < kItems; k++) {
        const int i = base + k * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];     // digit-major
}

__global__ __launch_bounds__(kThreads) void k_scatter(const uint32_t* keys, const int* vals, int n, int shift,
                                                       int tiles, const int* offs, uint32_t* keys2, int* vals2) {
    constexpr int NW = kThreads / 64;
    __shared__ int running[kRadix];
    __shared__ int wcnt[NW][kRadix];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    running[threadIdx.x] = offs[threadIdx.x * tiles + blockIdx.x];
    for (int j = 0; j < NW; j++) wcnt[j][threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int k = 0; k < kItems; k++) {
        const int i = base + k * kThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t d = (key >> shift) & 0xffu;
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t m = __ballot(valid && ((d >> bit) & 1u));
            peers &= ((d >> bit) & 1u) ? m : ~m;
        }
        const int before = __popcll(peers & ((1ull << lane) - 1ull));
        const bool leader = valid && before == 0;
        if (leader) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = running[d] + before;
            for (int j = 0; j < w; j++) pos += wcnt[j][d];
            keys2[pos] = key;
            vals2[pos] = vals[i];
        }
        __syncthreads();
        int add = 0;
        for (int j = 0; j < NW; j++) { add += wcnt[j][threadIdx.x]; wcnt[j][threadIdx.x] = 0; }
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

__global__ void k_gather(CudaRayPacket in, const int* perm, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    const size_t N = (size_t)n;
    out[i] = in.ox[j]; out[N + i] = in.oy[j]; out[2 * N + i] = in.oz[j];
    out[3 * N + i] = in.dx[j]; out[4 * N + i] = in.dy[j]; out[5 * N + i] = in.dz[j];
    out[6 * N + i] = in.tmin[j]; out[7 * N + i] = in.tmax[j];
}

__global__ void k_scatter_hits(const float* t, const int* id, const float* u, const float* v, const uint8_t* occ,
                               const int* perm, int n, CudaHitPacket hits, uint8_t* occOut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    if (occOut) { occOut[j] = occ[i]; return; }
    hits.t[j] = t[i];
    if (hits.primId) hits.primId[j] = id[i];
    if (hits.u) hits.u[j] = u[i];
    if (hits.v) hits.v[j] = v[i];
}

}  // namespace rts2

namespace rtk {

// Sorts rays by (octant, Morton of origin in the scene box), traces the sorted
// copy and writes hits back in the caller's order. Scratch comes from the
// handle's work buffer (ws, wsBytes must cover sort_scratch_bytes(n)).
size_t sort_scratch_bytes(int n) {
    const size_t N = (size_t)n;
    const size_t tiles = (N + rts2::kTile - 1) / rts2::kTile;
    return N * (4 * 4)                       // keys x2, vals x2
           + rts2::kRadix * tiles * 4 * 2    // histogram + offsets
           + ((rts2::kRadix * tiles + 2047) / 2048 + 16) * 4
           + N * 8 * 4                       // sorted rays
           + N * 4 * 4 + N;                  // sorted hits, occlusion
}

int trace_sorted(CudaBVHHandle* h, rt::TraceScratch& sc, int n, const CudaRayPacket& rays, const CudaHitPacket* hits,
                 uint8_t* occ, int mode, bool mt, void* ws, hipStream_t s) {
    const size_t N = (size_t)n;
    const int tiles = (int)((N + rts2::kTile - 1) / rts2::kTile);
    char* p = (char*)ws;
    uint32_t* k0 = (uint32_t*)p; p += N * 4;
    uint32_t* k1 = (uint32_t*)p; p += N * 4;
    int* v0 = (int*)p; p += N * 4;
    int* v1 = (int*)p; p += N * 4;
    int* hist = (int*)p; p += (size_t)rts2::kRadix * tiles * 4;
    int* offs = (int*)p; p += (size_t)rts2::kRadix * tiles * 4;
    int* stiles = (int*)p; p += (size_t)(((size_t)rts2::kRadix * tiles + 2047) / 2048 + 16) * 4;
    float* sr = (float*)p; p += N * 8 * 4;
    float* st = (float*)p; p += N * 4;
    int* sid = (int*)p; p += N * 4;
    float* su = (float*)p; p += N * 4;
    float* sv = (float*)p; p += N * 4;
    uint8_t* socc = (uint8_t*)p;
    float lo[3], inv[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = h->bboxMin[a];
        const float e = h->bboxMax[a] - h->bboxMin[a];
        inv[a] = e > 1e-8f ? 1.0f / e : 0.0f;
    }
    const dim3 g((unsigned)((N + 255) / 256)), b(256);
    hipLaunchKernelGGL(rts2::k_keys, g, b, 0, s, rays, n, lo[0], lo[1], lo[2], inv[0], inv[1], inv[2], k0, v0);
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(rts2::k_hist, dim3(tiles), dim3(rts2::kThreads), 0, s, k0, n, shift, tiles, hist);
        rt::scan_exclusive(hist, offs, rts2::kRadix * tiles, stiles + 8, stiles + 16, s);
        hipLaunchKernelGGL(rts2::k_scatter, dim3(tiles), dim3(rts2::kThreads), 0, s, k0, v0, n, shift, tiles, offs,
                           k1, v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    hipLaunchKernelGGL(rts2::k_gather, g, b, 0, s, rays, v0, n, sr);
    TraceArgs a;
    a.n = n;
    a.ox = sr; a.oy = sr + N; a.oz = sr + 2 * N; a.dx = sr + 3 * N; a.dy = sr + 4 * N; a.dz = sr + 5 * N;
    a.tmin = sr + 6 * N; a.tmax = sr + 7 * N;
    if (mode == kModeAny) a.occ = socc;
    else { a.t = st; a.id = sid; a.u = su; a.v = sv; }
    int rc = launch_trace(h, sc, a, mode, mt, false, false, s);
    if (rc != 0) return rc;
    CudaHitPacket hp{};
    if (hits) hp = *hits;
    hipLaunchKernelGGL(rts2::k_scatter_hits, g, b, 0, s, st, sid, su, sv, socc, v0, n, hp,
                       mode == kModeAny ? occ : (uint8_t*)nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("sorted trace launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

// Sort only (for tests and measurement): keys and the permutation in sorted order.
size_t radix_pairs_bytes(int n) {
    const size_t N = (size_t)n;
    const size_t tiles = (N + rts2::kTile - 1) / rts2::kTile;
    return N * 8 + rts2::kRadix * tiles * 4 * 2 + ((rts2::kRadix * tiles + 2047) / 2048 + 16) * 4 + 256;
}

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
// keys (8 bits per pass), in place in (keys, vals); ws holds
// radix_pairs_bytes(n) bytes.
int radix_sort_pairs(uint32_t* keys, int* vals, int n, int bits, void* ws, hipStream_t s) {
    const size_t N = (size_t)n;
    const int tiles = (int)((N + rts2::kTile - 1) / rts2::kTile);
    char* p = (char*)ws;
    uint32_t* k1 = (uint32_t*)p; p += N * 4;
    int* v1 = (int*)p; p += N * 4;
    int* hist = (int*)p; p += (size_t)rts2::kRadix * tiles * 4;
    int* offs = (int*)p; p += (size_t)rts2::kRadix * tiles * 4;
    int* stiles = (int*)p;
    uint32_t* k0 = keys;
    int* v0 = vals;
    const int passes = (bits + 7) / 8;
    for (int pass = 0; pass < passes; pass++) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(rts2::k_hist, dim3(tiles), dim3(rts2::kThreads), 0, s, k0, n, shift, tiles, hist);
        rt::scan_exclusive(hist, offs, rts2::kRadix * tiles, stiles + 8, stiles + 16, s);
        hipLaunchKernelGGL(rts2::k_scatter, dim3(tiles), dim3(rts2::kThreads), 0, s, k0, v0, n, shift, tiles, offs,
                           k1, v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    if (k0 != keys) {
        hipMemcpyAsync(keys, k0, N * 4, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(vals, v0, N * 4, hipMemcpyDeviceToDevice, s);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("radix sort launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

int sort_rays(CudaBVHHandle* h, int n, const CudaRayPacket& rays, uint32_t* keysOut, int* permOut, void* ws,
              hipStream_t s) {
    const size_t N = (size_t)n;
    const int tiles = (int)((N + rts2::kTile - 1) / rts2::kTile);
    char* p = (char*)ws;
    uint32_t* k0 = (uint32_t*)p; p += N * 4;
    uint32_t* k1 = (uint32_t*)p; p += N * 4;
    int* v0 = (int*)p; p += N * 4;
    int* v1 = (int*)p; p += N * 4;
    int* hist = (int*)p; p += (size_t)rts2::kRadix * tiles * 4;
    int* offs = (int*)p; p += (size_t)rts2::kRadix * tiles * 4;
    int* stiles = (int*)p;
    float lo[3], inv[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = h->bboxMin[a];
        const float e = h->bboxMax[a] - h->bboxMin[a];
        inv[a] = e > 1e-8f ? 1.0f / e : 0.0f;
    }
    const dim3 g((unsigned)((N + 255) / 256)), b(256);
    hipLaunchKernelGGL(rts2::k_keys, g, b, 0, s, rays, n, lo[0], lo[1], lo[2], inv[0], inv[1], inv[2], k0, v0);
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(rts2::k_hist, dim3(tiles), dim3(rts2::kThreads), 0, s, k0, n, shift, tiles, hist);
        rt::scan_exclusive(hist, offs, rts2::kRadix * tiles, stiles + 8, stiles + 16, s);
        hipLaunchKernelGGL(rts2::k_scatter, dim3(tiles), dim3(rts2::kThreads), 0, s, k0, v0, n, shift, tiles, offs,
                           k1, v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    hipMemcpyAsync(keysOut, k0, N * 4, hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(permOut, v0, N * 4, hipMemcpyDeviceToDevice, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("sort launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

}  // namespace rtk
