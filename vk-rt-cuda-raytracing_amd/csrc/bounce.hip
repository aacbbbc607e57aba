// bounce.hip — secondary rays from primary hits: one cosine-weighted diffuse
// bounce and one shadow ray toward a point light per hit pixel, compacted
// (stable) with wave64 scans in pixel order or, on request, in the tile walk
// of rtd::tile_strip_pixel (8x8 pixel tiles in 64-pixel strips), which keeps
// the rays that a trace takes together in a compact image block.
//
// Three launches, no host round trip: count hits per 2048-slot tile,
// exclusive-scan the tile counts (which also writes the total to device
// memory, where the queued traces read it), then emit rays at their compacted
// positions. The diffuse direction follows the reference's diffuse-ray
// generator (cuda_rt_v37.cu:514-540: per-ray seeded RNG, hemisphere around
// the surface normal) with PCG32 and Malley's method (rt_sample.h). Only IEEE
// +,-,*,/,sqrt are used, so a host restatement reproduces every ray bit for
// bit.
#include "rt_device.h"
#include "rt_sample.h"
#include "rt_scan.h"
#include "trace_launch.h"

namespace rtk {

constexpr int kTileThreads = 256;
constexpr int kTileItems = 8;
constexpr int kTile = kTileThreads * kTileItems;

struct BounceArgs {
    rtd::Cam cam;
    int W, H, n;
    const float* hitT;
    const int* hitId;
    const int* leafOfPrim;
    const float4* tris;
    uint32_t seed;
    float lx, ly, lz;
    float *dox, *doy, *doz, *ddx, *ddy, *ddz, *dtmin, *dtmax;
    float *sox, *soy, *soz, *sdx, *sdy, *sdz, *stmin, *stmax;
    int* src;
    int* count;
    int* tileCounts;
    int tileOrder;          // walk pixels in rtd::tile_strip_pixel order instead of row-major
};

// Pixel of compaction slot j.
__device__ __forceinline__ int bounce_pixel(const BounceArgs& a, int j) {
    return a.tileOrder ? rtd::tile_strip_pixel(j, a.W, a.H) : j;
}

// Slot j of tile b, step k: b*kTile + k*kTileThreads + threadIdx.x, so the
// lanes of a wave handle consecutive slots (consecutive pixels in row-major
// order, 8x8 pixel tiles in tile order).
__global__ __launch_bounds__(kTileThreads) void k_count_hits(BounceArgs a) {
    __shared__ int lds[kTileThreads / 64 + 1];
    const int base = blockIdx.x * kTile + threadIdx.x;
    int c = 0;
    for (int k = 0; k < kTileItems; k++) {
        const int j = base + k * kTileThreads;
        if (j < a.n && a.hitId[bounce_pixel(a, j)] >= 0) c++;
    }
    int total;
    rts::block_exclusive_scan<kTileThreads>(c, lds, &total);
    if (threadIdx.x == 0) a.tileCounts[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(int* counts, int nb, int* total) {
    __shared__ int lds[1024 / 64 + 1];
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? counts[i] : 0;
        int tot;
        const int ex = rts::block_exclusive_scan<1024>(v, lds, &tot);
        if (i < nb) counts[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__device__ __forceinline__ void emit(const BounceArgs& a, int i, int pos) {
    const int py = i / a.W, px = i - py * a.W;
    rtd::Ray r;
    rtd::pinhole_ray(a.cam, px, py, a.W, a.H, r);
    const float t = a.hitT[i];
    const int li = a.leafOfPrim[a.hitId[i]];
    const float4 p0 = a.tris[3 * li], p1 = a.tris[3 * li + 1], p2 = a.tris[3 * li + 2];
    const float hx = r.ox + t * r.dx, hy = r.oy + t * r.dy, hz = r.oz + t * r.dz;
    const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
    const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
    float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const float nl = 1.0f / sqrtf((nx * nx + ny * ny) + nz * nz);
    nx = nx * nl; ny = ny * nl; nz = nz * nl;
    if ((nx * r.dx + ny * r.dy) + nz * r.dz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }

    rtd::Pcg32 rng;
    rng.seed((uint64_t)a.seed, (uint64_t)i);
    float dx, dy, dz;
    rtd::cosine_dir(nx, ny, nz, rng, dx, dy, dz);
    a.dox[pos] = hx; a.doy[pos] = hy; a.doz[pos] = hz;
    a.ddx[pos] = dx; a.ddy[pos] = dy; a.ddz[pos] = dz;
    a.dtmin[pos] = 1e-3f; a.dtmax[pos] = RT_MISS_T;

    const float lx = a.lx - hx, ly = a.ly - hy, lz = a.lz - hz;
    const float dist = sqrtf((lx * lx + ly * ly) + lz * lz);
    const float il = 1.0f / dist;
    a.sox[pos] = hx; a.soy[pos] = hy; a.soz[pos] = hz;
    a.sdx[pos] = lx * il; a.sdy[pos] = ly * il; a.sdz[pos] = lz * il;
    a.stmin[pos] = 1e-3f; a.stmax[pos] = dist;
    a.src[pos] = i;
}

// Stable compaction: per step, a ballot gives each hit lane its rank within
// the wave, and the wave totals (in LDS) order the waves.
__global__ __launch_bounds__(kTileThreads) void k_emit(BounceArgs a) {
    constexpr int NW = kTileThreads / 64;
    __shared__ int waveCount[NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int running = a.tileCounts[blockIdx.x];
    for (int k = 0; k < kTileItems; k++) {
        const int j = blockIdx.x * kTile + k * kTileThreads + threadIdx.x;
        const int i = j < a.n ? bounce_pixel(a, j) : 0;
        const bool hit = j < a.n && a.hitId[i] >= 0;
        const uint64_t m = __ballot(hit);
        if (lane == 0) waveCount[w] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int j = 0; j < NW; j++) {
            const int cj = waveCount[j];
            before += j < w ? cj : 0;
            total += cj;
        }
        if (hit) emit(a, i, running + before + __popcll(m & ((1ull << lane) - 1ull)));
        running += total;
        __syncthreads();
    }
}

int launch_bounce(const CudaBVHHandle* h, rt::TraceScratch& sc, const rtd::Cam& cam, int W, int H,
                  const float* hitT, const int* hitId, uint32_t seed, const float light[3], const CudaRayPacket& d,
                  const CudaRayPacket& s, int* src, int* count, bool tileOrder, hipStream_t st) {
    const int n = W * H;
    const int tiles = (n + kTile - 1) / kTile;
    if (alloc_scratch(sc) != 0) return -1;
    if ((size_t)tiles > sc.tileCap) {
        hipFree(sc.tiles);
        sc.tiles = nullptr; sc.tileCap = 0;
        sc.gen++;
        if (hipMalloc((void**)&sc.tiles, (size_t)tiles * sizeof(int)) != hipSuccess) {
            rt::set_error("cannot allocate %d tile counters", tiles);
            return -1;
        }
        sc.tileCap = (size_t)tiles;
    }
    BounceArgs a;
    a.cam = cam; a.W = W; a.H = H; a.n = n;
    a.hitT = hitT; a.hitId = hitId;
    a.leafOfPrim = h->d_leafOfPrim;
    a.tris = (const float4*)h->d_tris;
    a.seed = seed; a.lx = light[0]; a.ly = light[1]; a.lz = light[2];
    a.dox = (float*)d.ox; a.doy = (float*)d.oy; a.doz = (float*)d.oz;
    a.ddx = (float*)d.dx; a.ddy = (float*)d.dy; a.ddz = (float*)d.dz;
    a.dtmin = (float*)d.tmin; a.dtmax = (float*)d.tmax;
    a.sox = (float*)s.ox; a.soy = (float*)s.oy; a.soz = (float*)s.oz;
    a.sdx = (float*)s.dx; a.sdy = (float*)s.dy; a.sdz = (float*)s.dz;
    a.stmin = (float*)s.tmin; a.stmax = (float*)s.tmax;
    a.src = src; a.count = count; a.tileCounts = sc.tiles;
    a.tileOrder = tileOrder ? 1 : 0;
    hipLaunchKernelGGL(k_count_hits, dim3(tiles), dim3(kTileThreads), 0, st, a);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, sc.tiles, tiles, count);
    hipLaunchKernelGGL(k_emit, dim3(tiles), dim3(kTileThreads), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("bounce launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

}  // namespace rtk
