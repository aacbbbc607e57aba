// pt.hip — wavefront path tracer over a two-level scene.
//
// Structure follows the reference's wavefront renderer (cuda_rt_v34.cu:
// RayState :62-97, generate/extend/shade/shadow kernels :1077-1545): per
// bounce, a closest-hit pass over the extension queue, a shade pass that
// appends next-bounce rays and shadow rays to queues, an any-hit pass over the
// shadow queue and a connect pass that adds the unoccluded light. Differences
// of this design:
//   * queues are SoA (ray arrays in the layout the traversal kernels read,
//     plus throughput, pixel and RNG state), appended with one atomic per
//     wave (wave64 ballot ranks);
//   * every queue count lives in device memory, the traversal kernels read it
//     there, and one sample is a fixed launch sequence that can be captured
//     once in a hipGraph and replayed (sample index and seed are read from
//     device memory as well);
//   * each pixel owns one path at a time, so contributions are added to the
//     pixel without atomics, in path order: the image is deterministic;
//   * lighting is a directional sun (next-event estimation with one shadow
//     ray per hit) plus a sky gradient on escape; surfaces are Lambertian with
//     per-instance albedo; cosine-weighted bounces; Russian roulette from a
//     chosen extension ray on.
// Sampling uses only IEEE +,-,*,/,sqrt (rt_sample.h), so a host restatement
// renders the same image bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "rt_device.h"
#include "rt_internal.h"
#include "rt_sample.h"
#include "scene.h"
#include "scene_dev.h"

namespace rtk {

constexpr int kPtBlock = 256;
constexpr int kPtMaxDepth = CUDABVH_PT_MAX_DEPTH;
constexpr float kInvPi = 0.318309886183790671538f;

struct PtQueue {
    float* r[8];          // ox oy oz dx dy dz tmin tmax
    float* thr[3];
    int* pix;
    uint64_t* rng;
};

struct PtArgs {
    PtQueue q[2];         // extension queues (ping-pong by bounce parity)
    PtQueue t;            // staging queue of the octant sort (shade appends here, the sort scatters to q)
    int* oct;             // [b * 8 + o] rays of bounce b + 1 in direction octant o; [128 + b * 8 + o] sort cursors
    int* camCost;         // camera-ray trace: step count of each 64-ray chunk in the last sample
    int* camOrder;        // and the chunk order built from them (slowest first)
    int sortOct;          // sort each bounce's next extension rays by direction octant
    float* s[8];          // shadow rays
    float* sc[3];         // shadow contribution
    int* spix;
    uint8_t* occ;
    float* ht;
    int* hinst;
    int* hprim;
    int* cnt;             // [b] extension rays of bounce b, [kPtMaxDepth + b] shadow rays
    unsigned long long* totals;
    int* frame;           // [0] sample index, [1] seed
    const InstanceDev* inst;
    const BlasDev* blas;
    float* rgb[3];
    float* alb[3];
    float* nrm[3];
    float* depth;
    rtd::Cam cam;
    int W, H, N, maxDepth, rrStart, tailPaths;
    float sun[3], sunRad[3], skyZ[3], skyH[3], clampMax;
    // tail kernel: the scene and a global stack of tailDepth entries per thread
    const uint4* tlasNodes;
    const float4* tlasTris;
    int2* tailStack;
    int tailStride, tailDepth;
    int* status;          // traversal status word (bit 0: stack overflow)
};

struct PtResources {
    int cap = 0;
    void* block = nullptr;            // one allocation for every per-ray buffer
    int* small = nullptr;             // counters, totals, frame
    int2* tailStack = nullptr;        // tail kernel: [depth][threads] stack entries
    size_t tailBytes = 0;
    hipStream_t own = nullptr;        // stream for graph capture when the scene has none
    hipStream_t side = nullptr;       // shadow traces, concurrent with the next bounce's extension trace
    hipEvent_t fork = nullptr, join = nullptr;
    rt::TraceScratch scratch2;        // traversal scratch of the side stream
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    PtArgs key{};                     // arguments the graph was captured with
    rt::TraceConfig keyTune[3];       // and the scene's traversal configuration
    uint64_t keyGen[2] = {0, 0};      // and the generations of the two traversal scratches
};

__device__ __forceinline__ void clamp_contrib(float c[3], float m) {
    const float mx = fmaxf(c[0], fmaxf(c[1], c[2]));
    if (m > 0.0f && mx > m) {
        const float s = m / mx;
        c[0] = c[0] * s; c[1] = c[1] * s; c[2] = c[2] * s;
    }
}

__global__ void k_pt_begin(PtArgs a, uint32_t seed) {
    const int t = threadIdx.x;
    if (t == 0) { a.frame[0] = 0; a.frame[1] = (int)seed; }
    for (int k = t; k < 2 * kPtMaxDepth; k += blockDim.x) a.totals[k] = 0ull;
}

// Camera rays, slot i holding pixel rtd::tile_strip_pixel(i) (jittered pinhole: the pixel
// offset is two uniforms instead of 0.5; the pixel's random stream is keyed
// by its index, not the slot), throughput 1; clears the image on sample 0
// and resets the queue counters.
__global__ __launch_bounds__(kPtBlock) void k_pt_generate(PtArgs a) {
    const int sample = a.frame[0];
    const uint32_t seed = (uint32_t)a.frame[1];
    const int stride = gridDim.x * kPtBlock;
    const int g = blockIdx.x * kPtBlock + threadIdx.x;
    if (g < 2 * kPtMaxDepth) a.cnt[g] = g == 0 ? a.N : 0;
    if (g < 16 * kPtMaxDepth) a.oct[g] = 0;
    for (int i = g; i < a.N; i += stride) {
        const int p = rtd::tile_strip_pixel(i, a.W, a.H);
        const int py = p / a.W, px = p - py * a.W;
        rtd::Pcg32 rng;
        rng.seed(((uint64_t)seed << 32) ^ (uint64_t)(uint32_t)sample, (uint64_t)p);
        const float jx = rng.uniform(), jy = rng.uniform();
        const float sx = (((float)px + jx) / (float)a.W) * 2.0f - 1.0f;
        const float sy = (((float)py + jy) / (float)a.H) * 2.0f - 1.0f;
        const float ax = sx * a.cam.sx, ay = sy * a.cam.sy;
        float dx = (a.cam.f[0] + ax * a.cam.r[0]) + ay * a.cam.u[0];
        float dy = (a.cam.f[1] + ax * a.cam.r[1]) + ay * a.cam.u[1];
        float dz = (a.cam.f[2] + ax * a.cam.r[2]) + ay * a.cam.u[2];
        const float il = 1.0f / sqrtf((dx * dx + dy * dy) + dz * dz);
        dx = dx * il; dy = dy * il; dz = dz * il;
        const PtQueue& q = a.q[0];
        q.r[0][i] = a.cam.o[0]; q.r[1][i] = a.cam.o[1]; q.r[2][i] = a.cam.o[2];
        q.r[3][i] = dx; q.r[4][i] = dy; q.r[5][i] = dz;
        q.r[6][i] = 0.0f; q.r[7][i] = RT_MISS_T;
        q.thr[0][i] = 1.0f; q.thr[1][i] = 1.0f; q.thr[2][i] = 1.0f;
        q.pix[i] = p;
        q.rng[i] = rng.state;
        if (sample == 0) { a.rgb[0][p] = 0.0f; a.rgb[1][p] = 0.0f; a.rgb[2][p] = 0.0f; }
    }
}

// Wave-aggregated append: one atomic per wave, ranks from the ballot.
__device__ __forceinline__ int wave_append(bool want, int* counter) {
    const uint64_t m = __ballot(want);
    if (!m) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    return want ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// One path event: extension ray b (origin o, direction d) has hit (t, inst,
// prim) or escaped (prim < 0). Adds the sky term of an escaped ray, writes the
// first-hit buffers when aov, and for a hit prepares the shadow ray (origin so,
// contribution c: wantShadow) and the next extension ray (origin so,
// direction nd, throughput thr updated: wantNext). rng continues the path's
// stream (state; the increment is fixed by the pixel).
struct ShadeOut {
    bool shadow, next;
    float so[3], c[3], nd[3];
};

__device__ __forceinline__ void shade_event(const PtArgs& a, int b, bool aov, int pix, const float o[3],
                                            const float d[3], float thr[3], rtd::Pcg32& rng, float t, int inst,
                                            int prim, ShadeOut& out) {
    out.shadow = false;
    out.next = false;
    if (prim < 0) {
        // escaped: sky gradient by elevation
        const float e = fmaxf(d[1], 0.0f);
        float s[3];
        for (int k = 0; k < 3; k++) s[k] = thr[k] * (a.skyH[k] + (a.skyZ[k] - a.skyH[k]) * e);
        clamp_contrib(s, a.clampMax);
        a.rgb[0][pix] += s[0]; a.rgb[1][pix] += s[1]; a.rgb[2][pix] += s[2];
        if (aov) {
            for (int k = 0; k < 3; k++) { a.alb[k][pix] = 0.0f; a.nrm[k][pix] = 0.0f; }
            a.depth[pix] = 0.0f;
        }
        return;
    }
    const InstanceDev& I = a.inst[inst];
    const BlasDev& B = a.blas[I.blas];
    const int li = B.leafOf[prim];
    float nx, ny, nz;
    world_normal(I, B, li, nx, ny, nz);
    if ((nx * d[0] + ny * d[1]) + nz * d[2] > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    const float hx = o[0] + t * d[0], hy = o[1] + t * d[1], hz = o[2] + t * d[2];
    const float eps = 1e-4f * fmaxf(1.0f, fmaxf(fabsf(hx), fmaxf(fabsf(hy), fabsf(hz))));
    out.so[0] = hx + nx * eps; out.so[1] = hy + ny * eps; out.so[2] = hz + nz * eps;
    const float alb[3] = {I.albedo[0], I.albedo[1], I.albedo[2]};
    if (aov) {
        for (int k = 0; k < 3; k++) a.alb[k][pix] = alb[k];
        a.nrm[0][pix] = nx; a.nrm[1][pix] = ny; a.nrm[2][pix] = nz;
        a.depth[pix] = t;
    }
    const float cosL = (nx * a.sun[0] + ny * a.sun[1]) + nz * a.sun[2];
    if (cosL > 0.0f) {
        const float k = cosL * kInvPi;
        for (int j = 0; j < 3; j++) out.c[j] = ((thr[j] * alb[j]) * a.sunRad[j]) * k;
        clamp_contrib(out.c, a.clampMax);
        out.shadow = true;
    }
    if (b + 1 < a.maxDepth) {
        rtd::cosine_dir(nx, ny, nz, rng, out.nd[0], out.nd[1], out.nd[2]);
        for (int j = 0; j < 3; j++) thr[j] = thr[j] * alb[j];
        out.next = true;
        if (b + 1 >= a.rrStart) {
            const float pm = fminf(fmaxf(thr[0], fmaxf(thr[1], thr[2])), 1.0f);
            const float u = rng.uniform();
            if (u >= pm) out.next = false;
            else for (int j = 0; j < 3; j++) thr[j] = thr[j] / pm;
        }
    }
}

// Wavefront mode runs bounce b while more than tailPaths paths are alive.
__device__ __forceinline__ bool wavefront_bounce(const PtArgs& a, int b) { return a.cnt[b] > a.tailPaths; }

__global__ __launch_bounds__(kPtBlock) void k_pt_shade(PtArgs a, int b) {
    if (!wavefront_bounce(a, b)) return;
    const int n = min(a.cnt[b], a.N);
    const bool aov = b == 0 && a.frame[0] == 0;
    const PtQueue& q = a.q[b & 1];
    const PtQueue& nq = a.sortOct ? a.t : a.q[(b + 1) & 1];
    const int stride = gridDim.x * kPtBlock;
    __shared__ int octCount[8];     // this block's next rays per direction octant
    if (threadIdx.x < 8) octCount[threadIdx.x] = 0;
    __syncthreads();
    for (int base = blockIdx.x * kPtBlock; base < n; base += stride) {
        const int i = base + (int)threadIdx.x;
        const bool act = i < n;
        ShadeOut so;
        so.shadow = so.next = false;
        float thr[3] = {0, 0, 0};
        int pix = 0;
        rtd::Pcg32 rng{0, 1};
        if (act) {
            pix = q.pix[i];
            const float o[3] = {q.r[0][i], q.r[1][i], q.r[2][i]}, d[3] = {q.r[3][i], q.r[4][i], q.r[5][i]};
            thr[0] = q.thr[0][i]; thr[1] = q.thr[1][i]; thr[2] = q.thr[2][i];
            rng.state = q.rng[i];
            rng.inc = ((uint64_t)pix << 1u) | 1u;
            shade_event(a, b, aov, pix, o, d, thr, rng, a.ht[i], a.hinst[i], a.hprim[i], so);
        }
        const int sp = wave_append(so.shadow, &a.cnt[kPtMaxDepth + b]);
        if (so.shadow) {
            a.s[0][sp] = so.so[0]; a.s[1][sp] = so.so[1]; a.s[2][sp] = so.so[2];
            a.s[3][sp] = a.sun[0]; a.s[4][sp] = a.sun[1]; a.s[5][sp] = a.sun[2];
            a.s[6][sp] = 0.0f; a.s[7][sp] = RT_MISS_T;
            a.sc[0][sp] = so.c[0]; a.sc[1][sp] = so.c[1]; a.sc[2][sp] = so.c[2];
            a.spix[sp] = pix;
        }
        const int np = wave_append(so.next, &a.cnt[b + 1]);
        if (so.next) {
            nq.r[0][np] = so.so[0]; nq.r[1][np] = so.so[1]; nq.r[2][np] = so.so[2];
            nq.r[3][np] = so.nd[0]; nq.r[4][np] = so.nd[1]; nq.r[5][np] = so.nd[2];
            nq.r[6][np] = 0.0f; nq.r[7][np] = RT_MISS_T;
            nq.thr[0][np] = thr[0]; nq.thr[1][np] = thr[1]; nq.thr[2][np] = thr[2];
            nq.pix[np] = pix;
            nq.rng[np] = rng.state;
        }
        if (a.sortOct) {
            // rays per direction octant: LDS counts per block
            const int o = (so.nd[0] < 0.0f ? 1 : 0) | (so.nd[1] < 0.0f ? 2 : 0) | (so.nd[2] < 0.0f ? 4 : 0);
            const int lane = threadIdx.x & 63;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint64_t m = __ballot(so.next && o == k);
                if (m && lane == 0) atomicAdd(&octCount[k], __popcll(m));
            }
        }
    }
    __syncthreads();
    if (a.sortOct && threadIdx.x < 8 && octCount[threadIdx.x])
        atomicAdd(&a.oct[b * 8 + threadIdx.x], octCount[threadIdx.x]);
}

// Octant sort of the next bounce's extension rays: scatters the staging
// queue into q[(b + 1) & 1] grouped by direction octant (octants in order).
// Each workgroup takes a contiguous range of the staging queue, counts it per
// octant in LDS, reserves its slots with one atomic per octant, and scatters
// with LDS cursors, so a wave traces rays of one octant whose origins keep
// roughly their queue (pixel) order. Paths are independent and each pixel's
// additions follow its own path, so the image does not depend on the queue
// order.
__global__ __launch_bounds__(kPtBlock) void k_pt_octsort(PtArgs a, int b) {
    if (!wavefront_bounce(a, b)) return;      // bounce b was shaded by the tail kernel
    const int n = min(a.cnt[b + 1], a.N);
    __shared__ int cnt8[8], base8[8];
    int start[8];
    int acc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) { start[k] = acc; acc += a.oct[b * 8 + k]; }
    const PtQueue& src = a.t;
    const PtQueue& dst = a.q[(b + 1) & 1];
    const int lane = threadIdx.x & 63;
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    const int lo = min(n, (int)blockIdx.x * per), hi = min(n, lo + per);
    auto octant = [&](int i) {
        return (src.r[3][i] < 0.0f ? 1 : 0) | (src.r[4][i] < 0.0f ? 2 : 0) | (src.r[5][i] < 0.0f ? 4 : 0);
    };
    if (threadIdx.x < 8) cnt8[threadIdx.x] = 0;
    __syncthreads();
    for (int j = lo; j < hi; j += kPtBlock) {
        const int i = j + (int)threadIdx.x;
        const bool act = i < hi;
        const int o = act ? octant(i) : 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint64_t m = __ballot(act && o == k);
            if (m && lane == 0) atomicAdd(&cnt8[k], __popcll(m));
        }
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        base8[threadIdx.x] = cnt8[threadIdx.x] ? atomicAdd(&a.oct[128 + b * 8 + threadIdx.x], cnt8[threadIdx.x]) : 0;
        cnt8[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int j = lo; j < hi; j += kPtBlock) {
        const int i = j + (int)threadIdx.x;
        const bool act = i < hi;
        const int o = act ? octant(i) : 0;
        int pos = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint64_t m = __ballot(act && o == k);
            if (!m) continue;
            int w = 0;
            if (lane == 0) w = atomicAdd(&cnt8[k], __popcll(m));
            w = __builtin_amdgcn_readfirstlane(w);
            if (act && o == k) pos = start[k] + base8[k] + w + __popcll(m & ((1ull << lane) - 1ull));
        }
        if (act) {
            for (int q = 0; q < 8; q++) dst.r[q][pos] = src.r[q][i];
            for (int q = 0; q < 3; q++) dst.thr[q][pos] = src.thr[q][i];
            dst.pix[pos] = src.pix[i];
            dst.rng[pos] = src.rng[i];
        }
    }
}

// Tail: at the first bounce b0 with at most tailPaths live paths, one thread
// per path runs all of its remaining bounces (closest hit, shading, shadow
// ray, next ray) with a one-ray traversal. Each pixel's contributions are
// added in the wavefront order, so the image is the same; the launch's
// duration is one longest path instead of one slowest ray per bounce and
// stage.
__global__ __launch_bounds__(64) void k_pt_tail(PtArgs a, int b0) {
    const int n0 = a.cnt[b0];
    if (n0 > a.tailPaths || (b0 > 0 && a.cnt[b0 - 1] <= a.tailPaths) || n0 == 0) return;
    const bool sample0 = a.frame[0] == 0;
    const PtQueue& q = a.q[b0 & 1];
    const int stride = gridDim.x * 64;
    const int g = blockIdx.x * 64 + threadIdx.x;
    int2* stk = a.tailStack + g;
    for (int i = g; i < n0; i += stride) {
        const int pix = q.pix[i];
        float o[3] = {q.r[0][i], q.r[1][i], q.r[2][i]}, d[3] = {q.r[3][i], q.r[4][i], q.r[5][i]};
        float thr[3] = {q.thr[0][i], q.thr[1][i], q.thr[2][i]};
        rtd::Pcg32 rng;
        rng.state = q.rng[i];
        rng.inc = ((uint64_t)pix << 1u) | 1u;
        for (int b = b0; b < a.maxDepth; b++) {
            if (b > b0) atomicAdd(&a.cnt[b], 1);
            rtd::Ray r;
            r.ox = o[0]; r.oy = o[1]; r.oz = o[2]; r.dx = d[0]; r.dy = d[1]; r.dz = d[2];
            r.tmin = 0.0f; r.tmax = RT_MISS_T;
            rtd::ray_setup(r);
            float t;
            int inst, prim;
            const bool hit = trace2l_one<false>(a.tlasNodes, a.tlasTris, a.inst, a.blas, r, stk, a.tailStride,
                                                a.tailDepth, a.status, t, inst, prim);
            ShadeOut so;
            shade_event(a, b, b == 0 && sample0, pix, o, d, thr, rng, hit ? t : RT_MISS_T, hit ? inst : -1,
                        hit ? prim : -1, so);
            if (so.shadow) {
                atomicAdd(&a.cnt[kPtMaxDepth + b], 1);
                rtd::Ray sr;
                sr.ox = so.so[0]; sr.oy = so.so[1]; sr.oz = so.so[2];
                sr.dx = a.sun[0]; sr.dy = a.sun[1]; sr.dz = a.sun[2];
                sr.tmin = 0.0f; sr.tmax = RT_MISS_T;
                rtd::ray_setup(sr);
                float st;
                int si, sprim;
                if (!trace2l_one<true>(a.tlasNodes, a.tlasTris, a.inst, a.blas, sr, stk, a.tailStride, a.tailDepth,
                                       a.status, st, si, sprim)) {
                    a.rgb[0][pix] += so.c[0]; a.rgb[1][pix] += so.c[1]; a.rgb[2][pix] += so.c[2];
                }
            }
            if (!so.next) break;
            for (int k = 0; k < 3; k++) { o[k] = so.so[k]; d[k] = so.nd[k]; }
        }
    }
}

__global__ __launch_bounds__(kPtBlock) void k_pt_connect(PtArgs a, int b) {
    if (!wavefront_bounce(a, b)) return;
    const int n = min(a.cnt[kPtMaxDepth + b], a.N);
    const int stride = gridDim.x * kPtBlock;
    for (int i = blockIdx.x * kPtBlock + threadIdx.x; i < n; i += stride) {
        if (a.occ[i]) continue;
        const int p = a.spix[i];
        a.rgb[0][p] += a.sc[0][i]; a.rgb[1][p] += a.sc[1][i]; a.rgb[2][p] += a.sc[2][i];
    }
}

__global__ void k_pt_end(PtArgs a) {
    const int t = threadIdx.x;
    for (int k = t; k < 2 * kPtMaxDepth; k += blockDim.x) a.totals[k] += (unsigned long long)a.cnt[k];
    __syncthreads();
    if (t == 0) a.frame[0] += 1;
}

__global__ __launch_bounds__(kPtBlock) void k_pt_scale(PtArgs a, float inv) {
    const int stride = gridDim.x * kPtBlock;
    for (int i = blockIdx.x * kPtBlock + threadIdx.x; i < a.N; i += stride) {
        a.rgb[0][i] = a.rgb[0][i] * inv; a.rgb[1][i] = a.rgb[1][i] * inv; a.rgb[2][i] = a.rgb[2][i] * inv;
    }
}

static int pt_grid(int n) {
    int dev = 0, cus = 1;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int want = (n + kPtBlock - 1) / kPtBlock;
    return max(1, min(want, cus * 8));
}

// Extension (closest-hit) trace of bounce b; the first bounce traces camera
// rays (coherent, scene class 0), later ones scattered rays (class 2).
static int pt_extend(CudaSceneHandle* sc, PtArgs& a, int b, hipStream_t s) {
    const PtQueue& q = a.q[b & 1];
    Trace2LArgs t;
    t.n = a.N; t.nDev = &a.cnt[b];
    t.ox = q.r[0]; t.oy = q.r[1]; t.oz = q.r[2]; t.dx = q.r[3]; t.dy = q.r[4]; t.dz = q.r[5];
    t.tmin = q.r[6]; t.tmax = q.r[7];
    t.t = a.ht; t.inst_out = a.hinst; t.id = a.hprim;
    t.gate = &a.cnt[b]; t.gateMax = a.tailPaths;
    if (b == 0) {
        // camera rays: every pixel, slot i holding its 8x8-tile-strip pixel,
        // so a 64-ray chunk is a tile; the batch kernels take the chunks in
        // the cost order of the previous sample's camera-ray trace
        if (launch_chunk_order(a.camCost, (a.N + 63) / 64, a.camOrder, s) != 0) {
            rt::set_error("path tracer: chunk order launch failed");
            return -1;
        }
        t.chunkOrder = a.camOrder;
        t.chunkCost = a.camCost;
    }
    return launch_trace2l(sc, sc->scratch, t, b == 0 ? 0 : 2, s);
}

// One sample: generate, then per bounce extend -> shade -> shadow -> connect,
// then fold the counts. Every size is read on the device. The shadow trace of
// bounce b needs only shade(b), and the extension trace of bounce b + 1 needs
// only shade(b) too, so they run side by side (the shadow trace on a second
// stream with its own traversal scratch); connect(b), the tail kernel and
// shade(b + 1) follow both, which keeps every pixel's additions in the
// one-stream order.
static int pt_sample(CudaSceneHandle* sc, PtResources& R, PtArgs& a, hipStream_t s) {
    const int g = pt_grid(a.N);
    hipLaunchKernelGGL(k_pt_generate, dim3(g), dim3(kPtBlock), 0, s, a);
    if (pt_extend(sc, a, 0, s) != 0) return -1;
    hipLaunchKernelGGL(k_pt_shade, dim3(g), dim3(kPtBlock), 0, s, a, 0);
    if (a.sortOct && a.maxDepth > 1) hipLaunchKernelGGL(k_pt_octsort, dim3(g), dim3(kPtBlock), 0, s, a, 0);
    for (int b = 0; b < a.maxDepth; b++) {
        Trace2LArgs u;
        u.n = a.N; u.nDev = &a.cnt[kPtMaxDepth + b];
        u.ox = a.s[0]; u.oy = a.s[1]; u.oz = a.s[2]; u.dx = a.s[3]; u.dy = a.s[4]; u.dz = a.s[5];
        u.tmin = a.s[6]; u.tmax = a.s[7];
        u.occ = a.occ;
        u.gate = &a.cnt[b]; u.gateMax = a.tailPaths;
        HIP_TRY(hipEventRecord(R.fork, s), return -1);
        HIP_TRY(hipStreamWaitEvent(R.side, R.fork, 0), return -1);
        if (launch_trace2l(sc, R.scratch2, u, 1, R.side) != 0) return -1;
        HIP_TRY(hipEventRecord(R.join, R.side), return -1);
        if (b + 1 < a.maxDepth && pt_extend(sc, a, b + 1, s) != 0) return -1;
        HIP_TRY(hipStreamWaitEvent(s, R.join, 0), return -1);
        hipLaunchKernelGGL(k_pt_connect, dim3(g), dim3(kPtBlock), 0, s, a, b);
        if (a.tailPaths > 0) hipLaunchKernelGGL(k_pt_tail, dim3(a.tailStride / 64), dim3(64), 0, s, a, b);
        if (b + 1 < a.maxDepth) hipLaunchKernelGGL(k_pt_shade, dim3(g), dim3(kPtBlock), 0, s, a, b + 1);
        if (a.sortOct && b + 2 < a.maxDepth)
            hipLaunchKernelGGL(k_pt_octsort, dim3(g), dim3(kPtBlock), 0, s, a, b + 1);
    }
    hipLaunchKernelGGL(k_pt_end, dim3(1), dim3(64), 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("path tracer launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

static int pt_alloc(PtResources& R, int N) {
    if (R.cap >= N && R.block) return 0;
    hipFree(R.block);
    R.block = nullptr; R.cap = 0;
    // per ray: 3 queues (two ping-pong, one octant-sort staging) x (8 + 3 floats, pix, rng 8 B)
    // + shadow (8 + 3 floats, pix, occ) + hit (t, inst, prim)
    const size_t per = 3 * (11 * 4 + 4 + 8) + (11 * 4 + 4 + 1) + 12;
    const size_t chunkBytes = 2 * ((size_t)(N + 63) / 64 * 4 + 256);     // camera-ray chunk costs and order
    const size_t bytes = per * (size_t)N + chunkBytes + 4096;
    if (hipMalloc(&R.block, bytes) != hipSuccess) {
        rt::set_error("path tracer: cannot allocate %zu bytes of queues", bytes);
        return -1;
    }
    if (!R.small && hipMalloc((void**)&R.small, 4096) != hipSuccess) {
        rt::set_error("path tracer: cannot allocate counters");
        return -1;
    }
    R.cap = N;
    return 0;
}

static void pt_layout(PtResources& R, PtArgs& a) {
    const size_t N = (size_t)R.cap;
    char* p = (char*)R.block;
    auto take = [&](size_t bytes) { char* r = p; p += (bytes + 255) & ~(size_t)255; return r; };
    for (int k = 0; k < 3; k++) {
        PtQueue& q = k < 2 ? a.q[k] : a.t;
        q.rng = (uint64_t*)take(8 * N);
        for (int j = 0; j < 8; j++) q.r[j] = (float*)take(4 * N);
        for (int j = 0; j < 3; j++) q.thr[j] = (float*)take(4 * N);
        q.pix = (int*)take(4 * N);
    }
    for (int j = 0; j < 8; j++) a.s[j] = (float*)take(4 * N);
    for (int j = 0; j < 3; j++) a.sc[j] = (float*)take(4 * N);
    a.spix = (int*)take(4 * N);
    a.ht = (float*)take(4 * N);
    a.hinst = (int*)take(4 * N);
    a.hprim = (int*)take(4 * N);
    a.occ = (uint8_t*)take(N);
    // the camera rays of one sample are traced slowest chunk first by the
    // previous sample's step counts (same camera, new jitter): any values
    // give a valid order, so a fresh allocation needs no clearing
    a.camCost = (int*)take((N + 63) / 64 * 4);
    a.camOrder = (int*)take((N + 63) / 64 * 4);
    a.cnt = R.small;                                           // 2 * kPtMaxDepth ints
    a.frame = R.small + 2 * kPtMaxDepth;                       // 2 ints
    a.totals = (unsigned long long*)(R.small + 4 * kPtMaxDepth);   // 8-byte aligned
    a.oct = R.small + 256;                                     // 16 * kPtMaxDepth ints
}

void pt_free(CudaSceneHandle* sc) {
    auto* R = (PtResources*)sc->pt;
    if (!R) return;
    if (R->exec) hipGraphExecDestroy(R->exec);
    if (R->graph) hipGraphDestroy(R->graph);
    if (R->own) hipStreamDestroy(R->own);
    if (R->side) hipStreamDestroy(R->side);
    if (R->fork) hipEventDestroy(R->fork);
    if (R->join) hipEventDestroy(R->join);
    free_scratch(R->scratch2);
    hipFree(R->tailStack);
    hipFree(R->block);
    hipFree(R->small);
    delete R;
    sc->pt = nullptr;
}

int pt_take_status(CudaSceneHandle* sc) {
    if (!sc->pt) return 0;
    PtResources& R = *(PtResources*)sc->pt;
    if (!R.side) return 0;
    return take_status(R.scratch2, R.side, "scenePathTrace");
}

static bool same_key(const PtArgs& x, const PtArgs& y) { return memcmp(&x, &y, sizeof(PtArgs)) == 0; }

int path_trace(CudaSceneHandle* sc, const CudaCamera& cam, const CudaPathParams& p, const CudaPathImage& img,
               long long* rayCounts, unsigned flags) {
    if (p.width <= 0 || p.height <= 0 || p.spp <= 0 || p.maxDepth <= 0 || p.maxDepth > kPtMaxDepth ||
        p.rrStart < 0) {
        rt::set_error("scenePathTrace: bad parameters (%dx%d, spp %d, depth %d, rr %d; depth <= %d)", p.width,
                      p.height, p.spp, p.maxDepth, p.rrStart, kPtMaxDepth);
        return -1;
    }
    if ((long long)p.width * p.height > (1 << 28)) { rt::set_error("scenePathTrace: image too large"); return -1; }
    if (!img.rgb) { rt::set_error("scenePathTrace: NULL rgb"); return -1; }
    if (!sc->pt) sc->pt = new PtResources();
    PtResources& R = *(PtResources*)sc->pt;
    const int N = p.width * p.height;
    if (pt_alloc(R, N) != 0) return -1;

    PtArgs a{};
    memset(&a, 0, sizeof a);      // the graph key compares the whole struct, padding included
    pt_layout(R, a);
    a.inst = sc->d_inst; a.blas = sc->d_blas;
    for (int k = 0; k < 3; k++) {
        a.rgb[k] = img.rgb + (size_t)k * N;
        a.alb[k] = img.albedo ? img.albedo + (size_t)k * N : nullptr;
        a.nrm[k] = img.normal ? img.normal + (size_t)k * N : nullptr;
    }
    a.depth = img.depth;
    if (!img.albedo || !img.normal || !img.depth) {
        rt::set_error("scenePathTrace: albedo, normal and depth buffers are required");
        return -1;
    }
    for (int k = 0; k < 3; k++) { a.cam.o[k] = cam.o[k]; a.cam.f[k] = cam.f[k]; a.cam.r[k] = cam.r[k]; a.cam.u[k] = cam.u[k]; }
    a.cam.sx = cam.scaleX; a.cam.sy = cam.scaleY;
    a.W = p.width; a.H = p.height; a.N = N; a.maxDepth = p.maxDepth; a.rrStart = p.rrStart;
    const double sl = std::sqrt((double)p.sunDir[0] * p.sunDir[0] + (double)p.sunDir[1] * p.sunDir[1] +
                                (double)p.sunDir[2] * p.sunDir[2]);
    if (!(sl > 0)) { rt::set_error("scenePathTrace: zero sun direction"); return -1; }
    for (int k = 0; k < 3; k++) {
        a.sun[k] = (float)(p.sunDir[k] / sl);
        a.sunRad[k] = p.sunRadiance[k]; a.skyZ[k] = p.skyZenith[k]; a.skyH[k] = p.skyHorizon[k];
    }
    a.clampMax = p.clampMax;
    a.sortOct = (flags & CUDABVH_FLAG_SORT_RAYS) ? 1 : 0;
    a.tailPaths = std::max(0, std::min(p.tailPaths < 0 ? sc->ptTail : p.tailPaths, N));
    a.tlasNodes = (const uint4*)sc->tlas->d_nodes4;
    a.tlasTris = (const float4*)sc->tlas->d_tris;
    if (a.tailPaths > 0) {
        a.tailStride = std::min(64 * 1024, (a.tailPaths + 63) / 64 * 64);
        a.tailDepth = trace2l_max_stack(sc);
        const size_t need = (size_t)a.tailStride * a.tailDepth * sizeof(int2);
        if (need > R.tailBytes) {
            hipFree(R.tailStack);
            R.tailStack = nullptr; R.tailBytes = 0;
            if (hipMalloc((void**)&R.tailStack, need) != hipSuccess) {
                rt::set_error("scenePathTrace: cannot allocate %zu stack bytes", need);
                return -1;
            }
            R.tailBytes = need;
        }
        a.tailStack = R.tailStack;
    }

    if (!R.side) {
        if (hipStreamCreateWithFlags(&R.side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&R.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&R.join, hipEventDisableTiming) != hipSuccess) {
            rt::set_error("scenePathTrace: cannot create the shadow stream");
            return -1;
        }
    }
    if (alloc_scratch(R.scratch2) != 0 || alloc_scratch(sc->scratch) != 0) return -1;
    a.status = sc->scratch.status;

    const bool graph = (flags & CUDABVH_FLAG_GRAPH) != 0;
    hipStream_t s = (hipStream_t)sc->stream;
    if (graph && !s) {
        // capture needs a stream of its own; order it after work on the null stream
        if (!R.own && hipStreamCreateWithFlags(&R.own, hipStreamNonBlocking) != hipSuccess) {
            rt::set_error("scenePathTrace: cannot create a stream");
            return -1;
        }
        HIP_TRY(hipStreamSynchronize(nullptr), return -1);
        s = R.own;
    }
    hipLaunchKernelGGL(k_pt_begin, dim3(1), dim3(64), 0, s, a, p.seed);
    int first = 0;
    if (graph) {
        // re-capture when the arguments, the launch configuration or any scratch
        // buffer the captured kernels address has changed since the capture
        if (!R.exec || !same_key(R.key, a) || memcmp(R.keyTune, sc->tune2l, sizeof(R.keyTune)) != 0 ||
            R.keyGen[0] != sc->scratch.gen || R.keyGen[1] != R.scratch2.gen) {
            if (R.exec) { hipGraphExecDestroy(R.exec); R.exec = nullptr; }
            if (R.graph) { hipGraphDestroy(R.graph); R.graph = nullptr; }
            // one eager sample allocates every scratch buffer the traces need
            if (pt_sample(sc, R, a, s) != 0) return -1;
            first = 1;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), return -1);
            const int rc = pt_sample(sc, R, a, s);
            hipGraph_t gr = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &gr);
            if (rc != 0 || ec != hipSuccess) {
                if (gr) hipGraphDestroy(gr);
                if (rc == 0) rt::set_error("scenePathTrace: capture failed: %s", hipGetErrorString(ec));
                return -1;
            }
            R.graph = gr;
            HIP_TRY(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0), return -1);
            R.key = a;
            memcpy(R.keyTune, sc->tune2l, sizeof(R.keyTune));
            R.keyGen[0] = sc->scratch.gen;
            R.keyGen[1] = R.scratch2.gen;
        }
        for (int k = first; k < p.spp; k++) HIP_TRY(hipGraphLaunch(R.exec, s), return -1);
    } else {
        for (int k = 0; k < p.spp; k++)
            if (pt_sample(sc, R, a, s) != 0) return -1;
    }
    if (p.spp > 1) hipLaunchKernelGGL(k_pt_scale, dim3(pt_grid(N)), dim3(kPtBlock), 0, s, a, 1.0f / (float)p.spp);
    if (rayCounts) {
        unsigned long long tot[2 * kPtMaxDepth];
        HIP_TRY(hipMemcpyAsync(tot, a.totals, sizeof(tot), hipMemcpyDeviceToHost, s), return -1);
        HIP_TRY(hipStreamSynchronize(s), return -1);
        for (int k = 0; k < p.maxDepth; k++) {
            rayCounts[k] = (long long)tot[k];
            rayCounts[p.maxDepth + k] = (long long)tot[kPtMaxDepth + k];
        }
    }
    if ((flags & CUDABVH_FLAG_SYNC) || s == R.own) {
        HIP_TRY(hipStreamSynchronize(R.side), return -1);
        if (take_status(sc->scratch, s, "scenePathTrace") != 0 || take_status(R.scratch2, s, "scenePathTrace") != 0)
            return -1;
    }
    return 0;
}

}  // namespace rtk
