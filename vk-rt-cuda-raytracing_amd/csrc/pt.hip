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

#include <cstring>

#include "rt_device.h"
#include "rt_internal.h"
#include "rt_sample.h"
#include "scene.h"

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
    int W, H, N, maxDepth, rrStart;
    float sun[3], sunRad[3], skyZ[3], skyH[3], clampMax;
};

struct PtResources {
    int cap = 0;
    void* block = nullptr;            // one allocation for every per-ray buffer
    int* small = nullptr;             // counters, totals, frame
    hipStream_t own = nullptr;        // stream for graph capture when the scene has none
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    PtArgs key{};                     // arguments the graph was captured with
    int keyDepth = 0;
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

// Camera rays in pixel order (jittered pinhole: the pixel offset is two
// uniforms instead of 0.5), throughput 1; clears the image on sample 0 and
// resets the queue counters.
__global__ __launch_bounds__(kPtBlock) void k_pt_generate(PtArgs a) {
    const int sample = a.frame[0];
    const uint32_t seed = (uint32_t)a.frame[1];
    const int stride = gridDim.x * kPtBlock;
    const int g = blockIdx.x * kPtBlock + threadIdx.x;
    if (g < 2 * kPtMaxDepth) a.cnt[g] = g == 0 ? a.N : 0;
    for (int i = g; i < a.N; i += stride) {
        const int py = i / a.W, px = i - py * a.W;
        rtd::Pcg32 rng;
        rng.seed(((uint64_t)seed << 32) ^ (uint64_t)(uint32_t)sample, (uint64_t)i);
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
        q.pix[i] = i;
        q.rng[i] = rng.state;
        if (sample == 0) { a.rgb[0][i] = 0.0f; a.rgb[1][i] = 0.0f; a.rgb[2][i] = 0.0f; }
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

__global__ __launch_bounds__(kPtBlock) void k_pt_shade(PtArgs a, int b) {
    const int n = min(a.cnt[b], a.N);
    const bool aov = b == 0 && a.frame[0] == 0;
    const PtQueue& q = a.q[b & 1];
    const PtQueue& nq = a.q[(b + 1) & 1];
    const int stride = gridDim.x * kPtBlock;
    for (int base = blockIdx.x * kPtBlock; base < n; base += stride) {
        const int i = base + (int)threadIdx.x;
        const bool act = i < n;
        bool wantShadow = false, wantNext = false;
        float ox = 0, oy = 0, oz = 0, nx = 0, ny = 0, nz = 0, ndx = 0, ndy = 0, ndz = 0;
        float c[3] = {0, 0, 0}, thr[3] = {0, 0, 0};
        int pix = 0;
        rtd::Pcg32 rng{0, 1};
        if (act) {
            pix = q.pix[i];
            const float dx = q.r[3][i], dy = q.r[4][i], dz = q.r[5][i];
            thr[0] = q.thr[0][i]; thr[1] = q.thr[1][i]; thr[2] = q.thr[2][i];
            const int hp = a.hprim[i];
            if (hp < 0) {
                // escaped: sky gradient by elevation
                const float e = fmaxf(dy, 0.0f);
                float s[3];
                for (int k = 0; k < 3; k++) s[k] = thr[k] * (a.skyH[k] + (a.skyZ[k] - a.skyH[k]) * e);
                clamp_contrib(s, a.clampMax);
                a.rgb[0][pix] += s[0]; a.rgb[1][pix] += s[1]; a.rgb[2][pix] += s[2];
                if (aov) {
                    for (int k = 0; k < 3; k++) { a.alb[k][pix] = 0.0f; a.nrm[k][pix] = 0.0f; }
                    a.depth[pix] = 0.0f;
                }
            } else {
                const float t = a.ht[i];
                const int ii = a.hinst[i];
                const InstanceDev& I = a.inst[ii];
                const BlasDev& B = a.blas[I.blas];
                const int li = B.leafOf[hp];
                const float4 p0 = B.tris[3 * li], p1 = B.tris[3 * li + 1], p2 = B.tris[3 * li + 2];
                const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
                const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
                const float mx = e1y * e2z - e1z * e2y, my = e1z * e2x - e1x * e2z, mz = e1x * e2y - e1y * e2x;
                const float* W = I.w2o;
                // normals map with the transpose of the world-to-object matrix
                nx = (W[0] * mx + W[4] * my) + W[8] * mz;
                ny = (W[1] * mx + W[5] * my) + W[9] * mz;
                nz = (W[2] * mx + W[6] * my) + W[10] * mz;
                const float nl = 1.0f / sqrtf((nx * nx + ny * ny) + nz * nz);
                nx = nx * nl; ny = ny * nl; nz = nz * nl;
                if ((nx * dx + ny * dy) + nz * dz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
                const float hx = q.r[0][i] + t * dx, hy = q.r[1][i] + t * dy, hz = q.r[2][i] + t * dz;
                const float eps = 1e-4f * fmaxf(1.0f, fmaxf(fabsf(hx), fmaxf(fabsf(hy), fabsf(hz))));
                ox = hx + nx * eps; oy = hy + ny * eps; oz = hz + nz * eps;
                const float alb[3] = {I.albedo[0], I.albedo[1], I.albedo[2]};
                if (aov) {
                    for (int k = 0; k < 3; k++) a.alb[k][pix] = alb[k];
                    a.nrm[0][pix] = nx; a.nrm[1][pix] = ny; a.nrm[2][pix] = nz;
                    a.depth[pix] = t;
                }
                const float cosL = (nx * a.sun[0] + ny * a.sun[1]) + nz * a.sun[2];
                if (cosL > 0.0f) {
                    const float k = cosL * kInvPi;
                    for (int j = 0; j < 3; j++) c[j] = ((thr[j] * alb[j]) * a.sunRad[j]) * k;
                    clamp_contrib(c, a.clampMax);
                    wantShadow = true;
                }
                if (b + 1 < a.maxDepth) {
                    rng.state = q.rng[i];
                    rng.inc = ((uint64_t)pix << 1u) | 1u;
                    rtd::cosine_dir(nx, ny, nz, rng, ndx, ndy, ndz);
                    for (int j = 0; j < 3; j++) thr[j] = thr[j] * alb[j];
                    wantNext = true;
                    if (b + 1 >= a.rrStart) {
                        const float pm = fminf(fmaxf(thr[0], fmaxf(thr[1], thr[2])), 1.0f);
                        const float u = rng.uniform();
                        if (u >= pm) wantNext = false;
                        else for (int j = 0; j < 3; j++) thr[j] = thr[j] / pm;
                    }
                }
            }
        }
        const int sp = wave_append(wantShadow, &a.cnt[kPtMaxDepth + b]);
        if (wantShadow) {
            a.s[0][sp] = ox; a.s[1][sp] = oy; a.s[2][sp] = oz;
            a.s[3][sp] = a.sun[0]; a.s[4][sp] = a.sun[1]; a.s[5][sp] = a.sun[2];
            a.s[6][sp] = 0.0f; a.s[7][sp] = RT_MISS_T;
            a.sc[0][sp] = c[0]; a.sc[1][sp] = c[1]; a.sc[2][sp] = c[2];
            a.spix[sp] = pix;
        }
        const int np = wave_append(wantNext, &a.cnt[b + 1]);
        if (wantNext) {
            nq.r[0][np] = ox; nq.r[1][np] = oy; nq.r[2][np] = oz;
            nq.r[3][np] = ndx; nq.r[4][np] = ndy; nq.r[5][np] = ndz;
            nq.r[6][np] = 0.0f; nq.r[7][np] = RT_MISS_T;
            nq.thr[0][np] = thr[0]; nq.thr[1][np] = thr[1]; nq.thr[2][np] = thr[2];
            nq.pix[np] = pix;
            nq.rng[np] = rng.state;
        }
    }
}

__global__ __launch_bounds__(kPtBlock) void k_pt_connect(PtArgs a, int b) {
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

// One sample: generate, then per bounce extend -> shade -> shadow -> connect,
// then fold the counts. Every size is read on the device.
static int pt_sample(CudaSceneHandle* sc, PtArgs& a, hipStream_t s) {
    const int g = pt_grid(a.N);
    hipLaunchKernelGGL(k_pt_generate, dim3(g), dim3(kPtBlock), 0, s, a);
    for (int b = 0; b < a.maxDepth; b++) {
        const PtQueue& q = a.q[b & 1];
        Trace2LArgs t;
        t.n = a.N; t.nDev = &a.cnt[b];
        t.ox = q.r[0]; t.oy = q.r[1]; t.oz = q.r[2]; t.dx = q.r[3]; t.dy = q.r[4]; t.dz = q.r[5];
        t.tmin = q.r[6]; t.tmax = q.r[7];
        t.t = a.ht; t.inst_out = a.hinst; t.id = a.hprim;
        if (launch_trace2l(sc, sc->scratch, t, false, s) != 0) return -1;
        hipLaunchKernelGGL(k_pt_shade, dim3(g), dim3(kPtBlock), 0, s, a, b);
        Trace2LArgs u;
        u.n = a.N; u.nDev = &a.cnt[kPtMaxDepth + b];
        u.ox = a.s[0]; u.oy = a.s[1]; u.oz = a.s[2]; u.dx = a.s[3]; u.dy = a.s[4]; u.dz = a.s[5];
        u.tmin = a.s[6]; u.tmax = a.s[7];
        u.occ = a.occ;
        if (launch_trace2l(sc, sc->scratch, u, true, s) != 0) return -1;
        hipLaunchKernelGGL(k_pt_connect, dim3(g), dim3(kPtBlock), 0, s, a, b);
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
    // per ray: 2 queues x (8 + 3 floats, pix, rng 8 B) + shadow (8 + 3 floats, pix, occ) + hit (t, inst, prim)
    const size_t per = 2 * (11 * 4 + 4 + 8) + (11 * 4 + 4 + 1) + 12;
    const size_t bytes = per * (size_t)N + 4096;
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
    for (int k = 0; k < 2; k++) {
        a.q[k].rng = (uint64_t*)take(8 * N);
        for (int j = 0; j < 8; j++) a.q[k].r[j] = (float*)take(4 * N);
        for (int j = 0; j < 3; j++) a.q[k].thr[j] = (float*)take(4 * N);
        a.q[k].pix = (int*)take(4 * N);
    }
    for (int j = 0; j < 8; j++) a.s[j] = (float*)take(4 * N);
    for (int j = 0; j < 3; j++) a.sc[j] = (float*)take(4 * N);
    a.spix = (int*)take(4 * N);
    a.ht = (float*)take(4 * N);
    a.hinst = (int*)take(4 * N);
    a.hprim = (int*)take(4 * N);
    a.occ = (uint8_t*)take(N);
    a.cnt = R.small;                                           // 2 * kPtMaxDepth ints
    a.frame = R.small + 2 * kPtMaxDepth;                       // 2 ints
    a.totals = (unsigned long long*)(R.small + 4 * kPtMaxDepth);   // 8-byte aligned
}

void pt_free(CudaSceneHandle* sc) {
    auto* R = (PtResources*)sc->pt;
    if (!R) return;
    if (R->exec) hipGraphExecDestroy(R->exec);
    if (R->graph) hipGraphDestroy(R->graph);
    if (R->own) hipStreamDestroy(R->own);
    hipFree(R->block);
    hipFree(R->small);
    delete R;
    sc->pt = nullptr;
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
        if (!R.exec || !same_key(R.key, a)) {
            if (R.exec) { hipGraphExecDestroy(R.exec); R.exec = nullptr; }
            if (R.graph) { hipGraphDestroy(R.graph); R.graph = nullptr; }
            // one eager sample allocates every scratch buffer the traces need
            if (pt_sample(sc, a, s) != 0) return -1;
            first = 1;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), return -1);
            const int rc = pt_sample(sc, a, s);
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
        }
        for (int k = first; k < p.spp; k++) HIP_TRY(hipGraphLaunch(R.exec, s), return -1);
    } else {
        for (int k = 0; k < p.spp; k++)
            if (pt_sample(sc, a, s) != 0) return -1;
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
    if ((flags & CUDABVH_FLAG_SYNC) || s == R.own) HIP_TRY(hipStreamSynchronize(s), return -1);
    return 0;
}

}  // namespace rtk
