// cwbvh.hip — closest-hit and any-hit traversal of the compressed 8-wide BVH
// (cwbvh_build.cpp; the reference's traceCWBVH, cuda_rt_v37.cu:146-373).
//
// One ray per lane, 64-ray batches per wave from a device counter. A node
// visit decodes the 8 quantized child boxes (p + q * 2^e in float, the
// arithmetic the builder verified its outward rounding against), runs the
// robust slab test on each, sorts the hit children by entry distance with a
// 19-comparator network and pushes them far to near, continuing with the
// nearest. Leaves hold 1-4 triangles of the CWBVH triangle order and use the
// watertight test of the BVH4 kernels, so hits (t, id, u, v) are identical to
// the BVH4 traversal's and the oracle's. Any hit stops at the first valid hit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rt_device.h"
#include "rt_internal.h"
#include "trace_launch.h"

namespace rtk {

using rtd::Ray;

constexpr int kCwBlock = 256;
constexpr int kCwLds = 8;

__device__ __forceinline__ float pow2f(int e) { return __int_as_float((e + 127) << 23); }   // e in [-126, 127]

template <bool FLAT>
__device__ __forceinline__ bool cw_slab(const Ray& r, const float lo[3], const float hi[3], float limit, float& tnear) {
    float t1x = (lo[0] - r.ox) * r.ix, t2x = (hi[0] - r.ox) * r.ix;
    float t1y = (lo[1] - r.oy) * r.iy, t2y = (hi[1] - r.oy) * r.iy;
    float t1z = (lo[2] - r.oz) * r.iz, t2z = (hi[2] - r.oz) * r.iz;
    if constexpr (FLAT) {
        t1x = t1x != t1x ? -r.ix : t1x; t2x = t2x != t2x ? r.ix : t2x;
        t1y = t1y != t1y ? -r.iy : t1y; t2y = t2y != t2y ? r.iy : t2y;
        t1z = t1z != t1z ? -r.iz : t1z; t2z = t2z != t2z ? r.iz : t2z;
    }
    const float tN = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    const float tF = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z)) * RT_ROBUST_TFAR;
    tnear = tN;
    return (tN <= tF) && (tF >= r.tmin) && (tN <= limit);
}

__device__ __forceinline__ uint32_t byte_of(uint32_t lo4, uint32_t hi4, int i) {
    return ((i < 4 ? lo4 : hi4) >> (8 * (i & 3))) & 0xFFu;
}

// Decodes the 8 children of node `ref`: entry (node index, leaf code, or -1)
// and entry distance (+inf when missed), tested against `limit`.
template <bool FLAT>
__device__ __forceinline__ void cw_children(const uint4* __restrict__ nodes, int ref, const Ray& r, float limit,
                                            int ent[8], float dist[8]) {
    const uint4* np = nodes + 5 * (size_t)ref;
    const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
    const float p[3] = {__uint_as_float(n0.x), __uint_as_float(n0.y), __uint_as_float(n0.z)};
    const float s[3] = {pow2f((int)(int8_t)(n0.w & 0xFFu)), pow2f((int)(int8_t)((n0.w >> 8) & 0xFFu)),
                        pow2f((int)(int8_t)((n0.w >> 16) & 0xFFu))};
    const int childBase = (int)n1.x, triBase = (int)n1.y;
    const uint4 q[3] = {n2, n3, n4};
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t m = byte_of(n1.z, n1.w, i);
        ent[i] = -1;
        dist[i] = INFINITY;
        if (m == 0xFFu) continue;
        float lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = p[a] + (float)byte_of(q[a].x, q[a].y, i) * s[a];
            hi[a] = p[a] + (float)byte_of(q[a].z, q[a].w, i) * s[a];
        }
        float tn;
        if (!cw_slab<FLAT>(r, lo, hi, limit, tn)) continue;
        dist[i] = tn;
        ent[i] = (m & 0x80u) ? childBase + (int)(m & 0x1Fu)
                             : -(((triBase + (int)(m & 0x1Fu)) << 3) | (int)((m >> 5) & 3u)) - 2;   // leaf code
    }
}

#define CW_SWAP(i, j)                                                                   \
    {                                                                                   \
        const bool sw = d[i] > d[j];                                                    \
        const float di = d[i], dj = d[j];                                               \
        const int ei = e[i], ej = e[j];                                                 \
        d[i] = sw ? dj : di; d[j] = sw ? di : dj; e[i] = sw ? ej : ei; e[j] = sw ? ei : ej; \
    }

__device__ __forceinline__ void sort8(float d[8], int e[8]) {
    CW_SWAP(0, 2) CW_SWAP(1, 3) CW_SWAP(4, 6) CW_SWAP(5, 7)
    CW_SWAP(0, 4) CW_SWAP(1, 5) CW_SWAP(2, 6) CW_SWAP(3, 7)
    CW_SWAP(0, 1) CW_SWAP(2, 3) CW_SWAP(4, 5) CW_SWAP(6, 7)
    CW_SWAP(2, 4) CW_SWAP(3, 5)
    CW_SWAP(1, 4) CW_SWAP(3, 6)
    CW_SWAP(1, 2) CW_SWAP(3, 4) CW_SWAP(5, 6)
}
#undef CW_SWAP

template <int MODE, bool FLAT, int LDSD>
__device__ __forceinline__ void cw_trace_ray(const TraceArgs& a, const Ray& r, int* s_ref, float* s_tn, int tid,
                                             int gtid, float& bt, int& bid, float& bu, float& bv, bool& found) {
    const uint4* __restrict__ nodes = a.nodes;
    const float4* __restrict__ tris = a.tris;
    int sp = 0, ref = 0;
    float tn = -INFINITY, best = r.tmax;
    bid = -1; bu = 0.f; bv = 0.f; found = false;
    auto push = [&](int rf, float t) {
        if (sp < LDSD) {
            s_ref[sp * kCwBlock + tid] = rf;
            s_tn[sp * kCwBlock + tid] = t;
        } else if (sp < a.maxStack) {
            a.spill[(size_t)(sp - LDSD) * a.spillStride + gtid] = make_int2(rf, __float_as_int(t));
        } else {
            atomicOr(a.status, 1);   // unreachable: maxStack bounds a DFS of this tree
            return;
        }
        sp++;
    };
    for (;;) {
        if (MODE == kModeAny || tn <= best) {
            if (ref >= 0) {
                int e[8];
                float d[8];
                cw_children<FLAT>(nodes, ref, r, MODE == kModeAny ? r.tmax : best, e, d);
                if constexpr (MODE != kModeAny) sort8(d, e);
                int nxt = -1;
                float nt = 0.f;
#pragma unroll
                for (int k = 7; k >= 0; k--) {
                    if (e[k] == -1) continue;
                    if (nxt != -1) push(nxt, nt);
                    nxt = e[k];
                    nt = d[k];
                }
                if (nxt != -1) { ref = nxt; tn = nt; continue; }
            } else {
                const int enc = -(ref + 2), ts = enc >> 3, tc = (enc & 7) + 1;
                for (int k = 0; k < tc; k++) {
                    const int ti = ts + k;
                    const float4 p0 = tris[3 * ti];
                    float t, u, v;
                    if (!rtd::isect_watertight(r, p0, tris[3 * ti + 1], tris[3 * ti + 2], t, u, v)) continue;
                    if (!(t > r.tmin) || !(t < r.tmax)) continue;
                    if constexpr (MODE == kModeAny) {
                        found = true;
                        return;
                    } else {
                        const int id = __float_as_int(p0.w);
                        if (t < best || (t == best && id < bid)) { best = t; bid = id; bu = u; bv = v; }
                    }
                }
            }
        }
        if (sp == 0) break;
        sp--;
        if (sp < LDSD) {
            ref = s_ref[sp * kCwBlock + tid];
            tn = s_tn[sp * kCwBlock + tid];
        } else {
            const int2 en = a.spill[(size_t)(sp - LDSD) * a.spillStride + gtid];
            ref = en.x;
            tn = __int_as_float(en.y);
        }
    }
    bt = best;
}

template <int MODE>
__global__ __launch_bounds__(kCwBlock) void k_trace_cw(TraceArgs a) {
    __shared__ int s_ref[kCwLds * kCwBlock];
    __shared__ float s_tn[kCwLds * kCwBlock];
    const int tid = threadIdx.x, lane = tid & 63;
    const int gtid = blockIdx.x * kCwBlock + tid;
    const int n = a.nDev ? min(*a.nDev, a.n) : a.n;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.counter, 64u);
        base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
        if (base >= (uint32_t)n) break;
        const int ri = (int)base + lane;
        const bool live = ri < n;
        Ray r;
        if (live) {
            r.ox = a.ox[ri]; r.oy = a.oy[ri]; r.oz = a.oz[ri];
            r.dx = a.dx[ri]; r.dy = a.dy[ri]; r.dz = a.dz[ri];
            r.tmin = a.tmin[ri]; r.tmax = a.tmax[ri];
            rtd::ray_setup(r);
        }
        const bool anyFlat = __any(live && rtd::ray_flat(r));
        if (!live) continue;
        float t, u, v;
        int id;
        bool found;
        if (anyFlat) cw_trace_ray<MODE, true, kCwLds>(a, r, s_ref, s_tn, tid, gtid, t, id, u, v, found);
        else cw_trace_ray<MODE, false, kCwLds>(a, r, s_ref, s_tn, tid, gtid, t, id, u, v, found);
        if constexpr (MODE == kModeAny) {
            a.occ[ri] = found ? 1 : 0;
        } else {
            const bool hit = id >= 0;
            a.t[ri] = hit ? t : RT_MISS_T;
            if (a.id) a.id[ri] = id;
            if (a.u) a.u[ri] = hit ? u : 0.f;
            if (a.v) a.v[ri] = hit ? v : 0.f;
        }
    }
}

__global__ void k_reset_counter_cw(uint32_t* c) { c[0] = 0u; }

template <int MODE>
static int launch_cw(rt::TraceScratch& sc, TraceArgs& a, int depth8, hipStream_t s) {
    static int maxBlocks = 0;
    auto kern = k_trace_cw<MODE>;
    if (!maxBlocks) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kCwBlock, 0) != hipSuccess || nb <= 0) nb = 1;
        maxBlocks = nb;
    }
    int dev = 0, cus = 1;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = maxBlocks * cus, threads = grid * kCwBlock;
    a.maxStack = 7 * depth8 + 8;
    const size_t need = (size_t)std::max(0, a.maxStack - kCwLds) * threads * sizeof(int2);
    if (alloc_scratch(sc) != 0) return -1;
    if (grow_spill(sc, need) != 0) return -1;
    a.spill = (int2*)sc.spill;
    a.spillStride = threads;
    a.counter = sc.counter;
    a.status = sc.status;
    hipLaunchKernelGGL(k_reset_counter_cw, dim3(1), dim3(1), 0, s, sc.counter);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kCwBlock), 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("CWBVH trace launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

int launch_trace_cw(const void* nodes, const void* tris, int depth8, rt::TraceScratch& sc, TraceArgs& a, int mode,
                    hipStream_t s) {
    if (a.n <= 0) return 0;
    a.nodes = (const uint4*)nodes;
    a.tris = (const float4*)tris;
    return mode == kModeAny ? launch_cw<kModeAny>(sc, a, depth8, s) : launch_cw<kModeClosest>(sc, a, depth8, s);
}

}  // namespace rtk
