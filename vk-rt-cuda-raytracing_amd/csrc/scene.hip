// scene.hip — two-level scenes: a TLAS (BVH4 over instance boxes, one
// instance per leaf) above per-mesh BLAS handles, and its closest-hit /
// any-hit traversal.
//
// One depth-first traversal per ray: reaching a TLAS leaf pushes a return
// marker, transforms the ray into the instance's object space with the
// world-to-object affine map (unnormalised, so t stays the world distance
// parameter) and continues into the BLAS; popping the marker restores the
// world ray. The closest hit is the lexicographic minimum of (t, instance,
// primitive). The schedule is the while-while one of trace.hip (node and leaf
// phases, idle lanes refilled from a wave-local pool of 64-ray batches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "rt_device.h"
#include "rt_internal.h"
#include "scene.h"
#include "scene_dev.h"
#include "trace_launch.h"

namespace rtk {

using rtd::Ray;

template <bool ANY, int LDSD, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace2l(Trace2LArgs a) {
    __shared__ int s_ref[LDSD * BLOCK];
    __shared__ float s_tn[LDSD * BLOCK];
    const int tid = threadIdx.x, lane = tid & 63;
    const int gtid = blockIdx.x * BLOCK + tid;
    if (a.gate && *a.gate <= a.gateMax) return;
    const int n = a.nDev ? min(*a.nDev, a.n) : a.n;

    Ray rw, ro;
    int ri = 0, inst = -1;
    const uint4* nodes = a.tlasNodes;
    const float4* tris = nullptr;
    bool has = false, exhausted = false, found = false, flatW = false, flatO = false;
    int poolNext = 0, poolEnd = 0;
    int band = blockIdx.x % rtd::kBands;   // wave-uniform: band of the work distribution
    int sp = 0, ref = -1, bid = -1, binst = -1, visits = 0, tests = 0;
    int pend = -1;             // postponed BLAS leaf (speculative node phase), -1: none
    float tn = 0.f, best = 0.f, bu = 0.f, bv = 0.f;
    // Uncounted launches postpone BLAS leaves speculatively, as the
    // single-level while-while kernel does; a lane holding a postponed leaf
    // stays inside its instance (it waits at the return marker), so the leaf
    // is tested with that instance's object-space ray.
    const bool spec = a.cntNodes == nullptr;

    auto push = [&](int rf, float t) {
        if (sp < LDSD) {
            s_ref[sp * BLOCK + tid] = rf;
            s_tn[sp * BLOCK + tid] = t;
        } else if (sp < a.maxStack) {
            a.spill[(size_t)(sp - LDSD) * a.spillStride + gtid] = make_int2(rf, __float_as_int(t));
        } else {
            atomicOr(a.status, 1);
            return;
        }
        sp++;
    };

    for (;;) {
        if (!exhausted || poolNext < poolEnd) {
            const uint64_t idle = __ballot(!has);
            if (idle) {
                const int cnt = __popcll(idle);
                const int avail = poolEnd - poolNext;
                int fresh = n;
                if (cnt > avail && !exhausted) {
                    fresh = rtd::take_chunk(a.counter, n, band);
                    if (fresh >= n) exhausted = true;
                }
                const int k = __popcll(idle & ((1ull << lane) - 1ull));
                int w = n;
                if (k < avail) w = poolNext + k;
                else if (fresh < n) w = fresh + (k - avail);
                if (cnt > avail) {
                    if (fresh < n) { poolNext = fresh + (cnt - avail); poolEnd = fresh + 64; }
                    else { poolNext = 0; poolEnd = 0; }
                } else {
                    poolNext += cnt;
                }
                if (!has && w < n) {
                    ri = w;
                    rw.ox = a.ox[ri]; rw.oy = a.oy[ri]; rw.oz = a.oz[ri];
                    rw.dx = a.dx[ri]; rw.dy = a.dy[ri]; rw.dz = a.dz[ri];
                    rw.tmin = a.tmin[ri]; rw.tmax = a.tmax[ri];
                    rtd::ray_setup(rw);
                    flatW = rtd::ray_flat(rw);
                    has = true;
                    sp = 0; ref = 0; tn = -INFINITY; inst = -1; nodes = a.tlasNodes;
                    best = rw.tmax; bu = 0.f; bv = 0.f; bid = -1; binst = -1;
                    visits = 0; tests = 0; found = false;
                }
            }
        }
        if (!__any(has)) {
            if (exhausted && poolNext >= poolEnd) break;
            continue;
        }
        bool fin = false;
        for (;;) {
            if (has && !fin) {
                for (;;) {
                    while (ref == -1 || ref == kReturnMark || (!ANY && tn > best)) {
                        if (ref == kReturnMark) {
                            if (pend != -1) break;           // wait inside the instance
                            inst = -1; nodes = a.tlasNodes; ref = -1;
                            continue;
                        }
                        if (sp == 0) { fin = true; break; }
                        sp--;
                        if (sp < LDSD) {
                            ref = s_ref[sp * BLOCK + tid];
                            tn = s_tn[sp * BLOCK + tid];
                        } else {
                            const int2 e = a.spill[(size_t)(sp - LDSD) * a.spillStride + gtid];
                            ref = e.x; tn = __int_as_float(e.y);
                        }
                    }
                    if (spec && !fin && inst >= 0 && ref < -1 && ref != kReturnMark && pend == -1) {
                        pend = ref; ref = -1;
                        continue;
                    }
                    break;
                }
            }
            // TLAS nodes, TLAS leaves (instance entry) and BLAS nodes are node
            // work; a lane waiting at the return marker with a postponed leaf
            // is not
            const bool atNode = has && !fin && ref != kReturnMark && (ref >= 0 || inst < 0);
            if (!__any(atNode && pend == -1)) break;
            if (atNode) {
                if (ref >= 0) {
                    visits++;
                    const bool w = inst < 0;
                    const Ray& r = w ? rw : ro;
                    visit2w<ANY>(nodes, r, w ? flatW : flatO, ANY ? r.tmax : best, ref, tn, push);
                } else {
                    // TLAS leaf: enter its instance
                    const int enc = -(ref + 2);
                    const int k = __float_as_int(a.tlasTris[3 * (enc >> 3)].w);
                    push(kReturnMark, -INFINITY);
                    inst = k;
                    const InstanceDev& I = a.inst[k];
                    xform(I.w2o, rw, ro);
                    flatO = rtd::ray_flat(ro);
                    const BlasDev& B = a.blas[I.blas];
                    nodes = B.nodes;
                    tris = B.tris;
                    ref = 0;
                    tn = -INFINITY;
                }
            }
        }
        const int leaf = spec ? pend : ((has && !fin) ? ref : -1);
        if (has && leaf < -1) {
            const int enc = -(leaf + 2), ts = enc >> 3, tc = (enc & 7) + 1;
            for (int q = 0; q < tc; q++) {
                const int ti = ts + q;
                tests++;
                const float4 p0 = rtd::ldg_f4(tris + 3 * ti);
                float t, u, v;
                if (!rtd::isect_watertight(ro, p0, rtd::ldg_f4(tris + 3 * ti + 1), rtd::ldg_f4(tris + 3 * ti + 2), t, u, v)) continue;
                if (!(t > ro.tmin) || !(t < ro.tmax)) continue;
                if constexpr (ANY) {
                    found = true;
                    fin = true;
                    break;
                } else {
                    const int id = __float_as_int(p0.w);
                    if (t < best || (t == best && (inst < binst || (inst == binst && id < bid)))) {
                        best = t; bid = id; binst = inst; bu = u; bv = v;
                    }
                }
            }
            if (spec) pend = -1;
            else ref = -1;
        }
        if (has && fin) {
            if constexpr (ANY) {
                a.occ[ri] = found ? 1 : 0;
            } else {
                const bool hit = bid >= 0;
                a.t[ri] = hit ? best : RT_MISS_T;
                if (a.inst_out) a.inst_out[ri] = binst;
                if (a.id) a.id[ri] = bid;
                if (a.u) a.u[ri] = hit ? bu : 0.f;
                if (a.v) a.v[ri] = hit ? bv : 0.f;
            }
            if (a.cntNodes) { a.cntNodes[ri] = visits; a.cntTris[ri] = tests; }
            has = false;
        }
    }
}

// Batch schedule: each wave takes 64 rays from its XCD band (rtd::take_chunk) and each
// lane runs its ray's whole traversal (trace2l_ray) with an LDS stack.
template <bool ANY, bool ORD, int LDSD, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace2l_batch(Trace2LArgs a) {
    __shared__ int s_ref[LDSD * BLOCK];
    __shared__ float s_tn[LDSD * BLOCK];
    const int tid = threadIdx.x, lane = tid & 63;
    const int gtid = blockIdx.x * BLOCK + tid;
    if (a.gate && *a.gate <= a.gateMax) return;
    const int n = a.nDev ? min(*a.nDev, a.n) : a.n;
    int band = blockIdx.x % rtd::kBands;
    const int nChunks = (n + 63) >> 6;
    for (;;) {
        const int base = rtd::take_chunk(a.counter, n, band);
        if (base >= n) break;
        // chunk c traces chunk c, or chunkOrder[c] in a cost-ordered launch
        int chunk = base >> 6;
        if (a.chunkOrder) chunk = a.chunkOrder[chunk];
        const bool chunkOk = (unsigned)chunk < (unsigned)nChunks;
        const int ri = chunk * 64 + lane;
        int steps = 0;
        if (chunkOk && ri < n) {
            Ray r;
            r.ox = a.ox[ri]; r.oy = a.oy[ri]; r.oz = a.oz[ri];
            r.dx = a.dx[ri]; r.dy = a.dy[ri]; r.dz = a.dz[ri];
            r.tmin = a.tmin[ri]; r.tmax = a.tmax[ri];
            rtd::ray_setup(r);
            LdsStack<LDSD, BLOCK> st{s_ref, s_tn, a.spill, (size_t)a.spillStride, tid, gtid, a.maxStack, 0, a.status};
            float t, u, v;
            int inst, id;
            const bool hit =
                trace2l_ray<ANY, ORD>(a.tlasNodes, a.tlasTris, a.inst, a.blas, r, st, t, inst, id, u, v, &steps);
            if constexpr (ANY) {
                a.occ[ri] = hit ? 1 : 0;
            } else {
                a.t[ri] = hit ? t : RT_MISS_T;
                if (a.inst_out) a.inst_out[ri] = hit ? inst : -1;
                if (a.id) a.id[ri] = hit ? id : -1;
                if (a.u) a.u[ri] = hit ? u : 0.f;
                if (a.v) a.v[ri] = hit ? v : 0.f;
            }
        }
        if (a.chunkCost) {
            const int m = rtd::wave_max(steps);
            if (lane == 0 && chunkOk) a.chunkCost[chunk] = m;
        }
    }
}

// Refill schedule: trace2l_ray's one-short-step loop (a node visit, an
// instance entry or one triangle of the current BLAS leaf per iteration),
// with lanes whose ray finished refilled from the wave's pool of 64 rays once
// at least kRefill2Min lanes are idle. Uncounted.
constexpr int kRefill2Min = 16;

template <bool ANY, int LDSD, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace2l_refill(Trace2LArgs a) {
    __shared__ int s_ref[LDSD * BLOCK];
    __shared__ float s_tn[LDSD * BLOCK];
    const int tid = threadIdx.x, lane = tid & 63;
    const int gtid = blockIdx.x * BLOCK + tid;
    if (a.gate && *a.gate <= a.gateMax) return;
    const int n = a.nDev ? min(*a.nDev, a.n) : a.n;

    Ray rw, ro;
    int ri = 0, inst = -1;
    const uint4* nodes = a.tlasNodes;
    const float4* tris = nullptr;
    bool has = false, exhausted = false, flatW = false, flatO = false;
    int poolNext = 0, poolEnd = 0;
    int band = blockIdx.x % rtd::kBands;   // wave-uniform: band of the work distribution
    int sp = 0, ref = -1, bid = -1, binst = -1, lts = 0, ltc = 0;
    float tn = 0.f, best = 0.f, bu = 0.f, bv = 0.f;

    auto push = [&](int rf, float t) {
        if (sp < LDSD) {
            s_ref[sp * BLOCK + tid] = rf;
            s_tn[sp * BLOCK + tid] = t;
        } else if (sp < a.maxStack) {
            a.spill[(size_t)(sp - LDSD) * a.spillStride + gtid] = make_int2(rf, __float_as_int(t));
        } else {
            atomicOr(a.status, 1);
            return;
        }
        sp++;
    };

    for (;;) {
        if (!exhausted || poolNext < poolEnd) {
            const uint64_t idle = __ballot(!has);
            if (idle && (__popcll(idle) >= kRefill2Min || !__any(has))) {
                const int cnt = __popcll(idle);
                const int avail = poolEnd - poolNext;
                int fresh = n;
                if (cnt > avail && !exhausted) {
                    fresh = rtd::take_chunk(a.counter, n, band);
                    if (fresh >= n) exhausted = true;
                }
                const int k = __popcll(idle & ((1ull << lane) - 1ull));
                int w = n;
                if (k < avail) w = poolNext + k;
                else if (fresh < n) w = fresh + (k - avail);
                if (cnt > avail) {
                    if (fresh < n) { poolNext = fresh + (cnt - avail); poolEnd = fresh + 64; }
                    else { poolNext = 0; poolEnd = 0; }
                } else {
                    poolNext += cnt;
                }
                if (!has && w < n) {
                    ri = w;
                    rw.ox = a.ox[ri]; rw.oy = a.oy[ri]; rw.oz = a.oz[ri];
                    rw.dx = a.dx[ri]; rw.dy = a.dy[ri]; rw.dz = a.dz[ri];
                    rw.tmin = a.tmin[ri]; rw.tmax = a.tmax[ri];
                    rtd::ray_setup(rw);
                    flatW = rtd::ray_flat(rw);
                    has = true;
                    sp = 0; ref = 0; tn = -INFINITY; inst = -1; nodes = a.tlasNodes; ltc = 0;
                    best = rw.tmax; bu = 0.f; bv = 0.f; bid = -1; binst = -1;
                }
            }
        }
        if (!__any(has)) {
            if (exhausted && poolNext >= poolEnd) break;
            continue;
        }
        if (!has) continue;
        bool fin = false, found = false;
        if (ltc == 0) {
            while (ref == -1 || ref == kReturnMark || (!ANY && tn > best)) {
                if (ref == kReturnMark) { inst = -1; nodes = a.tlasNodes; ref = -1; continue; }
                if (sp == 0) { fin = true; break; }
                sp--;
                if (sp < LDSD) {
                    ref = s_ref[sp * BLOCK + tid];
                    tn = s_tn[sp * BLOCK + tid];
                } else {
                    const int2 e = a.spill[(size_t)(sp - LDSD) * a.spillStride + gtid];
                    ref = e.x; tn = __int_as_float(e.y);
                }
            }
            if (!fin && ref < 0 && inst >= 0) {      // BLAS leaf: start its cursor
                const int enc = -(ref + 2);
                lts = enc >> 3; ltc = (enc & 7) + 1;
                ref = -1;
            }
        }
        if (!fin) {
            if (ltc > 0) {
                const int ti = lts;
                lts++; ltc--;
                const float4 p0 = rtd::ldg_f4(tris + 3 * ti);
                float t, u, v;
                if (rtd::isect_watertight(ro, p0, rtd::ldg_f4(tris + 3 * ti + 1), rtd::ldg_f4(tris + 3 * ti + 2), t, u, v) && t > ro.tmin &&
                    t < ro.tmax) {
                    if constexpr (ANY) {
                        found = true;
                        fin = true;
                    } else {
                        const int id = __float_as_int(p0.w);
                        if (t < best || (t == best && (inst < binst || (inst == binst && id < bid)))) {
                            best = t; bid = id; binst = inst; bu = u; bv = v;
                        }
                    }
                }
            } else if (ref >= 0) {
                const bool w = inst < 0;
                const Ray& r = w ? rw : ro;
                visit2w<ANY>(nodes, r, w ? flatW : flatO, ANY ? r.tmax : best, ref, tn, push);
            } else {
                // TLAS leaf: enter its instance
                const int enc = -(ref + 2);
                const int k = __float_as_int(a.tlasTris[3 * (enc >> 3)].w);
                push(kReturnMark, -INFINITY);
                inst = k;
                const InstanceDev& I = a.inst[k];
                xform(I.w2o, rw, ro);
                flatO = rtd::ray_flat(ro);
                const BlasDev& B = a.blas[I.blas];
                nodes = B.nodes;
                tris = B.tris;
                ref = 0;
                tn = -INFINITY;
            }
        }
        if (fin) {
            if constexpr (ANY) {
                a.occ[ri] = found ? 1 : 0;
            } else {
                const bool hit = bid >= 0;
                a.t[ri] = hit ? best : RT_MISS_T;
                if (a.inst_out) a.inst_out[ri] = hit ? binst : -1;
                if (a.id) a.id[ri] = hit ? bid : -1;
                if (a.u) a.u[ri] = hit ? bu : 0.f;
                if (a.v) a.v[ri] = hit ? bv : 0.f;
            }
            has = false;
        }
    }
}

constexpr int kBlock2 = 256;

// Zeroes the kBands work-distribution counters (launched with kBands threads).
__global__ void k_reset_counter_2l(uint32_t* c) { c[threadIdx.x * rtd::kBandStride] = 0u; }

template <bool ANY, int LDSD, int SCHED>
static int launch2(Trace2LArgs& a, rt::TraceScratch& sc, int maxStack, int blocksPerCU, hipStream_t s) {
    static rt::KernelInfo ki;
    // SCHED: 0 while-while, 1 batch, 2 batch with nearest-first any hit, 3 refill
    auto kern = SCHED == 0   ? k_trace2l<ANY, LDSD, kBlock2>
                : SCHED == 1 ? k_trace2l_batch<ANY, false, LDSD, kBlock2>
                : SCHED == 2 ? k_trace2l_batch<ANY, true, LDSD, kBlock2>
                             : k_trace2l_refill<ANY, LDSD, kBlock2>;
    rt::note_launch(kern, kBlock2, ki, 0, 0, nullptr);
    int dev = 0, cus = 1;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int perCU = blocksPerCU > 0 ? std::min(blocksPerCU, ki.maxBlocks) : ki.maxBlocks;
    const int grid = perCU * cus, threads = grid * kBlock2;
    rt::note_launch(kern, kBlock2, ki, perCU, grid, sc.launchInfo);
    a.maxStack = maxStack;
    // only the batch kernels take a chunk order
    if (SCHED != 1 && SCHED != 2) { a.chunkOrder = nullptr; a.chunkCost = nullptr; }
    const int spillDepth = maxStack > LDSD ? maxStack - LDSD : 0;
    const size_t need = (size_t)spillDepth * threads * sizeof(int2);
    if (alloc_scratch(sc) != 0) return -1;
    if (grow_spill(sc, need) != 0) return -1;
    a.spill = (int2*)sc.spill;
    a.spillStride = threads;
    a.counter = sc.counter;
    a.status = sc.status;
    hipLaunchKernelGGL(k_reset_counter_2l, dim3(1), dim3(rtd::kBands), 0, s, sc.counter);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock2), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("two-level trace launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

// Stack entries a two-level traversal can need: each visited node pushes at
// most 3 entries along a root-to-leaf path, plus the return marker.
int trace2l_max_stack(const CudaSceneHandle* sc) { return 3 * sc->tlas->depth4 + 4 + 1 + 3 * sc->maxBlasDepth + 4; }

int launch_trace2l(const CudaSceneHandle* sc, rt::TraceScratch& scratch, Trace2LArgs& a, int cls, hipStream_t s) {
    const bool any = cls == 1;
    if (a.n <= 0) return 0;
    a.tlasNodes = (const uint4*)sc->tlas->d_nodes4;
    a.tlasTris = (const float4*)sc->tlas->d_tris;
    a.inst = sc->d_inst;
    a.blas = sc->d_blas;
    const int maxStack = trace2l_max_stack(sc);
    const rt::TraceConfig& c = sc->tune2l[cls >= 0 && cls <= 2 ? cls : 0];
    // counted traces always use the while-while kernel (it keeps the counters)
    const int sched = (a.cntNodes || c.schedule < 0 || c.schedule > 3 || (!any && c.schedule == 2)) ? 0 : c.schedule;
#define RT_L2(D)                                                                                                  \
    if (sched == 0) return any ? launch2<true, D, 0>(a, scratch, maxStack, c.blocksPerCU, s)                     \
                               : launch2<false, D, 0>(a, scratch, maxStack, c.blocksPerCU, s);                   \
    if (sched == 1) return any ? launch2<true, D, 1>(a, scratch, maxStack, c.blocksPerCU, s)                     \
                               : launch2<false, D, 1>(a, scratch, maxStack, c.blocksPerCU, s);                   \
    if (sched == 3) return any ? launch2<true, D, 3>(a, scratch, maxStack, c.blocksPerCU, s)                     \
                               : launch2<false, D, 3>(a, scratch, maxStack, c.blocksPerCU, s);                   \
    return launch2<true, D, 2>(a, scratch, maxStack, c.blocksPerCU, s);
    switch (c.ldsDepth) {
        case 4: { RT_L2(4) }
        case 12: { RT_L2(12) }
        case 16: { RT_L2(16) }
        default: { RT_L2(8) }
    }
#undef RT_L2
}

// World-space geometric normals of closest hits (instance, original
// primitive ID); a miss or an ID outside the scene gives (0, 0, 0).
__global__ void k_hit_normals(const InstanceDev* inst, int numInst, const BlasDev* blas, int n, const int* instId,
                              const int* primId, float* nx, float* ny, float* nz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    const int k = instId[i];
    if (k >= 0 && k < numInst) {
        const InstanceDev& I = inst[k];
        const BlasDev& B = blas[I.blas];
        const int p = primId[i];
        if (p >= 0 && p < B.numTris) world_normal(I, B, B.leafOf[p], x, y, z);
    }
    nx[i] = x; ny[i] = y; nz[i] = z;
}

int launch_hit_normals(const CudaSceneHandle* sc, int n, const int* instId, const int* primId, float* nx, float* ny,
                       float* nz, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_hit_normals, dim3((n + 255) / 256), dim3(256), 0, s, sc->d_inst, (int)sc->inst.size(),
                       sc->d_blas, n, instId, primId, nx, ny, nz);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("hit normals launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

}  // namespace rtk
