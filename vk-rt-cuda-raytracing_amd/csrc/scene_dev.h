// scene_dev.h — device pieces of the two-level traversal shared by the
// persistent traversal kernel (scene.hip) and the path tracer's tail kernel
// (pt.hip): the instance transform, the node visit, and a one-ray traversal.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "rt_device.h"
#include "scene.h"

namespace rtk {

using rtd::Ray;

constexpr int kReturnMark = INT_MIN;   // return marker: leave the instance

// World-space unit geometric normal of leaf-order triangle li of instance I's
// BLAS B: the object-space e1 x e2 (the mesh's winding, not turned toward any
// ray) mapped with the transpose of the world-to-object 3x3 (the inverse
// transpose of the object-to-world map), then normalised.
__device__ __forceinline__ void world_normal(const InstanceDev& I, const BlasDev& B, int li, float& nx, float& ny,
                                             float& nz) {
    const float4 p0 = rtd::ldg_f4(B.tris + 3 * li), p1 = rtd::ldg_f4(B.tris + 3 * li + 1),
                 p2 = rtd::ldg_f4(B.tris + 3 * li + 2);
    const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
    const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
    const float mx = e1y * e2z - e1z * e2y, my = e1z * e2x - e1x * e2z, mz = e1x * e2y - e1y * e2x;
    const float* W = I.w2o;
    nx = (W[0] * mx + W[4] * my) + W[8] * mz;
    ny = (W[1] * mx + W[5] * my) + W[9] * mz;
    nz = (W[2] * mx + W[6] * my) + W[10] * mz;
    const float nl = 1.0f / sqrtf((nx * nx + ny * ny) + nz * nz);
    nx = nx * nl; ny = ny * nl; nz = nz * nl;
}

__device__ __forceinline__ void xform(const float* M, const Ray& w, Ray& o) {
    o.ox = ((M[0] * w.ox + M[1] * w.oy) + M[2] * w.oz) + M[3];
    o.oy = ((M[4] * w.ox + M[5] * w.oy) + M[6] * w.oz) + M[7];
    o.oz = ((M[8] * w.ox + M[9] * w.oy) + M[10] * w.oz) + M[11];
    o.dx = (M[0] * w.dx + M[1] * w.dy) + M[2] * w.dz;
    o.dy = (M[4] * w.dx + M[5] * w.dy) + M[6] * w.dz;
    o.dz = (M[8] * w.dx + M[9] * w.dy) + M[10] * w.dz;
    o.tmin = w.tmin;
    o.tmax = w.tmax;
    rtd::ray_setup(o);
}

template <bool ANY, bool FLAT, class Push>
__device__ __forceinline__ void visit2(const uint4* __restrict__ nodes, const Ray& r, float limit, int& ref, float& tn,
                                       Push&& push) {
    const uint4* np = nodes + 4 * (size_t)ref;
    const uint4 qc = rtd::ldg_u4(np + 3);
    const rtd::NodeBox<FLAT> b = rtd::node_box<FLAT>(r, rtd::ldg_u4(np), rtd::ldg_u4(np + 1), rtd::ldg_u4(np + 2));
    if constexpr (ANY) {
        float d0, d1, d2, d3;
        const bool h0 = rtd::child_hit<0, FLAT>(r, b, (int)qc.x, limit, d0);
        const bool h1 = rtd::child_hit<1, FLAT>(r, b, (int)qc.y, limit, d1);
        const bool h2 = rtd::child_hit<2, FLAT>(r, b, (int)qc.z, limit, d2);
        const bool h3 = rtd::child_hit<3, FLAT>(r, b, (int)qc.w, limit, d3);
        int nxt = -1;
        if (h0) { nxt = (int)qc.x; }
        if (h1) { if (nxt != -1) push(nxt, 0.f); nxt = (int)qc.y; }
        if (h2) { if (nxt != -1) push(nxt, 0.f); nxt = (int)qc.z; }
        if (h3) { if (nxt != -1) push(nxt, 0.f); nxt = (int)qc.w; }
        ref = nxt;
        tn = 0.f;
    } else {
        float d[4];
        int c[4];
        float tt;
        if (rtd::child_hit<0, FLAT>(r, b, (int)qc.x, limit, tt)) { d[0] = tt; c[0] = (int)qc.x; } else { d[0] = INFINITY; c[0] = -1; }
        if (rtd::child_hit<1, FLAT>(r, b, (int)qc.y, limit, tt)) { d[1] = tt; c[1] = (int)qc.y; } else { d[1] = INFINITY; c[1] = -1; }
        if (rtd::child_hit<2, FLAT>(r, b, (int)qc.z, limit, tt)) { d[2] = tt; c[2] = (int)qc.z; } else { d[2] = INFINITY; c[2] = -1; }
        if (rtd::child_hit<3, FLAT>(r, b, (int)qc.w, limit, tt)) { d[3] = tt; c[3] = (int)qc.w; } else { d[3] = INFINITY; c[3] = -1; }
#define RT_CSWAP(i, j)                                                                  \
    {                                                                                   \
        const bool s = d[i] > d[j];                                                     \
        const float di = d[i], dj = d[j];                                               \
        const int ci = c[i], cj = c[j];                                                 \
        d[i] = s ? dj : di; d[j] = s ? di : dj; c[i] = s ? cj : ci; c[j] = s ? ci : cj; \
    }
        RT_CSWAP(0, 1) RT_CSWAP(2, 3) RT_CSWAP(0, 2) RT_CSWAP(1, 3) RT_CSWAP(1, 2)
#undef RT_CSWAP
        int nxt = -1;
        float nt = 0.f;
        if (c[3] != -1) { nxt = c[3]; nt = d[3]; }
        if (c[2] != -1) { if (nxt != -1) push(nxt, nt); nxt = c[2]; nt = d[2]; }
        if (c[1] != -1) { if (nxt != -1) push(nxt, nt); nxt = c[1]; nt = d[1]; }
        if (c[0] != -1) { if (nxt != -1) push(nxt, nt); nxt = c[0]; nt = d[0]; }
        ref = nxt;
        tn = nt;
    }
}

// Node visit with the slab variant picked per wave: the one that resolves
// 0*inf when any active lane's current ray (world or object space) is flat,
// the plain one otherwise; both agree on every ray that is not flat.
template <bool ANY, class Push>
__device__ __forceinline__ void visit2w(const uint4* __restrict__ nodes, const Ray& r, bool flat, float limit, int& ref,
                                        float& tn, Push&& push) {
    if (__any(flat)) visit2<ANY, true>(nodes, r, limit, ref, tn, push);
    else visit2<ANY, false>(nodes, r, limit, ref, tn, push);
}


// Stack in global memory: entries base[k * stride], k < max.
struct GlobalStack {
    int2* base;
    int stride, max, sp;
    int* status;                        // bit 0 set on overflow (nullable)
    __device__ __forceinline__ void push(int rf, float t) {
        if (sp < max) { base[(size_t)sp * stride] = make_int2(rf, __float_as_int(t)); sp++; }
        else if (status) atomicOr(status, 1);
    }
    __device__ __forceinline__ bool pop(int& rf, float& t) {
        if (sp == 0) return false;
        sp--;
        const int2 e = base[(size_t)sp * stride];
        rf = e.x; t = __int_as_float(e.y);
        return true;
    }
};

// Stack in LDS ([LDSD][BLOCK]) spilling to global memory beyond LDSD entries.
template <int LDSD, int BLOCK>
struct LdsStack {
    int* sref;
    float* stn;
    int2* spill;
    size_t stride;
    int tid, gtid, max, sp;
    int* status;
    __device__ __forceinline__ void push(int rf, float t) {
        if (sp < LDSD) {
            sref[sp * BLOCK + tid] = rf;
            stn[sp * BLOCK + tid] = t;
        } else if (sp < max) {
            spill[(size_t)(sp - LDSD) * stride + gtid] = make_int2(rf, __float_as_int(t));
        } else {
            atomicOr(status, 1);   // unreachable: max bounds a DFS of the two levels
            return;
        }
        sp++;
    }
    __device__ __forceinline__ bool pop(int& rf, float& t) {
        if (sp == 0) return false;
        sp--;
        if (sp < LDSD) {
            rf = sref[sp * BLOCK + tid];
            t = stn[sp * BLOCK + tid];
        } else {
            const int2 e = spill[(size_t)(sp - LDSD) * stride + gtid];
            rf = e.x; t = __int_as_float(e.y);
        }
        return true;
    }
};

// One ray through TLAS and BLAS: the traversal of k_trace2l for a single ray
// (same culling and tie rule; any hit may visit children nearest first,
// ORD). Closest hit: returns whether something was hit and (t, inst, id, u,
// v); any hit: whether the ray is occluded. Each loop iteration is one short
// step — a node visit, an instance entry, or the next one or two triangles
// of the current BLAS leaf (leaf cursor; the two fetched together) — so
// lanes of a wave at different kinds of step do not wait for whole leaf
// loops. As in the one-level batch kernel, the wave's priority rises one
// level per 32 steps, so that long walks (grazing rays) issue ahead of short
// ones; it starts at 0 with each call.
template <bool ANY, bool ORD, class Stack>
__device__ __forceinline__ bool trace2l_ray(const uint4* tlasNodes, const float4* tlasTris, const InstanceDev* insts,
                                            const BlasDev* blas, const Ray& rw, Stack& st, float& bt, int& binst,
                                            int& bid, float& bu, float& bv, int* stepsOut = nullptr) {
    Ray ro;
    int ref = 0, inst = -1;
    int lts = 0, ltc = 0;                   // leaf cursor: next triangle, triangles left
    float tn = -INFINITY, best = rw.tmax;
    const bool flatW = rtd::ray_flat(rw);
    bool flatO = false;
    const uint4* nodes = tlasNodes;
    const float4* tris = nullptr;
    bid = -1; binst = -1; bt = rw.tmax; bu = 0.f; bv = 0.f;
    auto push = [&](int rf, float t) { st.push(rf, t); };
    constexpr int kPrio2 = 32;
    __builtin_amdgcn_s_setprio(0);
    int steps = 0;
    for (;;) {
        if ((++steps & 15) == 0) {
            const int s = __builtin_amdgcn_readfirstlane(steps);
            if (s == kPrio2) __builtin_amdgcn_s_setprio(1);
            else if (s == 2 * kPrio2) __builtin_amdgcn_s_setprio(2);
            else if (s == 3 * kPrio2) __builtin_amdgcn_s_setprio(3);
        }
        if (ltc == 0) {
            while (ref == -1 || ref == kReturnMark || (!ANY && tn > best)) {
                if (ref == kReturnMark) { inst = -1; nodes = tlasNodes; ref = -1; continue; }
                if (!st.pop(ref, tn)) {
                    if (stepsOut) *stepsOut = steps;
                    return ANY ? false : bid >= 0;
                }
            }
            if (ref < 0 && inst >= 0) {         // BLAS leaf: start its cursor
                const int enc = -(ref + 2);
                lts = enc >> 3; ltc = (enc & 7) + 1;
                ref = -1;
            }
        }
        if (ltc > 0) {
            const int ti = lts;
            const bool two = ltc > 1;
            const int tj = two ? ti + 1 : ti;
            const float4 p0 = rtd::ldg_f4(tris + 3 * ti), p1 = rtd::ldg_f4(tris + 3 * ti + 1), p2 = rtd::ldg_f4(tris + 3 * ti + 2);
            const float4 q0 = rtd::ldg_f4(tris + 3 * tj), q1 = rtd::ldg_f4(tris + 3 * tj + 1), q2 = rtd::ldg_f4(tris + 3 * tj + 2);
            lts += two ? 2 : 1;
            ltc -= two ? 2 : 1;
            // leaf order: the first triangle, then the second (a closest-hit
            // update of the first is seen by the second's tie rule)
            auto test = [&](const float4& a0, const float4& a1, const float4& a2) {
                float t, u, v;
                if (rtd::isect_watertight(ro, a0, a1, a2, t, u, v) && t > ro.tmin && t < ro.tmax) {
                    if constexpr (ANY) {
                        return true;
                    } else {
                        const int id = __float_as_int(a0.w);
                        if (t < best || (t == best && (inst < binst || (inst == binst && id < bid)))) {
                            best = t; bid = id; binst = inst; bt = t; bu = u; bv = v;
                        }
                    }
                }
                return false;
            };
            if ((test(p0, p1, p2)) || (two && test(q0, q1, q2))) {
                if (stepsOut) *stepsOut = steps;
                return true;
            }
        } else if (ref >= 0) {
            const bool w = inst < 0;
            const Ray& r = w ? rw : ro;
            if constexpr (ANY && ORD) visit2w<false>(nodes, r, w ? flatW : flatO, r.tmax, ref, tn, push);
            else visit2w<ANY>(nodes, r, w ? flatW : flatO, ANY ? r.tmax : best, ref, tn, push);
        } else {
            // TLAS leaf: enter its instance
            const int enc = -(ref + 2);
            const int k = __float_as_int(tlasTris[3 * (enc >> 3)].w);
            push(kReturnMark, -INFINITY);
            inst = k;
            const InstanceDev& I = insts[k];
            xform(I.w2o, rw, ro);
            flatO = rtd::ray_flat(ro);
            const BlasDev& B = blas[I.blas];
            nodes = B.nodes;
            tris = B.tris;
            ref = 0;
            tn = -INFINITY;
        }
    }
}

// Global-stack form used by the path tracer's tail kernel.
template <bool ANY>
__device__ __forceinline__ bool trace2l_one(const uint4* tlasNodes, const float4* tlasTris, const InstanceDev* insts,
                                            const BlasDev* blas, const Ray& rw, int2* stk, int stride, int maxStack,
                                            int* status, float& bt, int& binst, int& bid) {
    GlobalStack st{stk, stride, maxStack, 0, status};
    float bu, bv;
    return trace2l_ray<ANY, false>(tlasNodes, tlasTris, insts, blas, rw, st, bt, binst, bid, bu, bv);
}

}  // namespace rtk
