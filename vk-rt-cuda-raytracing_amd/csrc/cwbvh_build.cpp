// cwbvh_build.cpp — compressed 8-wide BVH (the reference's CWBVH,
// cuda_rt_v37.cu:641-900, after Ylitie, Karras and Laine 2017) built from the
// binned-SAH BVH2.
//
// Collapse: each 8-wide node starts from the two children of a BVH2 node and
// repeatedly opens the internal child with the largest surface area until it
// has 8 children or only leaves (cuda_rt_v37.cu:670-692). Encoding (80 bytes,
// 5 x uint4):
//   n0: origin p (3 floats), e.x | e.y << 8 | e.z << 16 | imask << 24
//   n1: childBase, triBase, meta[0..3], meta[4..7]
//   n2: qlo.x[0..7], qhi.x[0..7]   n3: the same for y   n4: for z
// Child boxes are p + q * 2^e per axis with 8-bit q. Unlike the reference,
// which rounds with floor/ceil in float and clamps (cuda_rt_v37.cu:836-841),
// every quantized bound is checked in the exact float arithmetic the kernel
// dequantizes with (p + q * 2^e, one rounding) and stepped outward until the
// child box is contained: no hit can be lost.
// meta: 0xFF empty slot; internal child: 0x80 | rank (its node is
// childBase + rank; internal children are stored contiguously); leaf:
// (count - 1) << 5 | offset (its triangles are triBase + offset .. + count - 1
// in the CWBVH triangle order, which this builder also emits).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rt_internal.h"

namespace rt {

namespace {

float area(const Box& b) {
    const float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

// Smallest e with 255 * 2^e >= extent (extent >= 0), within the normal range.
int exponent_for(double extent) {
    if (!(extent > 0)) return -126;
    int e = (int)std::ceil(std::log2(extent / 255.0));
    while (e > -126 && std::ldexp(255.0, e - 1) >= extent) e--;
    while (std::ldexp(255.0, e) < extent) e++;
    return std::max(-126, std::min(127, e));
}

// Quantized lower / upper bound of [lo, hi] relative to p with scale s = 2^e,
// stepped until p + q*s (float, one rounding) brackets the bound.
int q_lo(float p, float s, float lo) {
    int q = (int)std::floor(((double)lo - (double)p) / (double)s);
    q = std::max(0, std::min(255, q));
    while (q > 0 && p + (float)q * s > lo) q--;
    return q;
}
int q_hi(float p, float s, float hi) {
    int q = (int)std::ceil(((double)hi - (double)p) / (double)s);
    q = std::max(0, std::min(255, q));
    while (q < 255 && p + (float)q * s < hi) q++;
    return q;
}

}  // namespace

int build_cwbvh(const std::vector<Node2>& n2, std::vector<uint32_t>& nodes, std::vector<int>& triOrder,
                int& depth) {
    nodes.clear();
    triOrder.clear();
    depth = 0;
    if (n2.empty()) return -1;
    struct Job { int b2; int slot; int level; };   // BVH2 node to encode as 8-wide node `slot`
    std::vector<Job> jobs;
    nodes.resize(20);
    if (n2[0].triCount > 0) {
        // a single leaf: one node with one leaf child
        jobs.push_back({-1, 0, 1});
    } else {
        jobs.push_back({0, 0, 1});
    }
    size_t head = 0;
    while (head < jobs.size()) {
        const Job job = jobs[head++];
        depth = std::max(depth, job.level);
        int kids[8], nk = 0;
        if (job.b2 < 0) {
            kids[nk++] = 0;
        } else {
            kids[nk++] = n2[job.b2].left;
            kids[nk++] = n2[job.b2].right;
            while (nk < 8) {
                int best = -1;
                float bestA = -1.f;
                for (int i = 0; i < nk; i++) {
                    const Node2& c = n2[kids[i]];
                    if (c.triCount == 0 && area(c.box) > bestA) { bestA = area(c.box); best = i; }
                }
                if (best < 0) break;
                const Node2& o = n2[kids[best]];
                kids[best] = o.left;
                kids[nk++] = o.right;
            }
        }
        Box nb;
        for (int a = 0; a < 3; a++) { nb.mn[a] = INFINITY; nb.mx[a] = -INFINITY; }
        for (int i = 0; i < nk; i++)
            for (int a = 0; a < 3; a++) {
                nb.mn[a] = std::min(nb.mn[a], n2[kids[i]].box.mn[a]);
                nb.mx[a] = std::max(nb.mx[a], n2[kids[i]].box.mx[a]);
            }
        int e[3];
        float s[3];
        for (int a = 0; a < 3; a++) {
            e[a] = exponent_for((double)nb.mx[a] - (double)nb.mn[a]);
            s[a] = std::ldexp(1.0f, e[a]);
        }
        uint8_t meta[8], qlo[3][8], qhi[3][8];
        uint8_t imask = 0;
        int childBase = (int)(nodes.size() / 20), triBase = (int)triOrder.size();
        int rank = 0, off = 0;
        for (int i = 0; i < 8; i++) {
            if (i >= nk) {
                meta[i] = 0xFF;
                for (int a = 0; a < 3; a++) { qlo[a][i] = 255; qhi[a][i] = 0; }
                continue;
            }
            const Node2& c = n2[kids[i]];
            for (int a = 0; a < 3; a++) {
                qlo[a][i] = (uint8_t)q_lo(nb.mn[a], s[a], c.box.mn[a]);
                qhi[a][i] = (uint8_t)q_hi(nb.mn[a], s[a], c.box.mx[a]);
            }
            if (c.triCount > 0) {
                if (c.triCount > 4) return -1;              // the encoding holds 1-4 triangles
                meta[i] = (uint8_t)(((c.triCount - 1) << 5) | off);
                for (int k = 0; k < c.triCount; k++) triOrder.push_back(c.triStart + k);
                off += c.triCount;
            } else {
                meta[i] = (uint8_t)(0x80 | rank);
                imask |= (uint8_t)(1u << i);
                jobs.push_back({kids[i], childBase + rank, job.level + 1});
                rank++;
            }
        }
        if (off > 32) return -1;
        nodes.resize(nodes.size() + 20 * (size_t)rank);
        uint32_t* w = &nodes[20 * (size_t)job.slot];
        float p[3] = {nb.mn[0], nb.mn[1], nb.mn[2]};
        memcpy(&w[0], p, 12);
        w[3] = (uint32_t)(uint8_t)(int8_t)e[0] | (uint32_t)(uint8_t)(int8_t)e[1] << 8 |
               (uint32_t)(uint8_t)(int8_t)e[2] << 16 | (uint32_t)imask << 24;
        w[4] = (uint32_t)childBase;
        w[5] = (uint32_t)triBase;
        memcpy(&w[6], meta, 8);
        for (int a = 0; a < 3; a++) {
            memcpy(&w[8 + 4 * a], qlo[a], 8);
            memcpy(&w[10 + 4 * a], qhi[a], 8);
        }
    }
    return (int)(nodes.size() / 20);
}

}  // namespace rt
