// build.hip — binned-SAH BVH2 -> BVH4 build on the GPU.
//
// The algorithm is the reference's host builder (layer/cuda_bvh_backend.cu
// :201-251: 16 centroid bins per axis, bin = floor((c-cmin)/(cmax-cmin)*15),
// cost = nL*SA(L)/SA(P) + nR*SA(R)/SA(P) + 1, leaves of <= maxLeaf
// triangles) with two rules made exact: the partition is taken on the
// evaluated bin boundary and is stable, and a node whose centroids coincide on
// every axis splits at the middle of its range. Under those rules the tree is
// a pure function of the input, and this level-synchronous GPU build yields
// bit for bit the tree of a recursive host build:
//
//  * per level, every primitive of an unfinished node is binned on the three
//    axes with atomics (box min/max as order-preserving integers, which makes
//    the result independent of atomic order; LDS-privatised when a workgroup's
//    chunk lies in one node);
//  * one thread per node evaluates the 45 SAH candidates in the host order;
//  * a device-wide exclusive scan of the go-left flags gives every primitive
//    its stable position in the child range;
//  * child boxes and centroid bounds are min/max reductions over the moved
//    primitives;
//  * finally nodes are renumbered in depth-first pre-order (subtree sizes
//    bottom-up, offsets top-down), and the BVH2 is collapsed into the
//    reference's 64-byte BVH4 nodes (backend :31-34, :391-418), FP16 bounds
//    rounded outward. Which BVH2 nodes become BVH4 nodes is a surface-area
//    choice (dynamic programming over the BVH2, bottom-up costs then top-down
//    slot counts, one launch per BVH2 level each way): the least total area
//    of BVH4 nodes, i.e. the fewest expected node visits, with up to four
//    slots per node. The reference opens both children of every other level,
//    which leaves 45 % of the nodes 2-wide on the 3.4 M-triangle terrain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "rt_internal.h"
#include "rt_scan.h"

namespace rtb {

constexpr int kBins = 16;
constexpr int kBinInts = 7;                 // box min xyz, max xyz (ordered ints), count
constexpr int kAxisInts = kBins * kBinInts;
constexpr int kSlotInts = 3 * kAxisInts;
constexpr int kThreads = 256;
constexpr int kChunk = 2048;                // primitives per workgroup in the binning passes

__device__ __forceinline__ int f2o(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }
constexpr int kOrdPosInf = 0x7f800000;      // f2o(+inf)
constexpr int kOrdNegInf = (int)0x807fffff; // f2o(-inf)

__device__ __forceinline__ float area(float x0, float y0, float z0, float x1, float y1, float z1) {
    const float dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ int bin_of(float c, float cmin, float cmax) {
    int b = (int)((c - cmin) / (cmax - cmin) * (float)(kBins - 1));
    return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

struct Build {
    int n, maxLeaf;
    const float* pbox;      // [n][6] primitive boxes, original index
    const float4* cen;      // [n] centroids
    int* idx;               // [n] primitive at position
    int* pnode;             // [n] node of position
    int* idx2;
    int* pnode2;
    int* flag;              // [n] go-left flag
    int* rank;              // [n] exclusive scan of flag
    int* nbox;              // [cap][6] ordered ints
    int* ncen;              // [cap][6] ordered ints (centroid bounds)
    int* ns;                // [cap] range start
    int* ncnt;              // [cap] primitive count
    int* nleft;             // [cap] -1 for a leaf
    int* nright;
    int* ndepth;
    int* nodeSlot;          // [cap] slot of an active node
    int* act;               // [n] active nodes of the level
    int* bins;              // [m][3][16][7]
    int* sAxis;             // [m] split axis (-1: middle split)
    int* sBin;
    int* sNl;               // left count
    int* small;             // [cap] node whose level was done by k_small_level (bounds written there)
    int* bigSlot;           // [cap] bins slot of a node of more than kSmall primitives
    int* bigCount;          // bins slots handed out this level
};

// Nodes of at most kSmall primitives take one level per launch inside one
// wave (k_small_level): LDS bins, SAH, stable partition and child bounds
// without global atomics. Larger nodes use the device-wide passes.
constexpr int kSmall = 2048;

// ---------------------------------------------------------------------------
// Scene of primitive i: the last s with off[s] <= i (off: numScenes+1 ascending offsets).
__device__ __forceinline__ int scene_of(const int* off, int numScenes, int i) {
    int lo = 0, hi = numScenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_prims(const float* __restrict__ tris, int n, float* pbox, float4* cen, int* idx, int* pnode,
                        const int* off, int numScenes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* t = tris + 9 * (size_t)i;
    float mn[3], mx[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        mn[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]) + 0.0f;   // + 0: -0 -> +0
        mx[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]) + 0.0f;
        pbox[6 * (size_t)i + a] = mn[a];
        pbox[6 * (size_t)i + 3 + a] = mx[a];
    }
    cen[i] = make_float4((mn[0] + mx[0]) * .5f, (mn[1] + mx[1]) * .5f, (mn[2] + mx[2]) * .5f, 0.f);
    idx[i] = i;
    pnode[i] = numScenes == 1 ? 0 : scene_of(off, numScenes, i);
}

// Roots: node s spans scene s's primitives; slots[s] is its active slot (or -1).
__global__ void k_init_roots(Build b, const int* off, const int* slots, int numScenes) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= numScenes) return;
    b.ns[s] = off[s]; b.ncnt[s] = off[s + 1] - off[s];
    b.nleft[s] = -1; b.nright[s] = -1; b.ndepth[s] = 0;
    b.nodeSlot[s] = slots[s];
    if (slots[s] >= 0) b.act[slots[s]] = s;
}

__global__ void k_init_nodes(int* nbox, int* ncen, int* small, int first, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    small[first + i] = 0;
    const size_t o = 6 * (size_t)(first + i);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nbox[o + a] = kOrdPosInf; nbox[o + 3 + a] = kOrdNegInf;
        ncen[o + a] = kOrdPosInf; ncen[o + 3 + a] = kOrdNegInf;
    }
}

// Box and centroid bounds of every node at or above levelBase, reduced over
// its primitives (LDS first when the chunk lies in one node).
__global__ __launch_bounds__(kThreads) void k_node_bounds(Build b, int levelBase) {
    // The nodes bounded here (children of nodes of more than kSmall
    // primitives) are at most four per chunk: at most two such parents
    // overlap kChunk consecutive positions. Each gets an LDS accumulator,
    // claimed by node id; global atomics only if a fifth ever showed up.
    __shared__ int key[4];
    __shared__ int lb[4][12];
    const int c0 = blockIdx.x * kChunk, c1 = min(c0 + kChunk, b.n);
    if (threadIdx.x < 4) key[threadIdx.x] = -1;
    if (threadIdx.x < 48) lb[threadIdx.x / 12][threadIdx.x % 12] = (threadIdx.x % 6) < 3 ? kOrdPosInf : kOrdNegInf;
    __syncthreads();
    for (int i = c0 + threadIdx.x; i < c1; i += kThreads) {
        const int node = b.pnode[i];
        if (node < levelBase || b.small[node]) continue;
        int j = 0;
        for (; j < 4; j++) {
            const int k = key[j];
            if (k == node) break;
            if (k == -1) {
                const int prev = atomicCAS(&key[j], -1, node);
                if (prev == -1 || prev == node) break;
            }
        }
        const int p = b.idx[i];
        const float* pb = b.pbox + 6 * (size_t)p;
        const float4 c = b.cen[p];
        const float cv[3] = {c.x, c.y, c.z};
        int* nb = j < 4 ? &lb[j][0] : b.nbox + 6 * (size_t)node;
        int* nc = j < 4 ? &lb[j][6] : b.ncen + 6 * (size_t)node;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            atomicMin(&nb[a], f2o(pb[a])); atomicMax(&nb[3 + a], f2o(pb[3 + a]));
            atomicMin(&nc[a], f2o(cv[a])); atomicMax(&nc[3 + a], f2o(cv[a]));
        }
    }
    __syncthreads();
    if (threadIdx.x < 48) {
        const int j = threadIdx.x / 12, k = threadIdx.x % 12, u = key[j];
        if (u >= 0) {
            int* dst = (k < 6 ? b.nbox : b.ncen) + 6 * (size_t)u + (k % 6);
            if ((k % 6) < 3) atomicMin(dst, lb[j][k]); else atomicMax(dst, lb[j][k]);
        }
    }
}

// Bins slots of the level's nodes of more than kSmall primitives (the others
// bin in LDS): at most n / (kSmall + 1) + numScenes of them; which node gets
// which slot does not matter.
__global__ void k_big_slots(Build b, int m) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= m) return;
    const int node = b.act[slot];
    if (b.ncnt[node] > kSmall) b.bigSlot[node] = atomicAdd(b.bigCount, 1);
}

__global__ void k_zero_int(int* w) { *w = 0; }

__global__ void k_clear_bins(int* bins, int slots) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots * kSlotInts) return;
    const int f = i % kBinInts;
    bins[i] = f < 3 ? kOrdPosInf : (f < 6 ? kOrdNegInf : 0);
}

__device__ __forceinline__ void bin_prim(const Build& b, int node, int p, int* dst /*[3][16][7]*/) {
    const float* pb = b.pbox + 6 * (size_t)p;
    const float4 c = b.cen[p];
    const float cv[3] = {c.x, c.y, c.z};
    const int* nc = b.ncen + 6 * (size_t)node;
    int bx[6];
#pragma unroll
    for (int k = 0; k < 6; k++) bx[k] = f2o(pb[k]);
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const float cmin = o2f(nc[ax]), cmax = o2f(nc[3 + ax]);
        if (cmax - cmin < 1e-8f) continue;
        int* d = dst + ax * kAxisInts + bin_of(cv[ax], cmin, cmax) * kBinInts;
#pragma unroll
        for (int k = 0; k < 3; k++) { atomicMin(&d[k], bx[k]); atomicMax(&d[3 + k], bx[3 + k]); }
        atomicAdd(&d[6], 1);
    }
}

__global__ __launch_bounds__(kThreads) void k_bin(Build b) {
    // A node of more than kSmall (>= kChunk) primitives that overlaps this
    // chunk holds its first or its last position, so the chunk bins into at
    // most two LDS copies (those two nodes) and flushes them once.
    __shared__ int lb[2][kSlotInts];
    __shared__ int nodes2[2];
    const int c0 = blockIdx.x * kChunk, c1 = min(c0 + kChunk, b.n);
    if (threadIdx.x == 0) {
        const int a = b.pnode[c0], z = b.pnode[c1 - 1];
        nodes2[0] = b.ncnt[a] > kSmall ? a : -1;
        nodes2[1] = (z != a && b.ncnt[z] > kSmall) ? z : -1;
    }
    for (int i = threadIdx.x; i < 2 * kSlotInts; i += kThreads) {
        const int f = (i % kSlotInts) % kBinInts;
        lb[i / kSlotInts][i % kSlotInts] = f < 3 ? kOrdPosInf : (f < 6 ? kOrdNegInf : 0);
    }
    __syncthreads();
    const int na = nodes2[0], nz = nodes2[1];
    if (na < 0 && nz < 0) return;
    for (int i = c0 + threadIdx.x; i < c1; i += kThreads) {
        const int node = b.pnode[i];
        if (node == na) bin_prim(b, node, b.idx[i], lb[0]);
        else if (node == nz) bin_prim(b, node, b.idx[i], lb[1]);
    }
    __syncthreads();
    for (int q = 0; q < 2; q++) {
        const int u = nodes2[q];
        if (u < 0) continue;
        int* g = b.bins + (size_t)b.bigSlot[u] * kSlotInts;
        for (int i = threadIdx.x; i < kSlotInts; i += kThreads) {
            const int f = i % kBinInts;
            const int v = lb[q][i];
            if (f == 6) { if (v) atomicAdd(&g[i], v); }
            else if (f < 3) { if (v != kOrdPosInf) atomicMin(&g[i], v); }
            else if (v != kOrdNegInf) atomicMax(&g[i], v);
        }
    }
}

// SAH over 3 axes x 15 boundaries in the host order (axis-major, strict <)
// from the node's bins ([3][16][7], global or LDS): split axis (-1: middle
// split), last left bin and left count.
__device__ __forceinline__ void sah_choose(const int* bins, const int* nb, const int* nc, int cnt, int& bestAxis,
                                           int& bestBin, int& bestNl) {
    const float pA = area(o2f(nb[0]), o2f(nb[1]), o2f(nb[2]), o2f(nb[3]), o2f(nb[4]), o2f(nb[5]));
    float bestCost = 1e30f;
    bestAxis = -1; bestBin = -1; bestNl = 0;
    for (int ax = 0; ax < 3; ax++) {
        const float cmin = o2f(nc[ax]), cmax = o2f(nc[3 + ax]);
        if (cmax - cmin < 1e-8f) continue;
        const int* B = bins + ax * kAxisInts;
        float R[kBins][6];
        int rc[kBins];
        for (int k = 0; k < 6; k++) R[kBins - 1][k] = o2f(B[(kBins - 1) * kBinInts + k]);
        rc[kBins - 1] = B[(kBins - 1) * kBinInts + 6];
        for (int q = kBins - 2; q >= 0; q--) {
            for (int k = 0; k < 3; k++) {
                R[q][k] = fminf(R[q + 1][k], o2f(B[q * kBinInts + k]));
                R[q][3 + k] = fmaxf(R[q + 1][3 + k], o2f(B[q * kBinInts + 3 + k]));
            }
            rc[q] = rc[q + 1] + B[q * kBinInts + 6];
        }
        float L[6];
        for (int k = 0; k < 6; k++) L[k] = o2f(B[k]);
        int lc = B[6];
        for (int q = 0; q < kBins - 1; q++) {
            if (q > 0) {
                for (int k = 0; k < 3; k++) {
                    L[k] = fminf(L[k], o2f(B[q * kBinInts + k]));
                    L[3 + k] = fmaxf(L[3 + k], o2f(B[q * kBinInts + 3 + k]));
                }
                lc += B[q * kBinInts + 6];
            }
            if (lc == 0 || rc[q + 1] == 0) continue;
            const float cost = (float)lc * area(L[0], L[1], L[2], L[3], L[4], L[5]) / pA +
                               (float)rc[q + 1] * area(R[q + 1][0], R[q + 1][1], R[q + 1][2], R[q + 1][3],
                                                        R[q + 1][4], R[q + 1][5]) / pA + 1.f;
            if (cost < bestCost) { bestCost = cost; bestAxis = ax; bestBin = q; bestNl = lc; }
        }
    }
    if (bestAxis < 0) bestNl = cnt / 2;                      // coincident centroids: middle split
}

// Children of a split node: ids childBase + 2 slot (+1), ranges, depth.
__device__ __forceinline__ void make_children(const Build& b, int node, int slot, int childBase, int nl) {
    const int L = childBase + 2 * slot, R = L + 1;
    const int s = b.ns[node], d = b.ndepth[node] + 1, cnt = b.ncnt[node];
    b.nleft[node] = L; b.nright[node] = R;
    b.ns[L] = s; b.ncnt[L] = nl; b.ns[R] = s + nl; b.ncnt[R] = cnt - nl;
    b.ndepth[L] = d; b.ndepth[R] = d;
    b.nleft[L] = b.nright[L] = b.nleft[R] = b.nright[R] = -1;
}

// One thread per active node of more than kSmall primitives: SAH from the
// device-wide bins, then the two children.
__global__ void k_sah(Build b, int m, int childBase) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= m) return;
    const int node = b.act[slot];
    const int cnt = b.ncnt[node];
    if (cnt <= kSmall) return;
    int ax, bin, nl;
    sah_choose(b.bins + (size_t)b.bigSlot[node] * kSlotInts, b.nbox + 6 * (size_t)node, b.ncen + 6 * (size_t)node, cnt, ax,
               bin, nl);
    b.sAxis[slot] = ax;
    b.sBin[slot] = bin;
    b.sNl[slot] = nl;
    make_children(b, node, slot, childBase, nl);
}

// One level of every active node of at most kSmall primitives, one wave per
// node: bins in LDS, the SAH choice, the two children, the stable partition
// of its primitives into idx2 / pnode2 (ballot ranks, left ones first) and
// the children's box and centroid bounds, written directly (small[child] = 1
// tells k_node_bounds to leave them alone). Same results as the device-wide
// passes, which skip these nodes.
__global__ __launch_bounds__(64) void k_small_level(Build b, int m, int childBase) {
    __shared__ int lb[kSlotInts];
    __shared__ int cb[24];                   // left box, left centroids, right box, right centroids
    __shared__ int choice[3];
    const int slot = blockIdx.x;
    if (slot >= m) return;
    const int node = b.act[slot];
    const int cnt = b.ncnt[node];
    if (cnt > kSmall) return;                // uniform per workgroup
    const int lane = threadIdx.x;
    for (int i = lane; i < kSlotInts; i += 64) {
        const int f = i % kBinInts;
        lb[i] = f < 3 ? kOrdPosInf : (f < 6 ? kOrdNegInf : 0);
    }
    if (lane < 24) cb[lane] = (lane % 6) < 3 ? kOrdPosInf : kOrdNegInf;
    __syncthreads();
    const int s = b.ns[node];
    for (int i = lane; i < cnt; i += 64) bin_prim(b, node, b.idx[s + i], lb);
    __syncthreads();
    if (lane == 0) {
        int ax, bin, nl;
        sah_choose(lb, b.nbox + 6 * (size_t)node, b.ncen + 6 * (size_t)node, cnt, ax, bin, nl);
        choice[0] = ax; choice[1] = bin; choice[2] = nl;
        b.sAxis[slot] = ax; b.sBin[slot] = bin; b.sNl[slot] = nl;
        make_children(b, node, slot, childBase, nl);
    }
    __syncthreads();
    const int ax = choice[0], bin = choice[1], nl = choice[2];
    const int L = childBase + 2 * slot, R = L + 1;
    const int* nc = b.ncen + 6 * (size_t)node;
    const float cmin = ax >= 0 ? o2f(nc[ax]) : 0.f, cmax = ax >= 0 ? o2f(nc[3 + ax]) : 0.f;
    int leftBefore = 0;
    for (int base = 0; base < cnt; base += 64) {
        const int i = base + lane;
        const bool valid = i < cnt;
        int p = 0;
        bool left = false;
        if (valid) {
            p = b.idx[s + i];
            if (ax < 0) {
                left = i < nl;
            } else {
                const float4 c = b.cen[p];
                const float cv = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
                left = bin_of(cv, cmin, cmax) <= bin;
            }
        }
        const uint64_t lm = __ballot(valid && left);
        const int lr = __popcll(lm & ((1ull << lane) - 1ull));
        if (valid) {
            const int pos = left ? s + leftBefore + lr : s + nl + (i - leftBefore - lr);
            b.idx2[pos] = p;
            b.pnode2[pos] = left ? L : R;
            const float* pb = b.pbox + 6 * (size_t)p;
            const float4 c = b.cen[p];
            const float cv[3] = {c.x, c.y, c.z};
            int* d = cb + (left ? 0 : 12);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                atomicMin(&d[a], f2o(pb[a])); atomicMax(&d[3 + a], f2o(pb[3 + a]));
                atomicMin(&d[6 + a], f2o(cv[a])); atomicMax(&d[9 + a], f2o(cv[a]));
            }
        }
        leftBefore += __popcll(lm);
    }
    __syncthreads();
    if (lane < 24) {
        const int child = lane < 12 ? L : R, k = lane % 12;
        int* dst = (k < 6 ? b.nbox : b.ncen) + 6 * (size_t)child + (k % 6);
        *dst = cb[lane];
    }
    if (lane == 0) { b.small[L] = 1; b.small[R] = 1; }
}

__global__ void k_flags(Build b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    const int node = b.pnode[i];
    int f = 0;
    if (b.ncnt[node] > b.maxLeaf && b.ncnt[node] > kSmall) {
        const int slot = b.nodeSlot[node];
        const int ax = b.sAxis[slot];
        if (ax < 0) {
            f = (i - b.ns[node]) < b.sNl[slot];
        } else {
            const float4 c = b.cen[b.idx[i]];
            const float cv = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
            const int* nc = b.ncen + 6 * (size_t)node;
            f = bin_of(cv, o2f(nc[ax]), o2f(nc[3 + ax])) <= b.sBin[slot];
        }
    }
    b.flag[i] = f;
}

__global__ void k_scatter(Build b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    const int node = b.pnode[i];
    if (b.ncnt[node] <= b.maxLeaf) {
        b.idx2[i] = b.idx[i];
        b.pnode2[i] = node;
        return;
    }
    if (b.ncnt[node] <= kSmall) return;          // partitioned by k_small_level
    const int slot = b.nodeSlot[node];
    const int s = b.ns[node], nl = b.sNl[slot];
    const int lr = b.rank[i] - b.rank[s];
    int pos, child;
    if (b.flag[i]) { pos = s + lr; child = b.nleft[node]; }
    else { pos = s + nl + ((i - s) - lr); child = b.nright[node]; }
    b.idx2[pos] = b.idx[i];
    b.pnode2[pos] = child;
}

__global__ void k_next_active(Build b, int first, int count, const int* rank, int* act2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int node = first + i;
    if (b.ncnt[node] > b.maxLeaf) {
        act2[rank[i]] = node;
        b.nodeSlot[node] = rank[i];
    }
}

__global__ void k_active_flags(const int* ncnt, int first, int count, int maxLeaf, int* f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) f[i] = ncnt[first + i] > maxLeaf;
}

// ------------------------------- pre-order renumbering -----------------------
__global__ void k_sizes(const int* nleft, const int* nright, int* size, int first, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int v = first + i;
    size[v] = nleft[v] < 0 ? 1 : 1 + size[nleft[v]] + size[nright[v]];
}

// Pre-order positions of the roots: trees one after the other.
__global__ void k_root_pre(const int* size, int* pre, int numScenes) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int acc = 0;
    for (int s = 0; s < numScenes; s++) { pre[s] = acc; acc += size[s]; }
}

__global__ void k_root4(const int* pre, const int* b4id, int numScenes, int* root4) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < numScenes) root4[s] = b4id[pre[s]];
}

__global__ void k_preorder(const int* nleft, const int* nright, const int* size, int* pre, int first, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int v = first + i;
    if (nleft[v] < 0) return;
    pre[nleft[v]] = pre[v] + 1;
    pre[nright[v]] = pre[v] + 1 + size[nleft[v]];
}

// Pre-order BVH2 (rt::Node2 layout: box[6], left, right, triStart, triCount)
// and the collapse's starting state: every tree root presents itself as one
// slot (assign 1) and is the head of its BVH4 root at BVH4 level 0.
__global__ void k_emit_bvh2(Build b, const int* pre, int count, rt::Node2* out, int* head, int* assign, int* lev4) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= count) return;
    const int p = pre[v];
    rt::Node2 nd;
    for (int k = 0; k < 3; k++) {
        nd.box.mn[k] = o2f(b.nbox[6 * (size_t)v + k]);
        nd.box.mx[k] = o2f(b.nbox[6 * (size_t)v + 3 + k]);
    }
    const bool leaf = b.nleft[v] < 0;
    nd.left = leaf ? 0 : pre[b.nleft[v]];
    nd.right = leaf ? 0 : pre[b.nright[v]];
    nd.triStart = leaf ? b.ns[v] : 0;
    nd.triCount = leaf ? b.ncnt[v] : 0;
    out[p] = nd;
    head[p] = 0;
    assign[p] = b.ndepth[v] == 0 ? 1 : 0;
    lev4[p] = 0;
}

// --- BVH2 -> BVH4 collapse (same rule and float operation order as the
// oracle's collapse_costs / collapse_split / collapse_gather) ---
// cost[p] = (c1, c2, c3, c4): least total area of the BVH4 nodes needed to
// present subtree p as k slots; a leaf is one free slot; for an internal p,
// g_k = min over a of c_a(left) + c_{k-a}(right), c1 = area(p) + min g_k.
constexpr float kCostInf = 3.0e38f;

__global__ void k_collapse_cost(const rt::Node2* n2, const int* pre, int first, int count, float4* cost) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int p = pre[first + i];
    const rt::Node2& nd = n2[p];
    if (nd.triCount > 0) { cost[p] = make_float4(0.0f, kCostInf, kCostInf, kCostInf); return; }
    const float4 L = cost[nd.left], R = cost[nd.right];
    const float l[4] = {L.x, L.y, L.z, L.w}, r[4] = {R.x, R.y, R.z, R.w};
    float g[5] = {kCostInf, kCostInf, kCostInf, kCostInf, kCostInf};
    for (int k = 2; k <= 4; k++)
        for (int a = 1; a < k; a++) {
            const float x = l[a - 1] + r[k - a - 1];
            if (x < g[k]) g[k] = x;
        }
    float best = g[2];
    if (g[3] < best) best = g[3];
    if (g[4] < best) best = g[4];
    const float ar = area(nd.box.mn[0], nd.box.mn[1], nd.box.mn[2], nd.box.mx[0], nd.box.mx[1], nd.box.mx[2]);
    cost[p] = make_float4(ar + best, g[2], g[3], g[4]);
}

// Left slot count of subtree p presented as m slots (ties: the smaller).
__device__ __forceinline__ int collapse_left(const rt::Node2& nd, const float4* cost, int m) {
    const float4 L = cost[nd.left], R = cost[nd.right];
    const float l[4] = {L.x, L.y, L.z, L.w}, r[4] = {R.x, R.y, R.z, R.w};
    int ba = 1;
    float bx = kCostInf;
    for (int a = 1; a < m; a++) {
        const float x = l[a - 1] + r[m - a - 1];
        if (x < bx) { bx = x; ba = a; }
    }
    return ba;
}

// Top-down, one BVH2 level per launch: an internal node presented as one
// slot heads a BVH4 node of its cheapest width k (ties: the narrower) and
// hands its children their slot counts; a node presented as m >= 2 slots is
// dissolved into its parent's node and splits m between its children. lev4:
// BVH4 level of a head / of the node whose slots a dissolved node fills.
__global__ void k_collapse_assign(const rt::Node2* n2, const int* pre, int first, int count, const float4* cost,
                                  int* assign, int* head, int* lev4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int p = pre[first + i];
    const rt::Node2& nd = n2[p];
    const int m = assign[p];
    if (nd.triCount > 0 || m == 0) return;
    int k = m, cl = lev4[p];
    if (m == 1) {
        const float4 c = cost[p];
        k = 2;
        if (c.z < c.y) k = 3;
        if (c.w < (k == 3 ? c.z : c.y)) k = 4;
        head[p] = 1;
        cl = lev4[p] + 1;
    }
    const int a = collapse_left(nd, cost, k);
    assign[nd.left] = a;
    assign[nd.right] = k - a;
    lev4[nd.left] = cl;
    lev4[nd.right] = cl;
}

__device__ __forceinline__ uint16_t h_bits(_Float16 h) { return __builtin_bit_cast(uint16_t, h); }
__device__ __forceinline__ float h_val(uint16_t u) { return (float)__builtin_bit_cast(_Float16, u); }

// Largest half <= x / smallest half >= x: round to nearest even, then step
// once when the rounding went the wrong way (the host rule of the oracle).
__device__ __forceinline__ uint16_t f2h_down(float x) {
    uint16_t h = h_bits((_Float16)x);
    if (h_val(h) > x) {
        if ((h & 0x7fffu) == 0) h = 0x8001u;
        else if (h & 0x8000u) { if (h != 0xfc00u) h++; }
        else h--;
    }
    return h;
}
__device__ __forceinline__ uint16_t f2h_up(float x) {
    uint16_t h = h_bits((_Float16)x);
    if (h_val(h) < x) {
        if ((h & 0x7fffu) == 0) h = 0x0001u;
        else if (h & 0x8000u) h--;
        else if (h != 0x7c00u) h++;
    }
    return h;
}

__device__ __forceinline__ int32_t leaf_code(int first, int count) { return -((first << 3) | (count - 1)) - 2; }

// One BVH4 node per head (pre-order position p with head[p]); b4id[p] is the
// exclusive count of heads before p, i.e. the node's pre-order BVH4 index.
// Its slots are the nodes of assign 1 below it, reached through dissolved
// nodes (assign >= 2), left to right, so they stay in BVH2 pre-order.
__global__ void k_emit_bvh4(const rt::Node2* n2, const int* head, const int* b4id, const int* assign,
                            const int* lev4, int count, uint4* out, int* depth4) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count || !head[p]) return;
    const rt::Node2& nd = n2[p];
    int gather[4], ng = 0;
    int stk[4], sp = 0;
    stk[sp++] = nd.right;
    stk[sp++] = nd.left;
    while (sp > 0 && ng < 4) {
        const int q = stk[--sp];
        const rt::Node2& qn = n2[q];
        if (qn.triCount > 0 || assign[q] <= 1) { gather[ng++] = q; continue; }
        if (sp + 2 > 4) break;                  // unreachable: at most four slots
        stk[sp++] = qn.right;
        stk[sp++] = qn.left;
    }
    uint16_t hx[8], hy[8], hz[8];
    int32_t child[4];
    for (int i = 0; i < 4; i++) {
        if (i < ng) {
            const rt::Node2& g = n2[gather[i]];
            hx[i] = f2h_down(g.box.mn[0]); hx[4 + i] = f2h_up(g.box.mx[0]);
            hy[i] = f2h_down(g.box.mn[1]); hy[4 + i] = f2h_up(g.box.mx[1]);
            hz[i] = f2h_down(g.box.mn[2]); hz[4 + i] = f2h_up(g.box.mx[2]);
            child[i] = g.triCount > 0 ? leaf_code(g.triStart, g.triCount) : b4id[gather[i]];
        } else {
            hx[i] = hy[i] = hz[i] = 0x7c00u;
            hx[4 + i] = hy[4 + i] = hz[4 + i] = 0xfc00u;
            child[i] = -1;
        }
    }
    uint4* o = out + 4 * (size_t)b4id[p];
    o[0] = make_uint4(hx[0] | (uint32_t)hx[1] << 16, hx[2] | (uint32_t)hx[3] << 16, hx[4] | (uint32_t)hx[5] << 16,
                      hx[6] | (uint32_t)hx[7] << 16);
    o[1] = make_uint4(hy[0] | (uint32_t)hy[1] << 16, hy[2] | (uint32_t)hy[3] << 16, hy[4] | (uint32_t)hy[5] << 16,
                      hy[6] | (uint32_t)hy[7] << 16);
    o[2] = make_uint4(hz[0] | (uint32_t)hz[1] << 16, hz[2] | (uint32_t)hz[3] << 16, hz[4] | (uint32_t)hz[5] << 16,
                      hz[6] | (uint32_t)hz[7] << 16);
    o[3] = make_uint4((uint32_t)child[0], (uint32_t)child[1], (uint32_t)child[2], (uint32_t)child[3]);
    atomicMax(depth4, lev4[p] + 1);
}

// A tree whose root is a leaf (n <= maxLeaf): one BVH4 node with one slot.
__global__ void k_emit_leaf_root(const rt::Node2* n2, uint4* out, int* depth4) {
    const rt::Node2& r = n2[0];
    uint16_t h[3][8];
    for (int a = 0; a < 3; a++) {
        for (int i = 0; i < 4; i++) { h[a][i] = 0x7c00u; h[a][4 + i] = 0xfc00u; }
        h[a][0] = f2h_down(r.box.mn[a]);
        h[a][4] = f2h_up(r.box.mx[a]);
    }
    for (int a = 0; a < 3; a++)
        out[a] = make_uint4(h[a][0] | (uint32_t)h[a][1] << 16, h[a][2] | (uint32_t)h[a][3] << 16,
                            h[a][4] | (uint32_t)h[a][5] << 16, h[a][6] | (uint32_t)h[a][7] << 16);
    out[3] = make_uint4((uint32_t)leaf_code(r.triStart, r.triCount), 0xffffffffu, 0xffffffffu, 0xffffffffu);
    *depth4 = 1;
}

// Leaf-order triangles {v0.xyz, bits(id)}, {v1.xyz, 0}, {v2.xyz, 0} and the
// inverse permutation.
// In a forest the id is the index within the triangle's scene.
__global__ void k_leaf_tris(const float* __restrict__ tris, const int* idx, int n, float4* out, int* leafOf,
                            const int* off, int numScenes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = idx[i];
    const float* s = tris + 9 * (size_t)p;
    const int id = numScenes == 1 ? p : p - off[scene_of(off, numScenes, p)];
    out[3 * (size_t)i] = make_float4(s[0], s[1], s[2], __int_as_float(id));
    out[3 * (size_t)i + 1] = make_float4(s[3], s[4], s[5], 0.f);
    out[3 * (size_t)i + 2] = make_float4(s[6], s[7], s[8], 0.f);
    leafOf[p] = i;
}

// ------------------------------- scan ---------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;

__global__ __launch_bounds__(kThreads) void k_tile_sums(const int* in, int n, int* tiles) {
    __shared__ int lds[kThreads / 64 + 1];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
    for (int k = 0; k < kScanItems; k++) s += (base + k < n) ? in[base + k] : 0;
    int total;
    rts::block_exclusive_scan<kThreads>(s, lds, &total);
    if (threadIdx.x == 0) tiles[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_tile_sums(int* tiles, int nb, int* total) {
    __shared__ int lds[1024 / 64 + 1];
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? tiles[i] : 0;
        int tot;
        const int ex = rts::block_exclusive_scan<1024>(v, lds, &tot);
        if (i < nb) tiles[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kThreads) void k_tile_scan(const int* in, int n, const int* tiles, int* out) {
    __shared__ int lds[kThreads / 64 + 1];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; k++) { v[k] = (base + k < n) ? in[base + k] : 0; s += v[k]; }
    int total;
    int run = tiles[blockIdx.x] + rts::block_exclusive_scan<kThreads>(s, lds, &total);
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// Triangle gather from device vertex / index buffers (one thread per
// triangle): corner k of triangle t is vertex idx[3t + k] (or 3t + k without
// an index buffer), three floats at byte offset index * stride. An index at or
// beyond numVertices sets *bad and writes zeros instead of reading.
__global__ void k_gather_tris(const uint8_t* verts, int numVertices, int stride, const void* idx, int indexBytes,
                              int numTris, float* out, int* bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= numTris) return;
    for (int k = 0; k < 3; k++) {
        const size_t c = 3 * (size_t)t + k;
        unsigned int v;
        if (indexBytes == 4) v = ((const uint32_t*)idx)[c];
        else if (indexBytes == 2) v = ((const uint16_t*)idx)[c];
        else v = (unsigned int)c;
        float x = 0.f, y = 0.f, z = 0.f;
        if (v < (unsigned int)numVertices) {
            const float* p = (const float*)(verts + (size_t)v * stride);
            x = p[0]; y = p[1]; z = p[2];
        } else {
            atomicOr(bad, 1);
        }
        out[9 * (size_t)t + 3 * k] = x;
        out[9 * (size_t)t + 3 * k + 1] = y;
        out[9 * (size_t)t + 3 * k + 2] = z;
    }
}

// One tree of a forest as a BVH of its own: BVH4 nodes [first, first+count)
// with inner children rebased by -first and leaf triangle starts by -triShift
// (trees and their leaf triangles lie one after another in a forest).
__global__ void k_slice_nodes4(const uint4* __restrict__ src, int first, int count, int triShift, uint4* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint4* a = src + 4 * (size_t)(first + i);
    uint4* o = dst + 4 * (size_t)i;
    o[0] = a[0]; o[1] = a[1]; o[2] = a[2];
    const uint4 c = a[3];
    uint32_t w[4] = {c.x, c.y, c.z, c.w};
    for (int k = 0; k < 4; k++) {
        const int v = (int)w[k];
        if (v >= 0) {
            w[k] = (uint32_t)(v - first);
        } else if (v < -1) {
            const int enc = -(v + 2);
            w[k] = (uint32_t)leaf_code((enc >> 3) - triShift, (enc & 7) + 1);
        }
    }
    o[3] = make_uint4(w[0], w[1], w[2], w[3]);
}

// dst[i] = src[i] - sub, i < n
__global__ void k_slice_sub(const int* __restrict__ src, int n, int sub, int* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] - sub;
}

}  // namespace rtb

namespace rtk {
int launch_slice_forest(const void* nodes4, int first4, int count4, const void* tris, const int* leafOf, int t0,
                        int ntris, void* dNodes4, void* dTris, int* dLeafOf, hipStream_t s) {
    const int T = 256;
    hipLaunchKernelGGL(rtb::k_slice_nodes4, dim3((count4 + T - 1) / T), dim3(T), 0, s, (const uint4*)nodes4, first4,
                       count4, t0, (uint4*)dNodes4);
    hipLaunchKernelGGL(rtb::k_slice_sub, dim3((ntris + T - 1) / T), dim3(T), 0, s, leafOf + t0, ntris, t0, dLeafOf);
    hipMemcpyAsync(dTris, (const char*)tris + (size_t)t0 * 48, (size_t)ntris * 48, hipMemcpyDeviceToDevice, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("forest slice launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

hipError_t launch_gather_tris(const void* verts, int numVertices, int stride, const void* idx, int indexBytes,
                              int numTris, float* out, int* bad, hipStream_t s) {
    hipLaunchKernelGGL(rtb::k_gather_tris, dim3((numTris + 255) / 256), dim3(256), 0, s, (const uint8_t*)verts,
                       numVertices, stride, idx, indexBytes, numTris, out, bad);
    return hipGetLastError();
}
}  // namespace rtk

namespace rt {

// Device-wide exclusive scan of n ints (out may not alias in); *total (device)
// receives the sum. tiles needs ceil(n / 2048) ints.
void scan_exclusive(const int* in, int* out, int n, int* total, int* tiles, hipStream_t s) {
    const int nb = (n + rtb::kScanTile - 1) / rtb::kScanTile;
    if (nb > 0) hipLaunchKernelGGL(rtb::k_tile_sums, dim3(nb), dim3(rtb::kThreads), 0, s, in, n, tiles);
    hipLaunchKernelGGL(rtb::k_scan_tile_sums, dim3(1), dim3(1024), 0, s, tiles, nb, total);
    if (nb > 0) hipLaunchKernelGGL(rtb::k_tile_scan, dim3(nb), dim3(rtb::kThreads), 0, s, in, n, tiles, out);
}

int scan_tiles_needed(int n) { return (n + rtb::kScanTile - 1) / rtb::kScanTile + 1; }

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct DevBuf {
    std::vector<void*> ptrs;
    template <class T> T* get(size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
    ~DevBuf() { for (void* p : ptrs) hipFree(p); }
};

#define BUILD_CHECK(cond, msg)              \
    do {                                    \
        if (!(cond)) { set_error("%s", msg); return -1; } \
    } while (0)

// Builds the BVH on the device from d_tris (n x 9 floats, original order).
// Outputs (device, allocated here): nodes4 (numNodes4 x 64 B), leaf-order
// triangles (n x 48 B), leafOf (n ints), ordered (n ints, leaf order ->
// original index), nodes2 (pre-order rt::Node2). Returns 0 on success.
int gpu_build(const float* d_tris, int n, int maxLeaf, hipStream_t s, GpuBVH& out, const int* sceneOff,
              int numScenes, std::vector<int>* roots4) {
    using namespace rtb;
    const double tBegin = verbose() ? now_ms() : 0.0;
    struct Report {
        double t0;
        ~Report() { if (verbose()) fprintf(stderr, "[CudaBVH] gpu_build: %.2f ms incl. workspace free\n", now_ms() - t0); }
    } report{tBegin};
    DevBuf tmp;
    const int R = numScenes;
    std::vector<int> hOff = sceneOff ? std::vector<int>(sceneOff, sceneOff + R + 1) : std::vector<int>{0, n};
    std::vector<int> hSlots(R, -1);
    int m = 0;
    for (int k = 0; k < R; k++) {
        const int c = hOff[k + 1] - hOff[k];
        if (c <= 0 || (R > 1 && c <= maxLeaf)) {
            set_error("gpu build: scene %d has %d primitives (a batch needs more than %d per scene)", k, c, maxLeaf);
            return -1;
        }
        if (c > maxLeaf) hSlots[k] = m++;
    }
    int* dOff = tmp.get<int>(R + 1);
    int* dSlots = tmp.get<int>(R);
    BUILD_CHECK(dOff && dSlots, "gpu build: out of device memory");
    if (hipMemcpyAsync(dOff, hOff.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dSlots, hSlots.data(), R * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) {
        set_error("gpu build: upload failed");
        return -1;
    }
    Build b{};
    b.n = n; b.maxLeaf = maxLeaf;
    const size_t cap = 2 * (size_t)n;
    float* pbox = tmp.get<float>(6 * (size_t)n);
    float4* cen = tmp.get<float4>(n);
    b.pbox = pbox; b.cen = cen;
    b.idx = tmp.get<int>(n); b.pnode = tmp.get<int>(n);
    b.idx2 = tmp.get<int>(n); b.pnode2 = tmp.get<int>(n);
    b.flag = tmp.get<int>(n + 1); b.rank = tmp.get<int>(n + 1);
    b.nbox = tmp.get<int>(6 * cap); b.ncen = tmp.get<int>(6 * cap);
    b.ns = tmp.get<int>(cap); b.ncnt = tmp.get<int>(cap);
    b.nleft = tmp.get<int>(cap); b.nright = tmp.get<int>(cap); b.ndepth = tmp.get<int>(cap);
    b.nodeSlot = tmp.get<int>(cap);
    b.small = tmp.get<int>(cap);
    b.bigSlot = tmp.get<int>(cap);
    b.bigCount = tmp.get<int>(1);
    const int maxBig = n / (kSmall + 1) + R + 1;
    b.bins = tmp.get<int>((size_t)maxBig * kSlotInts);
    b.act = tmp.get<int>(n); int* act2 = tmp.get<int>(n);
    b.sAxis = tmp.get<int>(n); b.sBin = tmp.get<int>(n); b.sNl = tmp.get<int>(n);
    int* tiles = tmp.get<int>(scan_tiles_needed((int)cap) + 8);
    int* dcount = tmp.get<int>(4);
    BUILD_CHECK(pbox && cen && b.idx && b.pnode && b.idx2 && b.pnode2 && b.flag && b.rank && b.nbox && b.ncen &&
                b.ns && b.ncnt && b.nleft && b.nright && b.ndepth && b.nodeSlot && b.small && b.bigSlot && b.bigCount &&
                b.bins && b.act && act2 && b.sAxis &&
                b.sBin && b.sNl && tiles && dcount, "gpu build: out of device memory");

    const int T = kThreads;
    auto grid = [](long long items) { return dim3((unsigned)((items + kThreads - 1) / kThreads)); };
    const int chunks = (n + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_prims, grid(n), dim3(T), 0, s, d_tris, n, pbox, cen, b.idx, b.pnode, dOff, R);
    hipLaunchKernelGGL(k_init_roots, grid(R), dim3(T), 0, s, b, dOff, dSlots, R);
    hipLaunchKernelGGL(k_init_nodes, grid(R), dim3(T), 0, s, b.nbox, b.ncen, b.small, 0, R);
    hipLaunchKernelGGL(k_node_bounds, dim3(chunks), dim3(T), 0, s, b, 0);

    std::vector<int> levelStart = {0, R};     // node ids of level L: [levelStart[L], levelStart[L+1])
    int nodeCount = R;
    while (m > 0) {
        const int childBase = nodeCount;
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, b.bigCount);
        hipLaunchKernelGGL(k_big_slots, grid(m), dim3(T), 0, s, b, m);
        hipLaunchKernelGGL(k_clear_bins, grid((long long)maxBig * kSlotInts), dim3(T), 0, s, b.bins, maxBig);
        hipLaunchKernelGGL(k_init_nodes, grid(2 * m), dim3(T), 0, s, b.nbox, b.ncen, b.small, childBase, 2 * m);
        hipLaunchKernelGGL(k_bin, dim3(chunks), dim3(T), 0, s, b);
        hipLaunchKernelGGL(k_sah, grid(m), dim3(T), 0, s, b, m, childBase);
        hipLaunchKernelGGL(k_small_level, dim3(m), dim3(64), 0, s, b, m, childBase);
        hipLaunchKernelGGL(k_flags, grid(n), dim3(T), 0, s, b);
        scan_exclusive(b.flag, b.rank, n, dcount, tiles, s);
        hipLaunchKernelGGL(k_scatter, grid(n), dim3(T), 0, s, b);
        std::swap(b.idx, b.idx2);
        std::swap(b.pnode, b.pnode2);
        hipLaunchKernelGGL(k_node_bounds, dim3(chunks), dim3(T), 0, s, b, childBase);
        // next level's active nodes: children with more than maxLeaf primitives, in id order
        hipLaunchKernelGGL(k_active_flags, grid(2 * m), dim3(T), 0, s, b.ncnt, childBase, 2 * m, maxLeaf, b.flag);
        scan_exclusive(b.flag, b.rank, 2 * m, dcount, tiles, s);
        hipLaunchKernelGGL(k_next_active, grid(2 * m), dim3(T), 0, s, b, childBase, 2 * m, b.rank, act2);
        std::swap(b.act, act2);
        nodeCount += 2 * m;
        levelStart.push_back(nodeCount);
        int next = 0;
        if (hipMemcpyAsync(&next, dcount, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            set_error("gpu build: level %zu failed", levelStart.size() - 2);
            return -1;
        }
        m = next;
    }
    // pre-order renumbering
    int* size = tmp.get<int>(nodeCount);
    int* pre = tmp.get<int>(nodeCount);
    int* head = tmp.get<int>(nodeCount + 1);
    int* b4id = tmp.get<int>(nodeCount + 1);
    int* assign = tmp.get<int>(nodeCount);
    int* lev4 = tmp.get<int>(nodeCount);
    float4* ccost = tmp.get<float4>(nodeCount);
    BUILD_CHECK(size && pre && head && b4id && assign && lev4 && ccost, "gpu build: out of device memory");
    const int levels = (int)levelStart.size() - 1;
    for (int L = levels - 1; L >= 0; L--) {
        const int f = levelStart[L], c = levelStart[L + 1] - f;
        hipLaunchKernelGGL(k_sizes, grid(c), dim3(T), 0, s, b.nleft, b.nright, size, f, c);
    }
    hipLaunchKernelGGL(k_root_pre, dim3(1), dim3(1), 0, s, size, pre, R);
    for (int L = 0; L < levels; L++) {
        const int f = levelStart[L], c = levelStart[L + 1] - f;
        hipLaunchKernelGGL(k_preorder, grid(c), dim3(T), 0, s, b.nleft, b.nright, size, pre, f, c);
    }
    out.numNodes2 = nodeCount;
    if (hipMalloc((void**)&out.nodes2, (size_t)nodeCount * sizeof(Node2)) != hipSuccess) {
        set_error("gpu build: out of device memory");
        return -1;
    }
    hipLaunchKernelGGL(k_emit_bvh2, grid(nodeCount), dim3(T), 0, s, b, pre, nodeCount, out.nodes2, head, assign, lev4);
    for (int L = levels - 1; L >= 0; L--) {
        const int f = levelStart[L], c = levelStart[L + 1] - f;
        hipLaunchKernelGGL(k_collapse_cost, grid(c), dim3(T), 0, s, out.nodes2, pre, f, c, ccost);
    }
    for (int L = 0; L < levels; L++) {
        const int f = levelStart[L], c = levelStart[L + 1] - f;
        hipLaunchKernelGGL(k_collapse_assign, grid(c), dim3(T), 0, s, out.nodes2, pre, f, c, ccost, assign, head, lev4);
    }
    scan_exclusive(head, b4id, nodeCount, dcount, tiles, s);
    int heads = 0;
    if (hipMemcpyAsync(&heads, dcount, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        set_error("gpu build: renumbering failed");
        return -1;
    }
    out.numNodes4 = std::max(heads, 1);
    int* dDepth = dcount + 1;
    hipMemsetAsync(dDepth, 0, sizeof(int), s);
    if (hipMalloc(&out.nodes4, (size_t)out.numNodes4 * 64) != hipSuccess ||
        hipMalloc(&out.tris, (size_t)n * 48) != hipSuccess || hipMalloc((void**)&out.leafOf, (size_t)n * 4) != hipSuccess ||
        hipMalloc((void**)&out.ordered, (size_t)n * 4) != hipSuccess) {
        set_error("gpu build: out of device memory");
        return -1;
    }
    if (heads > 0)
        hipLaunchKernelGGL(k_emit_bvh4, grid(nodeCount), dim3(T), 0, s, out.nodes2, head, b4id, assign, lev4,
                           nodeCount, (uint4*)out.nodes4, dDepth);
    else
        hipLaunchKernelGGL(k_emit_leaf_root, dim3(1), dim3(1), 0, s, out.nodes2, (uint4*)out.nodes4, dDepth);
    hipLaunchKernelGGL(k_leaf_tris, grid(n), dim3(T), 0, s, d_tris, b.idx, n, (float4*)out.tris, out.leafOf, dOff, R);
    if (roots4) {
        int* dRoot4 = tmp.get<int>(R);
        BUILD_CHECK(dRoot4, "gpu build: out of device memory");
        hipLaunchKernelGGL(k_root4, grid(R), dim3(T), 0, s, pre, b4id, R, dRoot4);
        roots4->assign(R, 0);
        if (hipMemcpyAsync(roots4->data(), dRoot4, R * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) {
            set_error("gpu build: root download failed");
            return -1;
        }
    }
    hipMemcpyAsync(out.ordered, b.idx, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(&out.depth4, dDepth, 4, hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) { set_error("gpu build failed: %s", hipGetErrorString(e)); return -1; }
    if (verbose()) fprintf(stderr, "[CudaBVH] gpu_build: %d levels, %.2f ms to the last kernel\n", levels, now_ms() - tBegin);
    out.levels = levels;
    return 0;
}

}  // namespace rt
