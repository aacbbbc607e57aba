// lbvh.hip — the TLAS fast path (cudaBVH_buildTLASFast) on the GPU: the
// reference's LBVH (layer/cuda_bvh_backend.cu:257-389, :1620-1698) — 10-bit
// per axis Morton codes of the box centroids normalised by their bounds, a
// stable sort by code, a median split of the sorted order with one box per
// leaf, and the stackless DFS export of the nodes.
//
// The tree's shape depends only on n: the node of range [s, e) with
// pre-order index i has its left child at i + 1 ([s, m), m = s + (e-s)/2)
// and its right child at i + 2(m - s), and its subtree holds 2(e-s) - 1
// nodes. So every node finds its range by descending from the root
// (one thread per node, O(log n) steps), boxes are merged bottom-up one tree
// level per launch, and each node writes its 32-byte DFS record (box, leaf
// code, skip = i + size, -1 past the end) in place. Box merges reproduce the
// host builder's fminf / fmaxf exactly (x86-64 libm: equal operands, e.g.
// -0 and +0, give the second operand; a NaN operand gives the other), so the
// output is byte-identical to a host build of the same algorithm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "rt_internal.h"
#include "trace_launch.h"

namespace rtl {

constexpr int kT = 256;

__device__ __forceinline__ int f2o(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__device__ __forceinline__ float host_fminf(float x, float y) {
    if (y != y) return x;
    if (x != x) return y;
    return x < y ? x : y;
}
__device__ __forceinline__ float host_fmaxf(float x, float y) {
    if (y != y) return x;
    if (x != x) return y;
    return x > y ? x : y;
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t morton3d(float x, float y, float z) {
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    return spread10((uint32_t)x) * 4u + spread10((uint32_t)y) * 2u + spread10((uint32_t)z);
}

// Centroids (mn + mx) * 0.5 and their bounds (ordered-int atomics; NaN
// centroids are left out, as the host's fminf / fmaxf ignore them). The sign
// of a zero bound does not matter: only c - min and max - min use it.
__global__ __launch_bounds__(kT) void k_centroids(const float* __restrict__ aabbs, int n, float* cen, int* bnd) {
    __shared__ int s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0x7fffffff : (int)0x80000000;
    __syncthreads();
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {(int)0x80000000, (int)0x80000000, (int)0x80000000};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int a = 0; a < 3; a++) {
            const float c = (aabbs[6 * (size_t)i + a] + aabbs[6 * (size_t)i + 3 + a]) * 0.5f;
            cen[3 * (size_t)i + a] = c;
            if (c == c) { lo[a] = min(lo[a], f2o(c)); hi[a] = max(hi[a], f2o(c)); }
        }
    }
    for (int a = 0; a < 3; a++) { atomicMin(&s[a], lo[a]); atomicMax(&s[3 + a], hi[a]); }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&bnd[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&bnd[threadIdx.x], s[threadIdx.x]);
}

__global__ void k_zero(int* w) { *w = 0; }

__global__ void k_init_bounds(int* bnd) {
    if (threadIdx.x < 6) bnd[threadIdx.x] = threadIdx.x < 3 ? 0x7fffffff : (int)0x80000000;
}

// Morton key and index of every box; lo / inv as the host computes them.
__global__ __launch_bounds__(kT) void k_keys(const float* cen, int n, const int* bnd, uint32_t* keys, int* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float q[3];
    for (int a = 0; a < 3; a++) {
        // with no finite centroid the host keeps its initial bounds (1e30, -1e30)
        const float lo = bnd[a] == 0x7fffffff ? 1e30f : o2f(bnd[a]);
        const float hi = bnd[3 + a] == (int)0x80000000 ? -1e30f : o2f(bnd[3 + a]);
        const float inv = (hi - lo > 1e-8f) ? 1.0f / (hi - lo) : 0.0f;
        q[a] = (cen[3 * (size_t)i + a] - lo) * inv;
    }
    keys[i] = morton3d(q[0], q[1], q[2]);
    vals[i] = i;
}

// Range [s, e) and tree level of pre-order node `id` of an n-leaf median tree.
__device__ __forceinline__ void node_range(int n, int id, int& s, int& e, int& level) {
    s = 0; e = n; level = 0;
    int cur = 0;
    while (cur != id) {
        const int m = s + (e - s) / 2;
        const int right = cur + 2 * (m - s);
        if (id < right) { cur = cur + 1; e = m; }
        else { cur = right; s = m; }
        level++;
    }
}

__global__ __launch_bounds__(kT) void k_levels(int n, int total, int* level, int* maxLevel) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    int s, e, l;
    node_range(n, id, s, e, l);
    level[id] = l;
    atomicMax(maxLevel, l);
}

// Boxes of the nodes at tree level L (children done by the previous launch):
// a leaf takes its box, an inner node merges its children's as the host does
// (fminf(left, right), fmaxf(left, right)). Writes the node's DFS record.
__global__ __launch_bounds__(kT) void k_level_boxes(const float* __restrict__ aabbs, const int* ordered, int n,
                                                    int total, const int* level, int L, uint32_t* out) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total || level[id] != L) return;
    int s, e, l;
    node_range(n, id, s, e, l);
    float mn[3], mx[3];
    int32_t enc = 0;
    if (e - s <= 1) {
        const float* b = aabbs + 6 * (size_t)ordered[s];
        for (int a = 0; a < 3; a++) { mn[a] = host_fminf(1e30f, b[a]); mx[a] = host_fmaxf(-1e30f, b[3 + a]); }
        enc = -((s << 3) | 0) - 2;                    // leaf of one box (backend :393 encoding)
    } else {
        const int m = s + (e - s) / 2;
        const float* L_ = reinterpret_cast<const float*>(out + 8 * (size_t)(id + 1));
        const float* R_ = reinterpret_cast<const float*>(out + 8 * (size_t)(id + 2 * (m - s)));
        for (int a = 0; a < 3; a++) { mn[a] = host_fminf(L_[a], R_[a]); mx[a] = host_fmaxf(L_[3 + a], R_[3 + a]); }
    }
    const int size = 2 * (e - s) - 1;
    const int32_t skip = id + size < total ? id + size : -1;
    uint32_t* o = out + 8 * (size_t)id;
    for (int a = 0; a < 3; a++) { o[a] = __float_as_uint(mn[a]); o[3 + a] = __float_as_uint(mx[a]); }
    o[6] = (uint32_t)enc;
    o[7] = (uint32_t)skip;
}

}  // namespace rtl

namespace rtk {

size_t lbvh_scratch_bytes(int n) {
    const size_t N = (size_t)n;
    return N * 12 + N * 4 + (2 * N) * 4 + 64 + radix_pairs_bytes(n) + 256;
}

int lbvh_build(const float* d_aabbs, int n, uint32_t* d_nodes, int* d_ordered, void* ws, hipStream_t s) {
    if (n <= 0) { rt::set_error("lbvh: no boxes"); return -1; }
    const size_t N = (size_t)n;
    const int total = 2 * n - 1;
    char* p = (char*)ws;
    float* cen = (float*)p; p += N * 12;
    uint32_t* keys = (uint32_t*)p; p += N * 4;
    int* level = (int*)p; p += (2 * N) * 4;
    int* small = (int*)p; p += 64;                   // bounds (6 ints) and the deepest level
    void* sortWs = p;
    const int T = rtl::kT;
    const int g = (int)std::min<size_t>((N + T - 1) / T, 1024);
    hipLaunchKernelGGL(rtl::k_init_bounds, dim3(1), dim3(64), 0, s, small);
    hipLaunchKernelGGL(rtl::k_centroids, dim3(g), dim3(T), 0, s, d_aabbs, n, cen, small);
    hipLaunchKernelGGL(rtl::k_keys, dim3((n + T - 1) / T), dim3(T), 0, s, cen, n, small, keys, d_ordered);
    if (radix_sort_pairs(keys, d_ordered, n, 30, sortWs, s) != 0) return -1;
    int* maxLevel = small + 8;
    hipLaunchKernelGGL(rtl::k_zero, dim3(1), dim3(1), 0, s, maxLevel);
    hipLaunchKernelGGL(rtl::k_levels, dim3((total + T - 1) / T), dim3(T), 0, s, n, total, level, maxLevel);
    int deepest = 0;
    if (hipMemcpyAsync(&deepest, maxLevel, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        rt::set_error("lbvh: level pass failed");
        return -1;
    }
    for (int L = deepest; L >= 0; L--)
        hipLaunchKernelGGL(rtl::k_level_boxes, dim3((total + T - 1) / T), dim3(T), 0, s, d_aabbs, d_ordered, n, total,
                           level, L, d_nodes);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("lbvh launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

}  // namespace rtk
