// api.hip — the extern "C" boundary of libcudabvh_hip.so. Part 1 keeps the
// reference's cudaBVH_* symbols (layer/cuda_bvh_backend.h:12-141); part 2
// adds the packet entry points the reference lacks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "rt_internal.h"
#include "trace_launch.h"
#include "scene.h"

namespace rt {
const char* last_error();
}

using rt::set_error;

static inline hipStream_t hstream(const CudaBVHHandle* h) { return (hipStream_t)h->stream; }

// Waits for stream s and checks the traversal status word of sc: 0 when every
// trace on it kept its whole stack, -1 (error recorded) otherwise.
static int finish_sync(rt::TraceScratch& sc, hipStream_t s, const char* fn) {
    return rtk::take_status(sc, s, fn) == 0 ? 0 : -1;
}
static int ensure_cwbvh(CudaBVH_t bvh);
static void async_release(CudaBVH_t bvh);

static void free_handle_device(CudaBVHHandle* h) {
    hipFree(h->d_nodes4); hipFree(h->d_tris); hipFree(h->d_leafOfPrim); hipFree(h->d_work); hipFree(h->d_sortWs);
    h->d_sortWs = nullptr; h->sortWsBytes = 0;
    hipFree(h->d_cwNodes); hipFree(h->d_cwTris);
    h->d_cwNodes = h->d_cwTris = nullptr;
    hipFree(h->d_nodes2); hipFree(h->d_ordered);
    h->d_nodes2 = nullptr; h->d_ordered = nullptr;
    rtk::free_scratch(h->scratch);
    hipHostFree(h->h_pinned);
    h->h_pinned = nullptr; h->pinnedFloats = 0;
    h->d_nodes4 = h->d_tris = h->d_work = nullptr;
    h->d_leafOfPrim = nullptr;
}

// Runs f(begin, end) over [0, n) split across up to 8 host threads.
template <class F>
static void host_parallel(size_t n, F&& f) {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nt = n < (size_t(1) << 16) ? 1u : std::max(1u, std::min(8u, hw));
    const size_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++)
        th.emplace_back([&, t] { f(std::min(n, t * chunk), std::min(n, (t + 1) * chunk)); });
    f(0, std::min(n, chunk));
    for (auto& x : th) x.join();
}

// Bounds of n x 9 host floats as a sequential std::min / std::max pass
// computes them (per-chunk passes combined in order give the same values,
// signed zeros included), centre and diagonal extent.
static void set_bounds(CudaBVHHandle* h, const float* tri9, int n) {
    const unsigned nt = 8;
    std::vector<float> cmn(3 * nt, 1e30f), cmx(3 * nt, -1e30f);
    const size_t chunk = (size_t(n) + nt - 1) / nt;
    std::vector<std::thread> th;
    auto work = [&](unsigned t) {
        const size_t b = std::min<size_t>(n, t * chunk), e = std::min<size_t>(n, (t + 1) * chunk);
        float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};   // in registers
        for (size_t i = b; i < e; i++)
            for (int k = 0; k < 9; k++) {
                const float x = tri9[9 * i + k];
                mn[k % 3] = std::min(mn[k % 3], x);
                mx[k % 3] = std::max(mx[k % 3], x);
            }
        for (int a = 0; a < 3; a++) { cmn[3 * t + a] = mn[a]; cmx[3 * t + a] = mx[a]; }
    };
    if (n < (1 << 16)) {
        for (unsigned t = 0; t < nt; t++) work(t);
    } else {
        for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
    }
    float bmin[3] = {1e30f, 1e30f, 1e30f}, bmax[3] = {-1e30f, -1e30f, -1e30f};
    for (unsigned t = 0; t < nt; t++)
        for (int a = 0; a < 3; a++) { bmin[a] = std::min(bmin[a], cmn[3 * t + a]); bmax[a] = std::max(bmax[a], cmx[3 * t + a]); }
    for (int a = 0; a < 3; a++) { h->bboxMin[a] = bmin[a]; h->bboxMax[a] = bmax[a]; h->bboxCenter[a] = (bmin[a] + bmax[a]) * 0.5f; }
    const float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
    h->bboxExtent = sqrtf(ex * ex + ey * ey + ez * ez);
}

// The host copies the reference's getters hand out (BVH4 nodes, SoA and
// packed triangles, the stackless DFS export of a single tree) and the CWBVH
// build reads, made once from the handle's device arrays. False on failure.
static bool ensure_exports(CudaBVHHandle* h) {
    if (h->exported) return true;
    const int n = h->numOrdered;
    std::vector<uint8_t> nodes4(size_t(h->numNodes4) * 64);
    std::vector<int> ordered(n);
    std::vector<float4> tris(size_t(n) * 3);
    std::vector<rt::Node2> nodes2(h->d_nodes2 ? h->numNodes2 : 0);
    hipStream_t s = hstream(h);
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(nodes4.data(), h->d_nodes4, nodes4.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ordered.data(), h->d_ordered, size_t(n) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(tris.data(), h->d_tris, size_t(n) * 48, hipMemcpyDeviceToHost);
    if (e == hipSuccess && !nodes2.empty())
        e = hipMemcpy(nodes2.data(), h->d_nodes2, nodes2.size() * sizeof(rt::Node2), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { set_error("copying the BVH to the host failed: %s", hipGetErrorString(e)); return false; }
    for (int a = 0; a < 9; a++) h->h_soa[a].resize(n);
    h->h_packed.assign(size_t(n) * 12, 0.f);
    host_parallel(size_t(n), [&](size_t b, size_t en) {
        for (size_t i = b; i < en; i++) {
            const float4* t = &tris[3 * i];
            const float v[9] = {t[0].x, t[0].y, t[0].z, t[1].x, t[1].y, t[1].z, t[2].x, t[2].y, t[2].z};
            float* pk = &h->h_packed[12 * i];
            for (int a = 0; a < 9; a++) { h->h_soa[a][i] = v[a]; pk[a] = v[a]; }
            // the original ID's int bits, which is how the layer and the
            // RT-IR executor read this word (VkLayer_CudaRT.cpp:4410,
            // layer/rt_ir_exec.cu:167); the reference writes float(ID)
            // there (backend :730-733), which those readers misinterpret
            const int32_t id = ordered[i];
            memcpy(&pk[9], &id, sizeof id);
        }
    });
    h->h_nodes4 = std::move(nodes4);
    h->h_ordered = std::move(ordered);
    if (!nodes2.empty()) {
        rt::export_dfs(nodes2, h->h_dfs);
        h->numDFS = int(nodes2.size());
        h->h_nodes2 = std::move(nodes2);
    }
    h->exported = true;
    return true;
}

// Completes a handle whose device structure (d_nodes4, d_tris, d_leafOfPrim,
// d_ordered, d_nodes2 for a single tree, numNodes4, depth4) is set: the
// bounds from the n x 9 host triangles (original order), the traversal
// scratch and the default launch configurations; the host exports follow on
// first use. Frees the handle and returns nullptr on failure.
static CudaBVHHandle* make_handle(CudaBVHHandle* h, const float* tri9, int n) {
    h->numTris = n;
    h->numOrdered = n;
    set_bounds(h, tri9, n);
    if (rtk::alloc_scratch(h->scratch) != 0) {
        free_handle_device(h);
        delete h;
        return nullptr;
    }
    // defaults: the configurations the autotuner picks most often on the
    // 1080p terrain frame (refill / 8: its 16 KB of stack plus 8 KB of
    // gathered results per workgroup leave room for six workgroups per CU)
    h->tune[rt::kClassIncoherent].schedule = rt::kScheduleRefill;
    h->tune[rt::kClassIncoherent].ldsDepth = 8;
    h->tune[rt::kClassCoherent].ldsDepth = 12;
    h->tune[rt::kClassAny].ldsDepth = 12;
    if (rt::verbose())
        fprintf(stderr, "[CudaBVH] built %d tris: %d BVH4 nodes (depth %d), %d BVH2 nodes\n", n, h->numNodes4,
                h->depth4, h->numNodes2);
    return h;
}

// Builds a handle from n x 9 floats (original order). d_tri9 (nullable): the
// same triangles already on the device (not freed here). With sceneOff the
// handle is a forest of numScenes trees (roots4: each tree's BVH4 root).
static CudaBVHHandle* build_impl(const float* tri9, int n, int maxLeaf, const int* sceneOff = nullptr,
                                 int numScenes = 1, std::vector<int>* roots4 = nullptr,
                                 const float* d_tri9 = nullptr, std::vector<rt::Node2>* forestNodes2 = nullptr,
                                 std::vector<int>* forestOrdered = nullptr) {
    rt::clear_error();
    if (!tri9 || n <= 0) { set_error("cudaBVH_build: need tris != NULL and numTris > 0 (got %d)", n); return nullptr; }
    if (n > (1 << 27)) { set_error("cudaBVH_build: %d triangles exceed the leaf encoding (2^27)", n); return nullptr; }
    auto* h = new CudaBVHHandle();
    hipStream_t s = hstream(h);
    float* d_in = const_cast<float*>(d_tri9);
    float* d_own = nullptr;
    rt::GpuBVH g;
    const auto clk = [] { return std::chrono::duration<double, std::milli>(
                                     std::chrono::steady_clock::now().time_since_epoch()).count(); };
    const double t0 = clk();
    double t1 = t0;
    if (!d_in) {
        HIP_TRY(hipMalloc((void**)&d_own, size_t(n) * 36), goto fail);
        HIP_TRY(hipMemcpyAsync(d_own, tri9, size_t(n) * 36, hipMemcpyHostToDevice, s), hipFree(d_own); goto fail);
        d_in = d_own;
    }
    hipStreamSynchronize(s);          // the build syncs anyway; this separates the upload time
    t1 = clk();
    if (rt::gpu_build(d_in, n, maxLeaf, s, g, sceneOff, numScenes, roots4) != 0) {
        hipFree(d_own); hipFree(g.nodes2); hipFree(g.nodes4); hipFree(g.tris); hipFree(g.leafOf); hipFree(g.ordered);
        goto fail;
    }
    hipFree(d_own);
    h->d_nodes4 = g.nodes4;
    h->d_tris = g.tris;
    h->d_leafOfPrim = g.leafOf;
    h->d_ordered = g.ordered;
    h->numNodes4 = g.numNodes4;
    h->depth4 = g.depth4;
    h->numNodes2 = g.numNodes2;
    if (numScenes == 1) {
        h->d_nodes2 = g.nodes2;
    } else {
        // a forest: its trees' BVH2 nodes and leaf order go to the caller
        hipError_t e = hipSuccess;
        if (forestNodes2) {
            forestNodes2->resize(g.numNodes2);
            e = hipMemcpy(forestNodes2->data(), g.nodes2, size_t(g.numNodes2) * sizeof(rt::Node2), hipMemcpyDeviceToHost);
        }
        if (e == hipSuccess && forestOrdered) {
            forestOrdered->resize(n);
            e = hipMemcpy(forestOrdered->data(), g.ordered, size_t(n) * 4, hipMemcpyDeviceToHost);
        }
        hipFree(g.nodes2);
        if (e != hipSuccess) { set_error("copying the forest back failed: %s", hipGetErrorString(e)); goto fail; }
    }
    {
        const double t2 = clk();
        CudaBVHHandle* r = make_handle(h, tri9, n);
        const double t3 = clk();
        if (r) {
            r->buildStats[0] = float(t1 - t0); r->buildStats[1] = float(t2 - t1);
            r->buildStats[2] = float(t3 - t2); r->buildStats[3] = float(g.levels);
        }
        if (rt::verbose())
            fprintf(stderr, "[CudaBVH] build of %d tris: upload %.2f ms, device build %.2f ms, handle %.2f ms\n", n,
                    t1 - t0, t2 - t1, t3 - t2);
        return r;
    }
fail:
    free_handle_device(h);
    delete h;
    return nullptr;
}

static bool check_handle(CudaBVH_t bvh, const char* fn) {
    if (!bvh) { set_error("%s: NULL handle", fn); return false; }
    return true;
}

// Device staging for calls that hand over host arrays.
// Pinned host staging of a handle's host-pointer calls (hipHostMalloc), grown
// on demand: the caller's pageable arrays are copied into it by host threads
// and moved over PCIe in one DMA, which is several times faster than DMA from
// pageable memory.
static float* pinned_buffer(CudaBVHHandle* h, size_t floats) {
    if (floats <= h->pinnedFloats) return h->h_pinned;
    hipHostFree(h->h_pinned);
    h->h_pinned = nullptr; h->pinnedFloats = 0;
    if (hipHostMalloc((void**)&h->h_pinned, floats * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        h->h_pinned = nullptr;
        set_error("cannot allocate %zu bytes of pinned host staging", floats * sizeof(float));
        return nullptr;
    }
    h->pinnedFloats = floats;
    return h->h_pinned;
}

// dst[k][0, n) = src[k][0, n) for k < parts, split over up to 8 host threads.
static void copy_host_arrays(float* const* dst, const float* const* src, int parts, size_t n) {
    const size_t total = n * size_t(parts);
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nt = total < (size_t(1) << 18) ? 1u : std::max(1u, std::min(8u, hw));
    const size_t chunk = (total + nt - 1) / nt;
    auto work = [&](unsigned t) {
        size_t b = size_t(t) * chunk;
        const size_t e = std::min(total, b + chunk);
        while (b < e) {
            const size_t k = b / n, off = b - k * n, len = std::min(e - b, n - off);
            memcpy(dst[k] + off, src[k] + off, len * sizeof(float));
            b += len;
        }
    };
    if (nt == 1) { work(0); return; }
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
}

// Per-ray interval [lo, hi) of n rays written on the device (no host arrays,
// no PCIe traffic for constants).
__global__ void k_fill_interval(float* tmin, float* tmax, size_t n, float lo, float hi) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        tmin[i] = lo;
        tmax[i] = hi;
    }
}

static void* work_buffer(CudaBVHHandle* h, size_t bytes) {
    if (bytes > h->workBytes) {
        hipFree(h->d_work);
        h->d_work = nullptr; h->workBytes = 0;
        if (hipMalloc(&h->d_work, bytes) != hipSuccess) { set_error("cannot allocate %zu staging bytes", bytes); return nullptr; }
        h->workBytes = bytes;
    }
    return h->d_work;
}

static void* sort_workspace(CudaBVHHandle* h, int n) {
    const size_t bytes = rtk::sort_scratch_bytes(n);
    if (bytes > h->sortWsBytes) {
        hipFree(h->d_sortWs);
        h->d_sortWs = nullptr; h->sortWsBytes = 0;
        if (hipMalloc(&h->d_sortWs, bytes) != hipSuccess) { set_error("cannot allocate %zu sort bytes", bytes); return nullptr; }
        h->sortWsBytes = bytes;
    }
    return h->d_sortWs;
}

static rtd::Cam to_cam(const CudaCamera& c) {
    rtd::Cam k;
    for (int a = 0; a < 3; a++) { k.o[a] = c.o[a]; k.f[a] = c.f[a]; k.r[a] = c.r[a]; k.u[a] = c.u[a]; }
    k.sx = c.scaleX; k.sy = c.scaleY;
    return k;
}

// Reference look-at camera toward the scene centre (backend :991-1016).
static rtd::Cam lookat_cam(const CudaBVHHandle* h, int width, int height, float cx, float cy, float cz) {
    float fx = h->bboxCenter[0] - cx, fy = h->bboxCenter[1] - cy, fz = h->bboxCenter[2] - cz;
    float fl = sqrtf(fx * fx + fy * fy + fz * fz);
    if (fl < 1e-6f) fl = 1.f;
    fx /= fl; fy /= fl; fz /= fl;
    float rx = fz, ry = 0.f, rz = -fx;
    const float rl = sqrtf(rx * rx + rz * rz);
    if (rl > 1e-6f) { rx /= rl; rz /= rl; } else { rx = 1.f; rz = 0.f; }
    const float ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
    const float fov = 0.6f, aspect = (float)width / (float)height;
    rtd::Cam k{};
    k.o[0] = cx; k.o[1] = cy; k.o[2] = cz;
    k.f[0] = fx; k.f[1] = fy; k.f[2] = fz;
    k.r[0] = rx; k.r[1] = ry; k.r[2] = rz;
    k.u[0] = ux; k.u[1] = uy; k.u[2] = uz;
    k.sx = aspect * fov; k.sy = fov;
    return k;
}

static int shade_to_device(CudaBVHHandle* h, rt::TraceScratch& sc, hipStream_t s, int width, int height, float cx,
                           float cy, float cz, void* dimg, int fmt) {
    if (width <= 0 || height <= 0 || (long long)width * height > (1ll << 30)) { set_error("bad image size %dx%d", width, height); return -1; }
    if (fmt < 0 || fmt > 3) { set_error("bad output format %d", fmt); return -1; }
    rtk::TraceArgs a;
    a.n = width * height; a.W = width; a.H = height;
    a.cam = lookat_cam(h, width, height, cx, cy, cz);
    a.img = dimg; a.fmt = fmt;
    // a frame-after-frame display call: its tiles start in the cost order of
    // the last frame of this size traced with sc (the image is the same in
    // any order)
    a.costOrder = true;
    return rtk::launch_trace(h, sc, a, rtk::kModeShade, false, true, false, s);
}

extern "C" {

const char* cudaBVH_lastError(void) { return rt::last_error(); }

CudaBVH_t cudaBVH_build(const CudaTri* tris, int numTris) {
    return build_impl(reinterpret_cast<const float*>(tris), numTris, 4);
}

CudaBVH_t cudaBVH_buildDevice(const void* d_vertices, int numVertices, int vertexStride, const void* d_indices,
                              int indexBytes, int numTris) {
    rt::clear_error();
    if (!d_vertices || numVertices <= 0 || numTris <= 0 || vertexStride < 12 || vertexStride % 4 ||
        (indexBytes != 0 && indexBytes != 2 && indexBytes != 4) || (indexBytes && !d_indices)) {
        set_error("cudaBVH_buildDevice: bad arguments (vertices %p x %d stride %d, indices %p x %d bytes, %d tris)",
                  d_vertices, numVertices, vertexStride, d_indices, indexBytes, numTris);
        return nullptr;
    }
    if (!indexBytes && (long long)numTris * 3 > numVertices) {
        set_error("cudaBVH_buildDevice: %d triangles need %lld vertices, got %d", numTris, 3LL * numTris, numVertices);
        return nullptr;
    }
    float* d_tri = nullptr;
    int* d_bad = nullptr;
    if (hipMalloc((void**)&d_tri, size_t(numTris) * 36) != hipSuccess || hipMalloc((void**)&d_bad, sizeof(int)) != hipSuccess) {
        hipFree(d_tri);
        set_error("cudaBVH_buildDevice: cannot allocate %zu bytes", size_t(numTris) * 36);
        return nullptr;
    }
    std::vector<float> host(size_t(numTris) * 9);
    int bad = 0;
    hipError_t e = hipMemset(d_bad, 0, sizeof(int));
    if (e == hipSuccess)
        e = rtk::launch_gather_tris(d_vertices, numVertices, vertexStride, d_indices, indexBytes, numTris, d_tri, d_bad,
                                    nullptr);
    // the host copy feeds the reference's host-side exports (SoA, packed, DFS)
    if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && !bad) e = hipMemcpy(host.data(), d_tri, host.size() * 4, hipMemcpyDeviceToHost);
    hipFree(d_bad);
    if (e != hipSuccess || bad) {
        hipFree(d_tri);
        if (bad) set_error("cudaBVH_buildDevice: an index is >= numVertices (%d)", numVertices);
        else set_error("cudaBVH_buildDevice: gather failed: %s", hipGetErrorString(e));
        return nullptr;
    }
    CudaBVHHandle* h = build_impl(host.data(), numTris, 4, nullptr, 1, nullptr, d_tri);
    hipFree(d_tri);
    return h;
}

// Levels of a host BVH4 (64-byte nodes, root 0): the length of its longest
// root-to-leaf path of nodes.
static int bvh4_depth(const std::vector<uint8_t>& nodes) {
    const int nn = int(nodes.size() / 64);
    if (nn == 0) return 0;
    std::vector<int> lev(nn, 0);
    lev[0] = 1;
    int d = 1;
    for (int i = 0; i < nn; i++) {                 // children follow their parent in pre-order
        int32_t ch[4];
        memcpy(ch, nodes.data() + 64 * size_t(i) + 48, 16);
        for (int c = 0; c < 4; c++)
            if (ch[c] >= 0 && ch[c] < nn) { lev[ch[c]] = lev[i] + 1; d = std::max(d, lev[ch[c]]); }
    }
    return d;
}

static bool geometry_ok(const CudaGeometry& g) {
    return g.vertices && g.numVertices > 0 && g.numTris > 0 && g.vertexStride >= 12 && g.vertexStride % 4 == 0 &&
           ((g.indices && (g.indexBytes == 2 || g.indexBytes == 4)) || (!g.indices && g.indexBytes == 0)) &&
           (g.indices || (long long)g.numTris * 3 <= g.numVertices);
}

int cudaBVH_buildManyDevice(const CudaGeometry* geoms, const int* geomOffsets, int numBlas, CudaBVH_t* out) {
    const char* fn = "cudaBVH_buildManyDevice";
    rt::clear_error();
    if (!geoms || !geomOffsets || numBlas <= 0 || !out || geomOffsets[0] != 0) {
        set_error("%s: bad arguments", fn);
        return -1;
    }
    for (int k = 0; k < numBlas; k++) out[k] = nullptr;
    // triangles per BLAS; the forest takes the BLASes of more than four
    // triangles (contiguous, first), the smaller ones are built alone
    std::vector<long long> count(numBlas, 0);
    for (int k = 0; k < numBlas; k++) {
        if (geomOffsets[k + 1] <= geomOffsets[k]) { set_error("%s: BLAS %d has no geometry", fn, k); return -1; }
        for (int g = geomOffsets[k]; g < geomOffsets[k + 1]; g++) {
            if (!geometry_ok(geoms[g])) { set_error("%s: bad geometry %d of BLAS %d", fn, g, k); return -1; }
            count[k] += geoms[g].numTris;
        }
        if (count[k] > (1 << 27)) { set_error("%s: BLAS %d exceeds the leaf encoding", fn, k); return -1; }
    }
    std::vector<int> big, small, start(numBlas);
    for (int k = 0; k < numBlas; k++) (count[k] > 4 ? big : small).push_back(k);
    long long total = 0;
    for (int pass = 0; pass < 2; pass++)
        for (int k : (pass == 0 ? big : small)) { start[k] = (int)total; total += count[k]; }
    if (total > (1ll << 28)) { set_error("%s: %lld triangles in all", fn, total); return -1; }
    float* dTri = nullptr;
    int* dBad = nullptr;
    if (hipMalloc((void**)&dTri, size_t(total) * 36) != hipSuccess || hipMalloc((void**)&dBad, sizeof(int)) != hipSuccess) {
        hipFree(dTri);
        set_error("%s: cannot allocate %lld triangles", fn, total);
        return -1;
    }
    std::vector<float> host(size_t(total) * 9);
    int bad = 0;
    hipError_t e = hipMemset(dBad, 0, sizeof(int));
    for (int k = 0; k < numBlas && e == hipSuccess; k++) {
        long long off = start[k];
        for (int g = geomOffsets[k]; g < geomOffsets[k + 1] && e == hipSuccess; g++) {
            const CudaGeometry& G = geoms[g];
            e = rtk::launch_gather_tris(G.vertices, G.numVertices, G.vertexStride, G.indices, G.indexBytes, G.numTris,
                                        dTri + 9 * size_t(off), dBad, nullptr);
            off += G.numTris;
        }
    }
    if (e == hipSuccess) e = hipMemcpy(&bad, dBad, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && !bad) e = hipMemcpy(host.data(), dTri, host.size() * 4, hipMemcpyDeviceToHost);
    hipFree(dBad);
    if (e != hipSuccess || bad) {
        hipFree(dTri);
        if (bad) set_error("%s: an index is >= its geometry's numVertices", fn);
        else set_error("%s: gather failed: %s", fn, hipGetErrorString(e));
        return -1;
    }
    auto fail_all = [&]() {
        for (int k = 0; k < numBlas; k++) { cudaBVH_destroy(out[k]); out[k] = nullptr; }
        hipFree(dTri);
        return -1;
    };
    for (int k : small) {
        out[k] = build_impl(host.data() + 9 * size_t(start[k]), (int)count[k], 4, nullptr, 1, nullptr,
                            dTri + 9 * size_t(start[k]));
        if (!out[k]) return fail_all();
    }
    if (!big.empty()) {
        const int R = (int)big.size();
        std::vector<int> off(R + 1, 0);
        for (int i = 0; i < R; i++) off[i + 1] = off[i] + (int)count[big[i]];
        std::vector<int> roots4, ford;
        std::vector<rt::Node2> fn2;
        CudaBVHHandle* forest = build_impl(host.data(), off[R], 4, off.data(), R, &roots4, dTri, &fn2, &ford);
        if (!forest) return fail_all();
        // BVH2 subtree sizes: the trees lie one after another in pre-order
        std::vector<int> sz(fn2.size(), 1);
        for (int v = (int)fn2.size() - 1; v >= 0; v--)
            if (fn2[v].triCount == 0) sz[v] = 1 + sz[fn2[v].left] + sz[fn2[v].right];
        int root2 = 0;
        for (int i = 0; i < R; i++) {
            const int k = big[i];
            const int first4 = roots4[i], end4 = i + 1 < R ? roots4[i + 1] : forest->numNodes4;
            const int t0 = off[i], nt = (int)count[k];
            auto* h = new CudaBVHHandle();
            h->numNodes4 = end4 - first4;
            std::vector<uint8_t> hn4(size_t(h->numNodes4) * 64);
            // the tree's BVH2 nodes and leaf order, rebased like its BVH4 nodes
            std::vector<rt::Node2> sub(fn2.begin() + root2, fn2.begin() + root2 + sz[root2]);
            for (rt::Node2& nd : sub) {
                if (nd.triCount == 0) { nd.left -= root2; nd.right -= root2; }
                else nd.triStart -= t0;
            }
            std::vector<int> ord(ford.begin() + t0, ford.begin() + t0 + nt);
            for (int& x : ord) x -= t0;
            h->numNodes2 = int(sub.size());
            const bool ok = hipMalloc(&h->d_nodes4, hn4.size()) == hipSuccess &&
                            hipMalloc(&h->d_nodes2, sub.size() * sizeof(rt::Node2)) == hipSuccess &&
                            hipMalloc((void**)&h->d_ordered, size_t(nt) * 4) == hipSuccess &&
                            hipMemcpy(h->d_nodes2, sub.data(), sub.size() * sizeof(rt::Node2), hipMemcpyHostToDevice) == hipSuccess &&
                            hipMemcpy(h->d_ordered, ord.data(), size_t(nt) * 4, hipMemcpyHostToDevice) == hipSuccess &&
                            hipMalloc(&h->d_tris, size_t(nt) * 48) == hipSuccess &&
                            hipMalloc((void**)&h->d_leafOfPrim, size_t(nt) * 4) == hipSuccess &&
                            rtk::launch_slice_forest(forest->d_nodes4, first4, h->numNodes4, forest->d_tris,
                                                     forest->d_leafOfPrim, t0, nt, h->d_nodes4, h->d_tris,
                                                     h->d_leafOfPrim, nullptr) == 0 &&
                            hipMemcpy(hn4.data(), h->d_nodes4, hn4.size(), hipMemcpyDeviceToHost) == hipSuccess;
            if (!ok) {
                free_handle_device(h);
                delete h;
                cudaBVH_destroy(forest);
                set_error("%s: cannot split tree %d out of the forest", fn, i);
                return fail_all();
            }
            h->depth4 = bvh4_depth(hn4);
            root2 += sz[root2];
            out[k] = make_handle(h, host.data() + 9 * size_t(t0), nt);
            if (!out[k]) { cudaBVH_destroy(forest); return fail_all(); }
        }
        cudaBVH_destroy(forest);
    }
    hipFree(dTri);
    return 0;
}

void cudaBVH_destroy(CudaBVH_t bvh) {
    if (!bvh) return;
    async_release(bvh);          // the background tracer no longer renders from it
    if (bvh->stream) hipStreamSynchronize(hstream(bvh)); else hipDeviceSynchronize();
    free_handle_device(bvh);
    delete bvh;
}

float cudaBVH_tracePrimary(CudaBVH_t bvh, int side, float camOx, float camOy, float camOz, float* outHitT) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_tracePrimary")) return 0.f;
    if (side <= 0 || (long long)side * side > (1ll << 30)) { set_error("cudaBVH_tracePrimary: bad side %d", side); return 0.f; }
    const int n = side * side;
    float* dt = (float*)work_buffer(bvh, size_t(n) * sizeof(float));
    if (!dt) return 0.f;
    rtk::TraceArgs a;
    a.n = n; a.W = side; a.H = side;
    const float o[3] = {camOx, camOy, camOz};
    for (int k = 0; k < 3; k++) { a.cam.o[k] = o[k]; a.cam.f[k] = 0.f; a.cam.r[k] = 0.f; a.cam.u[k] = 0.f; }
    a.cam.f[2] = 1.f; a.cam.r[0] = 1.f; a.cam.u[1] = 1.f;   // looking +z, 90 degrees (backend :69-74)
    a.cam.sx = 1.f; a.cam.sy = 1.f;
    a.t = dt;
    a.costOrder = true;              // repeated per frame: tiles in the last frame's cost order
    hipEvent_t e0, e1;
    hipEventCreate(&e0); hipEventCreate(&e1);
    hipEventRecord(e0, hstream(bvh));
    int rc = rtk::launch_trace(bvh, bvh->scratch, a, rtk::kModeClosest, false, true, false, hstream(bvh));
    hipEventRecord(e1, hstream(bvh));
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0); hipEventDestroy(e1);
    if (rc != 0 || finish_sync(bvh->scratch, hstream(bvh), "cudaBVH_tracePrimary") != 0) return 0.f;
    if (outHitT) {
        float* pin = pinned_buffer(bvh, size_t(n));
        if (!pin) return 0.f;
        HIP_TRY(hipMemcpyAsync(pin, dt, size_t(n) * sizeof(float), hipMemcpyDeviceToHost, hstream(bvh)), return 0.f);
        HIP_TRY(hipStreamSynchronize(hstream(bvh)), return 0.f);
        float* o[1] = {outHitT};
        const float* i[1] = {pin};
        copy_host_arrays(o, i, 1, size_t(n));
    }
    const float mrs = ms > 0.f ? (float)n / (ms * 1000.f) : 0.f;
    if (rt::verbose()) fprintf(stderr, "[CudaBVH] Primary trace: %dx%d = %d rays in %.3fms -> %.0f MR/s\n", side, side, n, ms, mrs);
    return mrs;
}

float cudaBVH_traceDiffuse(CudaBVH_t bvh, int numRays, const float* rayOx, const float* rayOy, const float* rayOz,
                           const float* rayDx, const float* rayDy, const float* rayDz, float* outHitT) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_traceDiffuse")) return 0.f;
    if (numRays <= 0 || !rayOx || !rayOy || !rayOz || !rayDx || !rayDy || !rayDz) {
        set_error("cudaBVH_traceDiffuse: bad arguments (numRays %d)", numRays);
        return 0.f;
    }
    const size_t n = size_t(numRays);
    float* w = (float*)work_buffer(bvh, n * 9 * sizeof(float));
    if (!w) return 0.f;
    const float* src[6] = {rayOx, rayOy, rayOz, rayDx, rayDy, rayDz};
    HIP_TRY(hipStreamSynchronize(hstream(bvh)), return 0.f);   // no earlier copy still reads the staging
    float* pin = pinned_buffer(bvh, 6 * n);
    if (!pin) return 0.f;
    float* dst[6] = {pin, pin + n, pin + 2 * n, pin + 3 * n, pin + 4 * n, pin + 5 * n};
    copy_host_arrays(dst, src, 6, n);
    HIP_TRY(hipMemcpyAsync(w, pin, 6 * n * sizeof(float), hipMemcpyHostToDevice, hstream(bvh)), return 0.f);
    // t in (0, 1e30) (backend :594), written on the device
    hipLaunchKernelGGL(k_fill_interval, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0,
                       hstream(bvh), w + 6 * n, w + 7 * n, n, 0.0f, RT_MISS_T);
    // Rays are sorted on the GPU by octant + Morton code of the origin before
    // the trace (the reference sorts by Morton code on the host, backend
    // :1387-1433); hits come back in the caller's order.
    void* ws = sort_workspace(bvh, numRays);
    if (!ws) return 0.f;
    CudaRayPacket rp{w, w + n, w + 2 * n, w + 3 * n, w + 4 * n, w + 5 * n, w + 6 * n, w + 7 * n};
    CudaHitPacket hp{w + 8 * n, nullptr, nullptr, nullptr};
    hipEvent_t e0, e1;
    hipEventCreate(&e0); hipEventCreate(&e1);
    hipEventRecord(e0, hstream(bvh));
    int rc = rtk::trace_sorted(bvh, bvh->scratch, numRays, rp, &hp, nullptr, rtk::kModeClosest, false, ws, hstream(bvh));
    hipEventRecord(e1, hstream(bvh));
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0); hipEventDestroy(e1);
    if (rc != 0 || finish_sync(bvh->scratch, hstream(bvh), "cudaBVH_traceDiffuse") != 0) return 0.f;
    if (outHitT) {
        HIP_TRY(hipMemcpyAsync(pin, w + 8 * n, n * sizeof(float), hipMemcpyDeviceToHost, hstream(bvh)), return 0.f);
        HIP_TRY(hipStreamSynchronize(hstream(bvh)), return 0.f);
        float* o[1] = {outHitT};
        const float* i[1] = {pin};
        copy_host_arrays(o, i, 1, n);
    }
    return ms > 0.f ? (float)numRays / (ms * 1000.f) : 0.f;
}

int cudaBVH_traceToGPUPtr(CudaBVH_t bvh, int width, int height, float camOx, float camOy, float camOz,
                          void* d_outRGBA, int outFmt) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_traceToGPUPtr")) return -1;
    if (!d_outRGBA) { set_error("cudaBVH_traceToGPUPtr: NULL output"); return -1; }
    return shade_to_device(bvh, bvh->scratch, hstream(bvh), width, height, camOx, camOy, camOz, d_outRGBA, outFmt);
}

int cudaBVH_traceToRGBA(CudaBVH_t bvh, int width, int height, float camOx, float camOy, float camOz,
                        uint32_t* outRGBA_host) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_traceToRGBA")) return -1;
    if (!outRGBA_host) { set_error("cudaBVH_traceToRGBA: NULL output"); return -1; }
    if (width <= 0 || height <= 0) { set_error("bad image size %dx%d", width, height); return -1; }
    const size_t bytes = size_t(width) * height * 4;
    void* d = work_buffer(bvh, bytes);
    if (!d) return -1;
    if (shade_to_device(bvh, bvh->scratch, hstream(bvh), width, height, camOx, camOy, camOz, d, 0) != 0) return -1;
    HIP_TRY(hipMemcpyAsync(outRGBA_host, d, bytes, hipMemcpyDeviceToHost, hstream(bvh)), return -1);
    return finish_sync(bvh->scratch, hstream(bvh), "cudaBVH_traceToRGBA");
}

// ---------------------------------------------------------------------------
// Background tracer (backend :1093-1336): one process-wide worker thread with
// double-buffered host output, as in the reference. Calls are serialised.
// ---------------------------------------------------------------------------
namespace {
struct AsyncTracer {
    std::mutex mu;
    std::condition_variable req, ready;
    std::thread th;
    bool running = false, hasRequest = false;
    bool hasHostResult = false;          // host[readIdx] holds a finished host frame
    uint64_t reqSeq = 0, doneSeq = 0;    // requests submitted / the latest one finished
    int device = 0;
    CudaBVH_t bvh = nullptr;             // handle of the pending request
    CudaBVH_t busy = nullptr;            // handle of the frame being rendered
    int w = 0, h = 0, fmt = 0;
    float cx = 0, cy = 0, cz = 0;
    void* gpuTarget = nullptr;
    std::vector<uint8_t> host[2];
    int readIdx = 0, writeIdx = 1;
    uint64_t bufSize = 0;
    void* dbuf = nullptr; size_t dbytes = 0;
    hipStream_t stream = nullptr;
    rt::TraceScratch scratch;            // the worker's own traversal scratch
    std::mutex startMu;

    void loop() {
        hipSetDevice(device);
        hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        std::unique_lock<std::mutex> lk(mu);
        while (running) {
            req.wait(lk, [&] { return hasRequest || !running; });
            if (!running) break;
            CudaBVH_t b = bvh; int W = w, H = h, F = fmt; float X = cx, Y = cy, Z = cz;
            void* tgt = gpuTarget; int wi = writeIdx; uint64_t sz = bufSize;
            const uint64_t seq = reqSeq;     // requests not yet taken are merged into the latest
            hasRequest = false;
            busy = b;
            lk.unlock();
            if (tgt) {
                shade_to_device(b, scratch, stream, W, H, X, Y, Z, tgt, F);
                hipStreamSynchronize(stream);
            } else {
                if (dbytes < sz) { hipFree(dbuf); dbuf = nullptr; dbytes = 0; if (hipMalloc(&dbuf, sz) == hipSuccess) dbytes = sz; }
                if (dbuf && shade_to_device(b, scratch, stream, W, H, X, Y, Z, dbuf, F) == 0) {
                    hipMemcpyAsync(host[wi].data(), dbuf, sz, hipMemcpyDeviceToHost, stream);
                    hipStreamSynchronize(stream);
                }
            }
            lk.lock();
            if (!tgt) { readIdx = wi; writeIdx = 1 - wi; hasHostResult = true; }
            doneSeq = seq;
            busy = nullptr;
            ready.notify_all();
        }
        // The worker only stops when the process exits (static destruction),
        // when the HIP runtime may already be torn down: its stream, staging
        // and traversal scratch are left to the process teardown.
        lk.unlock();
    }
    void start(uint64_t bytes) {
        std::lock_guard<std::mutex> sg(startMu);
        if (running) return;
        hipGetDevice(&device);
        bufSize = bytes;
        host[0].assign(bytes, 0); host[1].assign(bytes, 0);
        running = true;
        th = std::thread([this] { loop(); });
    }
    void stop() {
        std::lock_guard<std::mutex> sg(startMu);
        { std::lock_guard<std::mutex> g(mu); if (!running) return; running = false; }
        req.notify_all();
        if (th.joinable()) th.join();
        hasHostResult = false; hasRequest = false;
    }
    ~AsyncTracer() { stop(); }
};
AsyncTracer g_async;
std::atomic<int> g_async_calls_host{0}, g_async_calls_gpu{0};
}  // namespace

// Waits until the background tracer neither renders from nor holds a pending
// request for bvh (a pending one is dropped), so the handle can be freed.
static void async_release(CudaBVH_t bvh) {
    std::unique_lock<std::mutex> lk(g_async.mu);
    if (!g_async.running) return;
    if (g_async.hasRequest && g_async.bvh == bvh) {
        g_async.hasRequest = false;                 // dropped: nobody waits for it any longer
        g_async.doneSeq = g_async.reqSeq;
        g_async.ready.notify_all();
    }
    if (g_async.bvh == bvh) g_async.bvh = nullptr;
    g_async.ready.wait(lk, [&] { return g_async.busy != bvh; });
}

int cudaBVH_traceToHostPtr(CudaBVH_t bvh, int width, int height, float camOx, float camOy, float camOz,
                           void* h_outRGBA, int outFmt, uint64_t bufSize) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_traceToHostPtr")) return -1;
    const uint64_t bpp = outFmt == 3 ? 16 : (outFmt == 2 ? 8 : 4);
    if (!h_outRGBA || width <= 0 || height <= 0 || bufSize < uint64_t(width) * height * bpp) {
        set_error("cudaBVH_traceToHostPtr: bad arguments");
        return -1;
    }
    const int call = ++g_async_calls_host;
    g_async.start(bufSize);
    std::unique_lock<std::mutex> lk(g_async.mu);
    if (bufSize > g_async.bufSize) {
        set_error("cudaBVH_traceToHostPtr: bufSize %llu exceeds the first call's %llu",
                  (unsigned long long)bufSize, (unsigned long long)g_async.bufSize);
        return -1;
    }
    g_async.bvh = bvh; g_async.w = width; g_async.h = height; g_async.fmt = outFmt;
    g_async.cx = camOx; g_async.cy = camOy; g_async.cz = camOz; g_async.gpuTarget = nullptr;
    g_async.hasRequest = true;
    const uint64_t my = ++g_async.reqSeq;
    g_async.req.notify_all();
    if (call == 1) g_async.ready.wait(lk, [my] { return g_async.doneSeq >= my; });   // first frame blocks
    if (g_async.hasHostResult) memcpy(h_outRGBA, g_async.host[g_async.readIdx].data(), bufSize);
    return 0;
}

void cudaBVH_traceToGPUAsync(CudaBVH_t bvh, int width, int height, float camOx, float camOy, float camOz,
                             void* d_outRGBA, int outFmt) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_traceToGPUAsync") || !d_outRGBA) return;
    const int call = ++g_async_calls_gpu;
    const uint64_t bpp = outFmt == 3 ? 16 : (outFmt == 2 ? 8 : 4);
    g_async.start(uint64_t(width) * height * bpp);
    std::unique_lock<std::mutex> lk(g_async.mu);
    g_async.bvh = bvh; g_async.w = width; g_async.h = height; g_async.fmt = outFmt;
    g_async.cx = camOx; g_async.cy = camOy; g_async.cz = camOz; g_async.gpuTarget = d_outRGBA;
    g_async.hasRequest = true;
    const uint64_t my = ++g_async.reqSeq;
    g_async.req.notify_all();
    if (call == 1) g_async.ready.wait(lk, [my] { return g_async.doneSeq >= my; });   // first frame blocks
}

void* cudaBVH_importBufferFd(int fd, uint64_t size) {
    rt::clear_error();
    if (fd < 0 || size == 0) { set_error("cannot import fd %d of %llu bytes", fd, (unsigned long long)size); return nullptr; }
    hipExternalMemoryHandleDesc desc;
    memset(&desc, 0, sizeof desc);
    desc.type = hipExternalMemoryHandleTypeOpaqueFd;
    desc.handle.fd = fd;
    desc.size = size;
    hipExternalMemory_t ext = nullptr;
    hipError_t e = hipImportExternalMemory(&ext, &desc);
    if (e != hipSuccess) { set_error("cannot import fd %d: %s", fd, hipGetErrorString(e)); return nullptr; }
    hipExternalMemoryBufferDesc bd;
    memset(&bd, 0, sizeof bd);
    bd.offset = 0; bd.size = size;
    void* p = nullptr;
    e = hipExternalMemoryGetMappedBuffer(&p, ext, &bd);
    if (e != hipSuccess) { set_error("cannot map imported fd %d: %s", fd, hipGetErrorString(e)); hipDestroyExternalMemory(ext); return nullptr; }
    return p;
}

int cudaBVH_readGPUMemoryFd(int fd, uint64_t allocationSize, uint64_t offset, uint64_t size, void* dst) {
    rt::clear_error();
    if (!dst) { set_error("cudaBVH_readGPUMemoryFd: NULL destination"); return -1; }
    if (fd < 0 || offset > allocationSize || size > allocationSize - offset) {
        set_error("cudaBVH_readGPUMemoryFd: fd %d, bytes [%llu, %llu) outside an allocation of %llu", fd,
                  (unsigned long long)offset, (unsigned long long)(offset + size), (unsigned long long)allocationSize);
        return -1;
    }
    hipExternalMemoryHandleDesc desc;
    memset(&desc, 0, sizeof desc);
    desc.type = hipExternalMemoryHandleTypeOpaqueFd;
    desc.handle.fd = fd;
    desc.size = allocationSize;
    hipExternalMemory_t ext = nullptr;
    hipError_t e = hipImportExternalMemory(&ext, &desc);
    if (e != hipSuccess) { set_error("cannot import fd %d: %s", fd, hipGetErrorString(e)); return -1; }
    // The range is mapped from the start of the allocation and the copy
    // starts at `offset`: this HIP runtime returns the allocation's base for
    // a mapped buffer whatever its descriptor's offset (measured on a
    // dma-buf import), so the offset is applied here.
    hipExternalMemoryBufferDesc bd;
    memset(&bd, 0, sizeof bd);
    bd.offset = 0; bd.size = offset + size;
    void* p = nullptr;
    e = hipExternalMemoryGetMappedBuffer(&p, ext, &bd);
    if (e == hipSuccess) e = hipMemcpy(dst, (const char*)p + offset, size, hipMemcpyDeviceToHost);
    if (p) hipFree(p);
    hipDestroyExternalMemory(ext);
    if (e != hipSuccess) { set_error("readGPUMemoryFd failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

void cudaBVH_getBounds(CudaBVH_t bvh, float* cx, float* cy, float* cz, float* extent) {
    if (!bvh) { set_error("cudaBVH_getBounds: NULL handle"); return; }
    if (cx) *cx = bvh->bboxCenter[0];
    if (cy) *cy = bvh->bboxCenter[1];
    if (cz) *cz = bvh->bboxCenter[2];
    if (extent) *extent = bvh->bboxExtent;
}

int cudaBVH_getNumTris(CudaBVH_t bvh) { return bvh ? bvh->numOrdered : 0; }
int cudaBVH_getNumBVH4Nodes(CudaBVH_t bvh) { return bvh ? bvh->numNodes4 : 0; }
int cudaBVH_getNumCWBVHNodes(CudaBVH_t bvh) { return bvh ? bvh->numCWBVH : 0; }

int cudaBVH_buildCWBVH(CudaBVH_t bvh) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_buildCWBVH")) return -1;
    return ensure_cwbvh(bvh) == 0 ? bvh->numCWBVH : -1;
}

int cudaBVH_getCWBVH(CudaBVH_t bvh, const uint32_t** nodes, const int** triOrder, int* depth) {
    if (!check_handle(bvh, "cudaBVH_getCWBVH") || !bvh->numCWBVH) return 0;
    if (nodes) *nodes = bvh->h_cw.data();
    if (triOrder) *triOrder = bvh->h_cwTriOrder.data();
    if (depth) *depth = bvh->cwDepth;
    return bvh->numCWBVH;
}
int cudaBVH_getDepth(CudaBVH_t bvh) { return bvh ? bvh->depth4 : 0; }

int cudaBVH_getBuildStats(CudaBVH_t bvh, float* stats) {
    rt::clear_error();
    if (!bvh || !stats) { set_error("cudaBVH_getBuildStats: NULL argument"); return -1; }
    memcpy(stats, bvh->buildStats, sizeof(bvh->buildStats));
    return 0;
}

const void* cudaBVH_getNodeData(CudaBVH_t bvh) {
    return (bvh && ensure_exports(bvh)) ? bvh->h_nodes4.data() : nullptr;
}

void cudaBVH_getTriData(CudaBVH_t bvh, const float* outPtrs[9]) {
    if (!outPtrs) return;
    const bool ok = bvh && ensure_exports(bvh);
    for (int i = 0; i < 9; i++) outPtrs[i] = ok ? bvh->h_soa[i].data() : nullptr;
}

int cudaBVH_getStacklessBVH2(CudaBVH_t bvh, uint32_t** outData) {
    if (!outData) return 0;
    if (!bvh || !ensure_exports(bvh) || bvh->h_dfs.empty()) { *outData = nullptr; return 0; }
    *outData = bvh->h_dfs.data();
    return bvh->numDFS;
}

int cudaBVH_getPackedTris(CudaBVH_t bvh, float** outData) {
    if (!outData) return 0;
    if (!bvh || !ensure_exports(bvh) || bvh->h_packed.empty()) { *outData = nullptr; return 0; }
    *outData = bvh->h_packed.data();
    return int(bvh->h_packed.size() / 4);
}

void cudaBVH_resetDevice(void) {
    g_async.stop();
    hipDeviceSynchronize();
}

CudaBVH_t cudaBVH_buildFromAABBs(const float* aabbs, int numAABBs) {
    rt::clear_error();
    if (!aabbs || numAABBs <= 0) { set_error("cudaBVH_buildFromAABBs: need aabbs and numAABBs > 0"); return nullptr; }
    // One degenerate triangle per box whose bounds equal the box (backend :1549-1563).
    std::vector<float> t(size_t(numAABBs) * 9);
    for (int i = 0; i < numAABBs; i++) {
        const float* b = aabbs + 6 * size_t(i);
        float* d = &t[9 * size_t(i)];
        d[0] = b[0]; d[1] = b[1]; d[2] = b[2];
        d[3] = b[3]; d[4] = b[4]; d[5] = b[5];
        d[6] = b[3]; d[7] = b[1]; d[8] = b[2];
    }
    return build_impl(t.data(), numAABBs, 4);
}

int cudaBVH_buildTLASFast(const float* aabbs, int numAABBs, uint32_t** outNodes, int** outOrdered) {
    rt::clear_error();
    if (outNodes) *outNodes = nullptr;
    if (outOrdered) *outOrdered = nullptr;
    if (!aabbs || numAABBs <= 0 || !outNodes) { set_error("cudaBVH_buildTLASFast: bad arguments"); return 0; }
    if (numAABBs > (1 << 27)) { set_error("cudaBVH_buildTLASFast: %d boxes exceed the leaf encoding", numAABBs); return 0; }
    // The LBVH is built on the GPU (lbvh.hip); boxes in, nodes and leaf
    // order out through host arrays the caller frees (backend :1620-1698).
    const size_t n = size_t(numAABBs), total = 2 * n - 1;
    const size_t inB = n * 24, nodeB = total * 32, ordB = n * 4, wsB = rtk::lbvh_scratch_bytes(numAABBs);
    char* d = nullptr;
    if (hipMalloc((void**)&d, inB + nodeB + ordB + wsB + 256) != hipSuccess) {
        set_error("cudaBVH_buildTLASFast: cannot allocate %zu device bytes", inB + nodeB + ordB + wsB);
        return 0;
    }
    float* dBoxes = (float*)d;
    uint32_t* dNodes = (uint32_t*)(d + inB);
    int* dOrd = (int*)(d + inB + nodeB);
    void* ws = d + ((inB + nodeB + ordB + 255) & ~size_t(255));
    uint32_t* p = (uint32_t*)malloc(nodeB);
    int* o = (int*)malloc(ordB);
    hipStream_t s = nullptr;
    int rc = (p && o) ? 0 : -1;
    if (rc != 0) set_error("out of host memory");
    if (rc == 0 && hipMemcpyAsync(dBoxes, aabbs, inB, hipMemcpyHostToDevice, s) != hipSuccess) {
        set_error("cudaBVH_buildTLASFast: upload failed");
        rc = -1;
    }
    if (rc == 0) rc = rtk::lbvh_build(dBoxes, numAABBs, dNodes, dOrd, ws, s);
    if (rc == 0 && (hipMemcpyAsync(p, dNodes, nodeB, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipMemcpyAsync(o, dOrd, ordB, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess)) {
        set_error("cudaBVH_buildTLASFast: download failed");
        rc = -1;
    }
    hipFree(d);
    if (rc != 0) { free(p); free(o); return 0; }
    *outNodes = p;
    if (outOrdered) *outOrdered = o; else free(o);
    return (int)total;
}

// ------------------------------- extensions -------------------------------

int cudaBVH_setStream(CudaBVH_t bvh, void* s) {
    if (!check_handle(bvh, "cudaBVH_setStream")) return -1;
    bvh->stream = s;
    return 0;
}
void* cudaBVH_getStream(CudaBVH_t bvh) { return bvh ? bvh->stream : nullptr; }
const void* cudaBVH_deviceNodes(CudaBVH_t bvh) { return bvh ? bvh->d_nodes4 : nullptr; }
const void* cudaBVH_deviceTris(CudaBVH_t bvh) { return bvh ? bvh->d_tris : nullptr; }

static bool packet_ok(const CudaRayPacket* r) {
    return r && r->ox && r->oy && r->oz && r->dx && r->dy && r->dz && r->tmin && r->tmax;
}

// Shared body of the packet entry points. Host pointers are staged through
// the handle's work buffer.
// Builds (once) the compressed 8-wide BVH and its triangle order from the BVH2.
static int ensure_cwbvh(CudaBVH_t bvh) {
    if (bvh->d_cwNodes) return 0;
    if (!ensure_exports(bvh)) return -1;
    if (bvh->h_nodes2.empty()) { set_error("CWBVH: this handle keeps no BVH2 (batched forest)"); return -1; }
    const int nn = rt::build_cwbvh(bvh->h_nodes2, bvh->h_cw, bvh->h_cwTriOrder, bvh->cwDepth);
    if (nn <= 0) { set_error("CWBVH build failed"); return -1; }
    const size_t nt = bvh->h_cwTriOrder.size();
    std::vector<float> tri(nt * 12, 0.f);
    for (size_t k = 0; k < nt; k++) {
        const int li = bvh->h_cwTriOrder[k];
        float* o = &tri[12 * k];
        for (int v = 0; v < 3; v++)
            for (int a = 0; a < 3; a++) o[4 * v + a] = bvh->h_soa[3 * v + a][li];
        const int id = bvh->h_ordered[li];
        memcpy(&o[3], &id, 4);
    }
    if (hipMalloc(&bvh->d_cwNodes, bvh->h_cw.size() * 4) != hipSuccess ||
        hipMalloc(&bvh->d_cwTris, tri.size() * 4) != hipSuccess ||
        hipMemcpy(bvh->d_cwNodes, bvh->h_cw.data(), bvh->h_cw.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(bvh->d_cwTris, tri.data(), tri.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(bvh->d_cwNodes); hipFree(bvh->d_cwTris);
        bvh->d_cwNodes = bvh->d_cwTris = nullptr;
        set_error("CWBVH: device allocation failed");
        return -1;
    }
    bvh->numCWBVH = nn;
    return 0;
}

static int trace_packet(CudaBVH_t bvh, int n, const CudaRayPacket* rays, const CudaHitPacket* hits, uint8_t* occ,
                        int* cn, int* ct, unsigned flags, int mode, const char* fn) {
    rt::clear_error();
    if (!check_handle(bvh, fn)) return -1;
    if (n == 0) return 0;                              // nothing to trace (buffers may be NULL)
    if (n < 0 || !packet_ok(rays)) { set_error("%s: bad ray packet", fn); return -1; }
    if (mode == rtk::kModeClosest && (!hits || !hits->t)) { set_error("%s: hit packet needs t", fn); return -1; }
    if (mode == rtk::kModeAny && !occ) { set_error("%s: NULL occlusion output", fn); return -1; }
    const bool sorted = (flags & CUDABVH_FLAG_SORT_RAYS) != 0;
    if (sorted && cn) { set_error("%s: counted traces do not sort", fn); return -1; }
    const unsigned isect = flags & CUDABVH_ISECT_MASK;
    if (isect > 1) { set_error("%s: unknown intersection mode %u", fn, isect); return -1; }
    const bool cw = (flags & CUDABVH_FLAG_CWBVH) != 0;
    if (cw && (sorted || cn || isect != 0)) {
        set_error("%s: the CWBVH traversal is watertight, unsorted and uncounted", fn);
        return -1;
    }
    if (cw && ensure_cwbvh(bvh) != 0) return -1;
    hipStream_t s = hstream(bvh);
    rtk::TraceArgs a;
    a.n = n;
    const bool host = (flags & CUDABVH_FLAG_HOST_POINTERS) != 0;
    const size_t N = size_t(n);
    if (!host) {
        a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
        a.tmin = rays->tmin; a.tmax = rays->tmax;
        if (hits) { a.t = hits->t; a.id = hits->primId; a.u = hits->u; a.v = hits->v; }
        a.occ = occ; a.cntNodes = cn; a.cntTris = ct;
        int rc;
        if (cw) {
            rc = rtk::launch_trace_cw(bvh->d_cwNodes, bvh->d_cwTris, bvh->cwDepth, bvh->scratch, a, mode, s);
        } else if (sorted) {
            void* ws = sort_workspace(bvh, n);
            if (!ws) return -1;
            rc = rtk::trace_sorted(bvh, bvh->scratch, n, *rays, hits, occ, mode, isect == 1, ws, s);
        } else {
            rc = rtk::launch_trace(bvh, bvh->scratch, a, mode, isect == 1, false, cn != nullptr, s);
        }
        if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) return finish_sync(bvh->scratch, s, fn);
        return rc;
    }
    // host pointers: 8 ray arrays + 4 hit arrays + 2 counters (+ occlusion bytes)
    float* w = (float*)work_buffer(bvh, N * 14 * 4 + N);
    if (!w) return -1;
    const float* src[8] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    for (int k = 0; k < 8; k++) HIP_TRY(hipMemcpyAsync(w + k * N, src[k], N * 4, hipMemcpyHostToDevice, s), return -1);
    a.ox = w; a.oy = w + N; a.oz = w + 2 * N; a.dx = w + 3 * N; a.dy = w + 4 * N; a.dz = w + 5 * N;
    a.tmin = w + 6 * N; a.tmax = w + 7 * N;
    a.t = w + 8 * N; a.id = (int*)(w + 9 * N); a.u = w + 10 * N; a.v = w + 11 * N;
    a.cntNodes = (int*)(w + 12 * N); a.cntTris = (int*)(w + 13 * N);
    a.occ = (uint8_t*)(w + 14 * N);
    int rc;
    if (cw) {
        rc = rtk::launch_trace_cw(bvh->d_cwNodes, bvh->d_cwTris, bvh->cwDepth, bvh->scratch, a, mode, s);
    } else if (sorted) {
        void* ws = sort_workspace(bvh, n);
        if (!ws) return -1;
        CudaRayPacket rp{a.ox, a.oy, a.oz, a.dx, a.dy, a.dz, a.tmin, a.tmax};
        CudaHitPacket hp{a.t, a.id, a.u, a.v};
        rc = rtk::trace_sorted(bvh, bvh->scratch, n, rp, &hp, a.occ, mode, isect == 1, ws, s);
    } else {
        rc = rtk::launch_trace(bvh, bvh->scratch, a, mode, isect == 1, false, cn != nullptr, s);
    }
    if (rc != 0) return rc;
    if (mode == rtk::kModeClosest) {
        HIP_TRY(hipMemcpyAsync(hits->t, a.t, N * 4, hipMemcpyDeviceToHost, s), return -1);
        if (hits->primId) HIP_TRY(hipMemcpyAsync(hits->primId, a.id, N * 4, hipMemcpyDeviceToHost, s), return -1);
        if (hits->u) HIP_TRY(hipMemcpyAsync(hits->u, a.u, N * 4, hipMemcpyDeviceToHost, s), return -1);
        if (hits->v) HIP_TRY(hipMemcpyAsync(hits->v, a.v, N * 4, hipMemcpyDeviceToHost, s), return -1);
    } else {
        HIP_TRY(hipMemcpyAsync(occ, a.occ, N, hipMemcpyDeviceToHost, s), return -1);
    }
    if (cn) HIP_TRY(hipMemcpyAsync(cn, a.cntNodes, N * 4, hipMemcpyDeviceToHost, s), return -1);
    if (ct) HIP_TRY(hipMemcpyAsync(ct, a.cntTris, N * 4, hipMemcpyDeviceToHost, s), return -1);
    return finish_sync(bvh->scratch, s, fn);
}

int cudaBVH_traceClosest(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, const CudaHitPacket* hits, unsigned flags) {
    return trace_packet(bvh, numRays, rays, hits, nullptr, nullptr, nullptr, flags, rtk::kModeClosest, "cudaBVH_traceClosest");
}
int cudaBVH_traceAny(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, uint8_t* occluded, unsigned flags) {
    return trace_packet(bvh, numRays, rays, nullptr, occluded, nullptr, nullptr, flags, rtk::kModeAny, "cudaBVH_traceAny");
}
// Closest hit through the static scratch of the current device (no device or
// host allocation): the ray class's launch configuration with at most as many
// resident workgroups per CU as the static spill region holds.
int cudaBVH_traceClosestNoAlloc(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, const CudaHitPacket* hits,
                                unsigned flags) {
    const char* fn = "cudaBVH_traceClosestNoAlloc";
    rt::clear_error();
    if (!check_handle(bvh, fn)) return -1;
    if (numRays == 0) return 0;
    if (numRays < 0 || !packet_ok(rays) || !hits || !hits->t) { set_error("%s: bad arguments", fn); return -1; }
    const unsigned isect = flags & CUDABVH_ISECT_MASK;
    if (isect > 1 || (flags & ~(unsigned)(CUDABVH_FLAG_SYNC | CUDABVH_ISECT_MASK))) {
        set_error("%s: device packets only (flags 0x%x: no host pointers, ray sort or CWBVH)", fn, flags);
        return -1;
    }
    rt::TraceScratch sc;
    if (rtk::noalloc_scratch(sc) != 0) return -1;
    hipStream_t s = hstream(bvh);
    rtk::TraceArgs a;
    a.n = numRays;
    a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
    a.tmin = rays->tmin; a.tmax = rays->tmax;
    a.t = hits->t; a.id = hits->primId; a.u = hits->u; a.v = hits->v;
    if (rtk::clear_status(sc, s) != 0) return -1;
    const int rc = rtk::launch_trace(bvh, sc, a, rtk::kModeClosest, isect == 1, false, false, s);
    if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) return finish_sync(sc, s, fn);
    return rc;
}

int cudaBVH_traceClosestCounted(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, const CudaHitPacket* hits,
                                int* nodeVisits, int* triTests, unsigned flags) {
    if (!nodeVisits || !triTests) { set_error("cudaBVH_traceClosestCounted: NULL counters"); return -1; }
    return trace_packet(bvh, numRays, rays, hits, nullptr, nodeVisits, triTests, flags, rtk::kModeClosest,
                        "cudaBVH_traceClosestCounted");
}
int cudaBVH_traceAnyCounted(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, uint8_t* occluded, int* nodeVisits,
                            int* triTests, unsigned flags) {
    if (!nodeVisits || !triTests) { set_error("cudaBVH_traceAnyCounted: NULL counters"); return -1; }
    return trace_packet(bvh, numRays, rays, nullptr, occluded, nodeVisits, triTests, flags, rtk::kModeAny,
                        "cudaBVH_traceAnyCounted");
}

int cudaBVH_genPinholeRays(int width, int height, const CudaCamera* cam, const CudaRayPacket* out, void* stream) {
    rt::clear_error();
    if (!cam || !packet_ok(out) || width <= 0 || height <= 0 || (long long)width * height > (1ll << 30)) {
        set_error("cudaBVH_genPinholeRays: bad arguments");
        return -1;
    }
    return rtk::launch_gen_pinhole(to_cam(*cam), width, height, *out, (hipStream_t)stream);
}

int cudaBVH_tracePinholeClosest(CudaBVH_t bvh, int width, int height, const CudaCamera* cam, const CudaHitPacket* hits,
                                unsigned flags) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_tracePinholeClosest")) return -1;
    if (!cam || !hits || !hits->t || width <= 0 || height <= 0 || (long long)width * height > (1ll << 30)) {
        set_error("cudaBVH_tracePinholeClosest: bad arguments");
        return -1;
    }
    const unsigned isect = flags & CUDABVH_ISECT_MASK;
    if (isect > 1 || (flags & (CUDABVH_FLAG_HOST_POINTERS | CUDABVH_FLAG_SORT_RAYS))) {
        set_error("cudaBVH_tracePinholeClosest: unsupported flags 0x%x", flags);
        return -1;
    }
    rtk::TraceArgs a;
    a.n = width * height; a.W = width; a.H = height;
    a.cam = to_cam(*cam);
    a.t = hits->t; a.id = hits->primId; a.u = hits->u; a.v = hits->v;
    a.costOrder = (flags & CUDABVH_FLAG_COST_ORDER) != 0;
    int rc = rtk::launch_trace(bvh, bvh->scratch, a, rtk::kModeClosest, isect == 1, true, false, hstream(bvh));
    if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) return finish_sync(bvh->scratch, hstream(bvh), "cudaBVH_tracePinholeClosest");
    return rc;
}

int cudaBVH_genBounceRays(CudaBVH_t bvh, int width, int height, const CudaCamera* cam, const CudaHitPacket* primary,
                          uint32_t seed, const float light[3], const CudaRayPacket* diffuseOut,
                          const CudaRayPacket* shadowOut, int* srcIndex, int* d_count, unsigned flags) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_genBounceRays")) return -1;
    if (!cam || !primary || !primary->t || !primary->primId || !light || !packet_ok(diffuseOut) ||
        !packet_ok(shadowOut) || !srcIndex || !d_count || width <= 0 || height <= 0 ||
        (long long)width * height > (1ll << 30)) {
        set_error("cudaBVH_genBounceRays: bad arguments");
        return -1;
    }
    if (flags & ~(unsigned)(CUDABVH_FLAG_SYNC | CUDABVH_FLAG_TILE_ORDER)) {
        set_error("cudaBVH_genBounceRays: unsupported flags 0x%x", flags);
        return -1;
    }
    int rc = rtk::launch_bounce(bvh, bvh->scratch, to_cam(*cam), width, height, primary->t, primary->primId, seed,
                                light, *diffuseOut, *shadowOut, srcIndex, d_count, (flags & CUDABVH_FLAG_TILE_ORDER) != 0,
                                hstream(bvh));
    if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) HIP_TRY(hipStreamSynchronize(hstream(bvh)), return -1);
    return rc;
}

static int trace_queued(CudaBVH_t bvh, int capacity, const int* d_count, const CudaRayPacket* rays,
                        const CudaHitPacket* hits, uint8_t* occ, unsigned flags, int mode, const char* fn) {
    rt::clear_error();
    if (!check_handle(bvh, fn)) return -1;
    if (capacity < 0 || !d_count || !packet_ok(rays) || (mode == rtk::kModeClosest && (!hits || !hits->t)) ||
        (mode == rtk::kModeAny && !occ)) {
        set_error("%s: bad arguments", fn);
        return -1;
    }
    const unsigned isect = flags & CUDABVH_ISECT_MASK;
    if (isect > 1 || (flags & (CUDABVH_FLAG_HOST_POINTERS | CUDABVH_FLAG_SORT_RAYS))) {
        set_error("%s: unsupported flags 0x%x", fn, flags);
        return -1;
    }
    if (capacity == 0) return 0;
    rtk::TraceArgs a;
    a.n = capacity; a.nDev = d_count;
    a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
    a.tmin = rays->tmin; a.tmax = rays->tmax;
    if (hits) { a.t = hits->t; a.id = hits->primId; a.u = hits->u; a.v = hits->v; }
    a.occ = occ;
    int rc = rtk::launch_trace(bvh, bvh->scratch, a, mode, isect == 1, false, false, hstream(bvh));
    if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) return finish_sync(bvh->scratch, hstream(bvh), fn);
    return rc;
}

int cudaBVH_traceClosestQueued(CudaBVH_t bvh, int capacity, const int* d_count, const CudaRayPacket* rays,
                               const CudaHitPacket* hits, unsigned flags) {
    return trace_queued(bvh, capacity, d_count, rays, hits, nullptr, flags, rtk::kModeClosest,
                        "cudaBVH_traceClosestQueued");
}

int cudaBVH_traceAnyQueued(CudaBVH_t bvh, int capacity, const int* d_count, const CudaRayPacket* rays,
                           uint8_t* occluded, unsigned flags) {
    return trace_queued(bvh, capacity, d_count, rays, nullptr, occluded, flags, rtk::kModeAny,
                        "cudaBVH_traceAnyQueued");
}

int cudaBVH_setTraceConfig(CudaBVH_t bvh, int rayClass, int schedule, int ldsDepth, int blocksPerCU) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_setTraceConfig")) return -1;
    if (rayClass < 0 || rayClass >= rt::kNumClasses || schedule < rt::kScheduleBatch ||
        schedule > rt::kScheduleRefill || (schedule == rt::kSchedulePacket && rayClass != rt::kClassCoherent) ||
        (ldsDepth != 4 && ldsDepth != 8 && ldsDepth != 12 &&
        ldsDepth != 16) || blocksPerCU < 0) {
        set_error("cudaBVH_setTraceConfig: bad class %d / schedule %d / ldsDepth %d / blocksPerCU %d", rayClass,
                  schedule, ldsDepth, blocksPerCU);
        return -1;
    }
    bvh->tune[rayClass].schedule = schedule;
    bvh->tune[rayClass].ldsDepth = ldsDepth;
    bvh->tune[rayClass].blocksPerCU = blocksPerCU;
    return 0;
}

int cudaBVH_getTraceConfig(CudaBVH_t bvh, int rayClass, int* schedule, int* ldsDepth, int* blocksPerCU) {
    if (!check_handle(bvh, "cudaBVH_getTraceConfig") || rayClass < 0 || rayClass >= rt::kNumClasses) return -1;
    if (schedule) *schedule = bvh->tune[rayClass].schedule;
    if (ldsDepth) *ldsDepth = bvh->tune[rayClass].ldsDepth;
    if (blocksPerCU) *blocksPerCU = bvh->tune[rayClass].blocksPerCU;
    return 0;
}

// Autotuner: times every candidate launch configuration of one ray class on
// the caller's workload (a pinhole frame for the coherent class, a device ray
// packet otherwise) and keeps the fastest in the handle. results (nullable,
// room for 4 floats x 36 candidates) receives {schedule, ldsDepth,
// blocksPerCU, milliseconds} per candidate.
int cudaBVH_autotune(CudaBVH_t bvh, int rayClass, int width, int height, const CudaCamera* cam, int numRays,
                     const CudaRayPacket* rays, int reps, float* results, int* numResults) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_autotune")) return -1;
    const bool coherent = rayClass == rt::kClassCoherent;
    if (rayClass < 0 || rayClass >= rt::kNumClasses || reps <= 0 || (coherent && (!cam || width <= 0 || height <= 0)) ||
        (!coherent && (numRays <= 0 || !packet_ok(rays)))) {
        set_error("cudaBVH_autotune: bad arguments");
        return -1;
    }
    const int n = coherent ? width * height : numRays;
    hipStream_t s = hstream(bvh);
    void* out = nullptr;
    HIP_TRY(hipMalloc(&out, size_t(n) * 16), return -1);
    rtk::TraceArgs base;
    base.n = n;
    // coherent candidates of the batch schedule run in the cost order, as
    // the frame-after-frame callers of tracePinholeClosest run them
    if (coherent) { base.W = width; base.H = height; base.cam = to_cam(*cam); base.costOrder = true; }
    else {
        base.ox = rays->ox; base.oy = rays->oy; base.oz = rays->oz; base.dx = rays->dx; base.dy = rays->dy;
        base.dz = rays->dz; base.tmin = rays->tmin; base.tmax = rays->tmax;
    }
    float* o = (float*)out;
    if (rayClass == rt::kClassAny) base.occ = (uint8_t*)out;
    else { base.t = o; base.id = (int*)(o + n); base.u = o + 2 * size_t(n); base.v = o + 3 * size_t(n); }
    const int mode = rayClass == rt::kClassAny ? rtk::kModeAny : rtk::kModeClosest;
    const int schedules[4] = {rt::kScheduleBatch, rt::kScheduleWhileWhile, rt::kSchedulePacket, rt::kScheduleRefill};
    const int depths[4] = {4, 8, 12, 16};
    const int bpcs[3] = {0, 4, 2};
    const rt::TraceConfig saved = bvh->tune[rayClass];
    rt::TraceConfig best = saved;
    float bestMs = 1e30f;
    int row = 0, rc = 0;
    hipEvent_t e0, e1;
    hipEventCreate(&e0); hipEventCreate(&e1);
    for (int sch : schedules) {
        if (sch == rt::kSchedulePacket && !coherent) continue;
        for (int d : depths) {
            for (int b : bpcs) {
                bvh->tune[rayClass].schedule = sch;
                bvh->tune[rayClass].ldsDepth = d;
                bvh->tune[rayClass].blocksPerCU = b;
                rtk::TraceArgs a = base;
                rc = rtk::launch_trace(bvh, bvh->scratch, a, mode, false, coherent, false, s);   // warm-up
                hipEventRecord(e0, s);
                for (int r = 0; r < reps && rc == 0; r++) {
                    a = base;
                    rc = rtk::launch_trace(bvh, bvh->scratch, a, mode, false, coherent, false, s);
                }
                hipEventRecord(e1, s);
                hipEventSynchronize(e1);
                if (rc != 0) break;
                float ms = 0.f;
                hipEventElapsedTime(&ms, e0, e1);
                ms /= reps;
                if (results) {
                    results[4 * row] = (float)sch; results[4 * row + 1] = (float)d;
                    results[4 * row + 2] = (float)b; results[4 * row + 3] = ms;
                }
                row++;
                if (ms < bestMs) { bestMs = ms; best = bvh->tune[rayClass]; }
            }
            if (rc != 0) break;
        }
        if (rc != 0) break;
    }
    hipEventDestroy(e0); hipEventDestroy(e1);
    hipFree(out);
    bvh->tune[rayClass] = rc == 0 ? best : saved;
    if (numResults) *numResults = row;
    return rc;
}

int cudaBVH_sortRays(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, uint32_t* keysOut, int* permOut) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_sortRays")) return -1;
    if (numRays <= 0 || !packet_ok(rays) || !keysOut || !permOut) { set_error("cudaBVH_sortRays: bad arguments"); return -1; }
    void* ws = sort_workspace(bvh, numRays);
    if (!ws) return -1;
    return rtk::sort_rays(bvh, numRays, *rays, keysOut, permOut, ws, hstream(bvh));
}

// ------------------------------- two-level scenes ----------------------------

static void invert_affine(const float* m, float* out) {
    // double-precision inverse of the 3x4 affine map, rounded to float
    double a[3][3], t[3];
    for (int r = 0; r < 3; r++) { for (int c = 0; c < 3; c++) a[r][c] = m[4 * r + c]; t[r] = m[4 * r + 3]; }
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                       a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv[3][3];
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int r = 0; r < 3; r++) {
        double tr = 0;
        for (int c = 0; c < 3; c++) { out[4 * r + c] = (float)inv[r][c]; tr -= inv[r][c] * t[c]; }
        out[4 * r + 3] = (float)tr;
    }
}

static bool scene_ok(CudaScene_t s, const char* fn) {
    if (!s) { set_error("%s: NULL scene", fn); return false; }
    return true;
}

CudaScene_t cudaBVH_sceneCreate(const CudaBVH_t* blas, int numBlas, const CudaInstance* inst, int numInst) {
    rt::clear_error();
    if (!blas || numBlas <= 0 || !inst || numInst <= 0) { set_error("cudaBVH_sceneCreate: bad arguments"); return nullptr; }
    for (int b = 0; b < numBlas; b++)
        if (!blas[b]) { set_error("cudaBVH_sceneCreate: BLAS %d is NULL", b); return nullptr; }
    auto* sc = new CudaSceneHandle();
    // defaults: batch / LDS 12, what the autotuner picks for all three
    // classes on the instanced 1080p scene (while-while / 8 measured 15-70 %
    // slower there)
    sc->tune2l[0].schedule = 1;
    sc->tune2l[0].ldsDepth = 12;
    sc->tune2l[1].schedule = 1;
    sc->tune2l[1].ldsDepth = 12;
    sc->tune2l[2] = sc->tune2l[0];
    sc->blas.assign(blas, blas + numBlas);
    sc->inst.resize(numInst);
    sc->o2w.resize(size_t(numInst) * 12);
    std::vector<float> aabbs(size_t(numInst) * 6);
    for (int i = 0; i < numInst; i++) {
        const CudaInstance& I = inst[i];
        if (I.blas < 0 || I.blas >= numBlas) { set_error("instance %d: bad BLAS index %d", i, I.blas); delete sc; return nullptr; }
        memcpy(&sc->o2w[12 * size_t(i)], I.objectToWorld, 48);
        invert_affine(I.objectToWorld, sc->inst[i].w2o);
        sc->inst[i].blas = I.blas;
        for (int a = 0; a < 3; a++) sc->inst[i].albedo[a] = I.albedo[a];
        // world box: the 8 transformed corners of the BLAS box (double), widened
        const CudaBVHHandle* B = blas[I.blas];
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int c = 0; c < 8; c++) {
            const double p[3] = {(c & 1) ? B->bboxMax[0] : B->bboxMin[0], (c & 2) ? B->bboxMax[1] : B->bboxMin[1],
                                 (c & 4) ? B->bboxMax[2] : B->bboxMin[2]};
            for (int r = 0; r < 3; r++) {
                const double w = I.objectToWorld[4 * r] * p[0] + I.objectToWorld[4 * r + 1] * p[1] +
                                 I.objectToWorld[4 * r + 2] * p[2] + I.objectToWorld[4 * r + 3];
                lo[r] = std::min(lo[r], w); hi[r] = std::max(hi[r], w);
            }
        }
        for (int r = 0; r < 3; r++) {
            const double pad = 1e-5 * (hi[r] - lo[r]) + 1e-6 * (std::fabs(lo[r]) + std::fabs(hi[r])) + 1e-30;
            aabbs[6 * size_t(i) + r] = (float)(lo[r] - pad);
            aabbs[6 * size_t(i) + 3 + r] = (float)(hi[r] + pad);
        }
        sc->maxBlasDepth = std::max(sc->maxBlasDepth, B->depth4);
    }
    std::vector<float> t(size_t(numInst) * 9);
    for (int i = 0; i < numInst; i++) {
        const float* b = &aabbs[6 * size_t(i)];
        float* d = &t[9 * size_t(i)];
        d[0] = b[0]; d[1] = b[1]; d[2] = b[2]; d[3] = b[3]; d[4] = b[4]; d[5] = b[5]; d[6] = b[3]; d[7] = b[1]; d[8] = b[2];
    }
    sc->tlas = build_impl(t.data(), numInst, 1);      // one instance per TLAS leaf
    if (!sc->tlas) { delete sc; return nullptr; }
    std::vector<rtk::BlasDev> bd(numBlas);
    for (int b = 0; b < numBlas; b++)
        bd[b] = rtk::BlasDev{(const uint4*)blas[b]->d_nodes4, (const float4*)blas[b]->d_tris, blas[b]->d_leafOfPrim,
                             blas[b]->depth4, blas[b]->numTris};
    if (hipMalloc((void**)&sc->d_inst, sizeof(rtk::InstanceDev) * numInst) != hipSuccess ||
        hipMalloc((void**)&sc->d_blas, sizeof(rtk::BlasDev) * numBlas) != hipSuccess ||
        hipMemcpy(sc->d_inst, sc->inst.data(), sizeof(rtk::InstanceDev) * numInst, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(sc->d_blas, bd.data(), sizeof(rtk::BlasDev) * numBlas, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("cudaBVH_sceneCreate: device allocation failed");
        hipFree(sc->d_inst); hipFree(sc->d_blas);
        cudaBVH_destroy(sc->tlas);
        delete sc;
        return nullptr;
    }
    return sc;
}

void cudaBVH_sceneDestroy(CudaScene_t sc) {
    if (!sc) return;
    if (sc->stream) hipStreamSynchronize((hipStream_t)sc->stream); else hipDeviceSynchronize();
    rtk::pt_free(sc);
    hipFree(sc->d_inst); hipFree(sc->d_blas);
    rtk::free_scratch(sc->scratch);
    cudaBVH_destroy(sc->tlas);
    delete sc;
}

int cudaBVH_sceneSetStream(CudaScene_t sc, void* stream) {
    if (!scene_ok(sc, "cudaBVH_sceneSetStream")) return -1;
    sc->stream = stream;
    sc->tlas->stream = stream;
    return 0;
}

CudaBVH_t cudaBVH_sceneTLAS(CudaScene_t sc) { return sc ? sc->tlas : nullptr; }

int cudaBVH_sceneWorldToObject(CudaScene_t sc, float* out) {
    if (!scene_ok(sc, "cudaBVH_sceneWorldToObject") || !out) return -1;
    for (size_t i = 0; i < sc->inst.size(); i++) memcpy(out + 12 * i, sc->inst[i].w2o, 48);
    return 0;
}

static int scene_trace(CudaScene_t sc, int n, const int* d_count, const CudaRayPacket* rays, const CudaSceneHit* hits,
                       uint8_t* occ, int* cn, int* ct, unsigned flags, bool any, const char* fn, int cls = -1) {
    rt::clear_error();
    if (!scene_ok(sc, fn)) return -1;
    if (n == 0) return 0;
    if (n < 0 || !packet_ok(rays) || (!any && (!hits || !hits->t)) || (any && !occ)) { set_error("%s: bad arguments", fn); return -1; }
    if (flags & ~(unsigned)CUDABVH_FLAG_SYNC) { set_error("%s: unsupported flags 0x%x", fn, flags); return -1; }
    rtk::Trace2LArgs a;
    a.n = n; a.nDev = d_count;
    a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
    a.tmin = rays->tmin; a.tmax = rays->tmax;
    if (hits) { a.t = hits->t; a.inst_out = hits->instId; a.id = hits->primId; a.u = hits->u; a.v = hits->v; }
    a.occ = occ; a.cntNodes = cn; a.cntTris = ct;
    hipStream_t s = (hipStream_t)sc->stream;
    int rc = rtk::launch_trace2l(sc, sc->scratch, a, cls >= 0 ? cls : (any ? 1 : 0), s);
    if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) return finish_sync(sc->scratch, s, fn);
    return rc;
}

int cudaBVH_sceneTraceClosest(CudaScene_t sc, int numRays, const CudaRayPacket* rays, const CudaSceneHit* hits,
                              unsigned flags) {
    return scene_trace(sc, numRays, nullptr, rays, hits, nullptr, nullptr, nullptr, flags, false, "cudaBVH_sceneTraceClosest");
}
int cudaBVH_sceneTraceAny(CudaScene_t sc, int numRays, const CudaRayPacket* rays, uint8_t* occluded, unsigned flags) {
    return scene_trace(sc, numRays, nullptr, rays, nullptr, occluded, nullptr, nullptr, flags, true, "cudaBVH_sceneTraceAny");
}
int cudaBVH_sceneTraceClosestCounted(CudaScene_t sc, int numRays, const CudaRayPacket* rays, const CudaSceneHit* hits,
                                     int* nodeVisits, int* triTests, unsigned flags) {
    if (!nodeVisits || !triTests) { set_error("cudaBVH_sceneTraceClosestCounted: NULL counters"); return -1; }
    return scene_trace(sc, numRays, nullptr, rays, hits, nullptr, nodeVisits, triTests, flags, false,
                       "cudaBVH_sceneTraceClosestCounted");
}
int cudaBVH_sceneTraceClosestQueued(CudaScene_t sc, int capacity, const int* d_count, const CudaRayPacket* rays,
                                    const CudaSceneHit* hits, unsigned flags) {
    if (!d_count) { set_error("cudaBVH_sceneTraceClosestQueued: NULL count"); return -1; }
    return scene_trace(sc, capacity, d_count, rays, hits, nullptr, nullptr, nullptr, flags, false,
                       "cudaBVH_sceneTraceClosestQueued");
}
int cudaBVH_getLaunchInfo(CudaBVH_t bvh, int* info) {
    rt::clear_error();
    if (!bvh || !info) { set_error("cudaBVH_getLaunchInfo: NULL argument"); return -1; }
    memcpy(info, bvh->scratch.launchInfo, sizeof(int) * rt::kLaunchInfoInts);
    return 0;
}

int cudaBVH_sceneGetLaunchInfo(CudaScene_t sc, int* info) {
    rt::clear_error();
    if (!sc || !info) { set_error("cudaBVH_sceneGetLaunchInfo: NULL argument"); return -1; }
    memcpy(info, sc->scratch.launchInfo, sizeof(int) * rt::kLaunchInfoInts);
    return 0;
}

int cudaBVH_getTraceStatus(CudaBVH_t bvh) {
    rt::clear_error();
    if (!check_handle(bvh, "cudaBVH_getTraceStatus")) return -1;
    return rtk::take_status(bvh->scratch, hstream(bvh), "cudaBVH_getTraceStatus");
}

int cudaBVH_sceneGetTraceStatus(CudaScene_t sc) {
    rt::clear_error();
    if (!scene_ok(sc, "cudaBVH_sceneGetTraceStatus")) return -1;
    const int a = rtk::take_status(sc->scratch, (hipStream_t)sc->stream, "cudaBVH_sceneGetTraceStatus");
    if (a < 0) return -1;
    const int b = rtk::pt_take_status(sc);
    if (b < 0) return -1;
    const int c = rtk::take_status(sc->tlas->scratch, (hipStream_t)sc->stream, "cudaBVH_sceneGetTraceStatus");
    if (c < 0) return -1;
    if (a | b | c) set_error("cudaBVH_sceneGetTraceStatus: traversal stack overflow (status 0x%x)", (unsigned)(a | b | c));
    return a | b | c;
}

int cudaBVH_sceneHitNormals(CudaScene_t sc, int numRays, const CudaSceneHit* hits, float* nx, float* ny, float* nz,
                            unsigned flags) {
    rt::clear_error();
    if (!scene_ok(sc, "cudaBVH_sceneHitNormals")) return -1;
    if (numRays == 0) return 0;
    if (numRays < 0 || !hits || !hits->instId || !hits->primId || !nx || !ny || !nz) {
        set_error("cudaBVH_sceneHitNormals: bad arguments");
        return -1;
    }
    if (flags & ~(unsigned)CUDABVH_FLAG_SYNC) { set_error("cudaBVH_sceneHitNormals: unsupported flags 0x%x", flags); return -1; }
    hipStream_t s = (hipStream_t)sc->stream;
    int rc = rtk::launch_hit_normals(sc, numRays, hits->instId, hits->primId, nx, ny, nz, s);
    if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) HIP_TRY(hipStreamSynchronize(s), return -1);
    return rc;
}

int cudaBVH_sceneTraceAnyQueued(CudaScene_t sc, int capacity, const int* d_count, const CudaRayPacket* rays,
                                uint8_t* occluded, unsigned flags) {
    if (!d_count) { set_error("cudaBVH_sceneTraceAnyQueued: NULL count"); return -1; }
    return scene_trace(sc, capacity, d_count, rays, nullptr, occluded, nullptr, nullptr, flags, true,
                       "cudaBVH_sceneTraceAnyQueued");
}

// ------------------------------- batched scenes ------------------------------

struct CudaBVHBatchHandle {
    CudaBVHHandle* forest = nullptr;     // every scene's nodes and triangles
    int numScenes = 0;
    std::vector<int> triOff, roots4;
    int* d_roots = nullptr;
    hipStream_t own = nullptr;
    // one captured launch per mode (closest hit, any hit), each valid for the
    // arguments and traversal-scratch generation it was captured with
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    rtk::TraceArgs key[2]{};
    uint64_t keyGen[2] = {0, 0};
};

static bool batch_ok(CudaBVHBatch_t b, const char* fn) {
    if (!b) { set_error("%s: NULL batch", fn); return false; }
    return true;
}

CudaBVHBatch_t cudaBVH_buildBatch(const CudaTri* tris, const int* triOffsets, int numScenes) {
    rt::clear_error();
    if (!tris || !triOffsets || numScenes <= 0) { set_error("cudaBVH_buildBatch: bad arguments"); return nullptr; }
    if (triOffsets[0] != 0) { set_error("cudaBVH_buildBatch: triOffsets[0] must be 0"); return nullptr; }
    for (int k = 0; k < numScenes; k++)
        if (triOffsets[k + 1] <= triOffsets[k]) { set_error("cudaBVH_buildBatch: scene %d is empty", k); return nullptr; }
    auto* b = new CudaBVHBatchHandle();
    b->numScenes = numScenes;
    b->triOff.assign(triOffsets, triOffsets + numScenes + 1);
    b->forest = build_impl(reinterpret_cast<const float*>(tris), triOffsets[numScenes], 4, triOffsets, numScenes,
                           &b->roots4);
    if (!b->forest) { delete b; return nullptr; }
    if (hipMalloc((void**)&b->d_roots, sizeof(int) * numScenes) != hipSuccess ||
        hipMemcpy(b->d_roots, b->roots4.data(), sizeof(int) * numScenes, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("cudaBVH_buildBatch: device allocation failed");
        cudaBVH_destroy(b->forest);
        hipFree(b->d_roots);
        delete b;
        return nullptr;
    }
    return b;
}

void cudaBVH_batchDestroy(CudaBVHBatch_t b) {
    if (!b) return;
    if (b->forest && b->forest->stream) hipStreamSynchronize((hipStream_t)b->forest->stream);
    if (b->own) hipStreamSynchronize(b->own);
    for (int m = 0; m < 2; m++) {
        if (b->exec[m]) hipGraphExecDestroy(b->exec[m]);
        if (b->graph[m]) hipGraphDestroy(b->graph[m]);
    }
    if (b->own) hipStreamDestroy(b->own);
    hipFree(b->d_roots);
    cudaBVH_destroy(b->forest);
    delete b;
}

int cudaBVH_batchSetStream(CudaBVHBatch_t b, void* stream) {
    if (!batch_ok(b, "cudaBVH_batchSetStream")) return -1;
    b->forest->stream = stream;
    return 0;
}

CudaBVH_t cudaBVH_batchForest(CudaBVHBatch_t b) { return b ? b->forest : nullptr; }

int cudaBVH_batchRoots(CudaBVHBatch_t b, int* roots4) {
    if (!batch_ok(b, "cudaBVH_batchRoots") || !roots4) return -1;
    memcpy(roots4, b->roots4.data(), sizeof(int) * b->numScenes);
    return 0;
}

static int batch_trace(CudaBVHBatch_t b, int totalRays, const int* d_rayOffsets, const CudaRayPacket* rays,
                       const CudaHitPacket* hits, uint8_t* occ, unsigned flags, bool any, const char* fn) {
    rt::clear_error();
    if (!batch_ok(b, fn)) return -1;
    if (totalRays == 0) return 0;
    if (totalRays < 0 || !d_rayOffsets || !packet_ok(rays) || (!any && (!hits || !hits->t)) || (any && !occ)) {
        set_error("%s: bad arguments", fn);
        return -1;
    }
    if (flags & ~(unsigned)(CUDABVH_FLAG_SYNC | CUDABVH_FLAG_GRAPH)) { set_error("%s: unsupported flags 0x%x", fn, flags); return -1; }
    rtk::TraceArgs a;
    memset(&a, 0, sizeof a);      // the graph key compares the whole struct, padding included
    a.n = totalRays;
    a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
    a.tmin = rays->tmin; a.tmax = rays->tmax;
    if (hits) { a.t = hits->t; a.id = hits->primId; a.u = hits->u; a.v = hits->v; }
    a.occ = occ;
    a.rayOff = d_rayOffsets; a.roots = b->d_roots; a.numScenes = b->numScenes;
    const int mode = any ? rtk::kModeAny : rtk::kModeClosest;
    hipStream_t s = hstream(b->forest);
    if (!(flags & CUDABVH_FLAG_GRAPH)) {
        const int rc = rtk::launch_trace(b->forest, b->forest->scratch, a, mode, false, false, false, s);
        if (rc == 0 && (flags & CUDABVH_FLAG_SYNC)) return finish_sync(b->forest->scratch, s, fn);
        return rc;
    }
    // graph mode: one captured launch (counter reset + persistent trace), replayed
    if (!s) {
        if (!b->own && hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) != hipSuccess) {
            set_error("%s: cannot create a stream", fn);
            return -1;
        }
        HIP_TRY(hipStreamSynchronize(nullptr), return -1);
        s = b->own;
    }
    const int m = any ? 1 : 0;
    rt::TraceScratch& scr = b->forest->scratch;
    if (!b->exec[m] || b->keyGen[m] != scr.gen || memcmp(&b->key[m], &a, sizeof(a)) != 0) {
        if (b->exec[m]) { hipGraphExecDestroy(b->exec[m]); b->exec[m] = nullptr; }
        if (b->graph[m]) { hipGraphDestroy(b->graph[m]); b->graph[m] = nullptr; }
        rtk::TraceArgs e = a;      // eager launch allocates the traversal scratch
        if (rtk::launch_trace(b->forest, scr, e, mode, false, false, false, s) != 0) return -1;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), return -1);
        rtk::TraceArgs c = a;
        const int rc = rtk::launch_trace(b->forest, scr, c, mode, false, false, false, s);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &gr);
        if (rc != 0 || ec != hipSuccess) {
            if (gr) hipGraphDestroy(gr);
            if (rc == 0) set_error("%s: capture failed: %s", fn, hipGetErrorString(ec));
            return -1;
        }
        b->graph[m] = gr;
        HIP_TRY(hipGraphInstantiate(&b->exec[m], b->graph[m], nullptr, nullptr, 0), return -1);
        b->key[m] = a;
        b->keyGen[m] = scr.gen;
    }
    HIP_TRY(hipGraphLaunch(b->exec[m], s), return -1);
    if ((flags & CUDABVH_FLAG_SYNC) || s == b->own) return finish_sync(scr, s, fn);
    return 0;
}

int cudaBVH_batchTraceClosest(CudaBVHBatch_t b, int totalRays, const int* d_rayOffsets, const CudaRayPacket* rays,
                              const CudaHitPacket* hits, unsigned flags) {
    return batch_trace(b, totalRays, d_rayOffsets, rays, hits, nullptr, flags, false, "cudaBVH_batchTraceClosest");
}

int cudaBVH_batchTraceAny(CudaBVHBatch_t b, int totalRays, const int* d_rayOffsets, const CudaRayPacket* rays,
                          uint8_t* occluded, unsigned flags) {
    return batch_trace(b, totalRays, d_rayOffsets, rays, nullptr, occluded, flags, true, "cudaBVH_batchTraceAny");
}

int cudaBVH_sceneSetTraceConfig(CudaScene_t sc, int rayClass, int schedule, int ldsDepth, int blocksPerCU) {
    rt::clear_error();
    if (!scene_ok(sc, "cudaBVH_sceneSetTraceConfig")) return -1;
    if (rayClass < 0 || rayClass > 2 || schedule < 0 || schedule > 3 || (schedule == 2 && rayClass != 1) ||
        (ldsDepth != 4 && ldsDepth != 8 && ldsDepth != 12 && ldsDepth != 16) || blocksPerCU < 0) {
        set_error("cudaBVH_sceneSetTraceConfig: class %d, schedule %d, LDS depth %d, blocks/CU %d not supported",
                  rayClass, schedule, ldsDepth, blocksPerCU);
        return -1;
    }
    sc->tune2l[rayClass].schedule = schedule;
    sc->tune2l[rayClass].ldsDepth = ldsDepth;
    sc->tune2l[rayClass].blocksPerCU = blocksPerCU;
    return 0;
}

int cudaBVH_sceneGetTraceConfig(CudaScene_t sc, int rayClass, int* schedule, int* ldsDepth, int* blocksPerCU) {
    if (!scene_ok(sc, "cudaBVH_sceneGetTraceConfig") || rayClass < 0 || rayClass > 2) return -1;
    if (schedule) *schedule = sc->tune2l[rayClass].schedule;
    if (ldsDepth) *ldsDepth = sc->tune2l[rayClass].ldsDepth;
    if (blocksPerCU) *blocksPerCU = sc->tune2l[rayClass].blocksPerCU;
    return 0;
}

// Times every (schedule, LDS depth, blocks per CU) candidate on the caller's
// rays (best of reps launches each) and keeps the fastest.
int cudaBVH_sceneAutotune(CudaScene_t sc, int rayClass, int numRays, const CudaRayPacket* rays, int reps,
                          float* results, int* numResults) {
    rt::clear_error();
    if (!scene_ok(sc, "cudaBVH_sceneAutotune")) return -1;
    if (rayClass < 0 || rayClass > 2 || numRays <= 0 || !packet_ok(rays) || reps <= 0) {
        set_error("cudaBVH_sceneAutotune: bad arguments");
        return -1;
    }
    hipStream_t s = (hipStream_t)sc->stream;
    void* buf = nullptr;
    if (hipMalloc(&buf, (size_t)numRays * 20) != hipSuccess) { set_error("cudaBVH_sceneAutotune: no memory"); return -1; }
    CudaSceneHit hit{(float*)buf, (int*)buf + numRays, (int*)buf + 2 * (size_t)numRays,
                     (float*)buf + 3 * (size_t)numRays, (float*)buf + 4 * (size_t)numRays};
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const int depths[4] = {4, 8, 12, 16}, bpcs[3] = {0, 4, 2};
    // schedules: 0 while-while, 1 batch, 2 batch nearest-first (any hit), 3 refill
    const int closestScheds[3] = {0, 1, 3}, anyScheds[4] = {0, 1, 2, 3};
    const int* scheds = rayClass == 1 ? anyScheds : closestScheds;
    const int nsched = rayClass == 1 ? 4 : 3;
    rt::TraceConfig keep = sc->tune2l[rayClass], best = keep;
    float bestMs = 1e30f;
    int k = 0, rc = 0;
    for (int si = 0; si < nsched && rc == 0; si++)
        for (int di = 0; di < 4 && rc == 0; di++)
            for (int bi = 0; bi < 3 && rc == 0; bi++) {
                sc->tune2l[rayClass].schedule = scheds[si];
                sc->tune2l[rayClass].ldsDepth = depths[di];
                sc->tune2l[rayClass].blocksPerCU = bpcs[bi];
                float ms = 1e30f;
                for (int r = 0; r <= reps && rc == 0; r++) {      // the first launch warms up
                    hipEventRecord(e0, s);
                    rc = rayClass == 1 ? scene_trace(sc, numRays, nullptr, rays, nullptr, (uint8_t*)buf, nullptr,
                                                     nullptr, 0, true, "cudaBVH_sceneAutotune")
                                       : scene_trace(sc, numRays, nullptr, rays, &hit, nullptr, nullptr, nullptr, 0,
                                                     false, "cudaBVH_sceneAutotune", rayClass);
                    hipEventRecord(e1, s);
                    hipEventSynchronize(e1);
                    float t = 0.f;
                    hipEventElapsedTime(&t, e0, e1);
                    if (r > 0) ms = std::min(ms, t);
                }
                if (results) {
                    results[4 * k] = (float)scheds[si];
                    results[4 * k + 1] = (float)depths[di];
                    results[4 * k + 2] = (float)bpcs[bi];
                    results[4 * k + 3] = ms;
                }
                k++;
                if (ms < bestMs) { bestMs = ms; best = sc->tune2l[rayClass]; }
            }
    sc->tune2l[rayClass] = rc == 0 ? best : keep;
    if (numResults) *numResults = k;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(buf);
    return rc;
}

int cudaBVH_sceneAutotunePathTracer(CudaScene_t sc, const CudaCamera* cam, const CudaPathParams* params,
                                    const CudaPathImage* image, int reps, float* results, int* numResults) {
    rt::clear_error();
    if (!scene_ok(sc, "cudaBVH_sceneAutotunePathTracer")) return -1;
    if (!cam || !params || !image || reps <= 0) { set_error("cudaBVH_sceneAutotunePathTracer: bad arguments"); return -1; }
    const int cands[6] = {0, 4096, 16384, 32768, 65536, 131072};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("cudaBVH_sceneAutotunePathTracer: cannot create events");
        return -1;
    }
    hipStream_t s = (hipStream_t)sc->stream;
    float best = 1e30f;
    int bestTail = sc->ptTail, rc = 0;
    for (int k = 0; k < 6 && rc == 0; k++) {
        CudaPathParams p = *params;
        p.tailPaths = cands[k];
        p.spp = 1;
        rc = rtk::path_trace(sc, *cam, p, *image, nullptr, CUDABVH_FLAG_GRAPH | CUDABVH_FLAG_SYNC);   // capture
        p.spp = reps;
        if (rc == 0) {
            hipEventRecord(e0, s);
            rc = rtk::path_trace(sc, *cam, p, *image, nullptr, CUDABVH_FLAG_GRAPH);
            hipEventRecord(e1, s);
            hipEventSynchronize(e1);
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)reps;
        if (results) { results[2 * k] = (float)cands[k]; results[2 * k + 1] = ms; }
        if (rc == 0 && ms < best) { best = ms; bestTail = cands[k]; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (numResults) *numResults = 6;
    if (rc == 0) sc->ptTail = bestTail;
    return rc;
}

int cudaBVH_scenePathTrace(CudaScene_t sc, const CudaCamera* cam, const CudaPathParams* params,
                           const CudaPathImage* image, long long* rayCounts, unsigned flags) {
    rt::clear_error();
    if (!scene_ok(sc, "cudaBVH_scenePathTrace")) return -1;
    if (!cam || !params || !image) { set_error("cudaBVH_scenePathTrace: NULL argument"); return -1; }
    if (flags & ~(unsigned)(CUDABVH_FLAG_SYNC | CUDABVH_FLAG_GRAPH | CUDABVH_FLAG_SORT_RAYS)) {
        set_error("cudaBVH_scenePathTrace: unsupported flags 0x%x", flags);
        return -1;
    }
    return rtk::path_trace(sc, *cam, *params, *image, rayCounts, flags);
}

}  // extern "C"
