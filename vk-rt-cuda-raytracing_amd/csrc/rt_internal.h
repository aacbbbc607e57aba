// rt_internal.h — internal structures of libcudabvh_hip.so (host side).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

// hipStream_t without the HIP headers (host-only files include this header).
typedef struct ihipStream_t* hipStream_t;

#include "../../include/cudabvh.h"

namespace rt {

// Records the failure message for cudaBVH_lastError() on this thread and
// prints it with the reference's "[CudaBVH]" prefix (backend :25).
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
bool verbose();

// ---------------------------------------------------------------------------
// BVH2 node of the build (pre-order numbering; children follow parents).
// ---------------------------------------------------------------------------
struct Box { float mn[3], mx[3]; };
struct Node2 { Box box; int left, right, triStart, triCount; };

// BVH4 node format shared with the reference (backend :31-34).
constexpr int kNode4Bytes = 64;
constexpr int kMaxLeaf = 8;             // leaf encoding holds count-1 in 3 bits

// Leaf child encoding (backend :393).
inline int32_t encode_leaf(int first, int count) { return -((first << 3) | (count - 1)) - 2; }

// Host FP16 conversions (IEEE binary16, round-to-nearest-even plus the
// outward-rounded variants used for child bounds).
uint16_t f2h_rne(float x);
uint16_t f2h_down(float x);
uint16_t f2h_up(float x);
float h2f(uint16_t h);

// Result of the device build (build.hip); all pointers are device memory
// owned by the caller afterwards.
struct GpuBVH {
    Node2* nodes2 = nullptr;  int numNodes2 = 0;   // pre-order BVH2
    void* nodes4 = nullptr;   int numNodes4 = 0;   // BVH4, 64 B per node, pre-order
    int depth4 = 0;
    void* tris = nullptr;                          // leaf order, 48 B per triangle
    int* leafOf = nullptr;                         // original index -> leaf position
    int* ordered = nullptr;                        // leaf position -> original index
    int levels = 0;
};
// Builds one BVH, or with sceneOff (numScenes+1 host offsets into d_tris) a
// forest of numScenes independent BVHs in one level-synchronous pass: node
// arrays shared, each tree contiguous in pre-order, leaf triangles carrying
// their index within their scene; roots4 receives each tree's BVH4 root.
int gpu_build(const float* d_tris, int n, int maxLeaf, hipStream_t s, GpuBVH& out, const int* sceneOff = nullptr,
              int numScenes = 1, std::vector<int>* roots4 = nullptr);
void scan_exclusive(const int* in, int* out, int n, int* total, int* tiles, hipStream_t s);
int scan_tiles_needed(int n);

// Compressed 8-wide BVH from the pre-order BVH2: 20 uint32 (80 B) per node,
// triOrder[k] = leaf-order triangle of CWBVH triangle k. Returns the node
// count (-1 if a leaf exceeds 4 triangles).
int build_cwbvh(const std::vector<Node2>& n2, std::vector<uint32_t>& nodes, std::vector<int>& triOrder, int& depth);

// Stackless DFS BVH2 export (backend :456-522): 8 uint32 per node.
void export_dfs(const std::vector<Node2>& nodes, std::vector<uint32_t>& out);



// Device scratch of one traversal context: the persistent-kernel work
// counter, a status word (bit 0: stack overflow, unreachable by
// construction) and the per-lane stack spill region (grown on demand).
// Launch configuration of one ray class, chosen by the autotuner: the wave
// schedule (batch: a wave keeps 64 rays until all finish; while-while: node
// and leaf phases with idle lanes refilled from a wave-local work pool;
// packet: one shared stack per wave of coherent pinhole rays), the
// LDS stack depth per lane (4, 8, 12 or 16 entries) and resident workgroups
// per CU (0 = as many as occupancy allows).
enum { kClassCoherent = 0, kClassIncoherent = 1, kClassAny = 2, kNumClasses = 3 };
enum { kScheduleBatch = 0, kScheduleWhileWhile = 1, kSchedulePacket = 2, kScheduleRefill = 3 };
struct TraceConfig {
    int schedule = kScheduleBatch;
    int ldsDepth = 16;
    int blocksPerCU = 0;
};

// Resources of one traversal kernel, read once per kernel (hipFuncGetAttributes
// and the occupancy calculator).
struct KernelInfo {
    int maxBlocks = 0;                  // resident workgroups per CU
    int vgprs = 0, ldsBytes = 0, scratchBytes = 0;
};

// What the last traversal launch of a handle (or scene) used: VGPRs per lane,
// static LDS bytes per workgroup, scratch bytes per lane, workgroups per CU
// launched, workgroups per CU that fit, total workgroups, threads per
// workgroup, waves per SIMD.
enum { kLaunchInfoInts = 8 };

struct TraceScratch {
    void* spill = nullptr;
    size_t spillBytes = 0;
    uint32_t* counter = nullptr;
    int* status = nullptr;
    int* tiles = nullptr;               // per-tile counts of the compaction scans
    size_t tileCap = 0;
    void* hits = nullptr;               // staged closest-hit records (16 bytes per ray)
    size_t hitsBytes = 0;
    // cost-ordered pinhole launches: per-tile step counts of the last launch
    // of a costW x costH image (row-major by tile position), the predicted
    // costs in chunk order and the order built from them
    int* tileCost = nullptr;
    int* tileOrder = nullptr;
    int* tilePred = nullptr;
    int costCap = 0, costW = 0, costH = 0;
    // Bumped whenever a buffer above is freed or reallocated: a captured graph
    // whose kernels hold these addresses is valid only while gen is unchanged.
    uint64_t gen = 0;
    // Borrowed fixed-size scratch (the no-alloc entry point): never grown or
    // freed; a launch whose spill would not fit runs fewer workgroups per CU.
    bool fixed = false;
    int launchInfo[kLaunchInfoInts] = {0, 0, 0, 0, 0, 0, 0, 0};
};

}  // namespace rt

// ---------------------------------------------------------------------------
// The opaque handle. All state is per handle (the reference keeps the top
// nodes and ray counters in process-global device symbols, backend :41-44).
// ---------------------------------------------------------------------------
struct CudaBVHHandle {
    int numTris = 0;                    // input triangles
    int numOrdered = 0;
    int numNodes4 = 0;
    int depth4 = 0;
    int numCWBVH = 0;

    // Device copies
    void* d_nodes4 = nullptr;           // numNodes4 * 64 B
    void* d_tris = nullptr;             // numOrdered * 48 B (3 x float4)
    int* d_leafOfPrim = nullptr;        // original index -> leaf-order position

    // Host copies for the Vulkan-side getters (borrowed by callers)
    std::vector<uint8_t> h_nodes4;
    std::vector<float> h_soa[9];
    std::vector<uint32_t> h_dfs;
    std::vector<float> h_packed;
    std::vector<int> h_ordered;
    int numDFS = 0;
    std::vector<rt::Node2> h_nodes2;    // pre-order BVH2 (source of the CWBVH)
    std::vector<uint32_t> h_cw;         // CWBVH nodes (20 words each), built on request
    std::vector<int> h_cwTriOrder;
    void* d_cwNodes = nullptr;
    void* d_cwTris = nullptr;
    int cwDepth = 0;

    float bboxMin[3], bboxMax[3], bboxCenter[3], bboxExtent = 0.f;

    void* stream = nullptr;             // hipStream_t, borrowed (NULL stream by default)

    rt::TraceScratch scratch;           // traversal scratch of calls on this handle
    rt::TraceConfig tune[rt::kNumClasses];  // per ray class launch configuration
    void* d_work = nullptr;   size_t workBytes = 0;   // staging for host-pointer calls
    float* h_pinned = nullptr; size_t pinnedFloats = 0; // pinned host staging of host-pointer calls
    void* d_sortWs = nullptr; size_t sortWsBytes = 0; // ray-sort workspace

    // The host copies above are made on the first call that needs them
    // (the reference's getters, the CWBVH build) from these device arrays,
    // so a build that is only traced never pays for them.
    bool exported = false;
    float buildStats[4] = {0, 0, 0, 0};  // upload ms, device build ms, handle ms, BVH2 levels
    void* d_nodes2 = nullptr;           // pre-order BVH2 (rt::Node2); NULL for a forest
    int numNodes2 = 0;
    int* d_ordered = nullptr;           // leaf position -> original index
};
