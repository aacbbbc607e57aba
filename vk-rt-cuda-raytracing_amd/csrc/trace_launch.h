// trace_launch.h — kernel argument block and launcher declarations shared by
// the traversal kernels (trace.hip) and the C-ABI layer (api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_device.h"
#include "rt_internal.h"

namespace rt {

// Fills ki on its first use for kernel kern (block threads per workgroup) and
// records a launch of blocksPerCU workgroups per CU in info.
template <class K>
inline void note_launch(K kern, int block, KernelInfo& ki, int blocksPerCU, int grid, int* info) {
    if (!ki.maxBlocks) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, block, 0) != hipSuccess || nb <= 0) nb = 1;
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, (const void*)kern) == hipSuccess) {
            ki.vgprs = fa.numRegs; ki.ldsBytes = (int)fa.sharedSizeBytes; ki.scratchBytes = (int)fa.localSizeBytes;
        }
        ki.maxBlocks = nb;
    }
    if (!info) return;
    info[0] = ki.vgprs; info[1] = ki.ldsBytes; info[2] = ki.scratchBytes;
    info[3] = blocksPerCU; info[4] = ki.maxBlocks; info[5] = grid; info[6] = block;
    info[7] = blocksPerCU * (block / 64) / 4;
}

}  // namespace rt

namespace rtk {

enum { kModeClosest = 0, kModeAny = 1, kModeShade = 2 };

struct TraceArgs {
    const uint4* nodes = nullptr;       // BVH4, 4 x uint4 per node
    const float4* tris = nullptr;       // leaf order, 3 x float4 per triangle
    int n = 0;                          // rays (capacity when nDev is set)
    const int* nDev = nullptr;          // device-resident ray count (queued traces)
    // ray source: SoA packet, or a pinhole camera when generated inline
    const float *ox = nullptr, *oy = nullptr, *oz = nullptr, *dx = nullptr, *dy = nullptr, *dz = nullptr;
    const float *tmin = nullptr, *tmax = nullptr;
    rtd::Cam cam{};
    int W = 0, H = 0;
    int tilesX = 0, nTiles = 0;         // pinhole: 8x8 pixel tiles, one per wave
    // pinhole batch launches with a cost order: chunk c traces tile
    // tileOrder[c], and each tile's step count goes to tileCost[tile]
    const int* tileOrder = nullptr; int* tileCost = nullptr;
    bool costOrder = false;             // requested by the caller (set before launch_trace)
    // outputs
    float* t = nullptr; int* id = nullptr; float* u = nullptr; float* v = nullptr;
    uint4* hits = nullptr;              // closest hits staged as one (t, id, u, v) record per ray
    uint8_t* occ = nullptr;
    int* cntNodes = nullptr; int* cntTris = nullptr;
    void* img = nullptr; int fmt = 0;   // shading output (RGBA8, BGRA8, RGBA16F, RGBA32F)
    // scheduling and stack
    uint32_t* counter = nullptr;
    int2* spill = nullptr; int spillStride = 0; int maxStack = 0;
    int* status = nullptr;
    // batched scenes (a forest of BVHs): ray i belongs to the last scene s
    // with rayOff[s] <= i and starts at BVH4 node roots[s] (root 0 if unset)
    const int* rayOff = nullptr; const int* roots = nullptr; int numScenes = 0;
};

int launch_trace(const CudaBVHHandle* h, rt::TraceScratch& sc, TraceArgs& a, int mode, bool mt, bool pinhole,
                 bool counted, hipStream_t s);
int alloc_scratch(rt::TraceScratch& sc);
// Orders nChunks work chunks of 64 items slowest first by their recorded
// step counts (cost, any int values), per XCD band as the pinhole tile
// order does: order[c] is the chunk to trace c-th. Stream-ordered, no
// allocation (graph-capturable).
int launch_chunk_order(const int* cost, int nChunks, int* order, hipStream_t s);
void free_scratch(rt::TraceScratch& sc);
// Grows the staged hit records of sc to at least need bytes (bumps sc.gen
// when it reallocates); 0 on success.
int grow_hits(rt::TraceScratch& sc, size_t need);
// Grows the spill region of sc to at least need bytes (bumps sc.gen when it
// reallocates); 0 on success.
int grow_spill(rt::TraceScratch& sc, size_t need);
// Synchronizes s, then reads and clears the status word of sc (bit 0: a
// traversal stack overflowed, so hits may be missing). A set word is recorded
// as the error of fn. Returns the word, or -1 if it could not be read.
int take_status(rt::TraceScratch& sc, hipStream_t s, const char* fn);

// Maximum bytes of the library's static device scratch of the no-alloc entry
// point (per GPU, held until the process ends): band counters, status word
// and traversal stack spill.
constexpr size_t kNoAllocBytes = size_t(64) << 20;
// Fills sc with views of the current device's static scratch (its device
// address is looked up once per device); sc.fixed is set. 0 on success.
int noalloc_scratch(rt::TraceScratch& sc);
// Zeroes the status word of sc on stream s (a kernel, no memset).
int clear_status(rt::TraceScratch& sc, hipStream_t s);
int launch_gen_pinhole(const rtd::Cam& cam, int W, int H, const CudaRayPacket& p, hipStream_t s);
// n x 9 floats gathered from device vertex / index buffers (build.hip)
hipError_t launch_gather_tris(const void* verts, int numVertices, int stride, const void* idx, int indexBytes,
                              int numTris, float* out, int* bad, hipStream_t s);
// Tree of a forest as a BVH of its own (device copies): BVH4 nodes
// [first4, first4 + count4) rebased, leaf-order triangles [t0, t0 + ntris)
// and their leafOf entries rebased by -t0.
int launch_slice_forest(const void* nodes4, int first4, int count4, const void* tris, const int* leafOf, int t0,
                        int ntris, void* dNodes4, void* dTris, int* dLeafOf, hipStream_t s);
int launch_bounce(const CudaBVHHandle* h, rt::TraceScratch& sc, const rtd::Cam& cam, int W, int H,
                  const float* hitT, const int* hitId, uint32_t seed, const float light[3], const CudaRayPacket& d,
                  const CudaRayPacket& s, int* src, int* count, bool tileOrder, hipStream_t st);

int launch_trace_cw(const void* nodes, const void* tris, int depth8, rt::TraceScratch& sc, TraceArgs& a, int mode,
                    hipStream_t s);

size_t sort_scratch_bytes(int n);
// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits, in
// place; ws holds radix_pairs_bytes(n) bytes of device scratch.
size_t radix_pairs_bytes(int n);
int radix_sort_pairs(uint32_t* keys, int* vals, int n, int bits, void* ws, hipStream_t s);
// LBVH over n AABBs (6 floats each, device) on the GPU: the stackless DFS
// nodes (8 words each, 2n - 1 of them) and the leaf order, both device
// arrays of the caller; ws holds lbvh_scratch_bytes(n). Byte-identical to the
// host restatement of the reference's LBVH (the oracle's orc_lbvh_tlas).
size_t lbvh_scratch_bytes(int n);
int lbvh_build(const float* d_aabbs, int n, uint32_t* d_nodes, int* d_ordered, void* ws, hipStream_t s);
int trace_sorted(CudaBVHHandle* h, rt::TraceScratch& sc, int n, const CudaRayPacket& rays, const CudaHitPacket* hits,
                 uint8_t* occ, int mode, bool mt, void* ws, hipStream_t s);
int sort_rays(CudaBVHHandle* h, int n, const CudaRayPacket& rays, uint32_t* keysOut, int* permOut, void* ws,
              hipStream_t s);

}  // namespace rtk
