/* cudabvh.h — C ABI of the MI355X ray-tracing core (libcudabvh_hip.so).
 *
 * Part 1 keeps every symbol of the reference boundary
 * layer/cuda_bvh_backend.h:12-141 with the same C signature, so the
 * reference's Vulkan layer (layer/VkLayer_CudaRT.cpp) or bvh_bench.cu links
 * against this library unchanged. Part 2 adds entry points in the same style
 * that the reference lacks (hit IDs and barycentrics, any-hit, device-pointer
 * packets, error reporting). Part 2 is new API with no reference counterpart.
 *
 * Conventions (as the reference): handles are opaque; getters return borrowed
 * host memory valid until cudaBVH_destroy; int returns are 0 on success and
 * -1 on failure; float returns are MRays/s and 0 on failure. Unlike the
 * reference, every failure also records a message for cudaBVH_lastError().
 * A handle must not be used from two threads at once; distinct handles are
 * independent (no process-global BVH state).
 */
#ifndef CUDABVH_H
#define CUDABVH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================== Part 1: reference boundary ===================== */

/* layer/cuda_bvh_backend.h:13 */
typedef struct CudaBVHHandle* CudaBVH_t;

/* layer/cuda_bvh_backend.h:16-18 */
typedef struct {
    float v0[3], v1[3], v2[3];
} CudaTri;

/* layer/cuda_bvh_backend.h:21. Copies the input. Builds on the GPU: binned
 * SAH (16 bins, <= 4 triangles per leaf) -> BVH4 with FP16 child bounds
 * rounded outward. numTris <= 0 or tris == NULL: returns NULL (the reference
 * has undefined behaviour there). */
CudaBVH_t cudaBVH_build(const CudaTri* tris, int numTris);

/* layer/cuda_bvh_backend.h:24. Accepts NULL. */
void cudaBVH_destroy(CudaBVH_t bvh);

/* layer/cuda_bvh_backend.h:29-31. side x side pinhole rays from camO looking
 * +z (90 degree FOV, backend :69-74). outHitT: nullable HOST array of side^2
 * floats, row-major py*side+px; misses are 1e30. Returns MRays/s over the
 * trace kernel; 0 on error. */
float cudaBVH_tracePrimary(CudaBVH_t bvh, int side,
                           float camOx, float camOy, float camOz,
                           float* outHitT);

/* layer/cuda_bvh_backend.h:37-40. Host SoA rays (t in (0, 1e30) as backend
 * :594), sorted on the GPU by octant+Morton key, traced, hit distances
 * written to outHitT (host) in the original order. Returns MRays/s over the
 * sort and trace kernels. */
float cudaBVH_traceDiffuse(CudaBVH_t bvh, int numRays,
                           const float* rayOx, const float* rayOy, const float* rayOz,
                           const float* rayDx, const float* rayDy, const float* rayDz,
                           float* outHitT);

/* layer/cuda_bvh_backend.h:45-47. Look-at camera toward the scene centre,
 * shaded RGBA8 into a host buffer. 0 on success. This and the other display
 * calls (traceToGPUPtr, traceToGPUAsync, traceToHostPtr) and
 * cudaBVH_tracePrimary start their 8x8
 * tiles in the cost order of the last frame of the same size (as
 * CUDABVH_FLAG_COST_ORDER); the image does not depend on the order. */
int cudaBVH_traceToRGBA(CudaBVH_t bvh, int width, int height,
                        float camOx, float camOy, float camOz,
                        uint32_t* outRGBA_host);

/* layer/cuda_bvh_backend.h:54-56. d_outRGBA is device memory; outFmt 0 RGBA8,
 * 1 BGRA8, 2 RGBA16F, 3 RGBA32F. Asynchronous on the handle's stream. */
int cudaBVH_traceToGPUPtr(CudaBVH_t bvh, int width, int height,
                          float camOx, float camOy, float camOz,
                          void* d_outRGBA, int outFmt);

/* layer/cuda_bvh_backend.h:62-64. Fire-and-forget on a background thread;
 * the first call blocks until its frame is done. */
void cudaBVH_traceToGPUAsync(CudaBVH_t bvh, int width, int height,
                             float camOx, float camOy, float camOz,
                             void* d_outRGBA, int outFmt);

/* layer/cuda_bvh_backend.h:71-73. Background trace into double-buffered host
 * memory; copies the last finished frame into h_outRGBA. */
int cudaBVH_traceToHostPtr(CudaBVH_t bvh, int width, int height,
                           float camOx, float camOy, float camOz,
                           void* h_outRGBA, int outFmt, uint64_t bufSize);

/* layer/cuda_bvh_backend.h:78. Imports an opaque POSIX fd as device memory
 * (hipImportExternalMemory). NULL on failure. */
void* cudaBVH_importBufferFd(int fd, uint64_t size);

/* layer/cuda_bvh_backend.h:81. Out-pointers nullable. */
void cudaBVH_getBounds(CudaBVH_t bvh, float* centerX, float* centerY, float* centerZ, float* extent);

/* layer/cuda_bvh_backend.h:84-86. 0 for NULL. getNumCWBVHNodes returns the
 * real compressed-wide-BVH node count (0 until one is built). */
int cudaBVH_getNumTris(CudaBVH_t bvh);
int cudaBVH_getNumBVH4Nodes(CudaBVH_t bvh);
int cudaBVH_getNumCWBVHNodes(CudaBVH_t bvh);   /* 0 until the CWBVH is built (unlike the reference's BVH4 count) */

/* layer/cuda_bvh_backend.h:91. numNodes x 64 bytes: halves boundsX[8],
 * boundsY[8], boundsZ[8] (mins in 0-3, maxes in 4-7, rounded outward), then
 * int32 child[4] (-1 empty, >= 0 node, <= -2 leaf = -((first<<3)|(count-1))-2). */
const void* cudaBVH_getNodeData(CudaBVH_t bvh);

/* layer/cuda_bvh_backend.h:96. Nine leaf-order SoA float arrays. */
void cudaBVH_getTriData(CudaBVH_t bvh, const float* outPtrs[9]);

/* layer/cuda_bvh_backend.h:103. 8 uint32 per node: bmin[3], bmax[3] (fp32),
 * leaf_enc (0 inner), skip (-1 past the end). Returns the node count. */
int cudaBVH_getStacklessBVH2(CudaBVH_t bvh, uint32_t** outData);

/* layer/cuda_bvh_backend.h:111. 3 vec4 per leaf-order triangle:
 * {v0.xyz, v1.x} {v1.yz, v2.xy} {v2.z, originalIndex, 0, 0}, the index
 * stored as the int32 bit pattern (how the layer and its RT-IR executor read
 * that word; the reference stored float(originalIndex), which they misread).
 * Returns the vec4 count. */
int cudaBVH_getPackedTris(CudaBVH_t bvh, float** outData);

/* layer/cuda_bvh_backend.h:116. Waits for all library work and releases
 * library-global device resources. The device itself is not reset because
 * the library shares it with its host process. */
void cudaBVH_resetDevice(void);

/* layer/cuda_bvh_backend.h:123. One degenerate triangle per AABB (backend
 * :1549-1563), built like cudaBVH_build. NULL if numAABBs <= 0. */
CudaBVH_t cudaBVH_buildFromAABBs(const float* aabbs, int numAABBs);

/* layer/cuda_bvh_backend.h:130-131. LBVH (Morton, median split, one item per
 * leaf) -> stackless BVH2. *outNodes and *outOrdered are malloc'd; the caller
 * frees them with free(). Returns the node count (0 on error). */
int cudaBVH_buildTLASFast(const float* aabbs, int numAABBs, uint32_t** outNodes,
                          int** outOrdered
#ifdef __cplusplus
                          = nullptr
#endif
                          );

/* layer/cuda_bvh_backend.h:140-141. 0 on success, -1 on failure. */
int cudaBVH_readGPUMemoryFd(int fd, uint64_t allocationSize,
                            uint64_t offset, uint64_t size, void* dst);

/* ===================== Part 2: extensions (no reference counterpart) ===================== */

/* Structure-of-arrays rays: a hit is valid when tmin < t < tmax. */
typedef struct {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
} CudaRayPacket;

/* Closest-hit output: t (1e30 on a miss), ORIGINAL primitive index (-1 on a
 * miss), barycentrics u (weight of v1) and v (weight of v2). */
typedef struct {
    float* t;
    int*   primId;
    float* u;
    float* v;
} CudaHitPacket;

/* Pinhole camera: origin, forward, right, up (need not be unit), and the
 * image-plane scales; pixel (px,py) of W x H looks along
 * f + (sx*scaleX)*r + (sy*scaleY)*u, sx = ((px+0.5)/W)*2-1, sy likewise. */
typedef struct {
    float o[3], f[3], r[3], u[3];
    float scaleX, scaleY;
} CudaCamera;

enum {
    CUDABVH_ISECT_WATERTIGHT   = 0,       /* Woop-Benthin-Wald 2013, default */
    CUDABVH_ISECT_REFERENCE_MT = 1,       /* backend :145-156 Moller-Trumbore */
    CUDABVH_ISECT_MASK         = 0xF,
    CUDABVH_FLAG_HOST_POINTERS = 1u << 4, /* packets hold host pointers: copy in/out */
    CUDABVH_FLAG_SORT_RAYS     = 1u << 5, /* octant+Morton sort before tracing */
    CUDABVH_FLAG_SYNC          = 1u << 6, /* wait for completion before returning */
    CUDABVH_FLAG_GRAPH         = 1u << 7, /* capture the launches in a hipGraph once, replay them */
    CUDABVH_FLAG_CWBVH         = 1u << 8, /* trace the compressed 8-wide BVH (built on first use) */
    CUDABVH_FLAG_TILE_ORDER    = 1u << 9, /* genBounceRays: compact in 8x8-tile order, not pixel order */
    CUDABVH_FLAG_COST_ORDER    = 1u << 10 /* tracePinholeClosest: start the 8x8 tiles that were slowest
                                             in the previous launch of the same size first */
};

/* Build from geometry already in device memory (no counterpart in the
 * reference; it replaces the host read-back of the layer's BLAS capture,
 * VkLayer_CudaRT.cpp:3039-3234, which gathers vertex and index buffers on the
 * CPU before cudaBVH_build). Triangle t has corners idx[3t], idx[3t+1],
 * idx[3t+2] (indexBytes 2 or 4: 16- or 32-bit indices; 0: no index buffer,
 * corners 3t..3t+2), each corner three floats at byte offset
 * index * vertexStride of d_vertices. The triangles are gathered on the GPU
 * and built as cudaBVH_build builds them (same tree). Indices >= numVertices
 * make the call fail (NULL, cudaBVH_lastError()) without reading out of
 * bounds. vertexStride >= 12 and a multiple of 4. */
CudaBVH_t cudaBVH_buildDevice(const void* d_vertices, int numVertices, int vertexStride, const void* d_indices,
                              int indexBytes, int numTris);

/* One geometry of a BLAS in device memory (the fields of cudaBVH_buildDevice). */
typedef struct {
    const void* vertices;   /* device; position (3 floats) at byte offset index * vertexStride */
    int numVertices;
    int vertexStride;       /* bytes, >= 12, multiple of 4 */
    const void* indices;    /* device 16- or 32-bit indices, or NULL */
    int indexBytes;         /* 2 or 4; 0 without indices (corners 3t..3t+2) */
    int numTris;
} CudaGeometry;

/* Many BLASes from device geometry in one pass (no counterpart in the
 * reference; it replaces the per-BLAS loop of the layer's pending-BLAS
 * capture, VkLayer_CudaRT.cpp:3039-3234, which reads every geometry back to
 * the host and calls cudaBVH_build per BLAS). BLAS k is the triangles of
 * geometries [geomOffsets[k], geomOffsets[k+1]) in order. All are gathered on
 * the GPU and built as one forest (BLASes of more than 4 triangles; smaller
 * ones alone), then each tree becomes its own handle out[k], identical to
 * cudaBVH_buildDevice on the same triangles. 0 on success; on failure every
 * out[k] is NULL and cudaBVH_lastError() says why. */
int cudaBVH_buildManyDevice(const CudaGeometry* geoms, const int* geomOffsets, int numBlas, CudaBVH_t* out);

/* Closest hit for n rays. Device pointers unless CUDABVH_FLAG_HOST_POINTERS.
 * Asynchronous on the handle's stream unless CUDABVH_FLAG_SYNC or host
 * pointers. 0 on success. */
int cudaBVH_traceClosest(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays,
                         const CudaHitPacket* hits, unsigned flags);

/* Any hit: occluded[i] = 1 when some triangle has tmin < t < tmax. */
int cudaBVH_traceAny(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays,
                     uint8_t* occluded, unsigned flags);

/* cudaBVH_traceClosest without any allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. Device ray and hit
 * packets only (flags: CUDABVH_FLAG_SYNC and the intersection mode; host
 * pointers, ray sort and CWBVH are rejected); results are identical to
 * cudaBVH_traceClosest. Its work counters, status word and stack spill live
 * in one static device array of 64 MiB per GPU (held until the process
 * ends), so a launch runs at most as many workgroups per CU as that spill
 * region holds; a BVH whose stack would not fit one workgroup per CU is
 * rejected with an error. Calls that overlap on one GPU (two streams or two
 * threads) share that array and race: serialise them. */
int cudaBVH_traceClosestNoAlloc(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, const CudaHitPacket* hits,
                                unsigned flags);

/* Closest hit with per-ray counters (BVH4 nodes fetched, triangles tested);
 * the counters equal the CPU oracle's on the same BVH and rays. Device
 * pointers. Slower than cudaBVH_traceClosest; for measurement and tests. */
int cudaBVH_traceClosestCounted(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays,
                                const CudaHitPacket* hits, int* nodeVisits, int* triTests,
                                unsigned flags);
int cudaBVH_traceAnyCounted(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays,
                            uint8_t* occluded, int* nodeVisits, int* triTests, unsigned flags);

/* Writes W*H pinhole rays (tmin 0, tmax 1e30) into a device packet. */
int cudaBVH_genPinholeRays(int width, int height, const CudaCamera* cam,
                           const CudaRayPacket* out, void* stream);

/* Fused generation + closest hit for a pinhole camera (no ray buffers:
 * coherent primary rays). Device hit packet of W*H entries.
 * CUDABVH_FLAG_COST_ORDER (batch schedule): the handle keeps each 8x8 tile's
 * step count of the last launch of this image size, and the next launch
 * starts the tiles predicted slowest first (a renderer tracing consecutive
 * frames; the hits are the same in any order). Extension of this build:
 * the reference's primary trace (layer/cuda_bvh_backend.cu:761-800) has
 * no tile order. */
int cudaBVH_tracePinholeClosest(CudaBVH_t bvh, int width, int height, const CudaCamera* cam,
                                const CudaHitPacket* hits, unsigned flags);

/* Secondary rays of a pinhole frame (camera as in tracePinholeClosest): for
 * every pixel whose primary hit has primId >= 0, in pixel order (with
 * CUDABVH_FLAG_TILE_ORDER: 8x8 pixel tiles in 64-pixel-wide vertical strips,
 * row-major inside a tile, then the margins row-major, so that rays taken
 * together by the traversal come from a compact image block), one
 * cosine-weighted diffuse ray about the geometric normal facing the camera
 * (PCG32 stream seeded by (seed, pixel), t in (1e-3, 1e30)) and one shadow
 * ray toward light (t in (1e-3, distance)). Outputs are device packets of
 * W*H capacity; srcIndex[k] is the pixel of ray k; *d_count (device) the
 * number written. No host synchronisation. */
int cudaBVH_genBounceRays(CudaBVH_t bvh, int width, int height, const CudaCamera* cam,
                          const CudaHitPacket* primary, uint32_t seed, const float light[3],
                          const CudaRayPacket* diffuseOut, const CudaRayPacket* shadowOut,
                          int* srcIndex, int* d_count, unsigned flags);

/* Device work queue traces: the ray count is read from device memory
 * (*d_count, at most capacity) by the kernel itself, so a chain of
 * generate -> trace launches needs no host round trip. */
int cudaBVH_traceClosestQueued(CudaBVH_t bvh, int capacity, const int* d_count,
                               const CudaRayPacket* rays, const CudaHitPacket* hits, unsigned flags);
int cudaBVH_traceAnyQueued(CudaBVH_t bvh, int capacity, const int* d_count,
                           const CudaRayPacket* rays, uint8_t* occluded, unsigned flags);

/* Launch configuration per ray class (0 coherent pinhole primaries, 1
 * incoherent closest hit on explicit rays, 2 any hit): the wave schedule
 * (0 batch: a wave keeps 64 rays until all finish; 1 while-while: separate
 * node and leaf phases, idle lanes refilled from a wave-local work pool;
 * 2 packet, coherent class only: one 8x8 pixel tile per wave sharing one
 * traversal stack; 3 refill: the batch schedule's loop with idle lanes
 * refilled from the wave-local pool),
 * LDS stack depth per lane (4, 8, 12 or 16 entries; deeper entries spill to
 * HBM) and resident workgroups per CU (0 = occupancy limit). Defaults after a
 * build: batch / 12 / 0 for classes 0 and 2, refill / 8 / 0 for class 1. */
int cudaBVH_setTraceConfig(CudaBVH_t bvh, int rayClass, int schedule, int ldsDepth, int blocksPerCU);
int cudaBVH_getTraceConfig(CudaBVH_t bvh, int rayClass, int* schedule, int* ldsDepth, int* blocksPerCU);

/* Autotuner: times every candidate configuration of one ray class on the
 * caller's workload (coherent: a W x H pinhole frame; otherwise numRays
 * device rays), reps timed launches each, and keeps the fastest in the
 * handle. results (nullable, room for 192 floats: schedule, ldsDepth,
 * blocksPerCU, ms per candidate) and *numResults report the sweep. The
 * coherent class's batch candidates are timed in the tile cost order of
 * CUDABVH_FLAG_COST_ORDER. */
int cudaBVH_autotune(CudaBVH_t bvh, int rayClass, int width, int height, const CudaCamera* cam,
                     int numRays, const CudaRayPacket* rays, int reps, float* results, int* numResults);

/* Ray sort of the incoherent path (device arrays): key = direction octant
 * << 29 | low 29 bits of the 30-bit Morton code of the origin normalised by
 * the scene box (cuda_rt_v29.cu:84-106); stable LSD radix sort. keysOut
 * receives the sorted keys, permOut the ray index of each sorted slot. */
int cudaBVH_sortRays(CudaBVH_t bvh, int numRays, const CudaRayPacket* rays, uint32_t* keysOut, int* permOut);

/* Handle stream: all work of the handle is issued on this HIP stream
 * (NULL = the null stream). */
int   cudaBVH_setStream(CudaBVH_t bvh, void* hipStream);
void* cudaBVH_getStream(CudaBVH_t bvh);

/* Device views of the built structure (borrowed, valid until destroy):
 * BVH4 nodes (64 B each), leaf-order triangles (3 x float4 each:
 * {v0.xyz, bits(primId)}, {v1.xyz, 0}, {v2.xyz, 0}). */
const void* cudaBVH_deviceNodes(CudaBVH_t bvh);
const void* cudaBVH_deviceTris(CudaBVH_t bvh);
/* Depth of the BVH4 (node levels on the longest root-to-leaf path). */
int cudaBVH_getDepth(CudaBVH_t bvh);
/* Where the build's time went (4 floats): host-to-device upload of the
 * triangles (ms; 0 for device input), the device build including its level
 * synchronisations (ms), completing the handle (bounds, scratch; ms), and the
 * number of BVH2 levels (binning passes). Host exports are made on the first
 * getter call and are not included. No backend counterpart. */
int cudaBVH_getBuildStats(CudaBVH_t bvh, float* stats);

/* Compressed 8-wide BVH (80-byte nodes: origin, per-axis power-of-two
 * scales, 8-bit child boxes rounded outward, 8 child slots) collapsed from the
 * binned-SAH BVH2; traced with CUDABVH_FLAG_CWBVH, with the same hits as the
 * BVH4. buildCWBVH returns the node count (or -1); getCWBVH gives borrowed
 * host copies: 20 uint32 per node (p.xyz, e.x|e.y<<8|e.z<<16|imask<<24,
 * childBase, triBase, meta[8], qlo/qhi x, y, z) and the leaf-order index of
 * each CWBVH triangle. */
int cudaBVH_buildCWBVH(CudaBVH_t bvh);
int cudaBVH_getCWBVH(CudaBVH_t bvh, const uint32_t** nodes, const int** triOrder, int* depth);

/* ----- Two-level scenes (instances of BLAS handles under a TLAS) ----- */
typedef struct CudaSceneHandle* CudaScene_t;

/* objectToWorld: row-major 3x4 affine map (p_world = M p_object + t). */
typedef struct {
    int blas;
    float objectToWorld[12];
    float albedo[3];
} CudaInstance;

/* Closest hit in a scene: t, instance index, ORIGINAL primitive index of the
 * instance's BLAS, barycentrics. Misses: t 1e30, instId = primId = -1. */
typedef struct {
    float* t;
    int*   instId;
    int*   primId;
    float* u;
    float* v;
} CudaSceneHit;

/* Builds the TLAS (binned-SAH BVH4 on the GPU over the instances' world
 * boxes, one instance per leaf). BLAS handles are borrowed and must outlive
 * the scene. Rays are transformed into each instance's object space
 * (unnormalised, so t is the world parameter); the closest hit is the
 * lexicographic minimum of (t, instance, primitive). */
CudaScene_t cudaBVH_sceneCreate(const CudaBVH_t* blas, int numBlas, const CudaInstance* instances, int numInstances);
void cudaBVH_sceneDestroy(CudaScene_t scene);
int cudaBVH_sceneSetStream(CudaScene_t scene, void* hipStream);
CudaBVH_t cudaBVH_sceneTLAS(CudaScene_t scene);                 /* borrowed */
int cudaBVH_sceneWorldToObject(CudaScene_t scene, float* out);  /* 12 floats per instance */
int cudaBVH_sceneTraceClosest(CudaScene_t scene, int numRays, const CudaRayPacket* rays,
                              const CudaSceneHit* hits, unsigned flags);
int cudaBVH_sceneTraceAny(CudaScene_t scene, int numRays, const CudaRayPacket* rays,
                          uint8_t* occluded, unsigned flags);
int cudaBVH_sceneTraceClosestCounted(CudaScene_t scene, int numRays, const CudaRayPacket* rays,
                                     const CudaSceneHit* hits, int* nodeVisits, int* triTests, unsigned flags);
int cudaBVH_sceneTraceClosestQueued(CudaScene_t scene, int capacity, const int* d_count,
                                    const CudaRayPacket* rays, const CudaSceneHit* hits, unsigned flags);
int cudaBVH_sceneTraceAnyQueued(CudaScene_t scene, int capacity, const int* d_count,
                                const CudaRayPacket* rays, uint8_t* occluded, unsigned flags);

/* Resources of the last traversal launch on the handle / scene (8 ints):
 * VGPRs per lane, static LDS bytes per workgroup, scratch bytes per lane,
 * workgroups per CU launched, workgroups per CU that fit, total workgroups,
 * threads per workgroup, waves per SIMD. Zeros before the first launch.
 * No backend counterpart (the reference fixes __launch_bounds__(256, 5),
 * layer/cuda_bvh_backend.cu:104); it reports the occupancy that the launch
 * configuration and the autotuner's choice give. */
int cudaBVH_getLaunchInfo(CudaBVH_t bvh, int* info);
int cudaBVH_sceneGetLaunchInfo(CudaScene_t scene, int* info);

/* Traversal status of the handle / scene since the last check: waits for the
 * handle's (scene's) stream, returns the status word and clears it. Bit 0 set
 * means a traversal stack overflowed, so some hits may be missing (the stacks
 * are sized 3 x BVH4 depth + 4 entries per level, which bounds a depth-first
 * walk, so it reads 0); the message is also left for cudaBVH_lastError.
 * Returns -1 for a NULL handle or a HIP error. Traces that synchronise
 * (CUDABVH_FLAG_SYNC, the host-pointer calls) check it themselves and fail
 * with -1 (0 Mrays/s) on an overflow. No backend counterpart: the reference's
 * 16-entry stacks drop entries silently (layer/cuda_bvh_backend.cu:179). A
 * batch's status is that of cudaBVH_batchForest(batch). */
int cudaBVH_getTraceStatus(CudaBVH_t bvh);
int cudaBVH_sceneGetTraceStatus(CudaScene_t scene);

/* World-space unit geometric normal of each closest hit of sceneTraceClosest*
 * (reads hits->instId and hits->primId; device arrays nx, ny, nz of numRays).
 * The normal is the object-space e1 x e2 of the hit triangle (the mesh's
 * winding, not turned toward the ray) mapped with the inverse transpose of the
 * instance's objectToWorld, normalised; a miss or an ID outside the scene
 * gives (0, 0, 0). No backend counterpart: it serves callers of the scene
 * trace that shade hits, as the layer's RT-IR executor does inline after its
 * own TLAS/BLAS traversal (layer/rt_ir_exec.cu:167-179, which maps with the
 * transpose of the forward matrix: the right direction only when that matrix
 * is a uniform scale). */
int cudaBVH_sceneHitNormals(CudaScene_t scene, int numRays, const CudaSceneHit* hits,
                            float* nx, float* ny, float* nz, unsigned flags);

/* Launch configuration of the scene traversal per ray class (0 closest hit,
 * 1 any hit, 2 closest hit of the path tracer's bounces after the first — the
 * first bounce and sceneTraceClosest* use class 0): schedule (0 while-while with a wave pool, 1 one ray per lane in
 * 64-ray batches, 2 = 1 with nearest-first children for any hit, 3 = 1 with
 * idle lanes refilled from a wave-local pool), LDS stack depth (4, 8, 12, 16)
 * and resident workgroups per CU (0 = as many as fit). sceneAutotune times
 * every candidate on the caller's rays (results: schedule, depth, blocks/CU,
 * ms per candidate; up to 48) and keeps the fastest. Defaults after
 * sceneCreate: 1 / 12 / 0 for every class. */
int cudaBVH_sceneSetTraceConfig(CudaScene_t scene, int rayClass, int schedule, int ldsDepth, int blocksPerCU);
int cudaBVH_sceneGetTraceConfig(CudaScene_t scene, int rayClass, int* schedule, int* ldsDepth, int* blocksPerCU);
int cudaBVH_sceneAutotune(CudaScene_t scene, int rayClass, int numRays, const CudaRayPacket* rays, int reps,
                          float* results, int* numResults);

/* ----- Wavefront path tracer over a scene ----- */
#define CUDABVH_PT_MAX_DEPTH 16

typedef struct {
    int width, height;
    int spp;              /* samples per pixel; the image is their mean */
    int maxDepth;         /* extension rays per path, camera ray included (<= 16) */
    int rrStart;          /* first extension ray (0 = camera ray) subject to Russian roulette */
    uint32_t seed;
    float sunDir[3];      /* towards the sun (normalised by the call) */
    float sunRadiance[3]; /* irradiance on a surface facing the sun */
    float skyZenith[3];   /* escaped-ray radiance: horizon + (zenith - horizon) * max(d.y, 0) */
    float skyHorizon[3];
    float clampMax;       /* per-contribution max-channel clamp; <= 0 disables */
    int tailPaths;        /* from the first bounce with at most this many live paths on, one
                             thread per path runs all remaining bounces (0: wavefront only;
                             < 0: the scene's tuned value, 32768 until
                             cudaBVH_sceneAutotunePathTracer); the image does not depend on it */
} CudaPathParams;

/* Device buffers of width*height floats per plane. rgb: 3 planes (R, G, B).
 * albedo/normal (3 planes) and depth (1 plane) are the first hit of sample 0
 * (zeros where the camera ray escapes). */
typedef struct {
    float* rgb;
    float* albedo;
    float* normal;
    float* depth;
} CudaPathImage;

/* Renders spp samples per pixel with next-event estimation toward the sun,
 * cosine-weighted Lambertian bounces and Russian roulette. Deterministic for
 * a given seed. rayCounts (host, nullable, 2*maxDepth entries): extension rays
 * traced per bounce, then shadow rays per bounce, summed over samples (the
 * call then waits for completion). With CUDABVH_FLAG_GRAPH one sample's launch
 * sequence is captured once and replayed (re-captured when sizes, buffers or
 * parameters other than seed and spp change). CUDABVH_FLAG_SORT_RAYS groups
 * each bounce's next extension rays by direction octant before they are
 * traced (same image). */
int cudaBVH_scenePathTrace(CudaScene_t scene, const CudaCamera* camera, const CudaPathParams* params,
                           const CudaPathImage* image, long long* rayCounts, unsigned flags);

/* The path tracer's queue-size knob, tuned on the caller's frame: renders
 * `reps` graph-replayed samples (after one warm-up) with each tail threshold
 * of {0, 4096, 16384, 32768, 65536, 131072} paths and keeps the fastest as
 * the scene's value for params->tailPaths < 0 (the image is the same for
 * all). results (nullable): (threshold, ms per sample) pairs, numResults
 * receives their count. 0 on success. No backend counterpart: the
 * reference's queues run every bounce as one launch (cuda_rt_v34.cu:1077-1545). */
int cudaBVH_sceneAutotunePathTracer(CudaScene_t scene, const CudaCamera* camera, const CudaPathParams* params,
                                    const CudaPathImage* image, int reps, float* results, int* numResults);

/* ----- Many independent scenes in one build and one launch ----- */
typedef struct CudaBVHBatchHandle* CudaBVHBatch_t;

/* Builds the BVHs of numScenes scenes (triangles concatenated; scene k is
 * [triOffsets[k], triOffsets[k+1]), each with more than 4 triangles) in one
 * level-synchronous GPU pass: one forest, each tree identical to the scene's
 * own cudaBVH_build tree, leaf triangles carrying their index within the
 * scene. Host inputs. */
CudaBVHBatch_t cudaBVH_buildBatch(const CudaTri* tris, const int* triOffsets, int numScenes);
void cudaBVH_batchDestroy(CudaBVHBatch_t batch);
int cudaBVH_batchSetStream(CudaBVHBatch_t batch, void* hipStream);
CudaBVH_t cudaBVH_batchForest(CudaBVHBatch_t batch);         /* borrowed: the forest's nodes/triangles */
int cudaBVH_batchRoots(CudaBVHBatch_t batch, int* roots4);   /* BVH4 root of each scene (host) */
/* Rays of all scenes concatenated; ray i belongs to the last scene k with
 * d_rayOffsets[k] <= i (device array of numScenes+1 ints). All rays go
 * through one persistent launch fed by a device work counter; primId is the
 * index within the ray's scene. CUDABVH_FLAG_GRAPH captures the launch in a
 * hipGraph once (per buffers and size) and replays it. */
int cudaBVH_batchTraceClosest(CudaBVHBatch_t batch, int totalRays, const int* d_rayOffsets,
                              const CudaRayPacket* rays, const CudaHitPacket* hits, unsigned flags);
int cudaBVH_batchTraceAny(CudaBVHBatch_t batch, int totalRays, const int* d_rayOffsets,
                          const CudaRayPacket* rays, uint8_t* occluded, unsigned flags);

/* ----- Learned denoiser ----- */
/* A 4-level convolutional encoder-decoder (32/64/128/256 channels, two 3x3
 * convolutions + ReLU per level, 2x2 max pooling, 2x nearest upsampling, skip
 * concatenation, 1x1 head added to the log-radiance input), BF16 weights and
 * activations, FP32 accumulation on MFMA, FP32 master weights trained with
 * Adam by the library's own forward and backward kernels.
 * Pixels are NHWC with 16 BF16 channels: log(1+rgb), albedo, normal,
 * log(1+depth)/4, zeros (cudaBVH_denoiserPack builds them from planes).
 * Images: height and width multiples of 8; batch*height*width <= maxPixels.
 * Parameters: per layer, weights [cout][taps][cin] then biases [cout]. */
typedef struct CudaDenoiserHandle* CudaDenoiser_t;

CudaDenoiser_t cudaBVH_denoiserCreate(int maxPixels, uint64_t seed);
void cudaBVH_denoiserDestroy(CudaDenoiser_t dn);
int cudaBVH_denoiserSetStream(CudaDenoiser_t dn, void* hipStream);
long long cudaBVH_denoiserNumParams(CudaDenoiser_t dn);
/* which: 0 weights, 1 gradients of the last training step, 2/3 Adam moments */
int cudaBVH_denoiserGetParams(CudaDenoiser_t dn, float* hostOut, int which);
int cudaBVH_denoiserSetParams(CudaDenoiser_t dn, const float* hostIn);   /* resets Adam */
/* planes of numPixels floats (rgb/albedo/normal: 3 planes) -> 16-channel pixels */
int cudaBVH_denoiserPack(CudaDenoiser_t dn, int numPixels, const float* rgb, const float* albedo,
                         const float* normal, const float* depth, uint16_t* out);
/* head output (before the residual), FP32 [pixels][3]; input stays on the device */
int cudaBVH_denoiserForward(CudaDenoiser_t dn, int batch, int height, int width, const uint16_t* input,
                            float* headOut);
/* One Adam step (lr <= 0: gradients only) on the L1 loss between
 * head + log(1+rgb) and target (FP32 [pixels][3], log(1+radiance)).
 * lossOut (host, nullable) makes the call wait for the loss. */
int cudaBVH_denoiserTrainStep(CudaDenoiser_t dn, int batch, int height, int width, const uint16_t* input,
                              const float* target, float learningRate, float* lossOut);
/* Denoise one image of planes: out (3 planes) = exp(head + log(1+rgb)) - 1. */
int cudaBVH_denoise(CudaDenoiser_t dn, int width, int height, const float* rgb, const float* albedo,
                    const float* normal, const float* depth, float* out);

/* Message of the last failure on the calling thread ("" if none). */
const char* cudaBVH_lastError(void);

#ifdef __cplusplus
}
#endif

#endif /* CUDABVH_H */
