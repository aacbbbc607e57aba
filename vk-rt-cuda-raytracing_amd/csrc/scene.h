// scene.h — two-level scene structures shared by the traversal kernels
// (scene.hip), the path tracer (pt.hip) and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "rt_internal.h"
#include "trace_launch.h"

// Runs expr (a HIP call); on failure records the error and runs the rest.
#define HIP_TRY(expr, ...)                                                        \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            rt::set_error("%s failed: %s", #expr, hipGetErrorString(_e));         \
            __VA_ARGS__;                                                          \
        }                                                                         \
    } while (0)

namespace rtk {

// Instance as the kernels read it: world-to-object affine (row-major 3x4),
// BLAS index, Lambertian albedo.
struct InstanceDev {
    float w2o[12];
    int blas;
    float albedo[3];
};

struct BlasDev {
    const uint4* nodes;
    const float4* tris;
    const int* leafOf;
    int depth;
    int numTris;
};

struct Trace2LArgs {
    const uint4* tlasNodes = nullptr;
    const float4* tlasTris = nullptr;
    const InstanceDev* inst = nullptr;
    const BlasDev* blas = nullptr;
    int n = 0;
    const int* nDev = nullptr;
    const float *ox = nullptr, *oy = nullptr, *oz = nullptr, *dx = nullptr, *dy = nullptr, *dz = nullptr;
    const float *tmin = nullptr, *tmax = nullptr;
    float* t = nullptr; int* inst_out = nullptr; int* id = nullptr; float* u = nullptr; float* v = nullptr;
    uint8_t* occ = nullptr;
    int* cntNodes = nullptr; int* cntTris = nullptr;
    uint32_t* counter = nullptr;
    int2* spill = nullptr; int spillStride = 0; int maxStack = 0;
    int* status = nullptr;
    const int* gate = nullptr; int gateMax = -1;   // skip the launch when *gate <= gateMax
    // batch schedule: chunk c of 64 rays traces chunk chunkOrder[c], and each
    // chunk's step count (its slowest lane's) goes to chunkCost[chunk]
    const int* chunkOrder = nullptr; int* chunkCost = nullptr;
};

}  // namespace rtk

// The opaque scene handle: a TLAS over instance boxes (one instance per
// leaf) plus borrowed BLAS handles.
struct CudaSceneHandle {
    CudaBVHHandle* tlas = nullptr;
    std::vector<CudaBVHHandle*> blas;          // borrowed
    std::vector<rtk::InstanceDev> inst;
    std::vector<float> o2w;                    // 12 floats per instance
    rtk::InstanceDev* d_inst = nullptr;
    rtk::BlasDev* d_blas = nullptr;
    int maxBlasDepth = 0;
    rt::TraceScratch scratch;
    void* stream = nullptr;
    void* pt = nullptr;                        // path-tracer queues and graph (pt.hip)
    int ptTail = 32768;                        // path tracer's tail threshold for params.tailPaths < 0
    rt::TraceConfig tune2l[3];                 // launch configuration: closest hit, any hit, closest hit of
                                               // the path tracer's secondary bounces
};

namespace rtk {
// cls: 0 closest hit, 1 any hit, 2 closest hit with the secondary-bounce configuration
int launch_trace2l(const CudaSceneHandle* sc, rt::TraceScratch& scratch, Trace2LArgs& a, int cls, hipStream_t s);
// World-space geometric normals of n hits (device arrays), on stream s.
int launch_hit_normals(const CudaSceneHandle* sc, int n, const int* instId, const int* primId, float* nx, float* ny,
                       float* nz, hipStream_t s);
int trace2l_max_stack(const CudaSceneHandle* sc);
int path_trace(CudaSceneHandle* sc, const CudaCamera& cam, const CudaPathParams& p, const CudaPathImage& img,
               long long* rayCounts, unsigned flags);
void pt_free(CudaSceneHandle* sc);
// Status word of the path tracer's side-stream traversal scratch (read and
// cleared after waiting for that stream); 0 when the scene has not traced paths.
int pt_take_status(CudaSceneHandle* sc);
}
