/* rt_oracle.h — CPU oracle for the ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke
 * check and the benchmark's CPU-baseline leg call it to judge the HIP path.
 * The product library never links, loads or calls it.
 *
 * It restates, in plain C, the algorithms of the reference BVH backend
 * (layer/cuda_bvh_backend.cu) and the standalone engines it was extracted
 * from; every function cites the reference lines it follows. Where the
 * reference has a defect that the north-star contract forbids (FP16 bounds
 * rounded to nearest, a 16-entry stack that drops nodes, an unstable sort in
 * the SAH partition), the oracle states the corrected rule and says so.
 *
 * Arithmetic contract shared with the GPU kernels: FP32, no FMA contraction
 * (built with -ffp-contract=off), IEEE division and square root, the same
 * operation order. That makes hit IDs, t, u, v and the node/triangle visit
 * counters bit-identical between this oracle and the HIP traversal.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- IEEE binary16 conversion ---------------- */
uint16_t orc_f2h_rne(float x);   /* round to nearest even (cuda_fp16 __float2half) */
uint16_t orc_f2h_down(float x);  /* largest half <= x  (outward rounding for mins) */
uint16_t orc_f2h_up(float x);    /* smallest half >= x (outward rounding for maxes) */
float    orc_h2f(uint16_t h);    /* exact */

/* ---------------- BVH build ----------------
 * Binned-SAH BVH2 (layer/cuda_bvh_backend.cu:201-251) collapsed to BVH4
 * (:391-418). tris: n x 9 floats (v0.xyz v1.xyz v2.xyz), original order. */
typedef struct orc_bvh orc_bvh;

orc_bvh* orc_build(const float* tris, int n, int max_leaf);
void     orc_free(orc_bvh* b);
int      orc_num_bvh2(const orc_bvh* b);
int      orc_num_bvh4(const orc_bvh* b);
int      orc_num_ordered(const orc_bvh* b);
int      orc_bvh4_depth(const orc_bvh* b);   /* levels of BVH4 nodes on the longest path */
/* 64-byte BVH4 nodes in the reference format (layer/cuda_bvh_backend.cu:31-34),
 * bounds rounded outward. */
void     orc_get_bvh4(const orc_bvh* b, void* out /* num_bvh4*64 bytes */);
/* ordered_idx[i] = original index of the i-th triangle in leaf order. */
void     orc_get_order(const orc_bvh* b, int* ordered_idx);
/* BVH2 nodes, pre-order: box[6], left, right, triStart, triCount. */
void     orc_get_bvh2(const orc_bvh* b, float* boxes /*n*6*/, int* links /*n*4*/);
/* Stackless DFS BVH2, 8 x uint32 per node (layer/cuda_bvh_backend.cu:456-522). */
int      orc_num_dfs(const orc_bvh* b);
void     orc_get_dfs(const orc_bvh* b, uint32_t* out /* n*8 */);
/* Binned-SAH cost of the BVH2 (sum over inner nodes of SA(node)/SA(root)
 * plus leaves' triCount*SA/SA(root)); a tree-quality figure. */
double   orc_sah_cost(const orc_bvh* b);

/* LBVH over AABBs (layer/cuda_bvh_backend.cu:257-389, :1620-1698).
 * Returns node count; out (8 uint32 per node) and ordered (n ints) are
 * malloc'd and must be released with orc_release(). */
int      orc_lbvh_tlas(const float* aabbs, int n, uint32_t** out, int** ordered);
void     orc_release(void* p);
uint32_t orc_morton3d(float x, float y, float z);   /* :265-273 */

/* ---------------- Rays ----------------
 * Pinhole generator shared with the GPU kernels. For pixel (px,py):
 *   sx = ((px+0.5)/W)*2-1, sy = ((py+0.5)/H)*2-1            (:69-71)
 *   d  = f + (sx*scale_x)*r + (sy*scale_y)*u, then d *= 1/sqrt(d.d)
 * cam: o[3], f[3], r[3], u[3], scale_x, scale_y (14 floats). */
void orc_gen_pinhole(const float* cam, int W, int H, float* rays /* 8 x W*H SoA */);

/* ---------------- Traversal ----------------
 * nodes: BVH4 (64 B each); tris: n x 9 floats in leaf order; prim_id: the
 * original primitive ID of each leaf-order triangle.
 * rays: SoA, 8 arrays of n_rays: ox oy oz dx dy dz tmin tmax.
 * A hit is valid when tmin < t < tmax; closest = lexicographic min (t, id).
 * isect: 0 = watertight (Woop, Benthin, Wald 2013), 1 = the reference's
 * Moller-Trumbore (layer/cuda_bvh_backend.cu:145-156).
 * Counters (nullable): BVH4 nodes fetched and triangles tested per ray.
 * Misses: t = 1e30, id = -1, u = v = 0. threads <= 0 means all cores. */
void orc_trace_closest(const void* nodes, int num_nodes,
                       const float* tris, const int* prim_id, int num_tris,
                       const float* rays, int n_rays,
                       float* t, int* id, float* u, float* v,
                       int* node_visits, int* tri_tests,
                       int isect, int threads);

/* Any hit (v10_pipeline.cu:195-250 semantics: stop at the first valid hit,
 * children pushed unsorted in slot order 3..0 so slot 0 is popped first). */
void orc_trace_any(const void* nodes, int num_nodes,
                   const float* tris, const int* prim_id, int num_tris,
                   const float* rays, int n_rays, uint8_t* occluded,
                   int* node_visits, int* tri_tests, int isect, int threads);

/* Brute force over every triangle: watertight FP32 (same per-triangle test as
 * the traversal) and an FP64 Moller-Trumbore (independent arithmetic). */
void orc_brute_closest(const float* tris, const int* prim_id, int num_tris,
                       const float* rays, int n_rays,
                       float* t, int* id, float* u, float* v, int isect, int threads);
void orc_brute_closest_f64(const float* tris, const int* prim_id, int num_tris,
                           const float* rays, int n_rays, double* t, int* id, int threads);

/* Single-triangle tests (for known-answer tests). Return 1 on hit. */
int orc_isect_watertight(const float* ray /*8*/, const float* tri /*9*/,
                         float* t, float* u, float* v);
int orc_isect_mt(const float* ray /*8*/, const float* tri /*9*/,
                 float* t, float* u, float* v);

/* Ray sort key (cuda_rt_v29.cu:84-106): octant<<29 | (morton & 0x1FFFFFFF),
 * origin normalised by box (lo, inv). */
void orc_sort_keys(const float* rays, int n, const float* lo3, const float* inv3,
                   uint32_t* keys);

int orc_num_threads(int requested);

/* ---------------- Two-level scenes ----------------
 * TLAS: a BVH4 whose leaves hold one instance each (leaf-order index ->
 * instance id via tlas_prim). Instance i: BLAS blas_of[i] (nodes, leaf-order
 * tris, prim ids given per BLAS), world-to-object affine w2o[12*i] (row-major
 * 3x4: p' = M p + t). A ray is transformed with
 *   o'[r] = ((M[r][0]*ox + M[r][1]*oy) + M[r][2]*oz) + t[r],
 *   d'[r] = (M[r][0]*dx + M[r][1]*dy) + M[r][2]*dz
 * and traced unnormalised, so t is the world-space distance parameter.
 * Closest = lexicographic min of (t, instance, prim). */
typedef struct {
    const void* nodes; int num_nodes;
    const float* tris; const int* prim_id; int num_tris;
} orc_blas;

void orc_trace_closest_2l(const void* tlas_nodes, int tlas_num_nodes, const int* tlas_prim,
                          const orc_blas* blas, const int* blas_of, const float* w2o,
                          const float* rays, int n_rays, float* t, int* inst, int* id, float* u, float* v,
                          int* node_visits, int* tri_tests, int threads);
void orc_trace_any_2l(const void* tlas_nodes, int tlas_num_nodes, const int* tlas_prim,
                      const orc_blas* blas, const int* blas_of, const float* w2o,
                      const float* rays, int n_rays, uint8_t* occluded, int* node_visits, int* tri_tests,
                      int threads);

/* ---------------- Path tracer ----------------
 * Same layout as the library's CudaPathParams. */
#define ORC_PT_MAX_DEPTH 16
typedef struct {
    int width, height, spp, max_depth, rr_start;
    uint32_t seed;
    float sun_dir[3], sun_rad[3], sky_zenith[3], sky_horizon[3];
    float clamp_max;
    int tail_paths;      /* a scheduling choice of the HIP path tracer; the image does not depend on it */
} orc_pt_params;

/* FP32 restatement of the wavefront path tracer over a two-level scene
 * (albedo: 3 floats per instance; cam: 14 floats as orc_gen_pinhole).
 * rgb: 3 planes of W*H; aov (nullable): 7 planes (albedo 3, normal 3, depth)
 * of sample 0's first hit; counts (nullable): 2*max_depth ray counts
 * (extension rays per bounce, then shadow rays per bounce). */
void orc_path_trace_2l(const void* tlas_nodes, int tlas_num_nodes, const int* tlas_prim, const orc_blas* blas,
                       const int* blas_of, const float* w2o, const float* albedo, const float* cam,
                       const orc_pt_params* p, float* rgb, float* aov, long long* counts, int threads);

/* The same estimator in FP64 with brute-force Moller-Trumbore over
 * world-space triangles (wtris: m x 9 doubles; tri_inst: instance of each),
 * independent of the BVHs and of the FP32 operation order. */
void orc_path_trace_f64(const double* wtris, const int* tri_inst, int m, const float* albedo, const float* cam,
                        const orc_pt_params* p, float* rgb, int threads);

/* PCG32 (pcg32_srandom_r / pcg32_random_r): n outputs. */
void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out);

/* World-space unit geometric normals of scene hits (instance, ORIGINAL
 * primitive ID of the instance's BLAS): normalize(M^T (e1 x e2)) with M the
 * instance's world-to-object 3x3 and e1 x e2 in object space (the mesh's
 * winding, not turned toward the ray); (0, 0, 0) for inst < 0 or an unknown
 * ID. out: 3 SoA arrays of n. */
void orc_scene_hit_normals(const orc_blas* blas, const int* blas_of, const float* w2o, int num_inst, int n,
                           const int* inst, const int* prim, float* out);

/* ---------------- The layer's GLSL compute tracers ----------------
 * Restatements of layer/shaders/bvh4_trace.comp:24-79 (BVH4 nodes as
 * cudaBVH_getNodeData exports them, a 13-entry stack that drops overflow,
 * children pushed unsorted in slot order) and bvh2_stackless.comp:20-78
 * (32-B DFS nodes with skip links as cudaBVH_getStacklessBVH2 exports them),
 * both over packed triangles as cudaBVH_getPackedTris exports them, with the
 * shaders' Moller-Trumbore test (no det threshold, accept 0 < t < tHit, first
 * found wins). They check that the exported formats are what the layer's
 * Vulkan fallback consumes. pc: the push constants cam o[3], fwd[3],
 * right[3], up[3], fov (13 floats); GLSL fma() is taken as fused and
 * normalize(v) as v * (1 / sqrt(dot(v, v))). Per pixel: t (1e30 on a miss)
 * and the leaf-order triangle index (-1 on a miss). rays (nullable): the
 * pixel's ray as 8 SoA arrays (ox oy oz dx dy dz, tmin 0, tmax 1e30). */
void orc_glsl_rays(const float* pc, int W, int H, float* rays);
void orc_glsl_bvh4_trace(const uint32_t* nodes, const float* packed, const float* pc, int W, int H,
                         float* t, int* tri);
void orc_glsl_stackless_trace(const uint32_t* dfs, const float* packed, const float* pc, int W, int H,
                              float* t, int* tri);

#ifdef __cplusplus
}
#endif
#endif
