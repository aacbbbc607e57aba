// trace.hip — BVH4 traversal kernels for gfx950 (closest hit, any hit,
// fused pinhole generation, shading) and their launchers.
//
// Execution model: persistent wave64 workers. Each wave takes 64 consecutive
// rays (or one 8x8 pinhole tile) with one atomicAdd on its XCD's band of the
// work (rtd::take_chunk), traces them one ray per lane, and repeats until
// every band is used up. Each lane owns a traversal stack whose first
// LDS_DEPTH entries live in LDS (laid out [depth][thread] so a wave's lanes hit
// consecutive banks) and whose deeper entries spill to a per-lane HBM region
// sized from the BVH depth, so no entry is ever dropped (the reference's
// 16-entry register stack drops children on overflow, backend :179).
//
// Traversal order is the reference's (backend :104-183): children tested
// against the ray, kept when the robust slab test passes and tnear <= the
// current closest t, sorted near-first with the 5-compare network of :177 and
// pushed far-to-near; entries popped later are re-checked against the closest
// t. Any-hit pushes children unsorted in slot order (v10_pipeline.cu:241-244).
#include "rt_device.h"
#include "rt_internal.h"
#include "trace_launch.h"

#include <algorithm>
#include <mutex>

namespace rtk {

using rtd::Ray;

// Zeroes the kBands work-distribution counters (launched with kBands threads).
__global__ void k_reset_counter(uint32_t* c) { c[threadIdx.x * rtd::kBandStride] = 0u; }


// Face-normal Lambert + rim shading of a primary hit, sky gradient on a miss
// (backend :923-984; IEEE 1/sqrt and an explicit cube in place of rsqrtf and
// powf).
__device__ __forceinline__ void shade_pixel(const TraceArgs& a, const Ray& r, int ri, int py, bool hit, int ti) {
    float rf, gf, bf;
    if (!hit) {
        const float vv = (float)py / (float)a.H;
        rf = (20.f + 10.f * (1.f - vv)) / 255.f;
        gf = (30.f + 20.f * (1.f - vv)) / 255.f;
        bf = (60.f + 100.f * vv) / 255.f;
    } else {
        const float4 p0 = a.tris[3 * ti], p1 = a.tris[3 * ti + 1], p2 = a.tris[3 * ti + 2];
        const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
        const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
        float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
        const float nl = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz + 1e-20f);
        nx *= nl; ny *= nl; nz *= nl;
        if (nx * r.dx + ny * r.dy + nz * r.dz > 0.f) { nx = -nx; ny = -ny; nz = -nz; }
        float lx = 0.5f, ly = 0.7f, lz = 0.5f;
        const float ll = 1.0f / sqrtf(lx * lx + ly * ly + lz * lz);
        lx *= ll; ly *= ll; lz *= ll;
        const float ndotl = fmaxf(nx * lx + ny * ly + nz * lz, 0.f);
        const float ndotv = fabsf(nx * (-r.dx) + ny * (-r.dy) + nz * (-r.dz));
        const float k = 1.f - ndotv;
        const float shade = fminf(0.15f + 0.65f * ndotl + 0.2f * (k * k * k), 1.f);
        rf = shade * 0.92f + 0.08f;
        gf = shade * 0.88f + 0.07f;
        bf = shade * 0.82f + 0.06f;
    }
    if (a.fmt == 3) {
        reinterpret_cast<float4*>(a.img)[ri] = make_float4(rf, gf, bf, 1.0f);
    } else if (a.fmt == 2) {
        _Float16* o = reinterpret_cast<_Float16*>(a.img) + 4 * (size_t)ri;
        o[0] = (_Float16)rf; o[1] = (_Float16)gf; o[2] = (_Float16)bf; o[3] = (_Float16)1.0f;
    } else {
        const uint32_t R = (uint32_t)(uint8_t)(rf * 255.f), G = (uint32_t)(uint8_t)(gf * 255.f),
                       B = (uint32_t)(uint8_t)(bf * 255.f);
        reinterpret_cast<uint32_t*>(a.img)[ri] =
            a.fmt == 1 ? (255u << 24) | (R << 16) | (G << 8) | B : (255u << 24) | (B << 16) | (G << 8) | R;
    }
}

// Closest-hit node visit: tests the four children with the current closest
// distance as limit, orders the hits near-first with the 5-compare network of
// backend :177, pushes all but the nearest (far-to-near) and returns the
// nearest in (ref, tn); ref = -1 when no child is hit.
template <bool FLAT, class Push>
__device__ __forceinline__ void visit_closest(const uint4* __restrict__ nodes, const Ray& r, float best, int& ref,
                                              float& tn, Push&& push) {
    const uint4* np = nodes + 4 * (size_t)ref;
    const uint4 qc = np[3];
    const rtd::NodeBox<FLAT> b = rtd::node_box<FLAT>(r, np[0], np[1], np[2]);
    float d[4];
    int c[4];
    float tt;
    if (rtd::child_hit<0, FLAT>(r, b, (int)qc.x, best, tt)) { d[0] = tt; c[0] = (int)qc.x; } else { d[0] = INFINITY; c[0] = -1; }
    if (rtd::child_hit<1, FLAT>(r, b, (int)qc.y, best, tt)) { d[1] = tt; c[1] = (int)qc.y; } else { d[1] = INFINITY; c[1] = -1; }
    if (rtd::child_hit<2, FLAT>(r, b, (int)qc.z, best, tt)) { d[2] = tt; c[2] = (int)qc.z; } else { d[2] = INFINITY; c[2] = -1; }
    if (rtd::child_hit<3, FLAT>(r, b, (int)qc.w, best, tt)) { d[3] = tt; c[3] = (int)qc.w; } else { d[3] = INFINITY; c[3] = -1; }
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

// Any-hit node visit: children kept against tmax, pushed in slot order; the
// last one kept is taken next (v10_pipeline.cu:241-244 pops it first).
template <bool FLAT, class Push>
__device__ __forceinline__ void visit_any(const uint4* __restrict__ nodes, const Ray& r, int& ref, Push&& push) {
    const uint4* np = nodes + 4 * (size_t)ref;
    const uint4 qc = np[3];
    const rtd::NodeBox<FLAT> b = rtd::node_box<FLAT>(r, np[0], np[1], np[2]);
    float d0, d1, d2, d3;
    const bool h0 = rtd::child_hit<0, FLAT>(r, b, (int)qc.x, r.tmax, d0);
    const bool h1 = rtd::child_hit<1, FLAT>(r, b, (int)qc.y, r.tmax, d1);
    const bool h2 = rtd::child_hit<2, FLAT>(r, b, (int)qc.z, r.tmax, d2);
    const bool h3 = rtd::child_hit<3, FLAT>(r, b, (int)qc.w, r.tmax, d3);
    int nxt = -1;
    if (h0) { nxt = (int)qc.x; }
    if (h1) { if (nxt != -1) push(nxt, 0.f); nxt = (int)qc.y; }
    if (h2) { if (nxt != -1) push(nxt, 0.f); nxt = (int)qc.z; }
    if (h3) { if (nxt != -1) push(nxt, 0.f); nxt = (int)qc.w; }
    ref = nxt;
}

// Per-lane traversal result.
struct Result {
    float t, u, v;
    int id, ti;        // original primitive ID and leaf-order index (-1: miss)
    bool occluded;
    int visits, tests;
    int steps;         // loop steps of the ray (batch schedule)
};

// Closest-hit result of ray ri. The while-while kernel, whose lanes finish
// rays at different times, writes it as one 16-byte record into a staging
// array when the launch provides one, split into the four output arrays by
// k_hits_split afterwards: four 4-byte stores from a few lanes at a time
// each dirtied a 32-byte sector of its own (233 MB written per launch for
// 27 MB of hits on the 1080p diffuse rays; 73 MB staged, MI355X PMC). The
// refill kernel gathers its results per work chunk in LDS instead (its
// pool staging). Appending the records in finishing order (one reservation
// of 64 slots per wave) wrote more in total and ran 10 % slower.
__device__ __forceinline__ void put_hit(const TraceArgs& a, int ri, bool hit, float t, int id, float u, float v) {
    if (a.hits) {
        a.hits[ri] = make_uint4(__float_as_uint(hit ? t : RT_MISS_T), (uint32_t)id, __float_as_uint(hit ? u : 0.f),
                                __float_as_uint(hit ? v : 0.f));
        return;
    }
    a.t[ri] = hit ? t : RT_MISS_T;
    if (a.id) a.id[ri] = id;
    if (a.u) a.u[ri] = hit ? u : 0.f;
    if (a.v) a.v[ri] = hit ? v : 0.f;
}

// Staged records -> output arrays (coalesced; the ray count read from the
// device for queued traces).
__global__ __launch_bounds__(256) void k_hits_split(const uint4* __restrict__ hits, int n, const int* nDev, float* t,
                                                    int* id, float* u, float* v) {
    const int m = nDev ? min(*nDev, n) : n;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
        const uint4 h = hits[i];
        t[i] = __uint_as_float(h.x);
        if (id) id[i] = (int)h.y;
        if (u) u[i] = __uint_as_float(h.z);
        if (v) v[i] = __uint_as_float(h.w);
    }
}

// Slot of stack entry sp (>= LDSD) of thread gtid in the HBM spill region.
// Laid out [depth][thread] (a lane-major layout measured the same).
template <int LDSD>
__device__ __forceinline__ size_t spill_index(const TraceArgs& a, int sp, int gtid) {
    return (size_t)(sp - LDSD) * a.spillStride + gtid;
}

// First BVH4 node of ray ri: 0, or in a forest its scene's root.
__device__ __forceinline__ int ray_root(const TraceArgs& a, int ri) {
    if (!a.roots) return 0;
    int lo = 0, hi = a.numScenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.rayOff[mid] <= ri) lo = mid; else hi = mid - 1;
    }
    return a.roots[lo];
}

// 8x8 pixel tile (tx, ty) of pinhole work item tile index `tile`. Tiles are
// ordered in vertical strips kStripTiles tiles (64 pixels) wide, row-major
// inside a strip, strips left to right; the last strip takes the remainder
// of the width. With the XCD bands of the work distribution, each XCD then
// works on a few neighbouring strips and its waves in flight cover a compact
// block of the image rather than a full-width band of rows: 7-9 % faster
// coherent traces than row-major order on the 1080p terrain frame.
constexpr int kStripTiles = 8;

__device__ __forceinline__ void strip_tile(int tilesX, int nTiles, int tile, int& tx, int& ty) {
    const int tilesY = nTiles / tilesX;
    const int full = tilesX / kStripTiles, per = kStripTiles * tilesY;
    int strip = tile / per, r = tile - strip * per, w = kStripTiles;
    if (strip >= full) { strip = full; r = tile - full * per; w = tilesX - full * kStripTiles; }
    ty = r / w;
    tx = strip * kStripTiles + (r - ty * w);
}
__device__ __forceinline__ void pin_tile(const TraceArgs& a, int tile, int& tx, int& ty) {
    strip_tile(a.tilesX, a.nTiles, tile, tx, ty);
}

// Cost order of the pinhole tiles (batch schedule). The launch's end is set
// by its slowest 8x8 tiles (grazing rays: up to 7x the mean step count), and
// in the plain order they start wherever they fall in their band. Ordering
// each XCD band's tiles by the step counts the previous launch of the same
// image size recorded, slowest first (longest-processing-time-first list
// scheduling), starts those chains at the launch's beginning, where they
// overlap the rest of the work. A tile's predicted cost is the largest
// recorded count within kCostReach tiles of it (k_cost_dilate), so a camera
// that moved a few tiles since the last launch still finds its slow tiles
// in front: the slow tiles are thin features (silhouettes, grazing slopes)
// that a shift of one tile misses entirely. The order itself: one workgroup
// per band (the bands of rtd::take_chunk, so a band's tiles stay on its
// XCD): the tiles predicted at kCostHeavy steps or more first, by
// descending cost in kCostBuckets buckets of kCostShift-bit costs; every
// other tile follows in its strip order (a stable partition), which keeps
// the L2 locality of the cheap majority. Only the schedule changes: every tile is
// traced once and each ray's result is the same.
constexpr int kCostBuckets = 32, kCostShift = 3, kCostHeavy = 48, kCostReach = 2, kOrderBlock = 1024;

// pred[c] = the largest cost (row-major by tile position) within kCostReach
// tiles of chunk c's tile
__global__ void k_cost_dilate(const int* cost, int tilesX, int nTiles, int* pred) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nTiles) return;
    const int tilesY = nTiles / tilesX;
    int tx, ty;
    strip_tile(tilesX, nTiles, c, tx, ty);
    int m = 0;
    for (int y = max(ty - kCostReach, 0); y <= min(ty + kCostReach, tilesY - 1); y++)
        for (int x = max(tx - kCostReach, 0); x <= min(tx + kCostReach, tilesX - 1); x++) m = max(m, cost[y * tilesX + x]);
    pred[c] = m;
}

// Also zeroes band b's work counter for the trace launch that follows (in
// place of k_reset_counter), unless counter is null.
__global__ __launch_bounds__(kOrderBlock) void k_tile_order(const int* cost, int nTiles, int* order, uint32_t* counter) {
    constexpr int NW = kOrderBlock / 64, NH = kCostBuckets;
    __shared__ int hbase[NH];                     // first slot of each heavy bucket in the band
    __shared__ int hfill[NH];                     // heavy tiles of each bucket placed so far
    __shared__ int wl[2][NW];                     // light tiles per wave in a round (alternating buffers)
    __shared__ int nheavy, nlight;                // heavy tiles of the band; light tiles placed so far
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int lo = (int)(((long long)nTiles * b) / rtd::kBands);
    const int hi = (int)(((long long)nTiles * (b + 1)) / rtd::kBands);
    // heavy bucket of tile c (0: slowest), -1 for a light tile
    auto hbucket = [&](int c) {
        const int v = cost[c];
        return v >= kCostHeavy ? NH - 1 - min(v >> kCostShift, NH - 1) : -1;
    };
    if (tid < NH) hfill[tid] = 0;
    if (tid == 0) {
        nlight = 0;
        if (counter) counter[b * rtd::kBandStride] = 0u;
    }
    __syncthreads();
    for (int c = lo + tid; c < hi; c += kOrderBlock) {
        const int k = hbucket(c);
        if (k >= 0) atomicAdd(&hfill[k], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int k = 0; k < NH; k++) { hbase[k] = acc; acc += hfill[k]; hfill[k] = 0; }
        nheavy = acc;
    }
    __syncthreads();
    // heavy tiles: slots [0, nheavy) by bucket (any order inside a bucket);
    // light tiles: slots [nheavy, band size) in strip order (a stable
    // partition: one ballot per round, ranks from the waves before)
    const uint64_t below = (1ull << lane) - 1;
    int buf = 0;
    for (int c0 = lo; c0 < hi; c0 += kOrderBlock, buf ^= 1) {
        const int c = c0 + tid;
        const int k = c < hi ? hbucket(c) : -2;
        const uint64_t m = __ballot(k == -1);
        if (lane == 0) wl[buf][w] = __popcll(m);
        __syncthreads();
        if (k == -1) {
            int before = nlight + __popcll(m & below);
            for (int v = 0; v < w; v++) before += wl[buf][v];
            order[lo + nheavy + before] = c;
        } else if (k >= 0) {
            order[lo + hbase[k] + atomicAdd(&hfill[k], 1)] = c;
        }
        __syncthreads();                          // nlight and wl[buf] read by every wave
        if (tid == 0) {
            int t = 0;
            for (int v = 0; v < NW; v++) t += wl[buf][v];
            nlight += t;
        }
    }
}

__global__ void k_fill_i32(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Batch schedule: traverses one ray in one loop. Each iteration is one short
// step: one node visit, or the next one or two triangles of the current leaf
// (a leaf cursor holds the next triangle and the number left), then a pop
// when the lane has nothing to descend into. A lane at a leaf and a lane at a
// node thus run steps of similar length side by side, instead of the node
// lanes waiting for a whole leaf loop. Stack entries (node ref, entry
// distance) live in LDS (slots [0, LDSD)) and spill to HBM beyond. Visit
// order and node / triangle counters are those of a one-ray-at-a-time
// depth-first traversal.
//
// The slowest batches of a frame set the launch's end: on the 1080p terrain
// view a grazing ray takes up to 7x the mean step count (about 75 node visits
// and 140 triangle tests), a chain of dependent fetches that cannot be split.
// Two measures shorten it: a leaf's triangles go in pairs (half the memory
// round trips of a leaf walk), and the wave priority rises one level per
// kPrioSteps steps on the same batch, so that a long batch issues ahead of
// the short ones sharing its SIMD (on that view: 0.55 to 0.46 ms per launch).
constexpr int kPrioSteps = 32;   // loop steps per wave-priority level

template <int MODE, bool MT, bool FLAT, int LDSD, int BLOCK>
__device__ __forceinline__ void trace_ray(const TraceArgs& a, const Ray& r, int* s_ref, float* s_tn, int tid,
                                             int gtid, Result& res, int root, int pc) {
    const uint4* __restrict__ nodes = a.nodes;
    const float4* __restrict__ tris = a.tris;
    int sp = 0, visits = 0, tests = 0;
    int ref = root;
    float tn = -INFINITY;
    float best = r.tmax, bu = 0.f, bv = 0.f;
    int bid = -1, bti = -1;
    int lts = 0, ltc = 0;          // leaf cursor: next triangle, triangles left
    bool found = false;

    auto push = [&](int rf, float t) {
        if (sp < LDSD) {
            s_ref[sp * BLOCK + tid] = rf;
            s_tn[sp * BLOCK + tid] = t;
        } else if (sp < a.maxStack) {
            a.spill[spill_index<LDSD>(a, sp, gtid)] = make_int2(rf, __float_as_int(t));
        } else {
            atomicOr(a.status, 1);
            return;
        }
        sp++;
    };

    // Wave priority grows with the wave's step count on its batch: the
    // slowest batches of a frame (grazing rays, up to 7x the mean step count
    // on the terrain view) set the launch's end, and at the default priority
    // they advance at a shared issue rate while short batches fill the SIMD.
    int steps = 0;
    for (;;) {
        if ((++steps & 15) == 0) {
            const int s = __builtin_amdgcn_readfirstlane(steps);
            if (s == kPrioSteps) __builtin_amdgcn_s_setprio(1);
            else if (s == 2 * kPrioSteps) __builtin_amdgcn_s_setprio(2);
            else if (s == 3 * kPrioSteps) __builtin_amdgcn_s_setprio(3);
        }
        if (ltc == 0 && ref < 0 && (MODE == kModeAny || tn <= best)) {
            const int enc = -(ref + 2);
            lts = enc >> 3; ltc = (enc & 7) + 1;
        }
        if (ltc > 0) {
            // Two triangles of the leaf per step when it has two left: both
            // are fetched together and tested in leaf order, so a long leaf
            // walk waits for half as many memory round trips. The second is
            // counted and tested only when the first does not end an any-hit
            // ray, as in a one-at-a-time walk.
            const int ti = lts;
            const bool two = ltc > 1;
            const int tj = two ? ti + 1 : ti;
            const float4 p0 = tris[3 * ti], p1 = tris[3 * ti + 1], p2 = tris[3 * ti + 2];
            const float4 q0 = tris[3 * tj], q1 = tris[3 * tj + 1], q2 = tris[3 * tj + 2];
            auto test = [&](int k, const float4& a0, const float4& a1, const float4& a2) {
                float t, u, v;
                bool x;
                if constexpr (MT) x = rtd::isect_mt(r, a0, a1, a2, t, u, v);
                else x = rtd::isect_watertight_wave(r, pc, a0, a1, a2, t, u, v);
                if (x && t > r.tmin && t < r.tmax) {
                    if constexpr (MODE == kModeAny) {
                        return true;
                    } else {
                        const int id = __float_as_int(a0.w);
                        if (t < best || (t == best && id < bid)) { best = t; bid = id; bu = u; bv = v; bti = k; }
                    }
                }
                return false;
            };
            tests++;
            if (test(ti, p0, p1, p2)) { found = true; break; }
            if (two) {
                tests++;
                if (test(tj, q0, q1, q2)) { found = true; break; }
            }
            lts += two ? 2 : 1;
            ltc -= two ? 2 : 1;
            if (ltc > 0) continue;
        } else if (ref >= 0 && (MODE == kModeAny || tn <= best)) {
            visits++;
            int nxt = ref;
            float nt = 0.f;
            if constexpr (MODE == kModeAny) visit_any<FLAT>(nodes, r, nxt, push);
            else visit_closest<FLAT>(nodes, r, best, nxt, nt, push);
            if (nxt != -1) { ref = nxt; tn = nt; continue; }
        }
        if (sp == 0) break;
        sp--;
        if (sp < LDSD) {
            ref = s_ref[sp * BLOCK + tid];
            tn = s_tn[sp * BLOCK + tid];
        } else {
            // a non-temporal load: the spill read cannot be merged with the
            // LDS read above into one generic (flat) load, which would send
            // every pop through the vector memory path
            typedef int v2i __attribute__((ext_vector_type(2)));
            const v2i e = __builtin_nontemporal_load((const v2i*)&a.spill[spill_index<LDSD>(a, sp, gtid)]);
            ref = e.x; tn = __int_as_float(e.y);
        }
    }
    __builtin_amdgcn_s_setprio(0);
    res.t = best; res.u = bu; res.v = bv; res.id = bid; res.ti = bti;
    res.occluded = found; res.visits = visits; res.tests = tests; res.steps = steps;
}

// Refill schedule: the batch schedule's one-short-step loop (a node visit or
// one or two triangles per iteration, as trace_ray), but a lane whose ray finished
// takes the next work item from the wave's pool of 64 (one 8x8 tile for
// pinhole rays; one atomicAdd per 64 items) once at least kRefillMin lanes
// are idle, so a wave does not wait for the slowest ray of its batch and
// the ray set-up of new rays is shared by many lanes. Same depth-first order
// per ray as trace_ray; uncounted.
constexpr int kRefillMin = 8;   // 8 measured ~1 % faster than 4, 16 or 24 on the 1080p diffuse rays

template <int MODE, bool MT, bool PIN, int LDSD, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_refill(TraceArgs a) {
    __shared__ int s_ref[LDSD * BLOCK];
    __shared__ float s_tn[LDSD * BLOCK];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int gtid = blockIdx.x * BLOCK + tid;
    const int n = PIN ? a.nTiles * 64 : (a.nDev ? min(*a.nDev, a.n) : a.n);
    const uint4* __restrict__ nodes = a.nodes;
    const float4* __restrict__ tris = a.tris;
    Ray r;
    int ri = 0, py = 0;
    bool has = false, exhausted = false, flat = false;
    int poolNext = 0, poolEnd = 0;
    int band = blockIdx.x % rtd::kBands;
    int sp = 0, ref = -1, bid = -1, bti = -1, lts = 0, ltc = 0;
    float tn = 0.f, best = 0.f, bu = 0.f, bv = 0.f;

    // Closest hits of explicit rays are gathered per pool (the 64 consecutive
    // rays of one work chunk) in LDS and written out, coalesced, when the
    // pool's last ray finishes. Lanes finish rays at different times, and
    // written one by one, a ray's four results each dirtied an L2 sector of
    // its own. Two pools per wave are held; a third evicts the older one
    // (its gathered results go out, its later rays are written directly).
    constexpr bool POOL = MODE == kModeClosest && !PIN;
    constexpr int kSlots = 2;
    __shared__ float4 s_rec[POOL ? (BLOCK / 64) * kSlots * 64 : 1];
    const int rbase = (tid >> 6) * kSlots * 64;
    int sbase[kSlots] = {-1, -1}, scount[kSlots] = {0, 0};   // wave-uniform: pool base, results gathered
    int sold = 0;                                             // wave-uniform: the slot assigned earlier
    int fs = -1;                                              // slot that gathered this lane's last result
    // writes the gathered results of slot sl (entries with t not NaN)
    auto flush = [&](int sl) {
        const int pb = sbase[sl];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const float4 rec = s_rec[rbase + sl * 64 + lane];
        if (lane < n - pb && rec.x == rec.x) {
            const int q = pb + lane;
            a.t[q] = rec.x;
            if (a.id) a.id[q] = __float_as_int(rec.y);
            if (a.u) a.u[q] = rec.z;
            if (a.v) a.v[q] = rec.w;
        }
        sbase[sl] = -1;
        scount[sl] = 0;
    };

    auto push = [&](int rf, float t) {
        if (sp < LDSD) {
            s_ref[sp * BLOCK + tid] = rf;
            s_tn[sp * BLOCK + tid] = t;
        } else if (sp < a.maxStack) {
            a.spill[spill_index<LDSD>(a, sp, gtid)] = make_int2(rf, __float_as_int(t));
        } else {
            atomicOr(a.status, 1);
            return;
        }
        sp++;
    };

    for (;;) {
        if constexpr (POOL) {
            // the previous step's finished rays: a pool whose rays are all
            // gathered goes out (every lane is active here)
#pragma unroll
            for (int sl = 0; sl < kSlots; sl++) {
                if (sbase[sl] < 0) continue;
                scount[sl] += __popcll(__ballot(fs == sl));
                if (scount[sl] == min(64, n - sbase[sl])) flush(sl);
            }
            fs = -1;
        }
        if (!exhausted || poolNext < poolEnd) {
            const uint64_t idle = __ballot(!has);
            if (idle && (__popcll(idle) >= kRefillMin || !__any(has))) {
                const int cnt = __popcll(idle);
                const int avail = poolEnd - poolNext;
                int fresh = n;
                if (cnt > avail && !exhausted) {
                    fresh = rtd::take_chunk(a.counter, n, band);
                    if (fresh >= n) exhausted = true;
                    if constexpr (POOL) {
                        if (fresh < n) {
                            const int sl = sbase[0] < 0 ? 0 : (sbase[1] < 0 ? 1 : sold);
                            if (sbase[sl] >= 0) flush(sl);
                            sbase[sl] = fresh;
                            sold = 1 - sl;
                            s_rec[rbase + sl * 64 + lane].x = __int_as_float(0x7fc00000);   // NaN: not yet gathered
                        }
                    }
                }
                const int k = __popcll(idle & ((1ull << lane) - 1ull));
                int w = n;
                if (k < avail) w = poolNext + k;
                else if (fresh < n) w = fresh + (k - avail);
                if (cnt > avail) {
                    if (fresh < n) { poolNext = fresh + (cnt - avail); poolEnd = fresh + 64; }
                    else { poolNext = 0; poolEnd = 0; }
                } else {
                    poolNext += cnt;
                }
                if (!has) {
                    bool live = w < n;
                    if constexpr (PIN) {
                        const int tile = w >> 6, kk = w & 63;
                        int tx, ty;
            pin_tile(a, tile, tx, ty);
                        const int px = tx * 8 + (kk & 7);
                        py = ty * 8 + (kk >> 3);
                        live = live && px < a.W && py < a.H;
                        ri = py * a.W + px;
                        if (live) rtd::pinhole_ray(a.cam, px, py, a.W, a.H, r);
                    } else {
                        ri = w;
                        if (live) {
                            r.ox = a.ox[ri]; r.oy = a.oy[ri]; r.oz = a.oz[ri];
                            r.dx = a.dx[ri]; r.dy = a.dy[ri]; r.dz = a.dz[ri];
                            r.tmin = a.tmin[ri]; r.tmax = a.tmax[ri];
                        }
                    }
                    if (live) {
                        rtd::ray_setup(r);
                        flat = rtd::ray_flat(r);
                        has = true;
                        sp = 0; ref = ray_root(a, ri); tn = -INFINITY; ltc = 0;
                        best = r.tmax; bu = 0.f; bv = 0.f; bid = -1; bti = -1;
                    }
                }
            }
        }
        if (!__any(has)) {
            if (exhausted && poolNext >= poolEnd) break;
            continue;
        }
        const bool waveFlat = __any(has && flat);
        if (!has) continue;
        bool cont = false, fin = false, found = false;
        if (ltc == 0 && ref < 0 && (MODE == kModeAny || tn <= best)) {
            const int enc = -(ref + 2);
            lts = enc >> 3; ltc = (enc & 7) + 1;
        }
        if (ltc > 0) {
            // two triangles per step when the leaf has two left (as trace_ray)
            const int ti = lts;
            const bool two = ltc > 1;
            const int tj = two ? ti + 1 : ti;
            const float4 p0 = tris[3 * ti], p1 = tris[3 * ti + 1], p2 = tris[3 * ti + 2];
            const float4 q0 = tris[3 * tj], q1 = tris[3 * tj + 1], q2 = tris[3 * tj + 2];
            auto test = [&](int k, const float4& a0, const float4& a1, const float4& a2) {
                float t, u, v;
                if (rtd::isect<MT>(r, a0, a1, a2, t, u, v) && t > r.tmin && t < r.tmax) {
                    if constexpr (MODE == kModeAny) {
                        return true;
                    } else {
                        const int id = __float_as_int(a0.w);
                        if (t < best || (t == best && id < bid)) { best = t; bid = id; bu = u; bv = v; bti = k; }
                    }
                }
                return false;
            };
            if (test(ti, p0, p1, p2) || (two && test(tj, q0, q1, q2))) { found = true; fin = true; }
            lts += two ? 2 : 1;
            ltc -= two ? 2 : 1;
            if (!fin && ltc > 0) cont = true;
        } else if (ref >= 0 && (MODE == kModeAny || tn <= best)) {
            int nxt = ref;
            float nt = 0.f;
            if constexpr (MODE == kModeAny) {
                if (waveFlat) visit_any<true>(nodes, r, nxt, push);
                else visit_any<false>(nodes, r, nxt, push);
            } else {
                if (waveFlat) visit_closest<true>(nodes, r, best, nxt, nt, push);
                else visit_closest<false>(nodes, r, best, nxt, nt, push);
            }
            if (nxt != -1) { ref = nxt; tn = nt; cont = true; }
        }
        if (!cont && !fin) {
            if (sp == 0) {
                fin = true;
            } else {
                sp--;
                if (sp < LDSD) {
                    ref = s_ref[sp * BLOCK + tid];
                    tn = s_tn[sp * BLOCK + tid];
                } else {
                    const int2 e = a.spill[spill_index<LDSD>(a, sp, gtid)];
                    ref = e.x; tn = __int_as_float(e.y);
                }
            }
        }
        if (fin) {
            if constexpr (MODE == kModeAny) {
                a.occ[ri] = found ? 1 : 0;
            } else if constexpr (POOL) {
                const bool hit = bid >= 0;
                const float4 rec = make_float4(hit ? best : RT_MISS_T, __int_as_float(bid), hit ? bu : 0.f,
                                               hit ? bv : 0.f);
                const int pb = ri & ~63;
                fs = pb == sbase[0] ? 0 : (pb == sbase[1] ? 1 : -1);
                if (fs >= 0) {
                    s_rec[rbase + fs * 64 + (ri & 63)] = rec;
                } else {
                    a.t[ri] = rec.x;
                    if (a.id) a.id[ri] = bid;
                    if (a.u) a.u[ri] = rec.z;
                    if (a.v) a.v[ri] = rec.w;
                }
            } else if constexpr (MODE == kModeClosest) {
                put_hit(a, ri, bid >= 0, best, bid, bu, bv);
            } else {
                shade_pixel(a, r, ri, py, bid >= 0, bti);
            }
            has = false;
        }

    }
}

// Batch schedule: each wave takes 64 rays (one 8x8 pixel tile for pinhole
// rays) with one atomicAdd and runs until all 64 are done. Best for coherent
// rays, whose lanes stay in step. Compiled for seven waves per SIMD (72
// VGPRs, no scratch): with an LDS stack of 8 entries seven fit, and the
// 1080p primaries ran ~2 % faster than at six.
template <int MODE, bool MT, bool PIN, bool COUNT, int LDSD, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(7))) void k_trace_batch(TraceArgs a) {
    __shared__ int s_ref[LDSD * BLOCK];
    __shared__ float s_tn[LDSD * BLOCK];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int gtid = blockIdx.x * BLOCK + tid;
    // queued traces read their ray count from device memory (no host round trip)
    // pinhole work items cover 8x8 pixel tiles, one tile per wave
    const int n = PIN ? a.nTiles * 64 : (a.nDev ? min(*a.nDev, a.n) : a.n);
    int band = blockIdx.x % rtd::kBands;

    for (;;) {
        const int base = rtd::take_chunk(a.counter, n, band);
        if (base >= n) break;
        int ri = base + lane;
        int py = 0;
        bool live;
        int tile = 0, tx = 0, ty = 0;
        if constexpr (PIN) {
            // chunk c traces tile c, or tileOrder[c] in a cost-ordered launch
            tile = (int)base >> 6;
            if (a.tileOrder) tile = a.tileOrder[tile];
            pin_tile(a, tile, tx, ty);
            const int px = tx * 8 + (lane & 7);
            py = ty * 8 + (lane >> 3);
            live = (unsigned)tile < (unsigned)a.nTiles && px < a.W && py < a.H;
            ri = py * a.W + px;
        } else {
            live = ri < n;
        }

        Ray r;
        if (live) {
            if constexpr (PIN) {
                rtd::pinhole_ray(a.cam, ri - py * a.W, py, a.W, a.H, r);
            } else {
                r.ox = a.ox[ri]; r.oy = a.oy[ri]; r.oz = a.oz[ri];
                r.dx = a.dx[ri]; r.dy = a.dy[ri]; r.dz = a.dz[ri];
                r.tmin = a.tmin[ri]; r.tmax = a.tmax[ri];
            }
            rtd::ray_setup(r);
        }
        // A wave with any flat ray takes the slab variant that resolves
        // 0*inf; both variants agree on every other ray.
        const bool anyFlat = __any(live && rtd::ray_flat(r));
        // the live lanes' shared watertight axis permutation (-1: mixed)
        const int pc = MT ? -1 : rtd::wave_perm(r, live);
        int bsteps = 0;
        if (live) {
            Result res;
            const int root = ray_root(a, ri);
            if (anyFlat) trace_ray<MODE, MT, true, LDSD, BLOCK>(a, r, s_ref, s_tn, tid, gtid, res, root, pc);
            else trace_ray<MODE, MT, false, LDSD, BLOCK>(a, r, s_ref, s_tn, tid, gtid, res, root, pc);

            if constexpr (MODE == kModeAny) {
                a.occ[ri] = res.occluded ? 1 : 0;
            } else if constexpr (MODE == kModeClosest) {
                const bool hit = res.id >= 0;
                a.t[ri] = hit ? res.t : RT_MISS_T;
                if (a.id) a.id[ri] = res.id;
                if (a.u) a.u[ri] = hit ? res.u : 0.f;
                if (a.v) a.v[ri] = hit ? res.v : 0.f;
            } else {
                shade_pixel(a, r, ri, py, res.id >= 0, res.ti);
            }
            if constexpr (COUNT) {
                a.cntNodes[ri] = res.visits;
                a.cntTris[ri] = res.tests;
            }
            bsteps = res.steps;
        }
        // a cost-ordered launch records the batch's step count (its slowest
        // lane's) for the next launch's order, row-major by tile position
        if constexpr (PIN && !COUNT) {
            if (a.tileCost) {
                const int m = rtd::wave_max(bsteps);
                if (lane == 0 && (unsigned)tile < (unsigned)a.nTiles) a.tileCost[ty * a.tilesX + tx] = m;
            }
        }
    }
}


// Persistent traversal kernel, while-while with dynamic ray fetch.
//
// Each wave loops: (1) lanes without a ray take new ones — one atomicAdd per
// wave for all idle lanes, ranks from the ballot mask — so a wave stays full
// while rays of very different lengths finish; (2) node phase: every lane
// pops/culls its stack and visits internal nodes, in a wave-uniform loop that
// ends when no lane is at an internal node (each lane then waits at a leaf or
// has finished); (3) leaf phase: all lanes at a leaf test its triangles
// together; (4) finished rays write their result. Node and leaf work never
// share an iteration, so their instruction streams do not serialise against
// each other. In counted launches every ray follows the reference's
// depth-first order (pop, re-check against the closest t, visit), so its
// node / triangle counters are exactly those of a one-ray-at-a-time
// traversal; uncounted launches postpone leaves speculatively (see the node
// phase) and return the same hits.
template <int MODE, bool MT, bool PIN, bool COUNT, int LDSD, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_ww(TraceArgs a) {
    __shared__ int s_ref[LDSD * BLOCK];
    __shared__ float s_tn[LDSD * BLOCK];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int gtid = blockIdx.x * BLOCK + tid;
    const uint4* __restrict__ nodes = a.nodes;
    const float4* __restrict__ tris = a.tris;
    // queued traces read their ray count from device memory (no host round
    // trip); pinhole work items cover 8x8 pixel tiles, 64 items per tile
    const int n = PIN ? a.nTiles * 64 : (a.nDev ? min(*a.nDev, a.n) : a.n);

    Ray r;
    int ri = 0, py = 0;
    bool has = false;          // lane holds a ray
    bool exhausted = false;    // wave-uniform: the work counter has passed n
    int poolNext = 0, poolEnd = 0;   // wave-uniform: reserved, not yet taken work items
    int band = blockIdx.x % rtd::kBands;   // wave-uniform: band of the work distribution
    bool flat = false, found = false;
    int sp = 0, ref = -1, visits = 0, tests = 0, bid = -1, bti = -1;
    int pend = -1;             // postponed leaf (speculative node phase), -1: none
    float tn = 0.f, best = 0.f, bu = 0.f, bv = 0.f;

    auto push = [&](int rf, float t) {
        if (sp < LDSD) {
            s_ref[sp * BLOCK + tid] = rf;
            s_tn[sp * BLOCK + tid] = t;
        } else if (sp < a.maxStack) {
            a.spill[spill_index<LDSD>(a, sp, gtid)] = make_int2(rf, __float_as_int(t));
        } else {
            atomicOr(a.status, 1);   // unreachable: maxStack bounds a DFS of this BVH
            return;
        }
        sp++;
    };

    for (;;) {
        // ---- (1) refill idle lanes from the wave's pool of reserved work
        // items; the pool is refilled 64 items at a time with one atomicAdd ----
        if (!exhausted || poolNext < poolEnd) {
            const uint64_t idle = __ballot(!has);
            if (idle) {
                const int cnt = __popcll(idle);
                const int avail = poolEnd - poolNext;
                int fresh = n;                               // n: no fresh batch
                if (cnt > avail && !exhausted) {
                    fresh = rtd::take_chunk(a.counter, n, band);
                    if (fresh >= n) exhausted = true;
                }
                const int k = __popcll(idle & ((1ull << lane) - 1ull));
                int w = n;                                   // n: no item
                if (k < avail) w = poolNext + k;
                else if (fresh < n) w = fresh + (k - avail);
                if (cnt > avail) {
                    if (fresh < n) { poolNext = fresh + (cnt - avail); poolEnd = fresh + 64; }
                    else { poolNext = 0; poolEnd = 0; }
                } else {
                    poolNext += cnt;
                }
                if (!has) {
                    bool live = w < n;
                    if constexpr (PIN) {
                        const int tile = w >> 6, k = w & 63;
                        int tx, ty;
            pin_tile(a, tile, tx, ty);
                        const int px = tx * 8 + (k & 7);
                        py = ty * 8 + (k >> 3);
                        live = live && px < a.W && py < a.H;
                        ri = py * a.W + px;
                        if (live) rtd::pinhole_ray(a.cam, px, py, a.W, a.H, r);
                    } else {
                        ri = w;
                        if (live) {
                            r.ox = a.ox[ri]; r.oy = a.oy[ri]; r.oz = a.oz[ri];
                            r.dx = a.dx[ri]; r.dy = a.dy[ri]; r.dz = a.dz[ri];
                            r.tmin = a.tmin[ri]; r.tmax = a.tmax[ri];
                        }
                    }
                    if (live) {
                        rtd::ray_setup(r);
                        flat = rtd::ray_flat(r);
                        has = true;
                        sp = 0; ref = ray_root(a, ri); tn = -INFINITY;
                        best = r.tmax; bu = 0.f; bv = 0.f; bid = -1; bti = -1;
                        visits = 0; tests = 0; found = false;
                    }
                }
            }
        }
        if (!__any(has)) {
            if (exhausted && poolNext >= poolEnd) break;
            continue;
        }
        // A wave with any flat ray takes the slab variant that resolves 0*inf;
        // both variants agree on every other ray.
        const bool waveFlat = __any(has && flat);

        // ---- (2) node phase ----
        // Speculative (uncounted launches): a lane whose next entry is a leaf
        // postpones it (one at a time) and keeps visiting nodes; the phase
        // ends when every lane holds a postponed leaf or is done, and the
        // leaf phase tests them together (Aila and Laine, "Understanding the
        // Efficiency of Ray Traversal on GPUs", HPG 2009). Nodes visited past
        // a postponed leaf use a closest t that the leaf may still lower, so
        // such a launch can visit a few more nodes than the depth-first
        // order; every box test stays conservative and the hits are the same.
        // Counted launches keep the exact depth-first order and counters.
        constexpr bool SPEC = !COUNT;
        bool fin = false;
        for (;;) {
            if (has && !fin) {
                for (;;) {
                    // pop until an entry survives the closest-t re-check
                    while (ref == -1 || (MODE != kModeAny && tn > best)) {
                        if (sp == 0) { fin = true; break; }
                        sp--;
                        if (sp < LDSD) {
                            ref = s_ref[sp * BLOCK + tid];
                            tn = s_tn[sp * BLOCK + tid];
                        } else {
                            const int2 e = a.spill[spill_index<LDSD>(a, sp, gtid)];
                            ref = e.x; tn = __int_as_float(e.y);
                        }
                    }
                    if (SPEC && !fin && ref < -1 && pend == -1) { pend = ref; ref = -1; continue; }
                    break;
                }
            }
            const bool atNode = has && !fin && ref >= 0;
            if (!__any(atNode && (!SPEC || pend == -1))) break;
            if (atNode) {
                visits++;
                if constexpr (MODE == kModeAny) {
                    if (waveFlat) visit_any<true>(nodes, r, ref, push);
                    else visit_any<false>(nodes, r, ref, push);
                } else {
                    if (waveFlat) visit_closest<true>(nodes, r, best, ref, tn, push);
                    else visit_closest<false>(nodes, r, best, ref, tn, push);
                }
            }
        }

        // ---- (3) leaf phase ----
        const int leaf = SPEC ? pend : ((has && !fin) ? ref : -1);
        if (has && leaf < -1) {
            const int enc = -(leaf + 2), ts = enc >> 3, tc = (enc & 7) + 1;
            for (int k = 0; k < tc; k++) {
                const int ti = ts + k;
                tests++;
                const float4 p0 = tris[3 * ti];
                float t, u, v;
                if (!rtd::isect<MT>(r, p0, tris[3 * ti + 1], tris[3 * ti + 2], t, u, v)) continue;
                if (!(t > r.tmin) || !(t < r.tmax)) continue;
                if constexpr (MODE == kModeAny) {
                    found = true;
                    fin = true;
                    break;
                } else {
                    const int id = __float_as_int(p0.w);
                    if (t < best || (t == best && id < bid)) { best = t; bid = id; bu = u; bv = v; bti = ti; }
                }
            }
            if (SPEC) pend = -1;
            else ref = -1;
        }

        // ---- (4) retire finished rays ----
        if (has && fin) {
            if constexpr (MODE == kModeAny) {
                a.occ[ri] = found ? 1 : 0;
            } else if constexpr (MODE == kModeClosest) {
                put_hit(a, ri, bid >= 0, best, bid, bu, bv);
            } else {
                shade_pixel(a, r, ri, py, bid >= 0, bti);
            }
            if constexpr (COUNT) {
                a.cntNodes[ri] = visits;
                a.cntTris[ri] = tests;
            }
            has = false;
        }
    }
}

// Wave-packet schedule for coherent pinhole rays: one wave per 8x8 pixel
// tile shares one traversal stack. The node or leaf on top is wave-uniform,
// so it is fetched with scalar loads and every lane runs the same
// instructions; each stack slot keeps a per-lane entry distance, and a slot
// whose box no lane still needs (entry beyond every lane's closest t) is
// dropped without a fetch. Children are ordered near-first by the entry
// distance of the lowest lane that hits them. A lane accepts a triangle hit
// by the same rule as the per-ray traversal (tmin < t < tmax, lexicographic
// (t, id) minimum), so a lane's answer equals the per-ray answer.
template <int MODE, bool MT, int LDSD, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_packet(TraceArgs a) {
    constexpr int NW = BLOCK / 64;
    __shared__ int s_ref[NW][LDSD];
    __shared__ float s_tn[NW][LDSD][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int gwave = blockIdx.x * NW + w;
    // Read-only structure seen through the constant address space, so that
    // wave-uniform node and triangle reads become scalar loads.
    typedef __attribute__((address_space(4))) const unsigned int cu32;
    typedef __attribute__((address_space(4))) const float cf32;
    const cu32* nodes = (const cu32*)a.nodes;
    const cf32* tris = (const cf32*)a.tris;
    const int n = a.nTiles * 64;

    int band = blockIdx.x % rtd::kBands;

    for (;;) {
        const int base = rtd::take_chunk(a.counter, n, band);
        if (base >= n) break;
        const int tile = base >> 6;
        int tx, ty;
            pin_tile(a, tile, tx, ty);
        const int px = tx * 8 + (lane & 7), py = ty * 8 + (lane >> 3);
        const bool live = px < a.W && py < a.H;
        const int ri = py * a.W + px;
        Ray r;
        rtd::pinhole_ray(a.cam, live ? px : 0, live ? py : 0, a.W, a.H, r);
        rtd::ray_setup(r);
        const bool waveFlat = __any(live && rtd::ray_flat(r));
        // a dead lane keeps best = -inf: it culls every box and accepts nothing
        float best = live ? r.tmax : -INFINITY, bu = 0.f, bv = 0.f;
        int bid = -1, bti = -1;
        int sp = 0, ref = 0;
        float tn = live ? -INFINITY : INFINITY;

        for (;;) {
            if (__any(tn <= best)) {
                if (ref >= 0) {
                    const cu32* np = nodes + 16 * (size_t)ref;
                    const uint4 qx = make_uint4(np[0], np[1], np[2], np[3]);
                    const uint4 qy = make_uint4(np[4], np[5], np[6], np[7]);
                    const uint4 qz = make_uint4(np[8], np[9], np[10], np[11]);
                    const uint4 qc = make_uint4(np[12], np[13], np[14], np[15]);
                    const bool need = tn <= best;
                    float d[4];
                    int c[4];
                    uint64_t m[4];
                    float tt;
                    bool h;
#define RT_PCHILD(K, CH)                                                                                \
    h = need && (int)CH != -1 &&                                                                        \
        (waveFlat ? rtd::slab_child<K, true>(r, rtd::node_box<true>(r, qx, qy, qz), best, tt)           \
                  : rtd::slab_child<K, false>(r, rtd::node_box<false>(r, qx, qy, qz), best, tt));       \
    d[K] = h ? tt : INFINITY;                                                                           \
    m[K] = __ballot(h);                                                                                 \
    c[K] = (int)CH;
                    RT_PCHILD(0, qc.x) RT_PCHILD(1, qc.y) RT_PCHILD(2, qc.z) RT_PCHILD(3, qc.w)
#undef RT_PCHILD
                    // wave-uniform keys: entry distance of the lowest hitting lane
                    float key[4];
                    int slot[4] = {0, 1, 2, 3};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int l = m[k] ? (int)__builtin_ctzll(m[k]) : 0;
                        key[k] = m[k] ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d[k]), l)) : INFINITY;
                    }
#define RT_USWAP(i, j)                                                                          \
    {                                                                                           \
        if (key[i] > key[j]) { float tk = key[i]; key[i] = key[j]; key[j] = tk;                 \
                               int ts = slot[i]; slot[i] = slot[j]; slot[j] = ts; }             \
    }
                    RT_USWAP(0, 1) RT_USWAP(2, 3) RT_USWAP(0, 2) RT_USWAP(1, 3) RT_USWAP(1, 2)
#undef RT_USWAP
                    int nxt = -1;
                    float ntn = INFINITY;
#pragma unroll
                    for (int k = 3; k >= 0; k--) {
                        const int sl = slot[k];
                        const uint64_t mk = sl == 0 ? m[0] : sl == 1 ? m[1] : sl == 2 ? m[2] : m[3];
                        if (mk) {
                            const float dk = sl == 0 ? d[0] : sl == 1 ? d[1] : sl == 2 ? d[2] : d[3];
                            const int ck = sl == 0 ? c[0] : sl == 1 ? c[1] : sl == 2 ? c[2] : c[3];
                            if (nxt != -1) {
                                if (sp < LDSD) {
                                    if (lane == 0) s_ref[w][sp] = nxt;
                                    s_tn[w][sp][lane] = ntn;
                                } else if (sp < a.maxStack) {
                                    const size_t e = ((size_t)(sp - LDSD) * a.spillStride + gwave) * 64 + lane;
                                    a.spill[e] = make_int2(nxt, __float_as_int(ntn));
                                } else if (lane == 0) {
                                    atomicOr(a.status, 1);
                                }
                                sp = sp < a.maxStack ? sp + 1 : sp;
                            }
                            nxt = ck;
                            ntn = dk;
                        }
                    }
                    if (nxt != -1) {
                        ref = __builtin_amdgcn_readfirstlane(nxt);
                        tn = ntn;
                        continue;
                    }
                } else {
                    const int enc = -(ref + 2), ts = enc >> 3, tc = (enc & 7) + 1;
                    for (int k = 0; k < tc; k++) {
                        const int ti = ts + k;
                        const cf32* tp = tris + 12 * (size_t)ti;
                        const float4 p0 = make_float4(tp[0], tp[1], tp[2], tp[3]);
                        const float4 p1 = make_float4(tp[4], tp[5], tp[6], tp[7]);
                        const float4 p2 = make_float4(tp[8], tp[9], tp[10], tp[11]);
                        float t, u, v;
                        if (!rtd::isect<MT>(r, p0, p1, p2, t, u, v)) continue;
                        if (!(t > r.tmin) || !(t < r.tmax) || !(t <= best)) continue;
                        const int id = __float_as_int(p0.w);
                        if (t < best || id < bid) { best = t; bid = id; bu = u; bv = v; bti = ti; }
                    }
                }
            }
            if (sp == 0) break;
            sp--;
            if (sp < LDSD) {
                ref = __builtin_amdgcn_readfirstlane(s_ref[w][sp]);
                tn = s_tn[w][sp][lane];
            } else {
                const int2 e = a.spill[((size_t)(sp - LDSD) * a.spillStride + gwave) * 64 + lane];
                ref = __builtin_amdgcn_readfirstlane(e.x);
                tn = __int_as_float(e.y);
            }
        }
        if (live) {
            if constexpr (MODE == kModeClosest) {
                const bool hit = bid >= 0;
                a.t[ri] = hit ? best : RT_MISS_T;
                if (a.id) a.id[ri] = bid;
                if (a.u) a.u[ri] = hit ? bu : 0.f;
                if (a.v) a.v[ri] = hit ? bv : 0.f;
            } else {
                shade_pixel(a, r, ri, py, bid >= 0, bti);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Launch plumbing
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;
constexpr int kDefaultLds = 16;

static int num_cus() {
    static int n = 0;
    if (!n) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        if (n <= 0) n = 1;
    }
    return n;
}

// Shared launch path: grid = resident workgroups per CU x CUs (capped by the
// kernel's occupancy), per-lane spill region of (maxStack - LDSD) entries.
// Wave-packet kernels index the spill region per wave (spillStride = waves).
template <class K>
static int launch_kernel(const CudaBVHHandle* h, rt::TraceScratch& sc, TraceArgs& a, K kern, int ldsd, bool perWave,
                         int blocksPerCU, rt::KernelInfo& ki, hipStream_t s, bool stageHits = false) {
    rt::note_launch(kern, kBlock, ki, 0, 0, nullptr);
    int bpc = (blocksPerCU > 0 && blocksPerCU < ki.maxBlocks) ? blocksPerCU : ki.maxBlocks;
    a.maxStack = 3 * h->depth4 + 4;
    const int spillDepth = a.maxStack > ldsd ? a.maxStack - ldsd : 0;
    if (sc.fixed && spillDepth > 0) {
        // a fixed scratch caps the resident workgroups so every lane's spill fits
        const size_t perBlockRow = (size_t)spillDepth * kBlock * sizeof(int2) * (size_t)num_cus();
        const size_t fit = sc.spillBytes / perBlockRow;
        if (fit < 1) {
            rt::set_error("BVH depth %d needs %zu bytes of stack spill for one workgroup per CU; the fixed scratch "
                          "holds %zu", h->depth4, perBlockRow, sc.spillBytes);
            return -1;
        }
        bpc = (int)std::min<size_t>((size_t)bpc, fit);
    }
    const int grid = bpc * num_cus();
    rt::note_launch(kern, kBlock, ki, bpc, grid, sc.launchInfo);
    const int threads = grid * kBlock;
    const size_t need = (size_t)spillDepth * threads * sizeof(int2);
    if (sc.fixed) {
        if (need > sc.spillBytes) { rt::set_error("fixed scratch too small (%zu < %zu)", sc.spillBytes, need); return -1; }
    } else if (grow_spill(sc, need) != 0) {
        return -1;
    }
    if (a.W > 0) {
        a.tilesX = (a.W + 7) / 8;
        a.nTiles = a.tilesX * ((a.H + 7) / 8);
    }
    a.tileOrder = nullptr;
    a.tileCost = nullptr;
    if (a.costOrder && a.W > 0 && !sc.fixed) {
        if (a.nTiles > sc.costCap) {
            hipFree(sc.tileCost); hipFree(sc.tileOrder); hipFree(sc.tilePred);
            sc.tileCost = sc.tileOrder = sc.tilePred = nullptr; sc.costCap = 0; sc.costW = sc.costH = 0;
            sc.gen++;
            if (hipMalloc((void**)&sc.tileCost, (size_t)a.nTiles * sizeof(int)) != hipSuccess ||
                hipMalloc((void**)&sc.tileOrder, (size_t)a.nTiles * sizeof(int)) != hipSuccess ||
                hipMalloc((void**)&sc.tilePred, (size_t)a.nTiles * sizeof(int)) != hipSuccess) {
                hipFree(sc.tileCost); hipFree(sc.tileOrder); hipFree(sc.tilePred);
                sc.tileCost = sc.tileOrder = sc.tilePred = nullptr;
                rt::set_error("cannot allocate the tile cost order of %d tiles", a.nTiles);
                return -1;
            }
            sc.costCap = a.nTiles;
        }
        if (sc.costW != a.W || sc.costH != a.H) {
            // no costs for this image size yet: the first launch takes the plain order
            hipLaunchKernelGGL(k_fill_i32, dim3((a.nTiles + 255) / 256), dim3(256), 0, s, sc.tileCost, a.nTiles, 0);
            sc.costW = a.W; sc.costH = a.H;
        }
        hipLaunchKernelGGL(k_cost_dilate, dim3((a.nTiles + 255) / 256), dim3(256), 0, s, sc.tileCost, a.tilesX,
                           a.nTiles, sc.tilePred);
        hipLaunchKernelGGL(k_tile_order, dim3(rtd::kBands), dim3(kOrderBlock), 0, s, sc.tilePred, a.nTiles,
                           sc.tileOrder, sc.counter);
        a.tileOrder = sc.tileOrder;
        a.tileCost = sc.tileCost;
    }
    a.spill = (int2*)sc.spill;
    a.spillStride = perWave ? threads / 64 : threads;
    a.counter = sc.counter;
    a.status = sc.status;
    a.nodes = (const uint4*)h->d_nodes4;
    a.tris = (const float4*)h->d_tris;
    // staged hit records (not with the borrowed fixed scratch, which has no room for them)
    a.hits = nullptr;
    if (stageHits && !sc.fixed && a.t && a.n > 0) {
        if (grow_hits(sc, (size_t)a.n * sizeof(uint4)) != 0) return -1;
        a.hits = (uint4*)sc.hits;
    }
    // (a cost-ordered launch's k_tile_order has zeroed the counters)
    if (!a.tileOrder) hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(rtd::kBands), 0, s, sc.counter);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, a);
    if (a.hits) {
        const int g = std::min((a.n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_hits_split, dim3(g), dim3(256), 0, s, (const uint4*)a.hits, a.n, a.nDev, a.t, a.id, a.u,
                           a.v);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("trace launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

// SCHED: rt::kScheduleBatch, kScheduleWhileWhile or kScheduleRefill (uncounted)
template <int MODE, bool MT, bool PIN, bool COUNT, int LDSD, int SCHED>
static int launch_one(const CudaBVHHandle* h, rt::TraceScratch& sc, TraceArgs& a, int blocksPerCU, hipStream_t s) {
    static rt::KernelInfo ki;
    auto kern = SCHED == rt::kScheduleWhileWhile ? k_trace_ww<MODE, MT, PIN, COUNT, LDSD, kBlock>
              : (SCHED == rt::kScheduleRefill && !COUNT) ? k_trace_refill<MODE, MT, PIN, LDSD, kBlock>
                                                         : k_trace_batch<MODE, MT, PIN, COUNT, LDSD, kBlock>;
    // the while-while kernel stages hit records for k_hits_split; the refill
    // kernel gathers them per pool in LDS
    constexpr bool stage = MODE == kModeClosest && !PIN && !COUNT && SCHED == rt::kScheduleWhileWhile;
    // the cost order applies to uncounted pinhole launches of the batch schedule
    if (!(PIN && !COUNT && SCHED == rt::kScheduleBatch)) a.costOrder = false;
    return launch_kernel(h, sc, a, kern, LDSD, false, blocksPerCU, ki, s, stage);
}

template <int MODE, bool MT, int LDSD>
static int launch_packet(const CudaBVHHandle* h, rt::TraceScratch& sc, TraceArgs& a, int blocksPerCU, hipStream_t s) {
    static rt::KernelInfo ki;
    a.costOrder = false;
    return launch_kernel(h, sc, a, k_trace_packet<MODE, MT, LDSD, kBlock>, LDSD, true, blocksPerCU, ki, s);
}

int launch_chunk_order(const int* cost, int nChunks, int* order, hipStream_t s) {
    if (nChunks <= 0) return 0;
    hipLaunchKernelGGL(k_tile_order, dim3(rtd::kBands), dim3(kOrderBlock), 0, s, cost, nChunks, order, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int alloc_scratch(rt::TraceScratch& sc) {
    if (sc.counter) return 0;
    if (hipMalloc((void**)&sc.counter, rtd::kBands * rtd::kBandStride * sizeof(uint32_t)) != hipSuccess || hipMalloc((void**)&sc.status, 64) != hipSuccess ||
        hipMemset(sc.status, 0, 64) != hipSuccess) {
        free_scratch(sc);
        rt::set_error("cannot allocate traversal scratch");
        return -1;
    }
    return 0;
}

void free_scratch(rt::TraceScratch& sc) {
    hipFree(sc.spill); hipFree(sc.counter); hipFree(sc.status); hipFree(sc.tiles); hipFree(sc.hits);
    hipFree(sc.tileCost); hipFree(sc.tileOrder); hipFree(sc.tilePred);
    sc.spill = nullptr; sc.spillBytes = 0; sc.counter = nullptr; sc.status = nullptr;
    sc.tiles = nullptr; sc.tileCap = 0; sc.hits = nullptr; sc.hitsBytes = 0;
    sc.tileCost = sc.tileOrder = sc.tilePred = nullptr; sc.costCap = 0; sc.costW = sc.costH = 0;
    sc.gen++;
}

int grow_hits(rt::TraceScratch& sc, size_t need) {
    if (need <= sc.hitsBytes) return 0;
    hipFree(sc.hits);
    sc.hits = nullptr; sc.hitsBytes = 0;
    sc.gen++;
    if (hipMalloc(&sc.hits, need) != hipSuccess) {
        rt::set_error("cannot allocate %zu bytes of staged hit records", need);
        return -1;
    }
    sc.hitsBytes = need;
    return 0;
}

int take_status(rt::TraceScratch& sc, hipStream_t s, const char* fn) {
    if (!sc.status) return 0;
    int w = 0;
    if (hipMemcpyAsync(&w, sc.status, sizeof w, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        rt::set_error("%s: cannot read the traversal status word", fn);
        return -1;
    }
    if (w) {
        const int zero = 0;
        if (hipMemcpyAsync(sc.status, &zero, sizeof zero, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rt::set_error("%s: cannot clear the traversal status word", fn);
            return -1;
        }
        rt::set_error("%s: traversal stack overflow (status 0x%x), hits may be missing", fn, (unsigned)w);
    }
    return w;
}

int grow_spill(rt::TraceScratch& sc, size_t need) {
    if (need <= sc.spillBytes) return 0;
    if (sc.fixed) {
        rt::set_error("fixed traversal scratch too small: %zu bytes of spill needed, %zu held", need, sc.spillBytes);
        return -1;
    }
    hipFree(sc.spill);
    sc.spill = nullptr; sc.spillBytes = 0;
    sc.gen++;
    if (hipMalloc(&sc.spill, need) != hipSuccess) {
        rt::set_error("cannot allocate %zu bytes of traversal stack spill", need);
        return -1;
    }
    sc.spillBytes = need;
    return 0;
}

// Tunable variants (LDS stack depth, resident blocks per CU) exist for the
// watertight closest/any kernels; the counted, Moller-Trumbore and shading
// kernels use the default configuration.
template <int MODE, bool PIN, int SCHED>
static int launch_depth(const CudaBVHHandle* h, rt::TraceScratch& sc, TraceArgs& a, const rt::TraceConfig& cfg,
                        hipStream_t s) {
    switch (cfg.ldsDepth) {
        case 4: return launch_one<MODE, false, PIN, false, 4, SCHED>(h, sc, a, cfg.blocksPerCU, s);
        case 8: return launch_one<MODE, false, PIN, false, 8, SCHED>(h, sc, a, cfg.blocksPerCU, s);
        case 12: return launch_one<MODE, false, PIN, false, 12, SCHED>(h, sc, a, cfg.blocksPerCU, s);
        default: return launch_one<MODE, false, PIN, false, kDefaultLds, SCHED>(h, sc, a, cfg.blocksPerCU, s);
    }
}

template <int MODE, bool PIN>
static int launch_tuned(const CudaBVHHandle* h, rt::TraceScratch& sc, TraceArgs& a, const rt::TraceConfig& cfg,
                        hipStream_t s) {
    if constexpr (PIN && MODE == kModeClosest) {
        if (cfg.schedule == rt::kSchedulePacket) {
            switch (cfg.ldsDepth) {
                case 4: return launch_packet<MODE, false, 4>(h, sc, a, cfg.blocksPerCU, s);
                case 8: return launch_packet<MODE, false, 8>(h, sc, a, cfg.blocksPerCU, s);
                case 12: return launch_packet<MODE, false, 12>(h, sc, a, cfg.blocksPerCU, s);
                default: return launch_packet<MODE, false, 16>(h, sc, a, cfg.blocksPerCU, s);
            }
        }
    }
    if (cfg.schedule == rt::kScheduleWhileWhile) return launch_depth<MODE, PIN, rt::kScheduleWhileWhile>(h, sc, a, cfg, s);
    if (cfg.schedule == rt::kScheduleRefill) return launch_depth<MODE, PIN, rt::kScheduleRefill>(h, sc, a, cfg, s);
    return launch_depth<MODE, PIN, rt::kScheduleBatch>(h, sc, a, cfg, s);
}

int launch_trace(const CudaBVHHandle* h, rt::TraceScratch& sc, TraceArgs& a, int mode, bool mt, bool pinhole,
                 bool counted, hipStream_t s) {
    if (alloc_scratch(sc) != 0) return -1;
    if (a.n <= 0) return 0;
    const rt::TraceConfig& cfg = h->tune[mode == kModeAny ? rt::kClassAny : (pinhole ? rt::kClassCoherent : rt::kClassIncoherent)];
    if (!mt && !counted) {
        if (mode == kModeClosest) return pinhole ? launch_tuned<kModeClosest, true>(h, sc, a, cfg, s)
                                                 : launch_tuned<kModeClosest, false>(h, sc, a, cfg, s);
        if (mode == kModeAny) return pinhole ? launch_tuned<kModeAny, true>(h, sc, a, cfg, s)
                                             : launch_tuned<kModeAny, false>(h, sc, a, cfg, s);
    }
#define RT_L(M, MTv, P, C) \
    launch_one<M, MTv, P, C, kDefaultLds, P ? rt::kScheduleBatch : rt::kScheduleWhileWhile>(h, sc, a, 0, s)
    if (mode == kModeClosest) {
        if (mt) return pinhole ? (counted ? RT_L(kModeClosest, true, true, true) : RT_L(kModeClosest, true, true, false))
                               : (counted ? RT_L(kModeClosest, true, false, true) : RT_L(kModeClosest, true, false, false));
        return pinhole ? RT_L(kModeClosest, false, true, true) : RT_L(kModeClosest, false, false, true);
    }
    if (mode == kModeAny) {
        if (mt) return pinhole ? (counted ? RT_L(kModeAny, true, true, true) : RT_L(kModeAny, true, true, false))
                               : (counted ? RT_L(kModeAny, true, false, true) : RT_L(kModeAny, true, false, false));
        return pinhole ? RT_L(kModeAny, false, true, true) : RT_L(kModeAny, false, false, true);
    }
    if (mode == kModeShade) return mt ? launch_packet<kModeShade, true, 16>(h, sc, a, 0, s)
                                      : launch_packet<kModeShade, false, 16>(h, sc, a, 0, s);
#undef RT_L
    rt::set_error("bad trace mode %d", mode);
    return -1;
}

// ---------------------------------------------------------------------------
// Pinhole ray generation into a packet (same arithmetic as the fused kernel)
// ---------------------------------------------------------------------------
__global__ void k_gen_pinhole(rtd::Cam cam, int W, int H, float* ox, float* oy, float* oz, float* dx, float* dy,
                              float* dz, float* tmin, float* tmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const int py = i / W, px = i - py * W;
    Ray r;
    rtd::pinhole_ray(cam, px, py, W, H, r);
    ox[i] = r.ox; oy[i] = r.oy; oz[i] = r.oz;
    dx[i] = r.dx; dy[i] = r.dy; dz[i] = r.dz;
    tmin[i] = r.tmin; tmax[i] = r.tmax;
}

int launch_gen_pinhole(const rtd::Cam& cam, int W, int H, const CudaRayPacket& p, hipStream_t s) {
    const int n = W * H;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_gen_pinhole, dim3((n + 255) / 256), dim3(256), 0, s, cam, W, H, (float*)p.ox, (float*)p.oy,
                       (float*)p.oz, (float*)p.dx, (float*)p.dy, (float*)p.dz, (float*)p.tmin, (float*)p.tmax);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("ray generation launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

}  // namespace rtk

// ---------------------------------------------------------------------------
// Static device scratch of the no-alloc entry point: one array per GPU,
// zero-initialised at module load and held until the process ends. Layout:
// [0, 1 KB) the kBands work-distribution counters (kBandStride words apart),
// [1 KB, 1 KB + 64) the status word, [4 KB, end) the traversal stack spill.
// Host code never uses the array's name as an address: its device address is
// looked up once per device with hipGetSymbolAddress and kept per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[rtk::kNoAllocBytes];

namespace rtk {
namespace {
constexpr size_t kNoAllocStatusOff = 1024;
constexpr size_t kNoAllocSpillOff = 4096;
static_assert(rtd::kBands * rtd::kBandStride * sizeof(uint32_t) <= kNoAllocStatusOff, "counters overlap the status");
constexpr int kMaxDevices = 64;
std::mutex g_noallocMu;
void* g_noallocPtr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch(rt::TraceScratch& sc) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        rt::set_error("no-alloc scratch: no current device (or device %d beyond %d)", dev, kMaxDevices);
        return -1;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_noallocMu);
        if (!g_noallocPtr[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                rt::set_error("no-alloc scratch: hipGetSymbolAddress failed on device %d: %s", dev,
                              hipGetErrorString(e));
                return -1;
            }
            g_noallocPtr[dev] = p;
        }
        base = g_noallocPtr[dev];
    }
    uint8_t* b = static_cast<uint8_t*>(base);
    sc.counter = reinterpret_cast<uint32_t*>(b);
    sc.status = reinterpret_cast<int*>(b + kNoAllocStatusOff);
    sc.spill = b + kNoAllocSpillOff;
    sc.spillBytes = kNoAllocBytes - kNoAllocSpillOff;
    sc.fixed = true;
    return 0;
}

__global__ void k_zero_word(int* w) { *w = 0; }

int clear_status(rt::TraceScratch& sc, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, sc.status);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rt::set_error("status clear launch failed: %s", hipGetErrorString(e)); return -1; }
    return 0;
}

}  // namespace rtk

