/* rt_oracle.c — CPU oracle for the ray-tracing hot path (TEST INFRASTRUCTURE).
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg use this
 * file, as the checker. Built with -O2 -ffp-contract=off (see Makefile) so the
 * FP32 arithmetic is exactly the operation sequence written here.
 *
 * Reference lines cited are in the reference snapshot; "backend" means
 * layer/cuda_bvh_backend.cu.
 */
#include "rt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 1 + 2*gamma(3) rounded to float: the tmax inflation of Ize's robust slab
 * test (T. Ize, "Robust BVH Ray Traversal", JCGT 2013). */
#define ROBUST_TFAR (1.0f + 3.0f * 0x1p-23f)
#define MISS_T 1e30f

int orc_num_threads(int requested) {
#ifdef _OPENMP
    return requested > 0 ? requested : omp_get_max_threads();
#else
    (void)requested;
    return 1;
#endif
}

void orc_release(void* p) { free(p); }

/* ===================================================================== */
/* IEEE binary16 (the conversion cuda_fp16's __float2half performs)       */
/* ===================================================================== */
static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

uint16_t orc_f2h_rne(float x) {
    uint32_t u = fbits(x);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) {                       /* inf or nan */
        return (uint16_t)(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
    }
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 -> inf */
    if (a < 0x33000001u) return (uint16_t)sign;   /* < 2^-25 (tie goes to even zero) */
    int e = (int)(a >> 23) - 127;                 /* unbiased exponent */
    uint32_t m = (a & 0x7fffffu) | 0x800000u;     /* 24-bit significand */
    if (e < -14) {                                /* subnormal half */
        int shift = -e - 14 + 13;                 /* bits to drop */
        uint32_t q = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t q = ((uint32_t)(e + 15) << 10) | ((m >> 13) & 0x3ffu);
    uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;   /* carry may reach inf: correct */
    return (uint16_t)(sign | q);
}

float orc_h2f(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0x1f) return bitsf(sign | 0x7f800000u | (m << 13));
    if (e == 0) {
        if (m == 0) return bitsf(sign);
        float f = (float)m * 0x1p-24f;            /* exact */
        return sign ? -f : f;
    }
    return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

/* Next half toward +inf / -inf in value order. */
static uint16_t h_next_up(uint16_t h) {
    if ((h & 0x7fffu) == 0) return 0x0001u;            /* +-0 -> smallest subnormal */
    if (h & 0x8000u) return (uint16_t)(h - 1u);        /* negative: toward zero */
    if (h == 0x7c00u) return h;                        /* +inf stays */
    return (uint16_t)(h + 1u);
}
static uint16_t h_next_down(uint16_t h) {
    if ((h & 0x7fffu) == 0) return 0x8001u;
    if (h & 0x8000u) { if (h == 0xfc00u) return h; return (uint16_t)(h + 1u); }
    return (uint16_t)(h - 1u);
}

/* Outward rounding: the corrected form of backend :405-407 (which rounds to
 * nearest and can shrink a box). */
uint16_t orc_f2h_down(float x) {
    uint16_t h = orc_f2h_rne(x);
    if (orc_h2f(h) > x) h = h_next_down(h);
    return h;
}
uint16_t orc_f2h_up(float x) {
    uint16_t h = orc_f2h_rne(x);
    if (orc_h2f(h) < x) h = h_next_up(h);
    return h;
}

/* ===================================================================== */
/* Binned-SAH BVH2 builder (backend :186-251)                             */
/* ===================================================================== */
typedef struct { float mn[3], mx[3]; } box3;
typedef struct { box3 box; int left, right, triStart, triCount; } bnode2;

struct orc_bvh {
    int n, max_leaf;
    bnode2* nodes; int num_nodes;
    int* ordered; int num_ordered;
    uint8_t* b4; int num_b4; int b4_depth;
};

static box3 box_empty(void) {
    box3 b; for (int a = 0; a < 3; a++) { b.mn[a] = 1e30f; b.mx[a] = -1e30f; } return b;
}
static box3 box_merge(box3 a, box3 b) {      /* backend :191-194 */
    box3 r;
    for (int k = 0; k < 3; k++) { r.mn[k] = fminf(a.mn[k], b.mn[k]); r.mx[k] = fmaxf(a.mx[k], b.mx[k]); }
    return r;
}
static float sa_area(box3 b) {               /* backend :195-198 */
    float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}
/* Triangle box (backend :187-190); "+ 0.0f" turns -0 into +0, so that the
 * box never depends on which of two equal zeros a min/max returns. */
static box3 tri_box(const float* t) {
    box3 b;
    for (int a = 0; a < 3; a++) {
        b.mn[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]) + 0.0f;
        b.mx[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]) + 0.0f;
    }
    return b;
}

#define NBINS 16

/* Bin index of backend :232: floor((c-cmin)/(cmax-cmin)*(NB-1)), clamped. */
static int bin_of(float c, float cmin, float cmax) {
    int b = (int)((c - cmin) / (cmax - cmin) * (float)(NBINS - 1));
    return b < 0 ? 0 : (b >= NBINS ? NBINS - 1 : b);
}

typedef struct { int s, e; int parent; int side; } btask;

/* SAH collapse of the BVH2 into BVH4 nodes (dynamic programming over the
 * BVH2, as in the wide-BVH collapse of Ylitie et al. 2017, here with four
 * slots): cost[n][k-1] is the least total surface area of the BVH4 nodes
 * needed to present BVH2 subtree n as k slots (k = 1: n's own slot, a leaf or
 * a BVH4 node; k >= 2: n is dissolved into k slots of its parent). Leaves
 * cost 0 as one slot; the choice is then the expected number of node visits
 * of a ray. Ties take the fewer slots, then the smaller left count. */
#define CINF 3.0e38f
static void collapse_costs(const bnode2* nodes, int num, float* cost) {
    for (int n = num - 1; n >= 0; n--) {           /* children follow parents in pre-order */
        float* c = cost + 4 * (size_t)n;
        const bnode2* nd = &nodes[n];
        if (nd->triCount > 0) { c[0] = 0.0f; c[1] = c[2] = c[3] = CINF; continue; }
        const float* l = cost + 4 * (size_t)nd->left;
        const float* r = cost + 4 * (size_t)nd->right;
        float g[5] = {CINF, CINF, CINF, CINF, CINF};
        for (int k = 2; k <= 4; k++)
            for (int a = 1; a < k; a++) {
                const float x = l[a - 1] + r[k - a - 1];
                if (x < g[k]) g[k] = x;
            }
        float best = g[2];
        if (g[3] < best) best = g[3];
        if (g[4] < best) best = g[4];
        c[0] = sa_area(nd->box) + best;
        c[1] = g[2]; c[2] = g[3]; c[3] = g[4];
    }
}
/* Slot count m of subtree n inside its head -> (left count, right count). */
static void collapse_split(const float* cost, const bnode2* nodes, int n, int m, int* al, int* ar) {
    const float* l = cost + 4 * (size_t)nodes[n].left;
    const float* r = cost + 4 * (size_t)nodes[n].right;
    int ba = 1; float bx = CINF;
    for (int a = 1; a < m; a++) {
        const float x = l[a - 1] + r[m - a - 1];
        if (x < bx) { bx = x; ba = a; }
    }
    *al = ba; *ar = m - ba;
}
/* Slots (BVH2 nodes, in pre-order) of the BVH4 node headed by internal n. */
static int collapse_gather(const bnode2* nodes, const float* cost, int p, int gather[4]) {
    const float* c = cost + 4 * (size_t)p;
    int k = 2;
    if (c[2] < c[k - 1]) k = 3;
    if (c[3] < c[k - 1]) k = 4;
    int stn[4], stm[4], sp = 0, ng = 0;
    stn[sp] = p; stm[sp] = k; sp++;
    while (sp > 0) {
        sp--;
        const int n = stn[sp], m = stm[sp];
        if (m == 1) { gather[ng++] = n; continue; }
        int al, ar;
        collapse_split(cost, nodes, n, m, &al, &ar);
        stn[sp] = nodes[n].right; stm[sp] = ar; sp++;
        stn[sp] = nodes[n].left; stm[sp] = al; sp++;
    }
    return ng;
}

orc_bvh* orc_build(const float* tris, int n, int max_leaf) {
    if (!tris || n <= 0 || max_leaf < 1 || max_leaf > 8) return NULL;
    orc_bvh* B = (orc_bvh*)calloc(1, sizeof(orc_bvh));
    B->n = n; B->max_leaf = max_leaf;
    box3* pb = (box3*)malloc(sizeof(box3) * (size_t)n);
    float* cen = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    int* idx = (int*)malloc(sizeof(int) * (size_t)n);
    int* tmp = (int*)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) {
        pb[i] = tri_box(tris + 9 * (size_t)i);
        for (int a = 0; a < 3; a++) cen[3 * i + a] = (pb[i].mn[a] + pb[i].mx[a]) * .5f;   /* :207-208 */
        idx[i] = i;
    }
    B->nodes = (bnode2*)malloc(sizeof(bnode2) * (size_t)(2 * n));
    B->ordered = (int*)malloc(sizeof(int) * (size_t)n);
    btask* st = (btask*)malloc(sizeof(btask) * (size_t)(2 * n + 2));
    int sp = 0;
    st[sp++] = (btask){0, n, -1, 0};
    while (sp > 0) {
        btask tk = st[--sp];
        int s = tk.s, e = tk.e, cnt = e - s;
        int id = B->num_nodes++;                       /* pre-order numbering, :248 */
        if (tk.parent >= 0) {
            if (tk.side == 0) B->nodes[tk.parent].left = id; else B->nodes[tk.parent].right = id;
        }
        bnode2 nd; nd.left = nd.right = nd.triStart = nd.triCount = 0;
        nd.box = pb[idx[s]];
        for (int i = s + 1; i < e; i++) nd.box = box_merge(nd.box, pb[idx[i]]);   /* :216 */
        if (cnt <= max_leaf) {                          /* :220 (reference: <=3) */
            nd.triStart = B->num_ordered; nd.triCount = cnt;
            for (int i = s; i < e; i++) B->ordered[B->num_ordered++] = idx[i];
            B->nodes[id] = nd;
            continue;
        }
        float bestCost = 1e30f; int bestAxis = -1, bestBin = -1;
        float pA = sa_area(nd.box);
        for (int ax = 0; ax < 3; ax++) {                /* :223-245 */
            float cmin = 1e30f, cmax = -1e30f;
            for (int i = s; i < e; i++) { float c = cen[3 * idx[i] + ax]; cmin = fminf(cmin, c); cmax = fmaxf(cmax, c); }
            if (cmax - cmin < 1e-8f) continue;
            box3 bb[NBINS]; int bc[NBINS];
            for (int b = 0; b < NBINS; b++) { bb[b] = box_empty(); bc[b] = 0; }
            for (int i = s; i < e; i++) {
                int b = bin_of(cen[3 * idx[i] + ax], cmin, cmax);
                bb[b] = box_merge(bb[b], pb[idx[i]]); bc[b]++;
            }
            box3 L[NBINS], R[NBINS]; int lc[NBINS], rc[NBINS];
            L[0] = bb[0]; lc[0] = bc[0];
            for (int b = 1; b < NBINS; b++) { L[b] = box_merge(L[b - 1], bb[b]); lc[b] = lc[b - 1] + bc[b]; }
            R[NBINS - 1] = bb[NBINS - 1]; rc[NBINS - 1] = bc[NBINS - 1];
            for (int b = NBINS - 2; b >= 0; b--) { R[b] = box_merge(R[b + 1], bb[b]); rc[b] = rc[b + 1] + bc[b]; }
            for (int b = 0; b < NBINS - 1; b++) {
                if (lc[b] == 0 || rc[b + 1] == 0) continue;
                float cost = (float)lc[b] * sa_area(L[b]) / pA + (float)rc[b + 1] * sa_area(R[b + 1]) / pA + 1.f;  /* :240 */
                if (cost < bestCost) { bestCost = cost; bestAxis = ax; bestBin = b; }
            }
        }
        int mid;
        if (bestAxis >= 0) {
            /* Stable partition on the evaluated bin boundary (the reference
             * re-derives a plane that does not match the bins, :242, and then
             * std::sorts, :246; both are replaced by this exact rule). */
            float cmin = 1e30f, cmax = -1e30f;
            for (int i = s; i < e; i++) { float c = cen[3 * idx[i] + bestAxis]; cmin = fminf(cmin, c); cmax = fmaxf(cmax, c); }
            int nl = 0, nr = 0;
            for (int i = s; i < e; i++) {
                int p = idx[i];
                if (bin_of(cen[3 * p + bestAxis], cmin, cmax) <= bestBin) idx[s + nl++] = p; else tmp[nr++] = p;
            }
            memcpy(idx + s + nl, tmp, sizeof(int) * (size_t)nr);
            mid = s + nl;
        } else {
            mid = s + cnt / 2;                          /* all centroids coincide: :247 */
        }
        B->nodes[id] = nd;
        st[sp++] = (btask){mid, e, id, 1};              /* right popped after the whole left subtree */
        st[sp++] = (btask){s, mid, id, 0};
    }
    free(st); free(pb); free(cen); free(idx); free(tmp);

    /* ---- collapse to BVH4 (node format of :31-34 and :391-418; which BVH2
     * nodes become BVH4 nodes is the SAH choice of collapse_costs, not the
     * reference's open-both-children rule), pre-order numbering ---- */
    B->b4 = (uint8_t*)calloc((size_t)(B->num_nodes + 1), 64);
    float* ccost = (float*)malloc(sizeof(float) * 4 * (size_t)B->num_nodes);
    collapse_costs(B->nodes, B->num_nodes, ccost);
    typedef struct { int n2; int parent; int slot; int depth; } ctask;
    ctask* cs = (ctask*)malloc(sizeof(ctask) * (size_t)(B->num_nodes + 4));
    int csp = 0, maxd = 0;
    cs[csp++] = (ctask){0, -1, 0, 1};
    while (csp > 0) {
        ctask ct = cs[--csp];
        int me = B->num_b4++;
        if (ct.depth > maxd) maxd = ct.depth;
        if (ct.parent >= 0) {
            int32_t v = me; memcpy(B->b4 + 64 * (size_t)ct.parent + 48 + 4 * ct.slot, &v, 4);
        }
        const bnode2* nn = &B->nodes[ct.n2];
        int gather[4], ng = 0;
        if (nn->triCount > 0) {
            gather[ng++] = ct.n2;                       /* leaf root: one-child node */
        } else {
            ng = collapse_gather(B->nodes, ccost, ct.n2, gather);
        }
        uint8_t* out = B->b4 + 64 * (size_t)me;
        uint16_t hx[8], hy[8], hz[8]; int32_t child[4];
        for (int i = 0; i < 4; i++) {
            if (i < ng) {
                const box3* bx = &B->nodes[gather[i]].box;
                hx[i] = orc_f2h_down(bx->mn[0]); hx[4 + i] = orc_f2h_up(bx->mx[0]);
                hy[i] = orc_f2h_down(bx->mn[1]); hy[4 + i] = orc_f2h_up(bx->mx[1]);
                hz[i] = orc_f2h_down(bx->mn[2]); hz[4 + i] = orc_f2h_up(bx->mx[2]);
            } else {                                    /* :409-411: empty = (+inf,-inf) */
                hx[i] = hy[i] = hz[i] = 0x7c00u;
                hx[4 + i] = hy[4 + i] = hz[4 + i] = 0xfc00u;
            }
            child[i] = -1;
        }
        for (int i = 0; i < ng; i++) {
            const bnode2* cn = &B->nodes[gather[i]];
            if (cn->triCount > 0) child[i] = -((cn->triStart << 3) | (cn->triCount - 1)) - 2;   /* :393 */
        }
        memcpy(out, hx, 16); memcpy(out + 16, hy, 16); memcpy(out + 32, hz, 16); memcpy(out + 48, child, 16);
        for (int i = ng - 1; i >= 0; i--)
            if (B->nodes[gather[i]].triCount == 0) cs[csp++] = (ctask){gather[i], me, i, ct.depth + 1};
    }
    free(cs);
    free(ccost);
    B->b4_depth = maxd;
    return B;
}

void orc_free(orc_bvh* b) {
    if (!b) return;
    free(b->nodes); free(b->ordered); free(b->b4); free(b);
}
int orc_num_bvh2(const orc_bvh* b) { return b ? b->num_nodes : 0; }
int orc_num_bvh4(const orc_bvh* b) { return b ? b->num_b4 : 0; }
int orc_num_ordered(const orc_bvh* b) { return b ? b->num_ordered : 0; }
int orc_bvh4_depth(const orc_bvh* b) { return b ? b->b4_depth : 0; }
void orc_get_bvh4(const orc_bvh* b, void* out) { memcpy(out, b->b4, 64 * (size_t)b->num_b4); }
void orc_get_order(const orc_bvh* b, int* o) { memcpy(o, b->ordered, sizeof(int) * (size_t)b->num_ordered); }
void orc_get_bvh2(const orc_bvh* b, float* boxes, int* links) {
    for (int i = 0; i < b->num_nodes; i++) {
        const bnode2* nd = &b->nodes[i];
        for (int a = 0; a < 3; a++) { boxes[6 * i + a] = nd->box.mn[a]; boxes[6 * i + 3 + a] = nd->box.mx[a]; }
        links[4 * i] = nd->left; links[4 * i + 1] = nd->right;
        links[4 * i + 2] = nd->triStart; links[4 * i + 3] = nd->triCount;
    }
}

/* Subtree sizes of a pre-order BVH2 (children always follow the parent). */
static int* subtree_sizes(const bnode2* nodes, int nn) {
    int* sz = (int*)malloc(sizeof(int) * (size_t)nn);
    for (int i = nn - 1; i >= 0; i--) {
        sz[i] = 1;
        if (nodes[i].triCount == 0) sz[i] += sz[nodes[i].left] + sz[nodes[i].right];
    }
    return sz;
}

/* Stackless DFS BVH2 (:456-522). The pre-order numbering of the builder is
 * already the DFS order; skip = first node after the subtree, -1 at the end. */
static void write_dfs(const bnode2* nodes, int nn, uint32_t* out) {
    int* sz = subtree_sizes(nodes, nn);
    for (int i = 0; i < nn; i++) {
        const bnode2* nd = &nodes[i];
        uint32_t* p = out + 8 * (size_t)i;
        memcpy(p + 0, &nd->box.mn[0], 4); memcpy(p + 1, &nd->box.mn[1], 4); memcpy(p + 2, &nd->box.mn[2], 4);
        memcpy(p + 3, &nd->box.mx[0], 4); memcpy(p + 4, &nd->box.mx[1], 4); memcpy(p + 5, &nd->box.mx[2], 4);
        int32_t enc = nd->triCount > 0 ? -((nd->triStart << 3) | (nd->triCount - 1)) - 2 : 0;
        int32_t skip = (i + sz[i] < nn) ? i + sz[i] : -1;
        memcpy(p + 6, &enc, 4); memcpy(p + 7, &skip, 4);
    }
    free(sz);
}
int orc_num_dfs(const orc_bvh* b) { return b ? b->num_nodes : 0; }
void orc_get_dfs(const orc_bvh* b, uint32_t* out) { write_dfs(b->nodes, b->num_nodes, out); }

double orc_sah_cost(const orc_bvh* b) {
    double root = sa_area(b->nodes[0].box), c = 0;
    if (root <= 0) return 0;
    for (int i = 0; i < b->num_nodes; i++) {
        double a = sa_area(b->nodes[i].box) / root;
        c += b->nodes[i].triCount > 0 ? a * b->nodes[i].triCount : a;
    }
    return c;
}

/* ===================================================================== */
/* LBVH for TLAS (:257-389, :1620-1698)                                  */
/* ===================================================================== */
static uint32_t expand_bits(uint32_t v) {      /* spread 10 bits to every third bit */
    uint32_t r = 0;
    for (int i = 0; i < 10; i++) r |= ((v >> i) & 1u) << (3 * i);
    return r;
}
uint32_t orc_morton3d(float x, float y, float z) {   /* :265-273 */
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    return expand_bits((uint32_t)x) * 4u + expand_bits((uint32_t)y) * 2u + expand_bits((uint32_t)z);
}

int orc_lbvh_tlas(const float* aabbs, int n, uint32_t** out, int** ordered) {
    if (out) *out = NULL;
    if (ordered) *ordered = NULL;
    if (!aabbs || n <= 0 || !out) return 0;
    float smin[3] = {1e30f, 1e30f, 1e30f}, smax[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            float c = (aabbs[6 * i + a] + aabbs[6 * i + 3 + a]) * 0.5f;
            smin[a] = fminf(smin[a], c); smax[a] = fmaxf(smax[a], c);
        }
    float inv[3];
    for (int a = 0; a < 3; a++) inv[a] = (smax[a] - smin[a] > 1e-8f) ? 1.0f / (smax[a] - smin[a]) : 0.0f;
    uint32_t* code = (uint32_t*)malloc(4 * (size_t)n); int* id = (int*)malloc(4 * (size_t)n);
    uint32_t* code2 = (uint32_t*)malloc(4 * (size_t)n); int* id2 = (int*)malloc(4 * (size_t)n);
    for (int i = 0; i < n; i++) {
        float c[3];
        for (int a = 0; a < 3; a++) c[a] = (aabbs[6 * i + a] + aabbs[6 * i + 3 + a]) * 0.5f;
        code[i] = orc_morton3d((c[0] - smin[0]) * inv[0], (c[1] - smin[1]) * inv[1], (c[2] - smin[2]) * inv[2]);
        id[i] = i;
    }
    for (int pass = 0; pass < 4; pass++) {          /* stable LSD radix sort, :309-324 */
        int shift = 8 * pass, count[256] = {0}, pre[256];
        for (int i = 0; i < n; i++) count[(code[i] >> shift) & 0xff]++;
        pre[0] = 0;
        for (int i = 1; i < 256; i++) pre[i] = pre[i - 1] + count[i - 1];
        for (int i = 0; i < n; i++) { int b = (code[i] >> shift) & 0xff; code2[pre[b]] = code[i]; id2[pre[b]++] = id[i]; }
        uint32_t* tc = code; code = code2; code2 = tc; int* ti = id; id = id2; id2 = ti;
    }
    /* median-split tree, leaf = 1 item (:345-387), pre-order ids */
    bnode2* nodes = (bnode2*)malloc(sizeof(bnode2) * (size_t)(2 * n));
    int nn = 0;
    typedef struct { int s, e, parent, side, stage; } lt;
    lt* st = (lt*)malloc(sizeof(lt) * (size_t)(4 * n + 4));
    int sp = 0;
    st[sp++] = (lt){0, n, -1, 0, 0};
    while (sp > 0) {
        lt t = st[--sp];
        if (t.stage == 1) {                          /* post-order box propagation */
            bnode2* p = &nodes[t.parent];
            p->box = box_merge(nodes[p->left].box, nodes[p->right].box);
            continue;
        }
        int me = nn++;
        if (t.parent >= 0) { if (t.side == 0) nodes[t.parent].left = me; else nodes[t.parent].right = me; }
        bnode2* nd = &nodes[me];
        nd->left = nd->right = nd->triStart = nd->triCount = 0;
        nd->box = box_empty();
        int cnt = t.e - t.s;
        if (cnt <= 1) {
            nd->triStart = t.s; nd->triCount = cnt;
            for (int i = t.s; i < t.e; i++) {
                const float* b = aabbs + 6 * (size_t)id[i];
                for (int a = 0; a < 3; a++) { nd->box.mn[a] = fminf(nd->box.mn[a], b[a]); nd->box.mx[a] = fmaxf(nd->box.mx[a], b[3 + a]); }
            }
            continue;
        }
        int mid = t.s + cnt / 2;
        st[sp++] = (lt){0, 0, me, 0, 1};
        st[sp++] = (lt){mid, t.e, me, 1, 0};
        st[sp++] = (lt){t.s, mid, me, 0, 0};
    }
    free(st);
    uint32_t* o = (uint32_t*)malloc(32 * (size_t)nn);
    write_dfs(nodes, nn, o);
    *out = o;
    if (ordered) { *ordered = id; id = NULL; }
    free(nodes); free(code); free(code2); free(id); free(id2);
    return nn;
}

/* ===================================================================== */
/* Rays                                                                   */
/* ===================================================================== */
void orc_gen_pinhole(const float* cam, int W, int H, float* rays) {
    size_t n = (size_t)W * (size_t)H;
    const float *o = cam, *f = cam + 3, *r = cam + 6, *u = cam + 9;
    float sxs = cam[12], sys = cam[13];
    for (int py = 0; py < H; py++)
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            float sx = ((float)px + 0.5f) / (float)W * 2.0f - 1.0f;   /* backend :69-70 */
            float sy = ((float)py + 0.5f) / (float)H * 2.0f - 1.0f;
            float a = sx * sxs, b = sy * sys;
            float dx = (f[0] + a * r[0]) + b * u[0];
            float dy = (f[1] + a * r[1]) + b * u[1];
            float dz = (f[2] + a * r[2]) + b * u[2];
            float inv = 1.0f / sqrtf((dx * dx + dy * dy) + dz * dz);  /* IEEE, not rsqrtf (:72) */
            rays[0 * n + i] = o[0]; rays[1 * n + i] = o[1]; rays[2 * n + i] = o[2];
            rays[3 * n + i] = dx * inv; rays[4 * n + i] = dy * inv; rays[5 * n + i] = dz * inv;
            rays[6 * n + i] = 0.0f; rays[7 * n + i] = MISS_T;
        }
}

void orc_sort_keys(const float* rays, int n, const float* lo, const float* inv, uint32_t* keys) {
    for (int i = 0; i < n; i++) {
        float nx = (rays[i] - lo[0]) * inv[0];
        float ny = (rays[(size_t)n + i] - lo[1]) * inv[1];
        float nz = (rays[2 * (size_t)n + i] - lo[2]) * inv[2];
        uint32_t oct = (rays[3 * (size_t)n + i] < 0.f ? 4u : 0u) | (rays[4 * (size_t)n + i] < 0.f ? 2u : 0u) |
                       (rays[5 * (size_t)n + i] < 0.f ? 1u : 0u);
        keys[i] = (oct << 29) | (orc_morton3d(nx, ny, nz) & 0x1FFFFFFFu);
    }
}

/* ===================================================================== */
/* Ray-triangle tests                                                     */
/* ===================================================================== */
typedef struct {
    float o[3], d[3], tmin, tmax, inv[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
    int flat;
} xray;

static void setup_ray(xray* r, const float* rays, size_t n, size_t i) {
    for (int a = 0; a < 3; a++) { r->o[a] = rays[a * n + i]; r->d[a] = rays[(3 + a) * n + i]; }
    r->tmin = rays[6 * n + i]; r->tmax = rays[7 * n + i];
    /* A zero (or tiny) direction component gives an infinite reciprocal; such
     * a "flat" ray can produce (b - o) * inv = 0 * inf = NaN when its origin
     * lies exactly on a box plane, which slab4 resolves explicitly. */
    r->flat = 0;
    for (int a = 0; a < 3; a++) {
        r->inv[a] = 1.0f / r->d[a];
        if (isinf(r->inv[a])) r->flat = 1;
    }
    /* watertight shear (Woop et al. 2013, sec. 3) */
    float ax = fabsf(r->d[0]), ay = fabsf(r->d[1]), az = fabsf(r->d[2]);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (r->d[kz] < 0.0f) { int t = kx; kx = ky; ky = t; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    r->Sx = r->d[kx] / r->d[kz];
    r->Sy = r->d[ky] / r->d[kz];
    r->Sz = 1.0f / r->d[kz];
}

/* Watertight test; returns 1 with t,u,v (u = weight of v1, v = weight of v2)
 * when the ray line crosses the triangle with a non-zero determinant. The
 * caller applies the (tmin, tmax) and closest-hit rules. */
static int isect_wt(const xray* r, const float* tri, float* tt, float* uu, float* vv) {
    float A[3], Bv[3], C[3];
    for (int a = 0; a < 3; a++) { A[a] = tri[a] - r->o[a]; Bv[a] = tri[3 + a] - r->o[a]; C[a] = tri[6 + a] - r->o[a]; }
    float Ax = A[r->kx] - r->Sx * A[r->kz], Ay = A[r->ky] - r->Sy * A[r->kz];
    float Bx = Bv[r->kx] - r->Sx * Bv[r->kz], By = Bv[r->ky] - r->Sy * Bv[r->kz];
    float Cx = C[r->kx] - r->Sx * C[r->kz], Cy = C[r->ky] - r->Sy * C[r->kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {      /* FP64 fallback on an edge */
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    float det = U + V + W;
    if (det == 0.0f) return 0;
    float Az = r->Sz * A[r->kz], Bz = r->Sz * Bv[r->kz], Cz = r->Sz * C[r->kz];
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    *tt = T * rcp; *uu = V * rcp; *vv = W * rcp;
    return 1;
}

/* Moller-Trumbore exactly as backend :145-156 (IEEE division in place of the
 * fast-math reciprocal); the t > 0 and t < tHit checks are left to the caller. */
static int isect_mt(const xray* r, const float* tri, float* tt, float* uu, float* vv) {
    float dx = r->d[0], dy = r->d[1], dz = r->d[2];
    float e1x = tri[3] - tri[0], e1y = tri[4] - tri[1], e1z = tri[5] - tri[2];
    float e2x = tri[6] - tri[0], e2y = tri[7] - tri[1], e2z = tri[8] - tri[2];
    float ppx = dy * e2z - dz * e2y, ppy = dz * e2x - dx * e2z, ppz = dx * e2y - dy * e2x;
    float det = e1x * ppx + e1y * ppy + e1z * ppz;
    if (fabsf(det) < 1e-12f) return 0;
    float inv = 1.f / det;
    float tx = r->o[0] - tri[0], ty = r->o[1] - tri[1], tz = r->o[2] - tri[2];
    float u = inv * (tx * ppx + ty * ppy + tz * ppz);
    if (u < 0.f || u > 1.f) return 0;
    float qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    float v = inv * (dx * qx + dy * qy + dz * qz);
    if (v < 0.f || u + v > 1.f) return 0;
    *tt = inv * (e2x * qx + e2y * qy + e2z * qz); *uu = u; *vv = v;
    return 1;
}

static int isect(int mode, const xray* r, const float* tri, float* t, float* u, float* v) {
    return mode == 1 ? isect_mt(r, tri, t, u, v) : isect_wt(r, tri, t, u, v);
}

int orc_isect_watertight(const float* ray, const float* tri, float* t, float* u, float* v) {
    xray r; setup_ray(&r, ray, 1, 0);
    return isect_wt(&r, tri, t, u, v);
}
int orc_isect_mt(const float* ray, const float* tri, float* t, float* u, float* v) {
    xray r; setup_ray(&r, ray, 1, 0);
    return isect_mt(&r, tri, t, u, v);
}

/* ===================================================================== */
/* BVH4 traversal                                                         */
/* ===================================================================== */
/* Child slab test shared by both traversals: robust far distance (Ize 2013),
 * inclusive planes for flat rays. Returns 1 and the entry distance when the box may
 * hold a hit in [tmin, tlimit]. */
static int slab4(const uint8_t* node, int c, const xray* r, float tlimit, float* tnear) {
    uint16_t h[6];
    memcpy(&h[0], node + 2 * c, 2);       memcpy(&h[1], node + 2 * (4 + c), 2);
    memcpy(&h[2], node + 16 + 2 * c, 2);  memcpy(&h[3], node + 16 + 2 * (4 + c), 2);
    memcpy(&h[4], node + 32 + 2 * c, 2);  memcpy(&h[5], node + 32 + 2 * (4 + c), 2);
    float t1x = (orc_h2f(h[0]) - r->o[0]) * r->inv[0], t2x = (orc_h2f(h[1]) - r->o[0]) * r->inv[0];
    float t1y = (orc_h2f(h[2]) - r->o[1]) * r->inv[1], t2y = (orc_h2f(h[3]) - r->o[1]) * r->inv[1];
    float t1z = (orc_h2f(h[4]) - r->o[2]) * r->inv[2], t2z = (orc_h2f(h[5]) - r->o[2]) * r->inv[2];
    if (r->flat) {
        /* origin on a plane of an axis the ray runs parallel to: that axis
         * must not constrain the interval (inclusive slab) */
        if (t1x != t1x) t1x = -r->inv[0];
        if (t2x != t2x) t2x = r->inv[0];
        if (t1y != t1y) t1y = -r->inv[1];
        if (t2y != t2y) t2y = r->inv[1];
        if (t1z != t1z) t1z = -r->inv[2];
        if (t2z != t2z) t2z = r->inv[2];
    }
    float tN = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    float tF = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z)) * ROBUST_TFAR;
    *tnear = tN;
    return (tN <= tF) && (tF >= r->tmin) && (tN <= tlimit);
}

typedef struct { int ref; float tn; } sentry;

static void closest_one(const uint8_t* nodes, const float* tris, const int* pid, const xray* r, int mode,
                        sentry* stk, float* ot, int* oid, float* ou, float* ov, int* nv, int* nt) {
    float best = r->tmax, bu = 0.f, bv = 0.f;
    int bid = -1, visits = 0, tests = 0, sp = 0;
    stk[sp++] = (sentry){0, -INFINITY};
    while (sp > 0) {
        sentry e = stk[--sp];
        if (e.tn > best) continue;                     /* re-check against the current hit */
        if (e.ref < 0) {                               /* leaf, backend :141-157 */
            int enc = -(e.ref + 2), ts = enc >> 3, tc = (enc & 7) + 1;
            for (int k = 0; k < tc; k++) {
                int ti = ts + k; float t, u, v;
                tests++;
                if (!isect(mode, r, tris + 9 * (size_t)ti, &t, &u, &v)) continue;
                if (!(t > r->tmin) || !(t < r->tmax)) continue;
                int id = pid[ti];
                /* best starts at tmax and t < tmax, so the first valid hit wins */
                if (t < best || (t == best && id < bid)) { best = t; bid = id; bu = u; bv = v; }
            }
            continue;
        }
        const uint8_t* nd = nodes + 64 * (size_t)e.ref;
        visits++;
        int32_t ch[4]; memcpy(ch, nd + 48, 16);
        float dist[4]; int child[4];
        for (int c = 0; c < 4; c++) {
            float tn;
            if (ch[c] != -1 && slab4(nd, c, r, best, &tn)) { dist[c] = tn; child[c] = ch[c]; }
            else { dist[c] = INFINITY; child[c] = -1; }
        }
#define CSWAP(a, b) do { if (dist[a] > dist[b]) { float td = dist[a]; dist[a] = dist[b]; dist[b] = td; \
                         int tc2 = child[a]; child[a] = child[b]; child[b] = tc2; } } while (0)
        CSWAP(0, 1); CSWAP(2, 3); CSWAP(0, 2); CSWAP(1, 3); CSWAP(1, 2);      /* :177 */
#undef CSWAP
        for (int c = 3; c >= 0; c--)
            if (child[c] != -1) stk[sp++] = (sentry){child[c], dist[c]};
    }
    if (bid < 0) { *ot = MISS_T; *oid = -1; *ou = 0.f; *ov = 0.f; }
    else { *ot = best; *oid = bid; *ou = bu; *ov = bv; }
    if (nv) *nv = visits;
    if (nt) *nt = tests;
}

static int stack_cap(int num_nodes) {
    /* A DFS over a 4-ary tree holds at most 3 pending siblings per level plus
     * 4 children; bounding the depth by the node count keeps it unbounded in
     * effect (no silent drops, unlike backend :179). */
    (void)num_nodes;
    return 3 * num_nodes + 8;
}

void orc_trace_closest(const void* nodes, int num_nodes, const float* tris, const int* prim_id, int num_tris,
                       const float* rays, int n_rays, float* t, int* id, float* u, float* v,
                       int* node_visits, int* tri_tests, int isect_mode, int threads) {
    (void)num_tris;
    int nth = orc_num_threads(threads);
    int cap = stack_cap(num_nodes);
#pragma omp parallel num_threads(nth)
    {
        sentry* stk = (sentry*)malloc(sizeof(sentry) * (size_t)cap);
#pragma omp for schedule(dynamic, 512)
        for (int i = 0; i < n_rays; i++) {
            xray r; setup_ray(&r, rays, (size_t)n_rays, (size_t)i);
            closest_one((const uint8_t*)nodes, tris, prim_id, &r, isect_mode, stk, &t[i], &id[i], &u[i], &v[i],
                        node_visits ? &node_visits[i] : NULL, tri_tests ? &tri_tests[i] : NULL);
        }
        free(stk);
    }
}

void orc_trace_any(const void* nodes_, int num_nodes, const float* tris, const int* prim_id, int num_tris,
                   const float* rays, int n_rays, uint8_t* occluded, int* node_visits, int* tri_tests,
                   int isect_mode, int threads) {
    (void)num_tris; (void)prim_id;
    const uint8_t* nodes = (const uint8_t*)nodes_;
    int nth = orc_num_threads(threads);
    int cap = stack_cap(num_nodes);
#pragma omp parallel num_threads(nth)
    {
        int* stk = (int*)malloc(sizeof(int) * (size_t)cap);
#pragma omp for schedule(dynamic, 512)
        for (int i = 0; i < n_rays; i++) {
            xray r; setup_ray(&r, rays, (size_t)n_rays, (size_t)i);
            int sp = 0, found = 0, visits = 0, tests = 0;
            stk[sp++] = 0;
            while (sp > 0 && !found) {
                int ref = stk[--sp];
                if (ref < 0) {
                    int enc = -(ref + 2), ts = enc >> 3, tc = (enc & 7) + 1;
                    for (int k = 0; k < tc && !found; k++) {
                        float tt, uu, vv;
                        tests++;
                        if (isect(isect_mode, &r, tris + 9 * (size_t)(ts + k), &tt, &uu, &vv) &&
                            tt > r.tmin && tt < r.tmax) found = 1;
                    }
                    continue;
                }
                const uint8_t* nd = nodes + 64 * (size_t)ref;
                visits++;
                int32_t ch[4]; memcpy(ch, nd + 48, 16);
                for (int c = 0; c < 4; c++) {                 /* v10_pipeline.cu:241-244, unsorted */
                    float tn;
                    if (ch[c] != -1 && slab4(nd, c, &r, r.tmax, &tn)) stk[sp++] = ch[c];
                }
            }
            occluded[i] = (uint8_t)found;
            if (node_visits) node_visits[i] = visits;
            if (tri_tests) tri_tests[i] = tests;
        }
        free(stk);
    }
}

/* ===================================================================== */
/* Brute force                                                            */
/* ===================================================================== */
void orc_brute_closest(const float* tris, const int* prim_id, int num_tris, const float* rays, int n_rays,
                       float* t, int* id, float* u, float* v, int isect_mode, int threads) {
    int nth = orc_num_threads(threads);
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
    for (int i = 0; i < n_rays; i++) {
        xray r; setup_ray(&r, rays, (size_t)n_rays, (size_t)i);
        float best = r.tmax, bu = 0, bv = 0; int bid = -1;
        for (int k = 0; k < num_tris; k++) {
            float tt, uu, vv;
            if (!isect(isect_mode, &r, tris + 9 * (size_t)k, &tt, &uu, &vv)) continue;
            if (!(tt > r.tmin) || !(tt < r.tmax)) continue;
            int pk = prim_id[k];
            if (bid < 0 || tt < best || (tt == best && pk < bid)) { best = tt; bid = pk; bu = uu; bv = vv; }
        }
        if (bid < 0) { t[i] = MISS_T; id[i] = -1; u[i] = 0; v[i] = 0; }
        else { t[i] = best; id[i] = bid; u[i] = bu; v[i] = bv; }
    }
}

void orc_brute_closest_f64(const float* tris, const int* prim_id, int num_tris, const float* rays, int n_rays,
                           double* t, int* id, int threads) {
    int nth = orc_num_threads(threads);
    size_t n = (size_t)n_rays;
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
    for (int i = 0; i < n_rays; i++) {
        double o[3], d[3];
        for (int a = 0; a < 3; a++) { o[a] = rays[a * n + i]; d[a] = rays[(3 + a) * n + i]; }
        double tmin = rays[6 * n + i], tmax = rays[7 * n + i];
        double best = tmax; int bid = -1;
        for (int k = 0; k < num_tris; k++) {
            const float* q = tris + 9 * (size_t)k;
            double e1[3], e2[3], s[3], p[3], qq[3];
            for (int a = 0; a < 3; a++) { e1[a] = (double)q[3 + a] - q[a]; e2[a] = (double)q[6 + a] - q[a]; s[a] = o[a] - q[a]; }
            p[0] = d[1] * e2[2] - d[2] * e2[1]; p[1] = d[2] * e2[0] - d[0] * e2[2]; p[2] = d[0] * e2[1] - d[1] * e2[0];
            double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
            if (det == 0.0) continue;
            double inv = 1.0 / det;
            double uu = inv * (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]);
            if (uu < 0.0 || uu > 1.0) continue;
            qq[0] = s[1] * e1[2] - s[2] * e1[1]; qq[1] = s[2] * e1[0] - s[0] * e1[2]; qq[2] = s[0] * e1[1] - s[1] * e1[0];
            double vv = inv * (d[0] * qq[0] + d[1] * qq[1] + d[2] * qq[2]);
            if (vv < 0.0 || uu + vv > 1.0) continue;
            double tt = inv * (e2[0] * qq[0] + e2[1] * qq[1] + e2[2] * qq[2]);
            if (!(tt > tmin) || !(tt < tmax)) continue;
            if (bid < 0 || tt < best || (tt == best && prim_id[k] < bid)) { best = tt; bid = prim_id[k]; }
        }
        t[i] = bid < 0 ? (double)MISS_T : best; id[i] = bid;
    }
}

/* ===================================================================== */
/* Two-level traversal (TLAS over instances, BLAS per mesh)               */
/* ===================================================================== */
/* The TLAS is traversed like a BVH4; reaching a TLAS leaf (one instance)
 * pushes a return marker, transforms the ray into the instance's object
 * space and continues into the BLAS; popping the marker restores the world
 * ray. Node order and culling are those of closest_one. */
#define RET_MARK INT_MIN

static void xform_ray(const xray* w, const float* M, xray* o) {
    float r[8];
    for (int k = 0; k < 3; k++) {
        const float* m = M + 4 * k;
        r[k] = ((m[0] * w->o[0] + m[1] * w->o[1]) + m[2] * w->o[2]) + m[3];
        r[3 + k] = (m[0] * w->d[0] + m[1] * w->d[1]) + m[2] * w->d[2];
    }
    r[6] = w->tmin; r[7] = w->tmax;
    setup_ray(o, r, 1, 0);
}

static void two_level_one(const uint8_t* tlas, const int* tlas_prim, const orc_blas* blas, const int* blas_of,
                          const float* w2o, const xray* rw, int any, sentry* stk,
                          float* ot, int* oinst, int* oid, float* ou, float* ov, int* found_out, int* nv, int* nt,
                          int* oleaf) {
    float best = rw->tmax, bu = 0.f, bv = 0.f;
    int bid = -1, binst = -1, bleaf = -1, visits = 0, tests = 0, sp = 0, inst = -1, found = 0;
    xray ro;
    const xray* r = rw;
    const uint8_t* nodes = tlas;
    stk[sp++] = (sentry){0, -INFINITY};
    while (sp > 0 && !found) {
        sentry e = stk[--sp];
        if (e.ref == RET_MARK) { inst = -1; r = rw; nodes = tlas; continue; }
        if (!any && e.tn > best) continue;
        if (e.ref < 0) {
            int enc = -(e.ref + 2), ts = enc >> 3, tc = (enc & 7) + 1;
            if (inst < 0) {                                   /* TLAS leaf: enter the instance */
                int k = tlas_prim[ts];
                stk[sp++] = (sentry){RET_MARK, -INFINITY};
                inst = k;
                xform_ray(rw, w2o + 12 * (size_t)k, &ro);
                r = &ro;
                nodes = (const uint8_t*)blas[blas_of[k]].nodes;
                stk[sp++] = (sentry){0, -INFINITY};
                (void)tc;
                continue;
            }
            const orc_blas* B = &blas[blas_of[inst]];
            for (int q = 0; q < tc && !found; q++) {
                int ti = ts + q; float tt, uu, vv;
                tests++;
                if (!isect_wt(r, B->tris + 9 * (size_t)ti, &tt, &uu, &vv)) continue;
                if (!(tt > r->tmin) || !(tt < r->tmax)) continue;
                if (any) { found = 1; break; }
                int id = B->prim_id[ti];
                if (tt < best || (tt == best && (inst < binst || (inst == binst && id < bid)))) {
                    best = tt; bid = id; binst = inst; bu = uu; bv = vv; bleaf = ti;
                }
            }
            continue;
        }
        const uint8_t* nd = nodes + 64 * (size_t)e.ref;
        visits++;
        int32_t ch[4]; memcpy(ch, nd + 48, 16);
        float lim = any ? r->tmax : best;
        if (any) {
            for (int c = 0; c < 4; c++) {
                float tn;
                if (ch[c] != -1 && slab4(nd, c, r, lim, &tn)) stk[sp++] = (sentry){ch[c], 0.f};
            }
            continue;
        }
        float dist[4]; int child[4];
        for (int c = 0; c < 4; c++) {
            float tn;
            if (ch[c] != -1 && slab4(nd, c, r, lim, &tn)) { dist[c] = tn; child[c] = ch[c]; }
            else { dist[c] = INFINITY; child[c] = -1; }
        }
#define CSWAP(a, b) do { if (dist[a] > dist[b]) { float td = dist[a]; dist[a] = dist[b]; dist[b] = td; \
                         int tc2 = child[a]; child[a] = child[b]; child[b] = tc2; } } while (0)
        CSWAP(0, 1); CSWAP(2, 3); CSWAP(0, 2); CSWAP(1, 3); CSWAP(1, 2);
#undef CSWAP
        for (int c = 3; c >= 0; c--)
            if (child[c] != -1) stk[sp++] = (sentry){child[c], dist[c]};
    }
    if (found_out) *found_out = found;
    if (ot) {
        if (bid < 0) { *ot = MISS_T; *oinst = -1; *oid = -1; *ou = 0.f; *ov = 0.f; }
        else { *ot = best; *oinst = binst; *oid = bid; *ou = bu; *ov = bv; }
    }
    if (nv) *nv = visits;
    if (nt) *nt = tests;
    if (oleaf) *oleaf = bleaf;
}

static int max_nodes_2l(int tlas_num_nodes, const orc_blas* blas, const int* blas_of, int n_inst) {
    int m = 0;
    for (int i = 0; i < n_inst; i++) if (blas[blas_of[i]].num_nodes > m) m = blas[blas_of[i]].num_nodes;
    return m + tlas_num_nodes;
}

void orc_trace_closest_2l(const void* tlas_nodes, int tlas_num_nodes, const int* tlas_prim, const orc_blas* blas,
                          const int* blas_of, const float* w2o, const float* rays, int n_rays, float* t, int* inst,
                          int* id, float* u, float* v, int* node_visits, int* tri_tests, int threads) {
    int nth = orc_num_threads(threads);
    int cap = 3 * (tlas_num_nodes + 4096) + 16;
    (void)max_nodes_2l;
#pragma omp parallel num_threads(nth)
    {
        sentry* stk = (sentry*)malloc(sizeof(sentry) * (size_t)cap * 4);
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n_rays; i++) {
            xray r; setup_ray(&r, rays, (size_t)n_rays, (size_t)i);
            two_level_one((const uint8_t*)tlas_nodes, tlas_prim, blas, blas_of, w2o, &r, 0, stk, &t[i], &inst[i],
                          &id[i], &u[i], &v[i], NULL, node_visits ? &node_visits[i] : NULL,
                          tri_tests ? &tri_tests[i] : NULL, NULL);
        }
        free(stk);
    }
}

void orc_trace_any_2l(const void* tlas_nodes, int tlas_num_nodes, const int* tlas_prim, const orc_blas* blas,
                      const int* blas_of, const float* w2o, const float* rays, int n_rays, uint8_t* occluded,
                      int* node_visits, int* tri_tests, int threads) {
    int nth = orc_num_threads(threads);
    int cap = 3 * (tlas_num_nodes + 4096) + 16;
#pragma omp parallel num_threads(nth)
    {
        sentry* stk = (sentry*)malloc(sizeof(sentry) * (size_t)cap * 4);
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n_rays; i++) {
            xray r; setup_ray(&r, rays, (size_t)n_rays, (size_t)i);
            int f = 0;
            two_level_one((const uint8_t*)tlas_nodes, tlas_prim, blas, blas_of, w2o, &r, 1, stk, NULL, NULL, NULL,
                          NULL, NULL, &f, node_visits ? &node_visits[i] : NULL, tri_tests ? &tri_tests[i] : NULL,
                          NULL);
            occluded[i] = (uint8_t)f;
        }
        free(stk);
    }
}

/* ===================================================================== */
/* Path tracer                                                            */
/* ===================================================================== */
/* Structure of the reference's wavefront renderer (cuda_rt_v34.cu: state
 * :62-97, generate :1077-1150, shade with next-event estimation and Russian
 * roulette :1185-1455, shadow :1458-1480), with the lighting and sampling of
 * this project (directional sun, sky gradient, Lambertian per-instance albedo,
 * PCG32 + Malley cosine sampling). Per pixel and sample, one path: the
 * contributions are added in the order the wavefront kernels add them (each
 * bounce: an escaped ray's sky term, or the hit's unoccluded sun term). */

typedef struct { uint64_t state, inc; } pcg32;

static uint32_t pcg_next(pcg32* r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
static void pcg_seed(pcg32* r, uint64_t initstate, uint64_t initseq) {
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    pcg_next(r);
    r->state += initstate;
    pcg_next(r);
}
static float pcg_uniform(pcg32* r) { return (float)(pcg_next(r) >> 8) * 0x1p-24f; }

/* Malley's method in the basis of Duff et al. 2017 (float32). */
static void cosine_dir_f(const float n[3], pcg32* rng, float d[3]) {
    float x = 0.f, y = 0.f, s = 0.f;
    int ok = 0;
    for (int k = 0; k < 64 && !ok; k++) {
        x = 2.0f * pcg_uniform(rng) - 1.0f;
        y = 2.0f * pcg_uniform(rng) - 1.0f;
        s = x * x + y * y;
        ok = s < 1.0f;
    }
    if (!ok) { x = 0.f; y = 0.f; s = 0.f; }
    float z = sqrtf(1.0f - s);
    float sg = copysignf(1.0f, n[2]);
    float ca = -1.0f / (sg + n[2]);
    float cb = n[0] * n[1] * ca;
    float t1x = 1.0f + sg * n[0] * n[0] * ca, t1y = sg * cb, t1z = -sg * n[0];
    float t2x = cb, t2y = sg + n[1] * n[1] * ca, t2z = -n[1];
    d[0] = (x * t1x + y * t2x) + z * n[0];
    d[1] = (x * t1y + y * t2y) + z * n[1];
    d[2] = (x * t1z + y * t2z) + z * n[2];
    float dl = 1.0f / sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    d[0] = d[0] * dl; d[1] = d[1] * dl; d[2] = d[2] * dl;
}

static void sun_unit(const orc_pt_params* p, float out[3]) {
    double sl = sqrt((double)p->sun_dir[0] * p->sun_dir[0] + (double)p->sun_dir[1] * p->sun_dir[1] +
                     (double)p->sun_dir[2] * p->sun_dir[2]);
    for (int k = 0; k < 3; k++) out[k] = (float)(p->sun_dir[k] / sl);
}

static void clamp3f(float c[3], float m) {
    float mx = fmaxf(c[0], fmaxf(c[1], c[2]));
    if (m > 0.0f && mx > m) { float s = m / mx; c[0] = c[0] * s; c[1] = c[1] * s; c[2] = c[2] * s; }
}

static void make_xray(xray* r, const float o[3], const float d[3], float tmin, float tmax) {
    float a[8] = {o[0], o[1], o[2], d[0], d[1], d[2], tmin, tmax};
    setup_ray(r, a, 1, 0);
}

/* World-space unit geometric normal of a triangle (object space, 9 floats)
 * under an instance with world-to-object matrix M (3x4): the object-space
 * e1 x e2 (the mesh's winding) mapped with the transpose of M's 3x3, which is
 * the inverse transpose of the object-to-world map, then normalised. */
static void world_normal(const float* tri, const float* M, float n[3]) {
    float e1x = tri[3] - tri[0], e1y = tri[4] - tri[1], e1z = tri[5] - tri[2];
    float e2x = tri[6] - tri[0], e2y = tri[7] - tri[1], e2z = tri[8] - tri[2];
    float mx = e1y * e2z - e1z * e2y, my = e1z * e2x - e1x * e2z, mz = e1x * e2y - e1y * e2x;
    n[0] = (M[0] * mx + M[4] * my) + M[8] * mz;
    n[1] = (M[1] * mx + M[5] * my) + M[9] * mz;
    n[2] = (M[2] * mx + M[6] * my) + M[10] * mz;
    float nl = 1.0f / sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    for (int k = 0; k < 3; k++) n[k] = n[k] * nl;
}

void orc_scene_hit_normals(const orc_blas* blas, const int* blas_of, const float* w2o, int num_inst, int n,
                           const int* inst, const int* prim, float* out) {
    int nb = 0;
    for (int k = 0; k < num_inst; k++) if (blas_of[k] + 1 > nb) nb = blas_of[k] + 1;
    int** leaf_of = (int**)calloc((size_t)nb, sizeof(int*));   /* original ID -> leaf-order index */
    for (int b = 0; b < nb; b++) {
        leaf_of[b] = (int*)malloc(sizeof(int) * (size_t)(blas[b].num_tris > 0 ? blas[b].num_tris : 1));
        for (int li = 0; li < blas[b].num_tris; li++) leaf_of[b][blas[b].prim_id[li]] = li;
    }
    for (int i = 0; i < n; i++) {
        float nv[3] = {0.0f, 0.0f, 0.0f};
        const int k = inst[i];
        if (k >= 0 && k < num_inst) {
            const orc_blas* B = &blas[blas_of[k]];
            if (prim[i] >= 0 && prim[i] < B->num_tris)
                world_normal(B->tris + 9 * (size_t)leaf_of[blas_of[k]][prim[i]], w2o + 12 * (size_t)k, nv);
        }
        for (int c = 0; c < 3; c++) out[(size_t)c * n + i] = nv[c];
    }
    for (int b = 0; b < nb; b++) free(leaf_of[b]);
    free(leaf_of);
}

void orc_path_trace_2l(const void* tlas_nodes, int tlas_num_nodes, const int* tlas_prim, const orc_blas* blas,
                       const int* blas_of, const float* w2o, const float* albedo, const float* cam,
                       const orc_pt_params* p, float* rgb, float* aov, long long* counts, int threads) {
    const int W = p->width, H = p->height, N = W * H;
    const float INV_PI = 0.318309886183790671538f;
    float sun[3];
    sun_unit(p, sun);
    int nth = orc_num_threads(threads);
    int cap = 3 * (tlas_num_nodes + 4096) + 16;
    long long tot[2 * ORC_PT_MAX_DEPTH] = {0};
#pragma omp parallel num_threads(nth)
    {
        sentry* stk = (sentry*)malloc(sizeof(sentry) * (size_t)cap * 4);
        long long loc[2 * ORC_PT_MAX_DEPTH] = {0};
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < N; i++) {
            const int py = i / W, px = i - py * W;
            float acc[3] = {0.f, 0.f, 0.f};
            for (int s = 0; s < p->spp; s++) {
                pcg32 rng;
                pcg_seed(&rng, ((uint64_t)p->seed << 32) ^ (uint64_t)(uint32_t)s, (uint64_t)i);
                float jx = pcg_uniform(&rng), jy = pcg_uniform(&rng);
                float sx = (((float)px + jx) / (float)W) * 2.0f - 1.0f;
                float sy = (((float)py + jy) / (float)H) * 2.0f - 1.0f;
                float ax = sx * cam[12], ay = sy * cam[13];
                float d[3], o[3] = {cam[0], cam[1], cam[2]};
                for (int k = 0; k < 3; k++) d[k] = (cam[3 + k] + ax * cam[6 + k]) + ay * cam[9 + k];
                float il = 1.0f / sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
                for (int k = 0; k < 3; k++) d[k] = d[k] * il;
                float thr[3] = {1.f, 1.f, 1.f};
                for (int b = 0; b < p->max_depth; b++) {
                    loc[b]++;
                    xray r;
                    make_xray(&r, o, d, 0.0f, MISS_T);
                    float t, u, v; int inst, id, leaf;
                    two_level_one((const uint8_t*)tlas_nodes, tlas_prim, blas, blas_of, w2o, &r, 0, stk, &t, &inst,
                                  &id, &u, &v, NULL, NULL, NULL, &leaf);
                    const int aovw = aov && b == 0 && s == 0;
                    if (id < 0) {
                        float e = fmaxf(d[1], 0.0f), c[3];
                        for (int k = 0; k < 3; k++)
                            c[k] = thr[k] * (p->sky_horizon[k] + (p->sky_zenith[k] - p->sky_horizon[k]) * e);
                        clamp3f(c, p->clamp_max);
                        for (int k = 0; k < 3; k++) acc[k] += c[k];
                        if (aovw) for (int k = 0; k < 7; k++) aov[(size_t)k * N + i] = 0.0f;
                        break;
                    }
                    const orc_blas* B = &blas[blas_of[inst]];
                    float n[3];
                    world_normal(B->tris + 9 * (size_t)leaf, w2o + 12 * (size_t)inst, n);
                    if ((n[0] * d[0] + n[1] * d[1]) + n[2] * d[2] > 0.0f) for (int k = 0; k < 3; k++) n[k] = -n[k];
                    float h[3];
                    for (int k = 0; k < 3; k++) h[k] = o[k] + t * d[k];
                    float eps = 1e-4f * fmaxf(1.0f, fmaxf(fabsf(h[0]), fmaxf(fabsf(h[1]), fabsf(h[2]))));
                    float q[3];
                    for (int k = 0; k < 3; k++) q[k] = h[k] + n[k] * eps;
                    const float* alb = albedo + 3 * (size_t)inst;
                    if (aovw) {
                        for (int k = 0; k < 3; k++) {
                            aov[(size_t)k * N + i] = alb[k];
                            aov[(size_t)(3 + k) * N + i] = n[k];
                        }
                        aov[(size_t)6 * N + i] = t;
                    }
                    float cosL = (n[0] * sun[0] + n[1] * sun[1]) + n[2] * sun[2];
                    if (cosL > 0.0f) {
                        float kf = cosL * INV_PI, c[3];
                        for (int k = 0; k < 3; k++) c[k] = ((thr[k] * alb[k]) * p->sun_rad[k]) * kf;
                        clamp3f(c, p->clamp_max);
                        loc[ORC_PT_MAX_DEPTH + b]++;
                        xray sr;
                        make_xray(&sr, q, sun, 0.0f, MISS_T);
                        int occl = 0;
                        two_level_one((const uint8_t*)tlas_nodes, tlas_prim, blas, blas_of, w2o, &sr, 1, stk, NULL,
                                      NULL, NULL, NULL, NULL, &occl, NULL, NULL, NULL);
                        if (!occl) for (int k = 0; k < 3; k++) acc[k] += c[k];
                    }
                    if (b + 1 >= p->max_depth) break;
                    float nd[3];
                    cosine_dir_f(n, &rng, nd);
                    for (int k = 0; k < 3; k++) thr[k] = thr[k] * alb[k];
                    if (b + 1 >= p->rr_start) {
                        float pm = fminf(fmaxf(thr[0], fmaxf(thr[1], thr[2])), 1.0f);
                        float uu = pcg_uniform(&rng);
                        if (uu >= pm) break;
                        for (int k = 0; k < 3; k++) thr[k] = thr[k] / pm;
                    }
                    for (int k = 0; k < 3; k++) { o[k] = q[k]; d[k] = nd[k]; }
                }
            }
            if (p->spp > 1) {
                float inv = 1.0f / (float)p->spp;
                for (int k = 0; k < 3; k++) acc[k] = acc[k] * inv;
            }
            for (int k = 0; k < 3; k++) rgb[(size_t)k * N + i] = acc[k];
        }
#pragma omp critical
        for (int k = 0; k < 2 * ORC_PT_MAX_DEPTH; k++) tot[k] += loc[k];
        free(stk);
    }
    if (counts) {
        for (int k = 0; k < p->max_depth; k++) {
            counts[k] = tot[k];
            counts[p->max_depth + k] = tot[ORC_PT_MAX_DEPTH + k];
        }
    }
}

/* ---- FP64 path tracer: brute force over world-space triangles ---- */
static int mt_f64(const double* o, const double* d, const double* tri, double tmin, double tmax, double* tout) {
    double e1[3], e2[3], pv[3], tv[3], qv[3];
    for (int k = 0; k < 3; k++) { e1[k] = tri[3 + k] - tri[k]; e2[k] = tri[6 + k] - tri[k]; }
    pv[0] = d[1] * e2[2] - d[2] * e2[1]; pv[1] = d[2] * e2[0] - d[0] * e2[2]; pv[2] = d[0] * e2[1] - d[1] * e2[0];
    double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0.0) return 0;
    double id = 1.0 / det;
    for (int k = 0; k < 3; k++) tv[k] = o[k] - tri[k];
    double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * id;
    if (u < 0.0 || u > 1.0) return 0;
    qv[0] = tv[1] * e1[2] - tv[2] * e1[1]; qv[1] = tv[2] * e1[0] - tv[0] * e1[2]; qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
    double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * id;
    if (v < 0.0 || u + v > 1.0) return 0;
    double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * id;
    if (!(t > tmin) || !(t < tmax)) return 0;
    *tout = t;
    return 1;
}

void orc_path_trace_f64(const double* wtris, const int* tri_inst, int m, const float* albedo, const float* cam,
                        const orc_pt_params* p, float* rgb, int threads) {
    const int W = p->width, H = p->height, N = W * H;
    double sun[3];
    {
        double sl = sqrt((double)p->sun_dir[0] * p->sun_dir[0] + (double)p->sun_dir[1] * p->sun_dir[1] +
                         (double)p->sun_dir[2] * p->sun_dir[2]);
        for (int k = 0; k < 3; k++) sun[k] = p->sun_dir[k] / sl;
    }
    int nth = orc_num_threads(threads);
#pragma omp parallel for num_threads(nth) schedule(dynamic, 16)
    for (int i = 0; i < N; i++) {
        const int py = i / W, px = i - py * W;
        double acc[3] = {0, 0, 0};
        for (int s = 0; s < p->spp; s++) {
            pcg32 rng;
            pcg_seed(&rng, ((uint64_t)p->seed << 32) ^ (uint64_t)(uint32_t)s, (uint64_t)i);
            double jx = pcg_uniform(&rng), jy = pcg_uniform(&rng);
            double sx = ((px + jx) / W) * 2.0 - 1.0, sy = ((py + jy) / H) * 2.0 - 1.0;
            double ax = sx * cam[12], ay = sy * cam[13];
            double o[3] = {cam[0], cam[1], cam[2]}, d[3];
            for (int k = 0; k < 3; k++) d[k] = cam[3 + k] + ax * cam[6 + k] + ay * cam[9 + k];
            double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int k = 0; k < 3; k++) d[k] /= dl;
            double thr[3] = {1, 1, 1};
            for (int b = 0; b < p->max_depth; b++) {
                double best = 1e300; int bi = -1;
                for (int j = 0; j < m; j++) {
                    double t;
                    if (mt_f64(o, d, wtris + 9 * (size_t)j, 0.0, best, &t)) { best = t; bi = j; }
                }
                if (bi < 0) {
                    double e = d[1] > 0 ? d[1] : 0;
                    double c[3];
                    for (int k = 0; k < 3; k++)
                        c[k] = thr[k] * (p->sky_horizon[k] + (p->sky_zenith[k] - p->sky_horizon[k]) * e);
                    double mx = fmax(c[0], fmax(c[1], c[2]));
                    if (p->clamp_max > 0 && mx > p->clamp_max) for (int k = 0; k < 3; k++) c[k] *= p->clamp_max / mx;
                    for (int k = 0; k < 3; k++) acc[k] += c[k];
                    break;
                }
                const double* tri = wtris + 9 * (size_t)bi;
                double e1[3], e2[3], n[3];
                for (int k = 0; k < 3; k++) { e1[k] = tri[3 + k] - tri[k]; e2[k] = tri[6 + k] - tri[k]; }
                n[0] = e1[1] * e2[2] - e1[2] * e2[1]; n[1] = e1[2] * e2[0] - e1[0] * e2[2]; n[2] = e1[0] * e2[1] - e1[1] * e2[0];
                double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                for (int k = 0; k < 3; k++) n[k] /= nl;
                if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] > 0) for (int k = 0; k < 3; k++) n[k] = -n[k];
                double h[3], q[3];
                for (int k = 0; k < 3; k++) h[k] = o[k] + best * d[k];
                double eps = 1e-4 * fmax(1.0, fmax(fabs(h[0]), fmax(fabs(h[1]), fabs(h[2]))));
                for (int k = 0; k < 3; k++) q[k] = h[k] + n[k] * eps;
                const float* alb = albedo + 3 * (size_t)tri_inst[bi];
                double cosL = n[0] * sun[0] + n[1] * sun[1] + n[2] * sun[2];
                if (cosL > 0) {
                    double c[3];
                    for (int k = 0; k < 3; k++) c[k] = thr[k] * alb[k] * p->sun_rad[k] * cosL / 3.14159265358979323846;
                    double mx = fmax(c[0], fmax(c[1], c[2]));
                    if (p->clamp_max > 0 && mx > p->clamp_max) for (int k = 0; k < 3; k++) c[k] *= p->clamp_max / mx;
                    int occl = 0;
                    for (int j = 0; j < m && !occl; j++) {
                        double t;
                        occl = mt_f64(q, sun, wtris + 9 * (size_t)j, 0.0, 1e300, &t);
                    }
                    if (!occl) for (int k = 0; k < 3; k++) acc[k] += c[k];
                }
                if (b + 1 >= p->max_depth) break;
                /* Malley + Duff basis in double, same random numbers */
                double x = 0, y = 0, ss = 0;
                int ok = 0;
                for (int k = 0; k < 64 && !ok; k++) {
                    x = 2.0 * pcg_uniform(&rng) - 1.0;
                    y = 2.0 * pcg_uniform(&rng) - 1.0;
                    ss = x * x + y * y;
                    ok = ss < 1.0;
                }
                if (!ok) { x = 0; y = 0; ss = 0; }
                double z = sqrt(1.0 - ss), sg = n[2] >= 0 ? 1.0 : -1.0;
                double ca = -1.0 / (sg + n[2]), cb = n[0] * n[1] * ca;
                double t1[3] = {1.0 + sg * n[0] * n[0] * ca, sg * cb, -sg * n[0]};
                double t2[3] = {cb, sg + n[1] * n[1] * ca, -n[1]};
                double nd[3];
                for (int k = 0; k < 3; k++) nd[k] = x * t1[k] + y * t2[k] + z * n[k];
                double ndl = sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
                for (int k = 0; k < 3; k++) { nd[k] /= ndl; thr[k] *= alb[k]; }
                if (b + 1 >= p->rr_start) {
                    double pm = fmin(fmax(thr[0], fmax(thr[1], thr[2])), 1.0);
                    double uu = pcg_uniform(&rng);
                    if (uu >= pm) break;
                    for (int k = 0; k < 3; k++) thr[k] /= pm;
                }
                for (int k = 0; k < 3; k++) { o[k] = q[k]; d[k] = nd[k]; }
            }
        }
        for (int k = 0; k < 3; k++) rgb[(size_t)k * N + i] = (float)(acc[k] / p->spp);
    }
}

/* PCG32 outputs for a known-answer check of the generator. */
void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    pcg32 r;
    pcg_seed(&r, initstate, initseq);
    for (int i = 0; i < n; i++) out[i] = pcg_next(&r);
}

/* ---------------- The layer's GLSL compute tracers ---------------- */

/* Camera ray of pixel (px, py) as the shaders form it
 * (bvh4_trace.comp:28-36, bvh2_stackless.comp:24-32). */
static void glsl_ray(const float* pc, int W, int H, int px, int py, float ro[3], float rd[3], float invD[3],
                     float ood[3]) {
    float u = ((float)px + 0.5f) / (float)W * 2.0f - 1.0f;
    float v = ((float)py + 0.5f) / (float)H * 2.0f - 1.0f;
    u *= (float)W / (float)H * pc[12];
    v *= pc[12];
    float d[3];
    for (int k = 0; k < 3; k++) d[k] = (pc[3 + k] + u * pc[6 + k]) + v * pc[9 + k];
    const float il = 1.0f / sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    for (int k = 0; k < 3; k++) {
        ro[k] = pc[k];
        rd[k] = d[k] * il;
        /* mix(rd, sign(rd) * 1e-8, |rd| < 1e-8); sign(0) = 0 */
        const float sg = rd[k] > 0.0f ? 1.0f : (rd[k] < 0.0f ? -1.0f : 0.0f);
        const float c = fabsf(rd[k]) < 1e-8f ? sg * 1e-8f : rd[k];
        invD[k] = 1.0f / c;
        ood[k] = -ro[k] * invD[k];
    }
}

/* The shaders' triangle loop body (bvh4_trace.comp:46-63). */
static void glsl_tri(const float* packed, int ti, const float ro[3], const float rd[3], float* tHit, int* hitTri) {
    const float* p = packed + 12 * (size_t)ti;
    const float v0[3] = {p[0], p[1], p[2]};
    const float e1[3] = {p[3] - v0[0], p[4] - v0[1], p[5] - v0[2]};
    const float e2[3] = {p[6] - v0[0], p[7] - v0[1], p[8] - v0[2]};
    const float pp[3] = {rd[1] * e2[2] - rd[2] * e2[1], rd[2] * e2[0] - rd[0] * e2[2], rd[0] * e2[1] - rd[1] * e2[0]};
    const float det = (e1[0] * pp[0] + e1[1] * pp[1]) + e1[2] * pp[2];
    const float inv = 1.0f / det;
    const float tv[3] = {ro[0] - v0[0], ro[1] - v0[1], ro[2] - v0[2]};
    const float uu = inv * ((tv[0] * pp[0] + tv[1] * pp[1]) + tv[2] * pp[2]);
    if (uu < 0.0f || uu > 1.0f) return;
    const float qq[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const float vv = inv * ((rd[0] * qq[0] + rd[1] * qq[1]) + rd[2] * qq[2]);
    if (vv < 0.0f || uu + vv > 1.0f) return;
    const float tt = inv * ((e2[0] * qq[0] + e2[1] * qq[1]) + e2[2] * qq[2]);
    if (tt > 0.0f && tt < *tHit) { *tHit = tt; *hitTri = ti; }
}

void orc_glsl_rays(const float* pc, int W, int H, float* rays) {
    const size_t n = (size_t)W * H;
    for (int py = 0; py < H; py++)
        for (int px = 0; px < W; px++) {
            float ro[3], rd[3], invD[3], ood[3];
            glsl_ray(pc, W, H, px, py, ro, rd, invD, ood);
            const size_t i = (size_t)py * W + px;
            for (int k = 0; k < 3; k++) { rays[k * n + i] = ro[k]; rays[(3 + k) * n + i] = rd[k]; }
            rays[6 * n + i] = 0.0f; rays[7 * n + i] = MISS_T;
        }
}

void orc_glsl_bvh4_trace(const uint32_t* nodes, const float* packed, const float* pc, int W, int H,
                         float* tOut, int* triOut) {
    enum { STK = 13 };
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < W * H; i++) {
        const int px = i % W, py = i / W;
        float ro[3], rd[3], invD[3], ood[3];
        glsl_ray(pc, W, H, px, py, ro, rd, invD, ood);
        const int sX = invD[0] >= 0.0f, sY = invD[1] >= 0.0f, sZ = invD[2] >= 0.0f;
        float tHit = MISS_T;
        int hitTri = -1;
        int stk[STK], sp = 0;
        stk[sp++] = 0;
        while (sp > 0) {
            const int ni = stk[--sp];
            if (ni < 0) {
                const int enc = -(ni + 2), ts = enc >> 3, tc = (enc & 7) + 1;
                for (int t = 0; t < tc; t++) glsl_tri(packed, ts + t, ro, rd, &tHit, &hitTri);
                continue;
            }
            /* uvec4 n0..n3: x, y, z halves (mins in slots 0-3, maxes 4-7), then children */
            const uint32_t* w = nodes + 16 * (size_t)ni;
            const uint16_t* hx = (const uint16_t*)w;
            const uint16_t* hy = (const uint16_t*)(w + 4);
            const uint16_t* hz = (const uint16_t*)(w + 8);
            const int* ch = (const int*)(w + 12);
            for (int c = 0; c < 4; c++) {   /* ch.x, ch.y, ch.z, ch.w in order */
                if (ch[c] == -1) continue;
                const float lx = orc_h2f(hx[c]), hxv = orc_h2f(hx[4 + c]);
                const float ly = orc_h2f(hy[c]), hyv = orc_h2f(hy[4 + c]);
                const float lz = orc_h2f(hz[c]), hzv = orc_h2f(hz[4 + c]);
                const float nX = sX ? lx : hxv, fX = sX ? hxv : lx;
                const float nY = sY ? ly : hyv, fY = sY ? hyv : ly;
                const float nZ = sZ ? lz : hzv, fZ = sZ ? hzv : lz;
                const float tN = fmaxf(fmaxf(fmaf(nX, invD[0], ood[0]), fmaf(nY, invD[1], ood[1])), fmaf(nZ, invD[2], ood[2]));
                const float tF = fminf(fminf(fmaf(fX, invD[0], ood[0]), fmaf(fY, invD[1], ood[1])), fmaf(fZ, invD[2], ood[2]));
                if (tN <= tF && tF > 0.0f && tN < tHit && sp < STK) stk[sp++] = ch[c];
            }
        }
        tOut[i] = tHit;
        triOut[i] = hitTri;
    }
}

void orc_glsl_stackless_trace(const uint32_t* dfs, const float* packed, const float* pc, int W, int H,
                              float* tOut, int* triOut) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < W * H; i++) {
        const int px = i % W, py = i / W;
        float ro[3], rd[3], invD[3], ood[3];
        glsl_ray(pc, W, H, px, py, ro, rd, invD, ood);
        float tHit = MISS_T;
        int hitTri = -1;
        int ni = 0;
        while (ni >= 0) {
            const uint32_t* w = dfs + 8 * (size_t)ni;
            const int leaf = (int)w[6], skip = (int)w[7];
            if (leaf != 0) {
                const int enc = -(leaf + 2), ts = enc >> 3, tc = (enc & 7) + 1;
                for (int t = 0; t < tc; t++) glsl_tri(packed, ts + t, ro, rd, &tHit, &hitTri);
                ni = skip;
            } else {
                float b[6];
                memcpy(b, w, sizeof b);   /* bmin.xyz, bmax.xyz */
                const float t1x = fmaf(b[0], invD[0], ood[0]), t2x = fmaf(b[3], invD[0], ood[0]);
                const float t1y = fmaf(b[1], invD[1], ood[1]), t2y = fmaf(b[4], invD[1], ood[1]);
                const float t1z = fmaf(b[2], invD[2], ood[2]), t2z = fmaf(b[5], invD[2], ood[2]);
                const float tN = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
                const float tF = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
                ni = (tN <= tF && tF > 0.0f && tN < tHit) ? ni + 1 : skip;
            }
        }
        tOut[i] = tHit;
        triOut[i] = hitTri;
    }
}
