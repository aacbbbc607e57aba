// host_bvh.cpp — host-side parts of the BVH pipeline: error reporting, FP16
// conversion, BVH4 collapse and the wire-format exports the Vulkan layer
// uploads (stackless BVH2, packed triangles, LBVH TLAS).
#include "rt_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <algorithm>

namespace rt {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    fprintf(stderr, "[CudaBVH] error: %s\n", buf);
}
void clear_error() { g_last_error.clear(); }
const char* last_error() { return g_last_error.c_str(); }

bool verbose() {
    static int v = -1;
    if (v < 0) { const char* e = getenv("CUDABVH_VERBOSE"); v = (e && *e && *e != '0') ? 1 : 0; }
    return v == 1;
}

// ---------------------------------------------------------------------------
// IEEE binary16
// ---------------------------------------------------------------------------
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

uint16_t f2h_rne(float x) {
    uint32_t u = fbits(x), sign = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return uint16_t(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
    if (a >= 0x477ff000u) return uint16_t(sign | 0x7c00u);
    if (a <= 0x33000000u) return uint16_t(sign);
    int e = int(a >> 23) - 127;
    uint32_t m = (a & 0x7fffffu) | 0x800000u;
    if (e < -14) {
        int drop = -e - 1;                       // 13 + (-14 - e)
        uint32_t q = m >> drop, rem = m & ((1u << drop) - 1u), half = 1u << (drop - 1);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return uint16_t(sign | q);
    }
    uint32_t q = (uint32_t(e + 15) << 10) | ((m >> 13) & 0x3ffu), rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return uint16_t(sign | q);
}
float h2f(uint16_t h) {
    uint32_t sign = uint32_t(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0x1f) return bitsf(sign | 0x7f800000u | (m << 13));
    if (e == 0) { float f = float(m) * 0x1p-24f; return sign ? -f : f; }
    return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}
uint16_t f2h_down(float x) {
    uint16_t h = f2h_rne(x);
    if (h2f(h) > x) {                            // step one half toward -inf
        if ((h & 0x7fffu) == 0) h = 0x8001u;
        else if (h & 0x8000u) { if (h != 0xfc00u) h++; }
        else h--;
    }
    return h;
}
uint16_t f2h_up(float x) {
    uint16_t h = f2h_rne(x);
    if (h2f(h) < x) {                            // step one half toward +inf
        if ((h & 0x7fffu) == 0) h = 0x0001u;
        else if (h & 0x8000u) h--;
        else if (h != 0x7c00u) h++;
    }
    return h;
}

// ---------------------------------------------------------------------------
// BVH2 -> BVH4 collapse. A BVH4 node stands for each internal BVH2 node at
// even depth; its slots are the BVH2 grandchildren reached by opening every
// internal child (a leaf child stays a slot). Pre-order numbering matches a
// recursive collapse (backend :391-418).
// ---------------------------------------------------------------------------
void collapse_bvh4(HostBVH& b) {
    const std::vector<Node2>& n2 = b.nodes2;
    b.nodes4.assign(size_t(std::max<size_t>(n2.size(), 1)) * kNode4Bytes, 0);
    struct Task { int n2, parent, slot, depth; };
    std::vector<Task> st;
    st.push_back({0, -1, 0, 1});
    int cnt = 0, maxd = 0;
    while (!st.empty()) {
        Task t = st.back(); st.pop_back();
        int me = cnt++;
        maxd = std::max(maxd, t.depth);
        if (t.parent >= 0) {
            int32_t v = me;
            memcpy(&b.nodes4[size_t(t.parent) * kNode4Bytes + 48 + 4 * t.slot], &v, 4);
        }
        int gather[4], ng = 0;
        const Node2& nd = n2[t.n2];
        if (nd.triCount > 0) {
            gather[ng++] = t.n2;                 // single-leaf tree: one-slot root
        } else {
            for (int c : {nd.left, nd.right}) {
                if (n2[c].triCount > 0) gather[ng++] = c;
                else { gather[ng++] = n2[c].left; gather[ng++] = n2[c].right; }
            }
        }
        uint16_t hx[8], hy[8], hz[8];
        int32_t child[4];
        for (int i = 0; i < 4; i++) {
            if (i < ng) {
                const Box& bx = n2[gather[i]].box;
                hx[i] = f2h_down(bx.mn[0]); hx[4 + i] = f2h_up(bx.mx[0]);
                hy[i] = f2h_down(bx.mn[1]); hy[4 + i] = f2h_up(bx.mx[1]);
                hz[i] = f2h_down(bx.mn[2]); hz[4 + i] = f2h_up(bx.mx[2]);
                const Node2& c = n2[gather[i]];
                child[i] = c.triCount > 0 ? encode_leaf(c.triStart, c.triCount) : -1;
            } else {
                hx[i] = hy[i] = hz[i] = 0x7c00u;          // empty slot: (+inf, -inf)
                hx[4 + i] = hy[4 + i] = hz[4 + i] = 0xfc00u;
                child[i] = -1;
            }
        }
        uint8_t* out = &b.nodes4[size_t(me) * kNode4Bytes];
        memcpy(out, hx, 16); memcpy(out + 16, hy, 16); memcpy(out + 32, hz, 16); memcpy(out + 48, child, 16);
        for (int i = ng - 1; i >= 0; i--)
            if (n2[gather[i]].triCount == 0) st.push_back({gather[i], me, i, t.depth + 1});
    }
    b.nodes4.resize(size_t(cnt) * kNode4Bytes);
    b.numNodes4 = cnt;
    b.depth4 = maxd;
}

// Stackless export: the pre-order numbering is already the DFS order, and
// the skip target of node i is i + size(subtree(i)), or -1 past the end.
void export_dfs(const std::vector<Node2>& nodes, std::vector<uint32_t>& out) {
    const int nn = int(nodes.size());
    std::vector<int> sz(nn, 1);
    for (int i = nn - 1; i >= 0; i--)
        if (nodes[i].triCount == 0) sz[i] = 1 + sz[nodes[i].left] + sz[nodes[i].right];
    out.resize(size_t(nn) * 8);
    for (int i = 0; i < nn; i++) {
        const Node2& nd = nodes[i];
        uint32_t* p = &out[size_t(i) * 8];
        memcpy(p + 0, &nd.box.mn[0], 4); memcpy(p + 1, &nd.box.mn[1], 4); memcpy(p + 2, &nd.box.mn[2], 4);
        memcpy(p + 3, &nd.box.mx[0], 4); memcpy(p + 4, &nd.box.mx[1], 4); memcpy(p + 5, &nd.box.mx[2], 4);
        int32_t enc = nd.triCount > 0 ? encode_leaf(nd.triStart, nd.triCount) : 0;
        int32_t skip = (i + sz[i] < nn) ? i + sz[i] : -1;
        memcpy(p + 6, &enc, 4); memcpy(p + 7, &skip, 4);
    }
}

// ---------------------------------------------------------------------------
// LBVH (TLAS fast path)
// ---------------------------------------------------------------------------
static inline uint32_t spread10(uint32_t v) {     // bit i -> bit 3i
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
uint32_t morton3d(float x, float y, float z) {   // 10 bits per axis (backend :265-273)
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    return spread10(uint32_t(x)) * 4u + spread10(uint32_t(y)) * 2u + spread10(uint32_t(z));
}

int build_tlas_lbvh(const float* aabbs, int n, std::vector<uint32_t>& dfs, std::vector<int>& ordered) {
    float smin[3] = {1e30f, 1e30f, 1e30f}, smax[3] = {-1e30f, -1e30f, -1e30f};
    std::vector<float> cen(size_t(n) * 3);
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            float c = (aabbs[6 * size_t(i) + a] + aabbs[6 * size_t(i) + 3 + a]) * 0.5f;
            cen[3 * size_t(i) + a] = c;
            smin[a] = fminf(smin[a], c); smax[a] = fmaxf(smax[a], c);
        }
    float inv[3];
    for (int a = 0; a < 3; a++) inv[a] = (smax[a] - smin[a] > 1e-8f) ? 1.0f / (smax[a] - smin[a]) : 0.0f;
    // Stable sort by 30-bit Morton code; ties keep input order.
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; i++) {
        uint32_t c = morton3d((cen[3 * size_t(i)] - smin[0]) * inv[0], (cen[3 * size_t(i) + 1] - smin[1]) * inv[1],
                              (cen[3 * size_t(i) + 2] - smin[2]) * inv[2]);
        key[i] = (uint64_t(c) << 32) | uint32_t(i);
    }
    std::sort(key.begin(), key.end());
    ordered.resize(n);
    for (int i = 0; i < n; i++) ordered[i] = int(key[i] & 0xffffffffu);
    // Median split, one item per leaf, pre-order numbering; boxes bottom-up.
    std::vector<Node2> nodes;
    nodes.reserve(size_t(2) * n);
    struct T { int s, e, parent, side; bool post; };
    std::vector<T> st;
    st.push_back({0, n, -1, 0, false});
    while (!st.empty()) {
        T t = st.back(); st.pop_back();
        if (t.post) {
            Node2& p = nodes[t.parent];
            const Box &L = nodes[p.left].box, &R = nodes[p.right].box;
            for (int a = 0; a < 3; a++) { p.box.mn[a] = fminf(L.mn[a], R.mn[a]); p.box.mx[a] = fmaxf(L.mx[a], R.mx[a]); }
            continue;
        }
        int me = int(nodes.size());
        nodes.push_back(Node2{});
        if (t.parent >= 0) (t.side == 0 ? nodes[t.parent].left : nodes[t.parent].right) = me;
        Node2& nd = nodes[me];
        nd.left = nd.right = nd.triStart = nd.triCount = 0;
        for (int a = 0; a < 3; a++) { nd.box.mn[a] = 1e30f; nd.box.mx[a] = -1e30f; }
        int cnt = t.e - t.s;
        if (cnt <= 1) {
            nd.triStart = t.s; nd.triCount = cnt;
            const float* b = aabbs + 6 * size_t(ordered[t.s]);
            for (int a = 0; a < 3; a++) { nd.box.mn[a] = fminf(nd.box.mn[a], b[a]); nd.box.mx[a] = fmaxf(nd.box.mx[a], b[3 + a]); }
            continue;
        }
        int mid = t.s + cnt / 2;
        st.push_back({0, 0, me, 0, true});
        st.push_back({mid, t.e, me, 1, false});
        st.push_back({t.s, mid, me, 0, false});
    }
    export_dfs(nodes, dfs);
    return int(nodes.size());
}

}  // namespace rt

// ---------------------------------------------------------------------------
// Binned-SAH BVH2 on the host (backend :201-251 with the partition taken on
// the evaluated bin boundary and kept stable). Pre-order numbering.
// ---------------------------------------------------------------------------
namespace rt {

static inline Box merge(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; k++) { r.mn[k] = fminf(a.mn[k], b.mn[k]); r.mx[k] = fmaxf(a.mx[k], b.mx[k]); }
    return r;
}
static inline float area(const Box& b) {
    float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}
static constexpr int kBins = 16;
static inline int bin_index(float c, float cmin, float cmax) {
    int b = int((c - cmin) / (cmax - cmin) * float(kBins - 1));
    return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

void build_sah_host(const float* tris, int n, int maxLeaf, HostBVH& out) {
    std::vector<Box> pb(n);
    std::vector<float> cen(size_t(n) * 3);
    std::vector<int> idx(n), tmp(n);
    for (int i = 0; i < n; i++) {
        const float* t = tris + 9 * size_t(i);
        for (int a = 0; a < 3; a++) {
            pb[i].mn[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
            pb[i].mx[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
            cen[3 * size_t(i) + a] = (pb[i].mn[a] + pb[i].mx[a]) * .5f;
        }
        idx[i] = i;
    }
    out.numTris = n;
    out.maxLeaf = maxLeaf;
    out.nodes2.clear();
    out.nodes2.reserve(size_t(2) * n);
    out.ordered.clear();
    out.ordered.reserve(n);
    struct T { int s, e, parent, side; };
    std::vector<T> st;
    st.push_back({0, n, -1, 0});
    while (!st.empty()) {
        T t = st.back(); st.pop_back();
        const int s = t.s, e = t.e, cnt = e - s;
        const int id = int(out.nodes2.size());
        out.nodes2.push_back(Node2{});
        if (t.parent >= 0) (t.side == 0 ? out.nodes2[t.parent].left : out.nodes2[t.parent].right) = id;
        Node2 nd{};
        nd.box = pb[idx[s]];
        for (int i = s + 1; i < e; i++) nd.box = merge(nd.box, pb[idx[i]]);
        if (cnt <= maxLeaf) {
            nd.triStart = int(out.ordered.size());
            nd.triCount = cnt;
            for (int i = s; i < e; i++) out.ordered.push_back(idx[i]);
            out.nodes2[id] = nd;
            continue;
        }
        float bestCost = 1e30f, pA = area(nd.box);
        int bestAxis = -1, bestBin = -1;
        float bestMin = 0.f, bestMax = 0.f;
        for (int ax = 0; ax < 3; ax++) {
            float cmin = 1e30f, cmax = -1e30f;
            for (int i = s; i < e; i++) { float c = cen[3 * size_t(idx[i]) + ax]; cmin = fminf(cmin, c); cmax = fmaxf(cmax, c); }
            if (cmax - cmin < 1e-8f) continue;
            Box bb[kBins]; int bc[kBins] = {0};
            for (int b = 0; b < kBins; b++)
                for (int a = 0; a < 3; a++) { bb[b].mn[a] = 1e30f; bb[b].mx[a] = -1e30f; }
            for (int i = s; i < e; i++) {
                int b = bin_index(cen[3 * size_t(idx[i]) + ax], cmin, cmax);
                bb[b] = merge(bb[b], pb[idx[i]]); bc[b]++;
            }
            Box R[kBins]; int rc[kBins];
            R[kBins - 1] = bb[kBins - 1]; rc[kBins - 1] = bc[kBins - 1];
            for (int b = kBins - 2; b >= 0; b--) { R[b] = merge(R[b + 1], bb[b]); rc[b] = rc[b + 1] + bc[b]; }
            Box L = bb[0]; int lc = bc[0];
            for (int b = 0; b < kBins - 1; b++) {
                if (b > 0) { L = merge(L, bb[b]); lc += bc[b]; }
                if (lc == 0 || rc[b + 1] == 0) continue;
                float cost = float(lc) * area(L) / pA + float(rc[b + 1]) * area(R[b + 1]) / pA + 1.f;
                if (cost < bestCost) { bestCost = cost; bestAxis = ax; bestBin = b; bestMin = cmin; bestMax = cmax; }
            }
        }
        int mid;
        if (bestAxis >= 0) {
            int nl = 0, nr = 0;
            for (int i = s; i < e; i++) {
                int p = idx[i];
                if (bin_index(cen[3 * size_t(p) + bestAxis], bestMin, bestMax) <= bestBin) idx[s + nl++] = p;
                else tmp[nr++] = p;
            }
            std::copy(tmp.begin(), tmp.begin() + nr, idx.begin() + s + nl);
            mid = s + nl;
        } else {
            mid = s + cnt / 2;
        }
        out.nodes2[id] = nd;
        st.push_back({mid, e, id, 1});
        st.push_back({s, mid, id, 0});
    }
    collapse_bvh4(out);
}

}  // namespace rt
