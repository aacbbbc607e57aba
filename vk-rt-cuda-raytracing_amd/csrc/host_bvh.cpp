// host_bvh.cpp — host-side parts of the BVH pipeline: error reporting, FP16
// conversion and the wire-format exports the Vulkan layer uploads (stackless
// BVH2, LBVH TLAS).
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

