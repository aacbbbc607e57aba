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

}  // namespace rt

