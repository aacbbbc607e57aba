// dn.hip — learned denoiser: a 4-level convolutional encoder-decoder (U-Net)
// on NHWC BF16 activations, trained on the GPU by its own kernels.
//
// Network (the reference's U-Net shape, train_denoiser.py:13-54 and
// train_denoiser_v2.py:60-110, at 32/64/128/256 channels): per level two 3x3
// convolutions with ReLU, 2x2 max pooling down, parameter-free 2x nearest
// upsampling up, skip concatenation, and a 1x1 head whose 3 outputs are added
// to the log-radiance input channels (a residual in log space, as the
// reference's residual head). Input: 16 channels per pixel — log(1+rgb),
// albedo, normal, log(1+depth)/4, zero padding.
//
// Kernels:
//   * k_conv3r: every 3x3 convolution (forward and input gradients), an
//     implicit GEMM on MFMA 16x16x32 BF16 with FP32 accumulation over an LDS
//     input halo with row reuse; bias + ReLU fused; writes into a channel
//     slice of a wider buffer, so skip concatenation costs nothing; optional
//     fusions: 2x2 max pooling of the output, 2x nearest upsampling of part
//     of the input, and the 1x1 head with the residual and expm1.
//   * k_conv01: the first two 3x3 convolutions in one kernel, the first
//     one's output kept in LDS (recomputed on the halo).
//   * k_conv: the generic implicit GEMM, kept for the 1x1 head in training
//     and its input gradient.
//   * k_wgrad: weight gradients, M = output channels, N = input channels,
//     K = pixels; 32-pixel tiles staged through LDS transposed, split over
//     pixel ranges into partials reduced in a fixed order (deterministic).
//   * upsampling (training), pooling / upsampling backward, ReLU-mask, loss,
//     Adam and weight layout kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/cudabvh.h"
#include "rt_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace dn {

constexpr int kBlock = 256;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                              // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Network input channels 0-2: log(1 + max(rgb, 0)), rounded to BF16 (also
// the residual that the head's output is added to).
__device__ __forceinline__ uint16_t log_rgb_bf(float v) { return f2bf(log1pf(fmaxf(v, 0.0f))); }

// One 16-channel BF16 input pixel (two 16-byte halves) from its plane values
// v = rgb[3], albedo[3], normal[3], depth: log(1 + rgb), albedo, normal,
// log(1 + depth) / 4, zeros.
__device__ __forceinline__ void pack_pixel(const float (&v)[10], uint4& o0, uint4& o1) {
    uint32_t h[10];
#pragma unroll
    for (int c = 0; c < 3; c++) h[c] = log_rgb_bf(v[c]);
#pragma unroll
    for (int c = 3; c < 9; c++) h[c] = f2bf(v[c]);
    h[9] = f2bf(0.25f * log1pf(fmaxf(v[9], 0.0f)));
    o0 = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
    o1 = make_uint4(h[8] | (h[9] << 16), 0, 0, 0);
}
__device__ __forceinline__ void load_pixel(const float* rgb, const float* alb, const float* nrm, const float* dep,
                                           size_t P, size_t p, float (&v)[10]) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v[c] = rgb[c * P + p];
        v[3 + c] = alb[c * P + p];
        v[6 + c] = nrm[c * P + p];
    }
    v[9] = dep[p];
}

// ---------------------------------------------------------------- conv ----
struct ConvArgs {
    const uint16_t* x; int ldx; int cin;     // input pixels (channel stride ldx), channels per tap
    const uint16_t* w;                       // [rows >= gridDim.y*16*WN][taps*cin]
    const uint16_t* wfrag;                   // the same weights in MFMA fragment order (k_conv3r)
    const float* bias;                       // nullable
    uint16_t* y; int ldy;                    // BF16 output slice (nullable)
    float* yf; int ldyf;                     // FP32 output (nullable)
    int n;                                   // output channels written
    int B, H, W, relu;
    // k_conv3r only: input channels [0, cup) are read 2x nearest-upsampled
    // from xu (B x H/2 x W/2, channel stride ldxu) instead of from x; and an
    // optional 2x2 max-pooled copy of the BF16 output goes to yp (stride ldyp)
    const uint16_t* xu; int ldxu; int cup;
    uint16_t* yp; int ldyp;
    // k_conv3r with 32 output channels per workgroup only: the 1x1 head
    // (3 x 32 BF16 weights hw, bias hbias) applied to the BF16-rounded output,
    // written as FP32 head values (head, P x 3; nullable) and/or as denoised
    // planes rgbOut[c][p] = max(expm1(head + residual), 0) (nullable), the
    // residual being log(1 + rgb) rounded to BF16 (input channel c) from the
    // noisy radiance planes rgbIn[c][p]
    const uint16_t* hw; const float* hbias; float* head; const float* rgbIn; float* rgbOut;
    // k_conv3r only: x stores input channels [xc0, cin) (channel c at c - xc0)
    int xc0;
};

// Block: 4 waves x (16*WM) pixels; each wave (16*WM) pixels x (16*WN)
// channels (WM x WN MFMA tiles). A operand: the 8 (or 4) consecutive channels
// of one neighbour pixel per lane; B operand: 8 (or 4) consecutive K of one
// weight row per lane.
template <int TAPS, int KS, int WN, int WM>
__global__ __launch_bounds__(kBlock) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = a.B * a.H * a.W;
    const int m0 = blockIdx.x * (64 * WM) + wave * (16 * WM);
    const int n0 = blockIdx.y * (16 * WN);
    const int row = lane & 15, kg = lane >> 4;
    int pb[WM], py[WM], px[WM];
    bool pv[WM];
    #pragma unroll
    for (int r = 0; r < WM; r++) {
        const int p = m0 + r * 16 + row;
        pv[r] = p < P;
        const int pp = pv[r] ? p : 0;
        px[r] = pp % a.W;
        const int q = pp / a.W;
        py[r] = q % a.H;
        pb[r] = q / a.H;
    }
    f32x4 acc[WM][WN];
    #pragma unroll
    for (int r = 0; r < WM; r++)
        #pragma unroll
        for (int j = 0; j < WN; j++) acc[r][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = TAPS * a.cin;
    for (int t = 0; t < TAPS; t++) {
        const int dy = TAPS == 9 ? t / 3 - 1 : 0, dx = TAPS == 9 ? t % 3 - 1 : 0;
        const uint16_t* src[WM];
        #pragma unroll
        for (int r = 0; r < WM; r++) {
            const int yy = py[r] + dy, xx = px[r] + dx;
            const bool ok = pv[r] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
            src[r] = ok ? a.x + ((size_t)(pb[r] * a.H + yy) * a.W + xx) * a.ldx : nullptr;
        }
        for (int c0 = 0; c0 < a.cin; c0 += KS) {
            if constexpr (KS == 32) {
                bf16x8 av[WM], bv[WN];
                #pragma unroll
                for (int r = 0; r < WM; r++) {
                    uint4 v = make_uint4(0, 0, 0, 0);
                    if (src[r]) v = *(const uint4*)(src[r] + c0 + 8 * kg);
                    av[r] = __builtin_bit_cast(bf16x8, v);
                }
                #pragma unroll
                for (int j = 0; j < WN; j++) {
                    const uint4 v = *(const uint4*)(a.w + (size_t)(n0 + j * 16 + row) * K + t * a.cin + c0 + 8 * kg);
                    bv[j] = __builtin_bit_cast(bf16x8, v);
                }
                #pragma unroll
                for (int r = 0; r < WM; r++)
                    #pragma unroll
                    for (int j = 0; j < WN; j++)
                        acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r], bv[j], acc[r][j], 0, 0, 0);
            } else {
                s16x4 av[WM], bv[WN];
                #pragma unroll
                for (int r = 0; r < WM; r++) {
                    uint2 v = make_uint2(0, 0);
                    if (src[r]) v = *(const uint2*)(src[r] + c0 + 4 * kg);
                    av[r] = __builtin_bit_cast(s16x4, v);
                }
                #pragma unroll
                for (int j = 0; j < WN; j++) {
                    const uint2 v = *(const uint2*)(a.w + (size_t)(n0 + j * 16 + row) * K + t * a.cin + c0 + 4 * kg);
                    bv[j] = __builtin_bit_cast(s16x4, v);
                }
                #pragma unroll
                for (int r = 0; r < WM; r++)
                    #pragma unroll
                    for (int j = 0; j < WN; j++)
                        acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[r], bv[j], acc[r][j], 0, 0, 0);
            }
        }
    }
    // D[i][j]: lane j + 16*(i/4), item i%4
    #pragma unroll
    for (int r = 0; r < WM; r++)
        #pragma unroll
        for (int j = 0; j < WN; j++) {
            const int col = n0 + j * 16 + row;
            if (col >= a.n) continue;
            const float b = a.bias ? a.bias[col] : 0.0f;
            #pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = m0 + r * 16 + 4 * kg + i;
                if (p >= P) continue;
                float v = acc[r][j][i] + b;
                if (a.relu) v = fmaxf(v, 0.0f);
                if (a.y) a.y[(size_t)p * a.ldy + col] = f2bf(v);
                if (a.yf) a.yf[(size_t)p * a.ldyf + col] = v;
            }
        }
}


// 3x3 convolution with row reuse (cin a multiple of 32). A workgroup computes
// an R-row x 16*WGX-pixel block of one image for 32*WGC output channels; wave
// (wx, wc) owns the R x 16 pixels of column block wx for the 32 channels of
// block wc (R x 2 MFMA 16x16x32 tiles). Per 32-channel input chunk the
// workgroup stages the (R + 2) x (16*WGX + 2)-pixel input halo in LDS. A wave
// then walks the halo rows once per horizontal tap dx: each 16-pixel row
// fragment it reads feeds the MFMAs of every output row that uses it (up to
// three: vertical taps dy = 0, 1, 2), so a wave reads 3 (R + 2) A fragments
// per chunk for 9 R x 2 MFMAs instead of 9 R (two per MFMA pair without row
// reuse). The B fragments (the three vertical taps of one dx, two 16-channel
// tiles) come straight from memory into registers, from a copy of the
// weights in fragment order (one 1 KB run per wave load: twice as fast as
// 16 rows of 64 bytes), the next dx's fragments loading while the current
// dx's MFMAs run; the next chunk's halo loads into registers likewise. With
// BL (every wave of the workgroup on the same 32 channels) the chunk's
// fragments are copied to LDS once per workgroup instead.
//
// LDS layout: one 64-byte row (32 channels) per halo pixel hp, its four
// 16-byte quarters q stored at slots q ^ (hp & 2). Against 80-byte padded
// rows this cut the LDS bank-conflict cycles from 50 % to 27-45 % of the
// LDS-active cycles (MI355X counters) and the LDS footprint by a fifth.
//
// The weights are the MFMA's A operand and the pixels its B operand, so the
// accumulators are channels x pixels: a lane holds 4 consecutive channels of
// one pixel, which it stores as one 8-byte write.
// The value of lane l ^ 1 (a DPP quad permutation [1, 0, 3, 2]: no LDS
// traffic, unlike a general lane shuffle).
__device__ __forceinline__ float swap_pair(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// Scheduling groups of one halo row walk (R output rows): the MFMAs of row
// hy (2 per output row that uses it: 2, 4, 6, ..., 6, 4, 2), then the LDS
// read of row hy + 2.
template <int R, int PD, int HYI>
__device__ __forceinline__ void halo_row_order() {
    if constexpr (HYI < R + 2) {
        constexpr int first = HYI < 2 ? HYI : 2, last = HYI - (R - 1) > 0 ? HYI - (R - 1) : 0;
        __builtin_amdgcn_sched_group_barrier(0x008, 2 * (first - last + 1), 0);
        if constexpr (HYI + PD < R + 2) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        halo_row_order<R, PD, HYI + 1>();
    }
}

template <int R, int WGX, int WGC, int OCC, int PD, bool BL, int CS = 32, bool HD = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(OCC))) void k_conv3r(ConvArgs a) {
    static_assert(WGX * WGC == kBlock / 64, "four waves");
    static_assert(!BL || WGC == 1, "weights are shared through LDS only when every wave has the same channels");
    static_assert(CS == 32 || (CS == 16 && !BL), "32-channel chunks, or one 16-channel chunk");
    static_assert(!HD || WGC == 1, "the head needs all 32 channels of a pixel in one wave");
    constexpr int TX = 16 * WGX, HX = TX + 2, HY = R + 2, HP = HX * HY;
    constexpr int QP = CS / 8;                                 // 16-byte pieces per halo pixel
    constexpr int NPC = HP * QP;                               // 16-byte pieces per chunk
    constexpr int NA = (NPC + kBlock - 1) / kBlock;
    constexpr int FPB = CS == 32 ? 9 : 6;                      // B fragments per 16-channel block and chunk
    __shared__ uint4 hs[HP * QP];
    __shared__ uint4 hb[BL ? 18 * 64 : 1];          // BL: the chunk's 18 B fragments (2 tiles x 9 taps)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wx = wave % WGX, wc = wave / WGX;
    const int row = lane & 15, kg = lane >> 4;
    const int tilesX = (a.W + TX - 1) / TX, tilesY = (a.H + R - 1) / R;
    const int tx = blockIdx.x % tilesX, tq = blockIdx.x / tilesX;
    const int ty = tq % tilesY, bb = tq / tilesY;
    const int x0 = tx * TX, y0 = ty * R;
    const int n0 = blockIdx.y * (32 * WGC) + wc * 32;
    // B fragments: a.wfrag is [16-channel block][chunk][fragment][lane][8]
    // (fragment = tap for 32-channel chunks, (dy, tap pair) for the 16-channel
    // input), so a wave's fragment is 1 KB of consecutive memory
    const int nch = a.cin / CS;
    const uint16_t* wf0 = a.wfrag + (size_t)(n0 / 16) * nch * FPB * 512 + lane * 8;
    const uint16_t* wf1 = wf0 + (size_t)nch * FPB * 512;

    // halo piece k of this thread: pixel (y0 - 1 + hy, x0 - 1 + hx), quarter
    // e % QP of the chunk; zero outside the image. Channels below cup come
    // from the half-resolution xu at (y / 2, x / 2).
    uint4 ra[NA];
    auto fetch = [&](int c0) {
        const bool up = c0 < a.cup;
#pragma unroll
        for (int k = 0; k < NA; k++) {
            const int e = tid + k * kBlock, hp = e / QP;
            const int hy = hp / HX, hx = hp - hy * HX;
            const int y = y0 - 1 + hy, x = x0 - 1 + hx;
            ra[k] = make_uint4(0, 0, 0, 0);
            if (e < NPC && y >= 0 && y < a.H && x >= 0 && x < a.W) {
                const uint16_t* src =
                    up ? a.xu + (size_t)((bb * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1)) * a.ldxu
                       : a.x + (size_t)((bb * a.H + y) * a.W + x) * a.ldx;
                ra[k] = *(const uint4*)(src + (up ? c0 : c0 - a.xc0) + 8 * (e % QP));
            }
        }
    };
    auto bload = [&](bf16x8 (&b)[3][2], int c0, int dx) {
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
            if constexpr (BL) {
                b[dy][0] = __builtin_bit_cast(bf16x8, hb[(dy * 3 + dx) * 64 + lane]);
                b[dy][1] = __builtin_bit_cast(bf16x8, hb[(9 + dy * 3 + dx) * 64 + lane]);
            } else {
                const int o = ((c0 / CS) * FPB + dy * (FPB / 3) + dx) * 512;
                b[dy][0] = __builtin_bit_cast(bf16x8, *(const uint4*)(wf0 + o));
                b[dy][1] = __builtin_bit_cast(bf16x8, *(const uint4*)(wf1 + o));
            }
        }
    };
    // BL: the chunk's 18 fragments go straight from memory to LDS (1 KB per
    // wave instruction, fragments wave, wave + 4, ...)
    auto bstage = [&](int c0) {
#pragma unroll
        for (int f = wave; f < 18; f += 4) {
            const uint16_t* src = (f < 9 ? wf0 : wf1) + ((c0 >> 5) * 9 + (f % 9)) * 512;
            __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)&hb[f * 64], 16,
                                             0, 0);
        }
    };
    f32x4 acc[R][2];

    // One horizontal tap: the MFMAs of dx with the fragments in bc, after
    // issuing the loads of the next step's fragments into bn (kept ahead of
    // the MFMAs by the scheduling barrier, so their latency is hidden).
    auto dxstep = [&](const bf16x8 (&bc)[3][2], bf16x8 (&bn)[3][2], int dx, bool next, int nc0, int ndx) {
        if (next) bload(bn, nc0, ndx);
        __builtin_amdgcn_sched_barrier(0);
        // 32-channel chunks: lane (row, kg) reads quarter kg of pixel row + dx.
        // The 16-channel input pairs horizontal taps in one K = 32 step:
        // lanes kg = 0, 1 read pixel row + 2 dx (tap 2 dx), kg = 2, 3 pixel
        // row + 2 dx + 1 (tap 2 dx + 1). For dx = 1 that is tap 3, outside
        // the filter: those lanes read tap 2's pixel (in bounds) and use zero
        // (zero weights alone would turn an infinite input into NaN). Rows of
        // 32 bytes need no swizzle.
        const bool pad = CS == 16 && dx == 1 && kg >= 2;
        const int hp0 = CS == 32 ? wx * 16 + row + dx : wx * 16 + row + 2 * dx + (pad ? 0 : kg >> 1);
        const int s0 = CS == 32 ? hp0 * 4 + (kg ^ (hp0 & 2)) : hp0 * 2 + (kg & 1);
        const int s1 = CS == 32 ? hp0 * 4 + (kg ^ (hp0 & 2) ^ 2) : s0;
#pragma unroll
        for (int hy = 0; hy < HY; hy++) {
            // (hp0 + hy * HX) & 2 flips with bit 1 of hy * HX (HX is even)
            const int slot = hy * HX * QP + (((hy * HX) & 2) ? s1 : s0);
            const bf16x8 av = __builtin_bit_cast(bf16x8, pad ? make_uint4(0, 0, 0, 0) : hs[slot]);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const int ro = hy - dy;
                if (ro < 0 || ro >= R) continue;
                acc[ro][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bc[dy][0], av, acc[ro][0], 0, 0, 0);
                acc[ro][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bc[dy][1], av, acc[ro][1], 0, 0, 0);
            }
        }
        // issue order: the A fragment of row hy + 2 is read while row hy's
        // MFMAs run (two reads in flight keep the register count down)
        __builtin_amdgcn_sched_group_barrier(0x100, PD, 0);
        halo_row_order<R, PD, 0>();
    };
    // One channel chunk: stage its halo, start the next chunk's halo loads,
    // run the taps. The fragments alternate between two register sets (x: dx
    // 0 and 2, y: dx 1 and the next chunk's dx 0), so no copy waits for a
    // load in flight; chunks alternate the roles of the two sets.
    auto chunk = [&](int c0, bf16x8 (&x)[3][2], bf16x8 (&y)[3][2]) {
        __syncthreads();               // the previous chunk's halo (and weights) are no longer read
        if constexpr (BL) bstage(c0);
#pragma unroll
        for (int k = 0; k < NA; k++) {
            const int e = tid + k * kBlock, hp = e >> 2;
            if (e < NPC) hs[CS == 32 ? hp * 4 + ((e & 3) ^ (hp & 2)) : e] = ra[k];
        }
        if constexpr (BL) __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wave's weight copies landed
        __syncthreads();
        const bool more = c0 + CS < a.cin;
        if (more) fetch(c0 + CS);
        if constexpr (BL) bload(x, c0, 0);
        if constexpr (CS == 32) {
            dxstep(x, y, 0, true, c0, 1);
            dxstep(y, x, 1, true, c0, 2);
            dxstep(x, y, 2, more && !BL, c0 + 32, 0);
        } else {
            dxstep(x, y, 0, true, c0, 1);     // a single 16-channel chunk
            dxstep(y, x, 1, false, 0, 0);
        }
    };
    // Epilogue of one R-row tile.
    auto store_tile = [&]() {
        // The weights are the MFMA's A operand and the pixels its B operand, so
        // D is channels x pixels: lane (pixel l % 16, k = l / 16) holds channels
        // 4k .. 4k + 3 of its pixel, stored as one 8-byte write.
        // With yp, the 2x2 maximum of each output row pair (r, r + 1; y0 and R
        // are even) and lane pair (l, l ^ 1: pixels 2m, 2m + 1) is written too:
        // rounding to BF16 is monotonic, so this equals pooling the BF16 output.
        const int x = x0 + wx * 16 + row;
        float pool[2][4];
        // HD: the head as one MFMA per output row. K slots 8k .. 8k + 7 of lane
        // group k hold channels 4k .. 4k + 3 and 16 + 4k .. 16 + 4k + 3, which is
        // what a lane of D already holds (tiles j = 0, 1); the A operand row 4c
        // holds head output c's weights for the same channels (other rows
        // zero), so output c lands in item 0 of lane group c.
        bf16x8 hwa;
        if constexpr (HD) {
            uint2 w0 = make_uint2(0, 0), w1 = make_uint2(0, 0);
            if ((row & 3) == 0 && row < 12) {
                const int c = row >> 2;
                w0 = *(const uint2*)(a.hw + c * 32 + 4 * kg);
                w1 = *(const uint2*)(a.hw + c * 32 + 16 + 4 * kg);
            }
            hwa = __builtin_bit_cast(bf16x8, make_uint4(w0.x, w0.y, w1.x, w1.y));
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int y = y0 + r;
            const size_t p = (size_t)(bb * a.H + y) * a.W + x;
            float v[2][4];
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int col = n0 + j * 16 + 4 * kg;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    v[j][i] = acc[r][j][i] + (a.bias && col < a.n ? a.bias[col + i] : 0.0f);
                    if (a.relu) v[j][i] = fmaxf(v[j][i], 0.0f);
                }
                const bool ok = y < a.H && x < a.W && col < a.n;
                if (ok && a.y)
                    *(uint2*)(a.y + p * a.ldy + col) = make_uint2(f2bf(v[j][0]) | ((uint32_t)f2bf(v[j][1]) << 16),
                                                                  f2bf(v[j][2]) | ((uint32_t)f2bf(v[j][3]) << 16));
                if (ok && a.yf) *(float4*)(a.yf + p * a.ldyf + col) = make_float4(v[j][0], v[j][1], v[j][2], v[j][3]);
                if constexpr (R % 2 == 0) {
                    if (a.yp) {
                        // rows r - 1 and r are a pair when r is odd: keep the even row's values
                        float m[4];
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const float h = fmaxf(v[j][i], swap_pair(v[j][i]));
                            m[i] = (r & 1) ? fmaxf(h, pool[j][i]) : h;
                            pool[j][i] = m[i];
                        }
                        if ((r & 1) && ok && !(row & 1)) {
                            const size_t q = (size_t)(bb * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1);
                            *(uint2*)(a.yp + q * a.ldyp + col) = make_uint2(f2bf(m[0]) | ((uint32_t)f2bf(m[1]) << 16),
                                                                            f2bf(m[2]) | ((uint32_t)f2bf(m[3]) << 16));
                        }
                    }
                }
            }
            if constexpr (HD) {
                const bf16x8 act = __builtin_bit_cast(
                    bf16x8, make_uint4(f2bf(v[0][0]) | ((uint32_t)f2bf(v[0][1]) << 16), f2bf(v[0][2]) | ((uint32_t)f2bf(v[0][3]) << 16),
                                       f2bf(v[1][0]) | ((uint32_t)f2bf(v[1][1]) << 16), f2bf(v[1][2]) | ((uint32_t)f2bf(v[1][3]) << 16)));
                const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hwa, act, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                if (kg < 3 && y < a.H && x < a.W) {
                    const float hv = d[0] + a.hbias[kg];
                    if (a.head) a.head[p * 3 + kg] = hv;
                    if (a.rgbOut) {
                        const size_t pc = (size_t)kg * a.B * a.H * a.W + p;
                        a.rgbOut[pc] = fmaxf(expm1f(hv + bf2f(log_rgb_bf(a.rgbIn[pc]))), 0.0f);
                    }
                }
            }
        }
    };
    bf16x8 b0[3][2], b1[3][2];
    fetch(0);
    if constexpr (!BL) bload(b0, 0, 0);
#pragma unroll
    for (int r = 0; r < R; r++) acc[r][0] = acc[r][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < a.cin; c0 += 2 * CS) {
        chunk(c0, b0, b1);
        if (c0 + CS < a.cin) chunk(c0 + CS, b1, b0);
    }
    store_tile();
}

// The first two 3x3 convolutions (16 -> 32 and 32 -> 32 channels, ReLU after
// each) in one kernel: the first layer's output goes from registers to LDS
// and never to memory (at 1920x1080 it is 133 MB written and read back, with
// the second layer's halo, 211 MB). A workgroup produces the second layer's
// output rows y0 .. y0 + 3, pixels x0 .. x0 + 61 (x0 = 62 tx), 32 channels,
// with the 2x2 max-pooled copy. Stage 0 computes the first layer on the
// 6 x 64 pixels those outputs read (rows y0 - 1 .. y0 + 4, pixels x0 - 1 ..
// x0 + 62; wave w the 16-pixel column block w) from the 8 x 66-pixel input
// halo, with the MFMAs of k_conv3r's 16-channel path in the same order, and
// writes it (bias, ReLU, BF16; zero outside the image, which is the second
// layer's zero padding) into the second layer's halo in LDS in k_conv3r's
// swizzled 64-byte rows. Halo columns 64 and 65 feed only outputs 62 and 63
// of each row, which are not stored; they are zero. Stage 1 runs k_conv3r's
// 32-channel path on that halo. Each output is bit-identical to the two
// separate convolutions (same MFMA sequence per output, same rounding). The
// halo recomputes the first layer on 1.55x its pixels, a sixth more MFMAs for
// the two layers. With y0out (training) stage 0 also stores the first
// layer's output on the workgroup's own 4 x 62 pixels.
struct Conv01Args {
    const uint16_t* x;                      // B x H x W x 16 BF16 pixels
    const uint16_t* wf0; const float* b0;   // first layer: weights in 16-channel fragment order, bias
    const uint16_t* wf1; const float* b1;   // second layer: weights in fragment order, bias
    uint16_t* y0out;                        // first layer output (32 channels, dense), nullable
    uint16_t* y; int ldy;                   // second layer output (32 channels at stride ldy)
    uint16_t* yp; int ldyp;                 // its 2x2 max-pooled copy
    int B, H, W;
};
constexpr int kC01Rows = 4, kC01Px = 62;

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_conv01(Conv01Args a) {
    constexpr int R = kC01Rows, R0 = R + 2;        // output rows of stage 1 / stage 0
    constexpr int HX = 66, HY0 = R0 + 2, HY1 = R + 2;
    constexpr int NP0 = HY0 * HX * 2;              // 16-byte pieces of the input halo (32-byte pixels)
    constexpr int NA = (NP0 + kBlock - 1) / kBlock;
    constexpr int PD = 2;
    __shared__ uint4 h0[NP0];                      // input halo: pixel hp's pieces at 2 hp, 2 hp + 1
    __shared__ uint4 h1[HY1 * HX * 4];             // first layer output halo: quarter q of pixel hp at 4 hp + (q ^ (hp & 2))
    const int tid = threadIdx.x, lane = tid & 63, wx = tid >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int tilesX = (a.W + kC01Px - 1) / kC01Px, tilesY = (a.H + R - 1) / R;
    const int tx = blockIdx.x % tilesX, tq = blockIdx.x / tilesX;
    const int ty = tq % tilesY, bb = tq / tilesY;
    const int x0 = tx * kC01Px, y0 = ty * R;
    const uint16_t* w00 = a.wf0 + lane * 8;        // 16-channel blocks 0 and 1: [3 dy][2 tap pairs][lane][8]
    const uint16_t* w01 = w00 + 6 * 512;
    const uint16_t* w10 = a.wf1 + lane * 8;        // [9 taps][lane][8]
    const uint16_t* w11 = w10 + 9 * 512;
    auto load0 = [&](bf16x8 (&b)[3][2], int pr) {
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
            b[dy][0] = __builtin_bit_cast(bf16x8, *(const uint4*)(w00 + (dy * 2 + pr) * 512));
            b[dy][1] = __builtin_bit_cast(bf16x8, *(const uint4*)(w01 + (dy * 2 + pr) * 512));
        }
    };
    auto load1 = [&](bf16x8 (&b)[3][2], int dx) {
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
            b[dy][0] = __builtin_bit_cast(bf16x8, *(const uint4*)(w10 + (dy * 3 + dx) * 512));
            b[dy][1] = __builtin_bit_cast(bf16x8, *(const uint4*)(w11 + (dy * 3 + dx) * 512));
        }
    };
    bf16x8 bx[3][2], by[3][2];
    load0(bx, 0);
    // input halo: pixel (y0 - 2 + hy, x0 - 2 + hx), zero outside the image
    uint4 ra[NA];
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int e = tid + k * kBlock, hp = e >> 1;
        const int hy = hp / HX, hx = hp - hy * HX;
        const int y = y0 - 2 + hy, x = x0 - 2 + hx;
        ra[k] = make_uint4(0, 0, 0, 0);
        if (e < NP0 && y >= 0 && y < a.H && x >= 0 && x < a.W)
            ra[k] = *(const uint4*)(a.x + (size_t)((bb * a.H + y) * a.W + x) * 16 + 8 * (e & 1));
    }
    float bias0[2][4], bias1[2][4];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            bias0[j][i] = a.b0[j * 16 + 4 * kg + i];
            bias1[j][i] = a.b1[j * 16 + 4 * kg + i];
        }
    if (tid < HY1 * 2 * 4) h1[((tid >> 3) * HX + 64 + ((tid >> 2) & 1)) * 4 + (tid & 3)] = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int e = tid + k * kBlock;
        if (e < NP0) h0[e] = ra[k];
    }
    __syncthreads();

    // stage 0: tap pair pr (taps 2 pr, 2 pr + 1; tap 3 is zero weight and a
    // zero operand) of the first layer, as k_conv3r's 16-channel path
    f32x4 acc0[R0][2];
#pragma unroll
    for (int r = 0; r < R0; r++) acc0[r][0] = acc0[r][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto step0 = [&](const bf16x8 (&bc)[3][2], int pr) {
        const bool pad = pr == 1 && kg >= 2;
        const int hp0 = wx * 16 + row + 2 * pr + (pad ? 0 : kg >> 1);
        const int s0 = hp0 * 2 + (kg & 1);
#pragma unroll
        for (int hy = 0; hy < HY0; hy++) {
            const bf16x8 av = __builtin_bit_cast(bf16x8, pad ? make_uint4(0, 0, 0, 0) : h0[hy * HX * 2 + s0]);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const int ro = hy - dy;
                if (ro < 0 || ro >= R0) continue;
                acc0[ro][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bc[dy][0], av, acc0[ro][0], 0, 0, 0);
                acc0[ro][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bc[dy][1], av, acc0[ro][1], 0, 0, 0);
            }
        }
        __builtin_amdgcn_sched_group_barrier(0x100, PD, 0);
        halo_row_order<R0, PD, 0>();
    };
    load0(by, 1);
    __builtin_amdgcn_sched_barrier(0);
    step0(bx, 0);
    load1(bx, 0);
    __builtin_amdgcn_sched_barrier(0);
    step0(by, 1);
    // stage 0 epilogue: first layer row r, pixel c = x0 - 1 + (16 wx + row)
    // -> halo pixel r * HX + 16 wx + row of the second layer
    {
        const int c = wx * 16 + row, x = x0 - 1 + c;
#pragma unroll
        for (int r = 0; r < R0; r++) {
            const int y = y0 - 1 + r;
            const bool in = y >= 0 && y < a.H && x >= 0 && x < a.W;
            const int hp = r * HX + c;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = fmaxf(acc0[r][j][i] + bias0[j][i], 0.0f);
                const uint2 w = in ? make_uint2(f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                                                f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16))
                                   : make_uint2(0, 0);
                const int q = 2 * j + (kg >> 1);
                ((uint2*)h1)[(hp * 4 + (q ^ (hp & 2))) * 2 + (kg & 1)] = w;
                if (a.y0out && in && r >= 1 && r <= R && c >= 1 && c <= kC01Px)
                    *(uint2*)(a.y0out + ((size_t)(bb * a.H + y) * a.W + x) * 32 + j * 16 + 4 * kg) = w;
            }
        }
    }
    __syncthreads();

    // stage 1: k_conv3r's 32-channel path over h1
    f32x4 acc[R][2];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r][0] = acc[r][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto step1 = [&](const bf16x8 (&bc)[3][2], int dx) {
        const int hp0 = wx * 16 + row + dx;
        const int s0 = hp0 * 4 + (kg ^ (hp0 & 2));
        const int s1 = hp0 * 4 + (kg ^ (hp0 & 2) ^ 2);
#pragma unroll
        for (int hy = 0; hy < HY1; hy++) {
            const int slot = hy * HX * 4 + (((hy * HX) & 2) ? s1 : s0);
            const bf16x8 av = __builtin_bit_cast(bf16x8, h1[slot]);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const int ro = hy - dy;
                if (ro < 0 || ro >= R) continue;
                acc[ro][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bc[dy][0], av, acc[ro][0], 0, 0, 0);
                acc[ro][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bc[dy][1], av, acc[ro][1], 0, 0, 0);
            }
        }
        __builtin_amdgcn_sched_group_barrier(0x100, PD, 0);
        halo_row_order<R, PD, 0>();
    };
    load1(by, 1);
    __builtin_amdgcn_sched_barrier(0);
    step1(bx, 0);
    load1(bx, 2);
    __builtin_amdgcn_sched_barrier(0);
    step1(by, 1);
    __builtin_amdgcn_sched_barrier(0);
    step1(bx, 2);

    // stage 1 epilogue (k_conv3r's, with outputs 62 and 63 of the row masked)
    const int x = x0 + wx * 16 + row;
    const bool colOk = x < a.W && wx * 16 + row < kC01Px;
    float pool[2][4];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + r;
        const size_t p = (size_t)(bb * a.H + y) * a.W + x;
        const bool ok = colOk && y < a.H;
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int col = j * 16 + 4 * kg;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = fmaxf(acc[r][j][i] + bias1[j][i], 0.0f);
            if (ok)
                *(uint2*)(a.y + p * a.ldy + col) = make_uint2(f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                                                              f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
            float m[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float h = fmaxf(v[i], swap_pair(v[i]));
                m[i] = (r & 1) ? fmaxf(h, pool[j][i]) : h;
                pool[j][i] = m[i];
            }
            if ((r & 1) && ok && !(row & 1)) {
                const size_t q = (size_t)(bb * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1);
                *(uint2*)(a.yp + q * a.ldyp + col) = make_uint2(f2bf(m[0]) | ((uint32_t)f2bf(m[1]) << 16),
                                                                f2bf(m[2]) | ((uint32_t)f2bf(m[3]) << 16));
            }
        }
    }
}

// --------------------------------------------------------------- wgrad ----
struct WgradArgs {
    const uint16_t* g; int ldg; int cout;    // output gradient (ReLU mask applied)
    const uint16_t* x; int ldx; int cin;     // layer input
    float* part;                             // [splits][cout][taps*cin]
    int B, H, W, splits, chunk;              // pixels per split (multiple of 32)
};

constexpr int kLdsRow = 40;   // 32 pixels + 8 padding (bf16 elements)

// grid: (ceil(cout/64), ceil(cin/64), taps*splits). Tile 64 (co) x 64 (ci);
// wave w: co [32*(w>>1), +32) x ci [32*(w&1), +32).
template <int TAPS>
__global__ __launch_bounds__(kBlock) void k_wgrad(WgradArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t Gt[64 * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint16_t Xt[64 * kLdsRow];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int co0 = blockIdx.x * 64, ci0 = blockIdx.y * 64;
    const int t = blockIdx.z % TAPS, split = blockIdx.z / TAPS;
    const int dy = TAPS == 9 ? t / 3 - 1 : 0, dx = TAPS == 9 ? t % 3 - 1 : 0;
    const int P = a.B * a.H * a.W;
    const int pBeg = split * a.chunk, pEnd = min(P, pBeg + a.chunk);
    const int row = lane & 15, kg = lane >> 4;
    const int wco = 32 * (wave >> 1), wci = 32 * (wave & 1);
    // loader: thread -> (pixel lp = tid/8, channel chunk 8*(tid%8))
    const int lp = threadIdx.x >> 3, lc = 8 * (threadIdx.x & 7);
    f32x4 acc[2][2];
    for (int r = 0; r < 2; r++)
        for (int j = 0; j < 2; j++) acc[r][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int p0 = pBeg; p0 < pEnd; p0 += 32) {
        const int p = p0 + lp;
        uint4 gv = make_uint4(0, 0, 0, 0), xv = make_uint4(0, 0, 0, 0);
        if (p < pEnd) {
            if (co0 + lc < a.cout) gv = *(const uint4*)(a.g + (size_t)p * a.ldg + co0 + lc);
            const int xx = p % a.W, q = p / a.W, yy = q % a.H, bb = q / a.H;
            const int sy = yy + dy, sx = xx + dx;
            if (sy >= 0 && sy < a.H && sx >= 0 && sx < a.W && ci0 + lc < a.cin)
                xv = *(const uint4*)(a.x + ((size_t)(bb * a.H + sy) * a.W + sx) * a.ldx + ci0 + lc);
        }
        const uint16_t* ge = (const uint16_t*)&gv;
        const uint16_t* xe = (const uint16_t*)&xv;
        for (int e = 0; e < 8; e++) {
            Gt[(lc + e) * kLdsRow + lp] = ge[e];
            Xt[(lc + e) * kLdsRow + lp] = xe[e];
        }
        __syncthreads();
        bf16x8 av[2], bv[2];
        for (int r = 0; r < 2; r++)
            av[r] = __builtin_bit_cast(bf16x8, *(const uint4*)&Gt[(wco + r * 16 + row) * kLdsRow + 8 * kg]);
        for (int j = 0; j < 2; j++)
            bv[j] = __builtin_bit_cast(bf16x8, *(const uint4*)&Xt[(wci + j * 16 + row) * kLdsRow + 8 * kg]);
        for (int r = 0; r < 2; r++)
            for (int j = 0; j < 2; j++)
                acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r], bv[j], acc[r][j], 0, 0, 0);
        __syncthreads();
    }
    const int K = TAPS * a.cin;
    float* out = a.part + (size_t)split * a.cout * K;
    for (int r = 0; r < 2; r++)
        for (int j = 0; j < 2; j++) {
            const int ci = ci0 + wci + j * 16 + row;
            if (ci >= a.cin) continue;
            for (int i = 0; i < 4; i++) {
                const int co = co0 + wco + r * 16 + 4 * kg + i;
                if (co < a.cout) out[(size_t)co * K + t * a.cin + ci] = acc[r][j][i];
            }
        }
}

// grad[k] = sum over splits of part[s][k], in split order
__global__ __launch_bounds__(kBlock) void k_reduce_splits(const float* part, int splits, size_t n, float* grad) {
    for (size_t k = blockIdx.x * (size_t)kBlock + threadIdx.x; k < n; k += (size_t)gridDim.x * kBlock) {
        float s = 0.0f;
        for (int i = 0; i < splits; i++) s += part[(size_t)i * n + k];
        grad[k] = s;
    }
}

// bias gradient partials: block b sums pixels [b*chunk, ...) of 64 channels
__global__ __launch_bounds__(kBlock) void k_bias_part(const uint16_t* g, int ldg, int cout, int P, int chunk,
                                                      float* part) {
    __shared__ float red[kBlock];
    const int c = blockIdx.y * 64 + (threadIdx.x & 63), sub = threadIdx.x >> 6;
    const int pBeg = blockIdx.x * chunk, pEnd = min(P, pBeg + chunk);
    float s = 0.0f;
    if (c < cout)
        for (int p = pBeg + sub; p < pEnd; p += 4) s += bf2f(g[(size_t)p * ldg + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (sub == 0 && c < cout)
        part[(size_t)blockIdx.x * cout + c] = ((red[threadIdx.x] + red[threadIdx.x + 64]) + red[threadIdx.x + 128]) +
                                              red[threadIdx.x + 192];
}

// ---------------------------------------------------------- elementwise ----
// g *= (y > 0), 8 channels per thread
__global__ __launch_bounds__(kBlock) void k_relu_bwd(uint16_t* g, int ldg, const uint16_t* y, int ldy, int c, int P) {
    const int cg = c / 8;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < (size_t)P * cg; i += (size_t)gridDim.x * kBlock) {
        const size_t p = i / cg;
        const int k = 8 * (int)(i % cg);
        uint4 gv = *(uint4*)(g + p * ldg + k);
        const uint4 yv = *(const uint4*)(y + p * ldy + k);
        uint16_t* ge = (uint16_t*)&gv;
        const uint16_t* ye = (const uint16_t*)&yv;
        for (int e = 0; e < 8; e++)
            if (!(bf2f(ye[e]) > 0.0f)) ge[e] = 0;
        *(uint4*)(g + p * ldg + k) = gv;
    }
}

// pool backward: gin[argmax] += gout (first maximum in window order)
__global__ __launch_bounds__(kBlock) void k_pool_bwd(const uint16_t* in, int ldi, const uint16_t* gout, int ldgo,
                                                     uint16_t* gin, int ldgi, int c, int B, int H, int W) {
    const int Ho = H / 2, Wo = W / 2, cg = c / 8;
    const size_t n = (size_t)B * Ho * Wo * cg;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int k = 8 * (int)(i % cg);
        const size_t po = i / cg;
        const int xo = (int)(po % Wo), q = (int)(po / Wo), yo = q % Ho, b = q / Ho;
        float m[8];
        int arg[8];
        size_t pis[4];
        for (int e = 0; e < 8; e++) { m[e] = -INFINITY; arg[e] = 0; }
        for (int s = 0; s < 4; s++) {
            pis[s] = ((size_t)(b * H + 2 * yo + (s >> 1)) * W + 2 * xo + (s & 1));
            const uint4 v = *(const uint4*)(in + pis[s] * ldi + k);
            const uint16_t* ve = (const uint16_t*)&v;
            for (int e = 0; e < 8; e++) {
                const float f = bf2f(ve[e]);
                if (f > m[e]) { m[e] = f; arg[e] = s; }
            }
        }
        const uint4 gv = *(const uint4*)(gout + po * ldgo + k);
        const uint16_t* gve = (const uint16_t*)&gv;
        for (int s = 0; s < 4; s++) {
            uint4 d = *(uint4*)(gin + pis[s] * ldgi + k);
            uint16_t* de = (uint16_t*)&d;
            for (int e = 0; e < 8; e++)
                if (arg[e] == s) de[e] = f2bf(bf2f(de[e]) + bf2f(gve[e]));
            *(uint4*)(gin + pis[s] * ldgi + k) = d;
        }
    }
}

// 2x nearest upsampling: out [B][2H][2W] (slice, stride ldo) from in [B][H][W]
__global__ __launch_bounds__(kBlock) void k_up(const uint16_t* in, int ldi, uint16_t* out, int ldo, int c, int B, int H,
                                               int W) {
    const int Ho = 2 * H, Wo = 2 * W, cg = c / 8;
    const size_t n = (size_t)B * Ho * Wo * cg;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int k = 8 * (int)(i % cg);
        const size_t po = i / cg;
        const int xo = (int)(po % Wo), q = (int)(po / Wo), yo = q % Ho, b = q / Ho;
        const size_t pi = (size_t)(b * H + yo / 2) * W + xo / 2;
        *(uint4*)(out + po * ldo + k) = *(const uint4*)(in + pi * ldi + k);
    }
}

// upsampling backward: gin[p] = sum of the 4 gout children (window order)
__global__ __launch_bounds__(kBlock) void k_up_bwd(const uint16_t* gout, int ldgo, uint16_t* gin, int ldgi, int c,
                                                   int B, int H, int W) {
    const int cg = c / 8;
    const size_t n = (size_t)B * H * W * cg;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int k = 8 * (int)(i % cg);
        const size_t pi = i / cg;
        const int x = (int)(pi % W), q = (int)(pi / W), y = q % H, b = q / H;
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < 4; t++) {
            const size_t po = (size_t)(b * 2 * H + 2 * y + (t >> 1)) * (2 * W) + 2 * x + (t & 1);
            const uint4 v = *(const uint4*)(gout + po * ldgo + k);
            const uint16_t* ve = (const uint16_t*)&v;
            for (int e = 0; e < 8; e++) s[e] += bf2f(ve[e]);
        }
        uint4 o;
        uint16_t* oe = (uint16_t*)&o;
        for (int e = 0; e < 8; e++) oe[e] = f2bf(s[e]);
        *(uint4*)(gin + pi * ldgi + k) = o;
    }
}

// L1 loss on log radiance: pred = head + input log-rgb; g (16 channels,
// zero beyond 3) = sign(pred - target) / (3P); per-block |d| partial sums.
__global__ __launch_bounds__(kBlock) void k_loss(const float* head, const uint16_t* xin, const float* target, int P,
                                                 uint16_t* g, float* lossPart) {
    __shared__ float red[kBlock];
    const float sc = 1.0f / (3.0f * (float)P);
    float s = 0.0f;
    for (int p = blockIdx.x * kBlock + threadIdx.x; p < P; p += gridDim.x * kBlock) {
        uint4 o0 = make_uint4(0, 0, 0, 0), o1 = make_uint4(0, 0, 0, 0);
        uint16_t* oe = (uint16_t*)&o0;
        for (int c = 0; c < 3; c++) {
            const float pred = head[(size_t)p * 3 + c] + bf2f(xin[(size_t)p * 16 + c]);
            const float d = pred - target[(size_t)p * 3 + c];
            s += fabsf(d);
            oe[c] = f2bf(d > 0.0f ? sc : (d < 0.0f ? -sc : 0.0f));
        }
        *(uint4*)(g + (size_t)p * 16) = o0;
        *(uint4*)(g + (size_t)p * 16 + 8) = o1;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossPart[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ __launch_bounds__(kBlock) void k_adam(float* w, float* m, float* v, const float* g, size_t n, float lr,
                                                 float b1, float b2, float eps, float c1, float c2) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// master [cout][taps][cin] -> forward BF16 [rowsF][taps*cin] and the flipped
// transpose for input gradients [rowsB][taps*coutK]: wb[ci][t][co] = w[co][taps-1-t][ci]
__global__ __launch_bounds__(kBlock) void k_wlayout(const float* w, int cout, int taps, int cin, uint16_t* wf,
                                                    uint16_t* wb, int coutK) {
    const size_t n = (size_t)cout * taps * cin;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int ci = (int)(i % cin);
        const int t = (int)((i / cin) % taps);
        const int co = (int)(i / ((size_t)cin * taps));
        const uint16_t h = f2bf(w[i]);
        wf[i] = h;
        wb[((size_t)ci * taps + (taps - 1 - t)) * coutK + co] = h;
    }
}

// [rows][9*cin] (row-major 3x3 weights, cin a multiple of 32) -> MFMA
// fragment order [rows/16][cin/32][9][lane][8]: lane = row%16 + 16*kg holds
// the 8 consecutive input channels 32*chunk + 8*kg .. +7 of tap t of that row
// (the B operand of mfma_f32_16x16x32_bf16 for one tap and channel chunk).
__global__ __launch_bounds__(kBlock) void k_wfrag(const uint16_t* w, int rows, int cin, uint16_t* out) {
    const size_t n = (size_t)rows * 9 * cin;
    const int nch = cin / 32;
    for (size_t o = blockIdx.x * (size_t)kBlock + threadIdx.x; o < n; o += (size_t)gridDim.x * kBlock) {
        const int e = (int)(o & 7), lane = (int)((o >> 3) & 63);
        const size_t q = o >> 9;
        const int t = (int)(q % 9), c = (int)((q / 9) % nch), blk = (int)(q / 9 / nch);
        const int row = blk * 16 + (lane & 15), kg = lane >> 4;
        out[o] = w[(size_t)row * 9 * cin + t * cin + c * 32 + kg * 8 + e];
    }
}

// [rows][9*16] (3x3 weights of a 16-channel input) -> fragment order
// [rows/16][3 dy][2 tap pairs][lane][8]: lane = row%16 + 16*kg holds input
// channels 8 (kg & 1) .. +7 of tap (dy, 2 pair + (kg >> 1)); tap column 3
// (the second half of the last pair) is zero.
__global__ __launch_bounds__(kBlock) void k_wfrag16(const uint16_t* w, int rows, uint16_t* out) {
    const size_t n = (size_t)rows * 6 * 64;
    for (size_t o = blockIdx.x * (size_t)kBlock + threadIdx.x; o < n; o += (size_t)gridDim.x * kBlock) {
        const int e = (int)(o & 7), lane = (int)((o >> 3) & 63);
        const size_t q = o >> 9;
        const int pr = (int)(q % 2), dy = (int)((q / 2) % 3), blk = (int)(q / 6);
        const int row = blk * 16 + (lane & 15), kg = lane >> 4, dx = 2 * pr + (kg >> 1);
        out[o] = dx < 3 ? w[(size_t)row * 144 + (dy * 3 + dx) * 16 + 8 * (kg & 1) + e] : (uint16_t)0;
    }
}

// planes (rgb, albedo, normal: 3 x P; depth: P) -> 16-channel BF16 pixels
__global__ __launch_bounds__(kBlock) void k_pack(const float* rgb, const float* alb, const float* nrm, const float* dep,
                                                 int P, uint16_t* x) {
    for (int p = blockIdx.x * kBlock + threadIdx.x; p < P; p += gridDim.x * kBlock) {
        float v[10];
        uint4 o0, o1;
        load_pixel(rgb, alb, nrm, dep, (size_t)P, (size_t)p, v);
        pack_pixel(v, o0, o1);
        *(uint4*)(x + (size_t)p * 16) = o0;
        *(uint4*)(x + (size_t)p * 16 + 8) = o1;
    }
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) p[i] = v;
}

// ---------------------------------------------------------------- host ----
struct LayerDef { int cin, cout, taps; };
// c1a c1b | c2a c2b | c3a c3b | c4a c4b | c5a c5b | c6a c6b | c7a c7b | head
static const LayerDef kLayers[15] = {
    {16, 32, 9},  {32, 32, 9},   {32, 64, 9},   {64, 64, 9},   {64, 128, 9}, {128, 128, 9}, {128, 256, 9},
    {256, 256, 9}, {384, 128, 9}, {128, 128, 9}, {192, 64, 9}, {64, 64, 9},  {96, 32, 9},   {32, 32, 9},
    {32, 3, 1}};
constexpr int kNumLayers = 15;

struct Layer {
    LayerDef d;
    size_t wOff, bOff;       // into the master parameter array
    int coutK, rowsF, rowsB;
    uint16_t *wf, *wb;       // BF16 forward / input-gradient layouts
    uint16_t *wfF, *wfB;     // the same in MFMA fragment order (3x3 layers with 32-channel chunks)
};

static inline int rup(int x, int m) { return (x + m - 1) / m * m; }

struct Net {
    int maxB, maxH, maxW;
    Layer L[kNumLayers];
    size_t nParams = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;   // master params, Adam state, grads
    uint16_t* wbuf = nullptr;                                        // all BF16 weight layouts
    // activations (level l has (H>>l) x (W>>l) pixels)
    uint16_t *X, *A0, *CAT1, *A12, *D1, *P1, *A2, *CAT2, *A10, *D2, *P2, *A4, *CAT3, *A8, *D3, *P3, *A6, *E4;
    uint16_t *gHead, *gD1, *gA12, *gCAT1, *gD2, *gA10, *gCAT2, *gD3, *gA8, *gCAT3, *gE4, *gA6, *gP3, *gA4, *gP2, *gA2,
        *gP1, *gA0;
    float *head, *target, *part, *lossPart;
    size_t partFloats = 0;
    void* act = nullptr;
    int step = 0;
    hipStream_t stream = nullptr;
};

static int grid_for(size_t n) { return (int)std::min<size_t>((n + kBlock - 1) / kBlock, 8192); }

// Pixels per block: 256 (4 row groups per wave) on large images, 128 below.
template <int TAPS, int KS, int WN>
static void launch_conv(const ConvArgs& a, int P, int n, hipStream_t s) {
    if (P >= (1 << 16)) {
        dim3 grid((P + 255) / 256, (n + 16 * WN - 1) / (16 * WN));
        hipLaunchKernelGGL((k_conv<TAPS, KS, WN, 4>), grid, dim3(kBlock), 0, s, a);
    } else {
        dim3 grid((P + 127) / 128, (n + 16 * WN - 1) / (16 * WN));
        hipLaunchKernelGGL((k_conv<TAPS, KS, WN, 2>), grid, dim3(kBlock), 0, s, a);
    }
}

// 3x3 convolution, cin a multiple of 32, n output channels (a multiple of
// 32; weight rows are allocated up to a multiple of 64, and up to 128 where n
// is). Tile per workgroup, as measured fastest per layer at 1920x1080 on
// MI355X: 128-channel multiples take 8 rows x 16 pixels x 128 channels (each
// wave 32 channels; no two waves share weights, so nothing is gained by
// staging them); 64-channel multiples with inputs of 128 channels or more
// 8 rows x 32 pixels x 64 channels, with 64-channel inputs 4 rows x 32 pixels
// x 64 channels at three waves per SIMD (one pass over the input instead of
// one per 32 output channels: 1.5 % of the 1080p frame); the rest 4 rows x
// 64 pixels x 32 channels at three waves per SIMD (the full-resolution
// layers, where the smaller accumulator set buys occupancy), with the weights
// staged once per workgroup in LDS when there is more than one channel chunk
// (all four waves use the same 32 channels; 12 % faster on the 96-channel
// decoder layer).
template <int R, int WGX, int WGC, int OCC, int PD, bool BL, int CS = 32, bool HD = false>
static void launch_r(const ConvArgs& a, int chanTiles, hipStream_t s) {
    constexpr int TX = 16 * WGX;
    const dim3 grid(((a.W + TX - 1) / TX) * ((a.H + R - 1) / R) * a.B, chanTiles);
    hipLaunchKernelGGL((k_conv3r<R, WGX, WGC, OCC, PD, BL, CS, HD>), grid, dim3(kBlock), 0, s, a);
}

static void launch_conv3(const ConvArgs& a, int n, hipStream_t s) {
    if (n % 128 == 0) launch_r<8, 1, 4, 2, 3, false>(a, n / 128, s);
    else if (n % 64 == 0 && a.cin >= 128) launch_r<8, 2, 2, 2, 3, false>(a, n / 64, s);
    else if (n % 64 == 0 && a.cin == 64) launch_r<4, 2, 2, 3, 2, false>(a, n / 64, s);
    else if (a.cin > 32) launch_r<4, 4, 1, 3, 2, true>(a, (n + 31) / 32, s);
    else if (a.hw) launch_r<4, 4, 1, 2, 2, false, 32, true>(a, 1, s);
    else launch_r<4, 4, 1, 3, 2, false>(a, (n + 31) / 32, s);
}

// Optional fusions (3x3 layers with 32-channel chunks): `up` supplies input
// channels [0, up.cup) 2x upsampled from a half-resolution tensor; `yp`
// receives the 2x2 max-pooled output.
// `head` applies layer `head.li` (the 1x1 head) in the epilogue.
// With `skipOnly` the layer's x holds only its channels [up.cup, cin).
struct UpSrc { const uint16_t* x = nullptr; int ld = 0, cup = 0; bool skipOnly = false; };
struct HeadFuse { int li = -1; float* head = nullptr; const float* rgbIn = nullptr; float* rgbOut = nullptr; };
static void conv_fwd(Net& N, int li, const uint16_t* x, int ldx, uint16_t* y, int ldy, int B, int H, int W, bool relu,
                     float* yf = nullptr, int ldyf = 0, UpSrc up = UpSrc(), uint16_t* yp = nullptr, int ldyp = 0,
                     HeadFuse hf = HeadFuse()) {
    const Layer& L = N.L[li];
    const Layer* Lh = hf.li >= 0 ? &N.L[hf.li] : nullptr;
    ConvArgs a{x, ldx, L.d.cin, L.wf, L.wfF, N.w + L.bOff, y, ldy, yf, ldyf, L.d.cout, B, H, W, relu ? 1 : 0,
               up.x, up.ld, up.cup, yp, ldyp,
               Lh ? Lh->wf : nullptr, Lh ? N.w + Lh->bOff : nullptr, hf.head, hf.rgbIn, hf.rgbOut,
               up.skipOnly ? up.cup : 0};
    if (Lh && !(L.d.taps == 9 && L.d.cin == 32 && L.d.cout == 32 && Lh->d.cin == 32 && Lh->d.cout == 3)) {
        rt::set_error("denoiser: head fusion needs a 32 -> 32 3x3 layer and a 32 -> 3 head");
        return;
    }
    const int P = B * H * W;
    const int WN = L.d.cout >= 64 ? 4 : (L.d.cout >= 32 ? 2 : 1);
    hipStream_t s = N.stream;
    if ((L.d.taps == 1 || L.d.cin % 32) && (up.x || yp || Lh)) {
        rt::set_error("denoiser: fused up/pool on a layer without 32-channel chunks");
        return;
    }
    if (L.d.taps == 1) launch_conv<1, 32, 1>(a, P, L.d.cout, s);
    else if (L.d.cin == 16 && L.d.cout == 32) {
        // the 16-channel input layer: one chunk, tap pairs per MFMA, 4 x 64 x 32 tiles
        launch_r<4, 4, 1, 3, 2, false, 16>(a, 1, s);
    } else if (L.d.cin % 32) {
        if (WN == 2) launch_conv<9, 16, 2>(a, P, L.d.cout, s);
        else launch_conv<9, 16, 4>(a, P, L.d.cout, s);
    } else launch_conv3(a, L.d.cout, s);
}

// The first two layers (16 -> 32 -> 32 channels, ReLU after each) in one
// kernel from N.X; y0out (training) receives the first layer's output.
static void conv01_fwd(Net& N, uint16_t* y0out, uint16_t* y, int ldy, uint16_t* yp, int ldyp, int B, int H, int W) {
    const Layer& L0 = N.L[0];
    const Layer& L1 = N.L[1];
    const Conv01Args a{N.X, L0.wfF, N.w + L0.bOff, L1.wfF, N.w + L1.bOff, y0out, y, ldy, yp, ldyp, B, H, W};
    const dim3 grid(((W + kC01Px - 1) / kC01Px) * ((H + kC01Rows - 1) / kC01Rows) * B);
    hipLaunchKernelGGL(k_conv01, grid, dim3(kBlock), 0, N.stream, a);
}

// input gradient: dx = conv(g, flipped transposed weights)
static void conv_dx(Net& N, int li, const uint16_t* g, int ldg, uint16_t* dx, int lddx, int B, int H, int W) {
    const Layer& L = N.L[li];
    ConvArgs a{g, ldg, L.coutK, L.wb, L.wfB, nullptr, dx, lddx, nullptr, 0, L.d.cin, B, H, W, 0};
    const int P = B * H * W;
    const int WN = L.d.cin >= 64 ? 4 : 2;
    hipStream_t s = N.stream;
    if (L.d.taps == 1) launch_conv<1, 16, 2>(a, P, L.d.cin, s);
    else launch_conv3(a, L.d.cin, s);
}

static void conv_dw(Net& N, int li, const uint16_t* g, int ldg, const uint16_t* x, int ldx, int B, int H, int W) {
    const Layer& L = N.L[li];
    const int P = B * H * W;
    const size_t K = (size_t)L.d.taps * L.d.cin, n = (size_t)L.d.cout * K;
    int splits = std::max(1, std::min(64, P / 2048));
    while (splits > 1 && (size_t)splits * n > N.partFloats) splits--;
    const int chunk = rup((P + splits - 1) / splits, 32);
    splits = (P + chunk - 1) / chunk;
    WgradArgs a{g, ldg, L.d.cout, x, ldx, L.d.cin, N.part, B, H, W, splits, chunk};
    dim3 grid((L.d.cout + 63) / 64, (L.d.cin + 63) / 64, L.d.taps * splits);
    if (L.d.taps == 9) hipLaunchKernelGGL(k_wgrad<9>, grid, dim3(kBlock), 0, N.stream, a);
    else hipLaunchKernelGGL(k_wgrad<1>, grid, dim3(kBlock), 0, N.stream, a);
    hipLaunchKernelGGL(k_reduce_splits, dim3(grid_for(n)), dim3(kBlock), 0, N.stream, N.part, splits, n, N.g + L.wOff);
    // bias: partials over pixel chunks, then the same fixed-order reduction
    const int bchunk = 4096, nb = (P + bchunk - 1) / bchunk;
    hipLaunchKernelGGL(k_bias_part, dim3(nb, (L.d.cout + 63) / 64), dim3(kBlock), 0, N.stream, g, ldg, L.d.cout, P,
                       bchunk, N.part);
    hipLaunchKernelGGL(k_reduce_splits, dim3(grid_for(L.d.cout)), dim3(kBlock), 0, N.stream, N.part, nb,
                       (size_t)L.d.cout, N.g + L.bOff);
}

static void relu_bwd(Net& N, uint16_t* g, int ldg, const uint16_t* y, int ldy, int c, int P) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for((size_t)P * c / 8)), dim3(kBlock), 0, N.stream, g, ldg, y, ldy, c, P);
}
static void pool_bwd(Net& N, const uint16_t* in, int ldi, const uint16_t* go, int ldgo, uint16_t* gi, int ldgi, int c,
                     int B, int H, int W) {
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for((size_t)B * H * W / 4 * c / 8)), dim3(kBlock), 0, N.stream, in, ldi,
                       go, ldgo, gi, ldgi, c, B, H, W);
}
static void up(Net& N, const uint16_t* in, int ldi, uint16_t* out, int ldo, int c, int B, int H, int W) {
    hipLaunchKernelGGL(k_up, dim3(grid_for((size_t)B * H * W * 4 * c / 8)), dim3(kBlock), 0, N.stream, in, ldi, out,
                       ldo, c, B, H, W);
}
static void up_bwd(Net& N, const uint16_t* go, int ldgo, uint16_t* gi, int ldgi, int c, int B, int H, int W) {
    hipLaunchKernelGGL(k_up_bwd, dim3(grid_for((size_t)B * H * W * c / 8)), dim3(kBlock), 0, N.stream, go, ldgo, gi,
                       ldgi, c, B, H, W);
}

static void update_layouts(Net& N) {
    for (int i = 0; i < kNumLayers; i++) {
        const Layer& L = N.L[i];
        const size_t n = (size_t)L.d.cout * L.d.taps * L.d.cin;
        hipLaunchKernelGGL(k_wlayout, dim3(grid_for(n)), dim3(kBlock), 0, N.stream, N.w + L.wOff, L.d.cout, L.d.taps,
                           L.d.cin, L.wf, L.wb, L.coutK);
        if (L.wfF && L.d.cin == 16)
            hipLaunchKernelGGL(k_wfrag16, dim3(grid_for((size_t)L.rowsF * 6 * 64)), dim3(kBlock), 0, N.stream, L.wf,
                               L.rowsF, L.wfF);
        else if (L.wfF)
            hipLaunchKernelGGL(k_wfrag, dim3(grid_for((size_t)L.rowsF * 9 * L.d.cin)), dim3(kBlock), 0, N.stream, L.wf,
                               L.rowsF, L.d.cin, L.wfF);
        if (L.wfB)
            hipLaunchKernelGGL(k_wfrag, dim3(grid_for((size_t)L.rowsB * 9 * L.coutK)), dim3(kBlock), 0, N.stream,
                               L.wb, L.rowsB, L.coutK, L.wfB);
    }
}

// forward from N.X (B x H x W x 16) to N.head (P x 3 FP32, before the
// residual), or in inference with rgbOut straight to the denoised planes
// (residual from the noisy radiance planes rgbIn). The pooled encoder outputs come from the producing convolutions' epilogues; the
// decoder convolutions read their upsampled inputs straight from the
// half-resolution tensors. Training also stores the upsampled channels in the
// concatenation buffers (the weight gradients read them), with the skip
// channels after them in each pixel's row. Inference stores only the skip
// channels, each pixel's row dense (32, 64, 128 channels): a 32-channel skip
// row inside a 96-channel pixel (64 bytes at byte 128 of 192) fills half of
// a 128-byte line, and the decoder layer fetched twice its bytes.
static void forward(Net& N, int B, int H, int W, bool train, float* rgbOut = nullptr, const float* rgbIn = nullptr) {
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    // skip tensors E1, E2, E3: pointer and channel stride
    uint16_t* const e1 = train ? N.CAT1 + 64 : N.CAT1;
    uint16_t* const e2 = train ? N.CAT2 + 128 : N.CAT2;
    uint16_t* const e3 = train ? N.CAT3 + 256 : N.CAT3;
    const int l1 = train ? 96 : 32, l2 = train ? 192 : 64, l3 = train ? 384 : 128;
    conv01_fwd(N, train ? N.A0 : nullptr, e1, l1, N.P1, 32, B, H, W);                         // A0, E1, P1
    conv_fwd(N, 2, N.P1, 32, N.A2, 64, B, H1, W1, true);
    conv_fwd(N, 3, N.A2, 64, e2, l2, B, H1, W1, true, nullptr, 0, UpSrc(), N.P2, 64);      // E2, P2
    conv_fwd(N, 4, N.P2, 64, N.A4, 128, B, H2, W2, true);
    conv_fwd(N, 5, N.A4, 128, e3, l3, B, H2, W2, true, nullptr, 0, UpSrc(), N.P3, 128);    // E3, P3
    conv_fwd(N, 6, N.P3, 128, N.A6, 256, B, H3, W3, true);
    conv_fwd(N, 7, N.A6, 256, N.E4, 256, B, H3, W3, true);
    if (train) up(N, N.E4, 256, N.CAT3, 384, 256, B, H3, W3);
    conv_fwd(N, 8, train ? N.CAT3 : e3, l3, N.A8, 128, B, H2, W2, true, nullptr, 0, UpSrc{N.E4, 256, 256, !train});
    conv_fwd(N, 9, N.A8, 128, N.D3, 128, B, H2, W2, true);
    if (train) up(N, N.D3, 128, N.CAT2, 192, 128, B, H2, W2);
    conv_fwd(N, 10, train ? N.CAT2 : e2, l2, N.A10, 64, B, H1, W1, true, nullptr, 0, UpSrc{N.D3, 128, 128, !train});
    conv_fwd(N, 11, N.A10, 64, N.D2, 64, B, H1, W1, true);
    if (train) up(N, N.D2, 64, N.CAT1, 96, 64, B, H1, W1);
    conv_fwd(N, 12, train ? N.CAT1 : e1, l1, N.A12, 32, B, H, W, true, nullptr, 0, UpSrc{N.D2, 64, 64, !train});
    if (train) {
        conv_fwd(N, 13, N.A12, 32, N.D1, 32, B, H, W, true);
        conv_fwd(N, 14, N.D1, 32, nullptr, 0, B, H, W, false, N.head, 3);
    } else {
        // inference: the head (and with rgbOut the residual and expm1) in the
        // last convolution's epilogue; its 32-channel output is not stored
        HeadFuse hf;
        hf.li = 14;
        hf.head = rgbOut ? nullptr : N.head;
        hf.rgbIn = rgbIn;
        hf.rgbOut = rgbOut;
        conv_fwd(N, 13, N.A12, 32, nullptr, 0, B, H, W, true, nullptr, 0, UpSrc(), nullptr, 0, hf);
    }
}

// backward from N.gHead (P x 16, loss gradient in channels 0-2) to N.g
static void backward(Net& N, int B, int H, int W) {
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    const int P0 = B * H * W, P1 = B * H1 * W1, P2 = B * H2 * W2, P3 = B * H3 * W3;
    conv_dw(N, 14, N.gHead, 16, N.D1, 32, B, H, W);
    conv_dx(N, 14, N.gHead, 16, N.gD1, 32, B, H, W);
    relu_bwd(N, N.gD1, 32, N.D1, 32, 32, P0);
    conv_dw(N, 13, N.gD1, 32, N.A12, 32, B, H, W);
    conv_dx(N, 13, N.gD1, 32, N.gA12, 32, B, H, W);
    relu_bwd(N, N.gA12, 32, N.A12, 32, 32, P0);
    conv_dw(N, 12, N.gA12, 32, N.CAT1, 96, B, H, W);
    conv_dx(N, 12, N.gA12, 32, N.gCAT1, 96, B, H, W);
    up_bwd(N, N.gCAT1, 96, N.gD2, 64, 64, B, H1, W1);
    relu_bwd(N, N.gD2, 64, N.D2, 64, 64, P1);
    conv_dw(N, 11, N.gD2, 64, N.A10, 64, B, H1, W1);
    conv_dx(N, 11, N.gD2, 64, N.gA10, 64, B, H1, W1);
    relu_bwd(N, N.gA10, 64, N.A10, 64, 64, P1);
    conv_dw(N, 10, N.gA10, 64, N.CAT2, 192, B, H1, W1);
    conv_dx(N, 10, N.gA10, 64, N.gCAT2, 192, B, H1, W1);
    up_bwd(N, N.gCAT2, 192, N.gD3, 128, 128, B, H2, W2);
    relu_bwd(N, N.gD3, 128, N.D3, 128, 128, P2);
    conv_dw(N, 9, N.gD3, 128, N.A8, 128, B, H2, W2);
    conv_dx(N, 9, N.gD3, 128, N.gA8, 128, B, H2, W2);
    relu_bwd(N, N.gA8, 128, N.A8, 128, 128, P2);
    conv_dw(N, 8, N.gA8, 128, N.CAT3, 384, B, H2, W2);
    conv_dx(N, 8, N.gA8, 128, N.gCAT3, 384, B, H2, W2);
    up_bwd(N, N.gCAT3, 384, N.gE4, 256, 256, B, H3, W3);
    relu_bwd(N, N.gE4, 256, N.E4, 256, 256, P3);
    conv_dw(N, 7, N.gE4, 256, N.A6, 256, B, H3, W3);
    conv_dx(N, 7, N.gE4, 256, N.gA6, 256, B, H3, W3);
    relu_bwd(N, N.gA6, 256, N.A6, 256, 256, P3);
    conv_dw(N, 6, N.gA6, 256, N.P3, 128, B, H3, W3);
    conv_dx(N, 6, N.gA6, 256, N.gP3, 128, B, H3, W3);
    pool_bwd(N, N.CAT3 + 256, 384, N.gP3, 128, N.gCAT3 + 256, 384, 128, B, H2, W2);
    relu_bwd(N, N.gCAT3 + 256, 384, N.CAT3 + 256, 384, 128, P2);
    conv_dw(N, 5, N.gCAT3 + 256, 384, N.A4, 128, B, H2, W2);
    conv_dx(N, 5, N.gCAT3 + 256, 384, N.gA4, 128, B, H2, W2);
    relu_bwd(N, N.gA4, 128, N.A4, 128, 128, P2);
    conv_dw(N, 4, N.gA4, 128, N.P2, 64, B, H2, W2);
    conv_dx(N, 4, N.gA4, 128, N.gP2, 64, B, H2, W2);
    pool_bwd(N, N.CAT2 + 128, 192, N.gP2, 64, N.gCAT2 + 128, 192, 64, B, H1, W1);
    relu_bwd(N, N.gCAT2 + 128, 192, N.CAT2 + 128, 192, 64, P1);
    conv_dw(N, 3, N.gCAT2 + 128, 192, N.A2, 64, B, H1, W1);
    conv_dx(N, 3, N.gCAT2 + 128, 192, N.gA2, 64, B, H1, W1);
    relu_bwd(N, N.gA2, 64, N.A2, 64, 64, P1);
    conv_dw(N, 2, N.gA2, 64, N.P1, 32, B, H1, W1);
    conv_dx(N, 2, N.gA2, 64, N.gP1, 32, B, H1, W1);
    pool_bwd(N, N.CAT1 + 64, 96, N.gP1, 32, N.gCAT1 + 64, 96, 32, B, H, W);
    relu_bwd(N, N.gCAT1 + 64, 96, N.CAT1 + 64, 96, 32, P0);
    conv_dw(N, 1, N.gCAT1 + 64, 96, N.A0, 32, B, H, W);
    conv_dx(N, 1, N.gCAT1 + 64, 96, N.gA0, 32, B, H, W);
    relu_bwd(N, N.gA0, 32, N.A0, 32, 32, P0);
    conv_dw(N, 0, N.gA0, 32, N.X, 16, B, H, W);
}

// Seeded He-uniform initialisation (weights of the zero input-padding
// channels of the first layer stay 0), biases 0.
static void init_params(std::vector<float>& w, const Net& N, uint64_t seed) {
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        return (double)(s >> 11) * 0x1p-53;
    };
    for (int i = 0; i < kNumLayers; i++) {
        const Layer& L = N.L[i];
        const int fan = L.d.taps * (i == 0 ? 10 : L.d.cin);
        const double lim = std::sqrt(6.0 / fan) * (i == kNumLayers - 1 ? 0.1 : 1.0);
        for (int co = 0; co < L.d.cout; co++)
            for (int t = 0; t < L.d.taps; t++)
                for (int ci = 0; ci < L.d.cin; ci++) {
                    const double r = (2.0 * next() - 1.0) * lim;
                    w[L.wOff + ((size_t)co * L.d.taps + t) * L.d.cin + ci] = (i == 0 && ci >= 10) ? 0.0f : (float)r;
                }
    }
}

}  // namespace dn

struct CudaDenoiserHandle {
    dn::Net net;
};

using dn::Net;

static bool dn_ok(CudaDenoiser_t d, const char* fn) {
    if (!d) { rt::set_error("%s: NULL denoiser", fn); return false; }
    return true;
}

static bool dn_dims(const Net& N, int B, int H, int W, const char* fn) {
    if (B <= 0 || H <= 0 || W <= 0 || H % 8 || W % 8 || B * H * W > N.maxB * N.maxH * N.maxW) {
        rt::set_error("%s: batch %d x %dx%d: need H, W multiples of 8 and B*H*W <= %d", fn, B, H, W,
                      N.maxB * N.maxH * N.maxW);
        return false;
    }
    return true;
}

extern "C" {

CudaDenoiser_t cudaBVH_denoiserCreate(int maxPixels, uint64_t seed) {
    rt::clear_error();
    if (maxPixels <= 0 || maxPixels % 64) {
        rt::set_error("cudaBVH_denoiserCreate: maxPixels must be a positive multiple of 64 (got %d)", maxPixels);
        return nullptr;
    }
    auto* h = new CudaDenoiserHandle();
    Net& N = h->net;
    N.maxB = 1; N.maxH = 1; N.maxW = maxPixels;
    size_t off = 0, wbytes = 0;
    for (int i = 0; i < dn::kNumLayers; i++) {
        dn::Layer& L = N.L[i];
        L.d = dn::kLayers[i];
        L.wOff = off; off += (size_t)L.d.cout * L.d.taps * L.d.cin;
        L.bOff = off; off += (size_t)L.d.cout;
        L.coutK = dn::rup(L.d.cout, 16);
        L.rowsF = dn::rup(L.d.cout, 64);
        L.rowsB = dn::rup(L.d.cin, 64);
        wbytes += ((size_t)L.rowsF * L.d.taps * L.d.cin + (size_t)L.rowsB * L.d.taps * L.coutK) * 2;
        if (L.d.taps == 9 && L.d.cin % 32 == 0) wbytes += (size_t)L.rowsF * 9 * L.d.cin * 2;
        if (L.d.taps == 9 && L.d.cin == 16) wbytes += (size_t)L.rowsF * 6 * 64 * 2;
        if (L.d.taps == 9 && L.coutK % 32 == 0) wbytes += (size_t)L.rowsB * 9 * L.coutK * 2;
    }
    N.nParams = off;
    const size_t P = (size_t)maxPixels;
    // activation channel counts at each level (level l: P >> 2l pixels)
    const size_t lv0 = P, lv1 = P / 4, lv2 = P / 16, lv3 = P / 64;
    struct Buf { uint16_t** p; size_t elems; };
    std::vector<Buf> bufs = {
        {&N.X, lv0 * 16}, {&N.A0, lv0 * 32}, {&N.CAT1, lv0 * 96}, {&N.A12, lv0 * 32}, {&N.D1, lv0 * 32},
        {&N.P1, lv1 * 32}, {&N.A2, lv1 * 64}, {&N.CAT2, lv1 * 192}, {&N.A10, lv1 * 64}, {&N.D2, lv1 * 64},
        {&N.P2, lv2 * 64}, {&N.A4, lv2 * 128}, {&N.CAT3, lv2 * 384}, {&N.A8, lv2 * 128}, {&N.D3, lv2 * 128},
        {&N.P3, lv3 * 128}, {&N.A6, lv3 * 256}, {&N.E4, lv3 * 256},
        {&N.gHead, lv0 * 16}, {&N.gD1, lv0 * 32}, {&N.gA12, lv0 * 32}, {&N.gCAT1, lv0 * 96}, {&N.gD2, lv1 * 64},
        {&N.gA10, lv1 * 64}, {&N.gCAT2, lv1 * 192}, {&N.gD3, lv2 * 128}, {&N.gA8, lv2 * 128}, {&N.gCAT3, lv2 * 384},
        {&N.gE4, lv3 * 256}, {&N.gA6, lv3 * 256}, {&N.gP3, lv3 * 128}, {&N.gA4, lv2 * 128}, {&N.gP2, lv2 * 64},
        {&N.gA2, lv1 * 64}, {&N.gP1, lv1 * 32}, {&N.gA0, lv0 * 32}};
    size_t actBytes = 0;
    for (auto& b : bufs) actBytes += (b.elems * 2 + 255) & ~(size_t)255;
    N.partFloats = std::max<size_t>(16u << 20, (size_t)(P / 4096 + 1) * 256);
    const size_t fBytes = (lv0 * 3 * 2 + N.partFloats + 8192) * 4;
    if (hipMalloc(&N.act, actBytes + fBytes) != hipSuccess ||
        hipMalloc((void**)&N.w, N.nParams * 4 * 4) != hipSuccess || hipMalloc((void**)&N.wbuf, wbytes) != hipSuccess) {
        rt::set_error("cudaBVH_denoiserCreate: cannot allocate %zu bytes", actBytes + fBytes + N.nParams * 16 + wbytes);
        hipFree(N.act); hipFree(N.w); hipFree(N.wbuf);
        delete h;
        return nullptr;
    }
    char* p = (char*)N.act;
    for (auto& b : bufs) { *b.p = (uint16_t*)p; p += (b.elems * 2 + 255) & ~(size_t)255; }
    N.head = (float*)p; p += lv0 * 3 * 4;
    N.target = (float*)p; p += lv0 * 3 * 4;
    N.lossPart = (float*)p; p += 8192 * 4;
    N.part = (float*)p;
    N.m = N.w + N.nParams; N.v = N.m + N.nParams; N.g = N.v + N.nParams;
    uint16_t* wp = N.wbuf;
    for (int i = 0; i < dn::kNumLayers; i++) {
        dn::Layer& L = N.L[i];
        L.wf = wp; wp += (size_t)L.rowsF * L.d.taps * L.d.cin;
        L.wb = wp; wp += (size_t)L.rowsB * L.d.taps * L.coutK;
        L.wfF = L.wfB = nullptr;
        if (L.d.taps == 9 && L.d.cin % 32 == 0) { L.wfF = wp; wp += (size_t)L.rowsF * 9 * L.d.cin; }
        if (L.d.taps == 9 && L.d.cin == 16) { L.wfF = wp; wp += (size_t)L.rowsF * 6 * 64; }
        if (L.d.taps == 9 && L.coutK % 32 == 0) { L.wfB = wp; wp += (size_t)L.rowsB * 9 * L.coutK; }
    }
    std::vector<float> w(N.nParams, 0.0f);
    dn::init_params(w, N, seed);
    if (hipMemset(N.wbuf, 0, wbytes) != hipSuccess || hipMemset(N.m, 0, N.nParams * 12) != hipSuccess ||
        hipMemcpy(N.w, w.data(), N.nParams * 4, hipMemcpyHostToDevice) != hipSuccess) {
        rt::set_error("cudaBVH_denoiserCreate: initialisation failed");
        hipFree(N.act); hipFree(N.w); hipFree(N.wbuf);
        delete h;
        return nullptr;
    }
    dn::update_layouts(N);
    if (hipStreamSynchronize(N.stream) != hipSuccess) { rt::set_error("cudaBVH_denoiserCreate: layout failed"); }
    return h;
}

void cudaBVH_denoiserDestroy(CudaDenoiser_t d) {
    if (!d) return;
    hipStreamSynchronize(d->net.stream);
    hipFree(d->net.act); hipFree(d->net.w); hipFree(d->net.wbuf);
    delete d;
}

int cudaBVH_denoiserSetStream(CudaDenoiser_t d, void* stream) {
    if (!dn_ok(d, "cudaBVH_denoiserSetStream")) return -1;
    d->net.stream = (hipStream_t)stream;
    return 0;
}

long long cudaBVH_denoiserNumParams(CudaDenoiser_t d) { return d ? (long long)d->net.nParams : 0; }

int cudaBVH_denoiserGetParams(CudaDenoiser_t d, float* out, int which) {
    if (!dn_ok(d, "cudaBVH_denoiserGetParams") || !out || which < 0 || which > 3) {
        rt::set_error("cudaBVH_denoiserGetParams: bad arguments");
        return -1;
    }
    const float* src = which == 0 ? d->net.w : which == 1 ? d->net.g : which == 2 ? d->net.m : d->net.v;
    if (hipStreamSynchronize(d->net.stream) != hipSuccess ||
        hipMemcpy(out, src, d->net.nParams * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        rt::set_error("cudaBVH_denoiserGetParams: copy failed");
        return -1;
    }
    return 0;
}

int cudaBVH_denoiserSetParams(CudaDenoiser_t d, const float* in) {
    if (!dn_ok(d, "cudaBVH_denoiserSetParams") || !in) return -1;
    Net& N = d->net;
    if (hipStreamSynchronize(N.stream) != hipSuccess ||
        hipMemcpy(N.w, in, N.nParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(N.m, 0, N.nParams * 8) != hipSuccess) {
        rt::set_error("cudaBVH_denoiserSetParams: copy failed");
        return -1;
    }
    N.step = 0;
    dn::update_layouts(N);
    return 0;
}

int cudaBVH_denoiserPack(CudaDenoiser_t d, int numPixels, const float* rgb, const float* albedo, const float* normal,
                         const float* depth, uint16_t* out) {
    if (!dn_ok(d, "cudaBVH_denoiserPack")) return -1;
    if (numPixels <= 0 || !rgb || !albedo || !normal || !depth || !out) {
        rt::set_error("cudaBVH_denoiserPack: bad arguments");
        return -1;
    }
    hipLaunchKernelGGL(dn::k_pack, dim3(dn::grid_for(numPixels)), dim3(dn::kBlock), 0, d->net.stream, rgb, albedo,
                       normal, depth, numPixels, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cudaBVH_denoiserForward(CudaDenoiser_t d, int batch, int height, int width, const uint16_t* input,
                            float* headOut) {
    if (!dn_ok(d, "cudaBVH_denoiserForward") || !dn_dims(d->net, batch, height, width, "cudaBVH_denoiserForward"))
        return -1;
    if (!input) { rt::set_error("cudaBVH_denoiserForward: NULL input"); return -1; }
    Net& N = d->net;
    const size_t P = (size_t)batch * height * width;
    if (hipMemcpyAsync(N.X, input, P * 32, hipMemcpyDeviceToDevice, N.stream) != hipSuccess) {
        rt::set_error("cudaBVH_denoiserForward: input copy failed");
        return -1;
    }
    dn::forward(N, batch, height, width, false);
    if (headOut && hipMemcpyAsync(headOut, N.head, P * 12, hipMemcpyDeviceToDevice, N.stream) != hipSuccess) {
        rt::set_error("cudaBVH_denoiserForward: output copy failed");
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cudaBVH_denoise(CudaDenoiser_t d, int width, int height, const float* rgb, const float* albedo,
                    const float* normal, const float* depth, float* out) {
    if (!dn_ok(d, "cudaBVH_denoise") || !dn_dims(d->net, 1, height, width, "cudaBVH_denoise")) return -1;
    if (!rgb || !albedo || !normal || !depth || !out) { rt::set_error("cudaBVH_denoise: NULL buffer"); return -1; }
    Net& N = d->net;
    const int P = width * height;
    hipLaunchKernelGGL(dn::k_pack, dim3(dn::grid_for(P)), dim3(dn::kBlock), 0, N.stream, rgb, albedo, normal, depth, P,
                       N.X);
    dn::forward(N, 1, height, width, false, out, rgb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cudaBVH_denoiserTrainStep(CudaDenoiser_t d, int batch, int height, int width, const uint16_t* input,
                              const float* target, float learningRate, float* lossOut) {
    if (!dn_ok(d, "cudaBVH_denoiserTrainStep") || !dn_dims(d->net, batch, height, width, "cudaBVH_denoiserTrainStep"))
        return -1;
    if (!input || !target) { rt::set_error("cudaBVH_denoiserTrainStep: NULL buffer"); return -1; }
    Net& N = d->net;
    const int P = batch * height * width;
    if (hipMemcpyAsync(N.X, input, (size_t)P * 32, hipMemcpyDeviceToDevice, N.stream) != hipSuccess ||
        hipMemcpyAsync(N.target, target, (size_t)P * 12, hipMemcpyDeviceToDevice, N.stream) != hipSuccess) {
        rt::set_error("cudaBVH_denoiserTrainStep: input copy failed");
        return -1;
    }
    dn::forward(N, batch, height, width, true);
    const int lb = std::min(dn::grid_for(P), 8192);
    hipLaunchKernelGGL(dn::k_loss, dim3(lb), dim3(dn::kBlock), 0, N.stream, N.head, N.X, N.target, P, N.gHead,
                       N.lossPart);
    dn::backward(N, batch, height, width);
    if (learningRate > 0.0f) {
        N.step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - (float)std::pow(0.9, N.step), c2 = 1.0f - (float)std::pow(0.999, N.step);
        hipLaunchKernelGGL(dn::k_adam, dim3(dn::grid_for(N.nParams)), dim3(dn::kBlock), 0, N.stream, N.w, N.m, N.v,
                           N.g, N.nParams, learningRate, b1, b2, 1e-8f, c1, c2);
        dn::update_layouts(N);
    }
    if (lossOut) {
        std::vector<float> lp(lb);
        if (hipMemcpyAsync(lp.data(), N.lossPart, lb * 4, hipMemcpyDeviceToHost, N.stream) != hipSuccess ||
            hipStreamSynchronize(N.stream) != hipSuccess) {
            rt::set_error("cudaBVH_denoiserTrainStep: loss copy failed");
            return -1;
        }
        double s = 0;
        for (float x : lp) s += x;
        *lossOut = (float)(s / (3.0 * P));
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
