// fdsurr.hip - gfx950 BF16 MFMA 3x3 convolutions of the phase-field
// surrogate (declarations and conventions: include/fdsurr.h).
//
// All three passes are GEMMs on v_mfma_f32_16x16x32_bf16 (FP32 accumulate).
// Operand layout of one 16x16x32 MFMA (measured on MI355X):
//   A[i][k] (16 x 32): lane i + 16*(k/8), element k%8
//   B[k][j] (32 x 16): lane j + 16*(k/8), element k%8
//   D[i][j] (16 x 16): lane j + 16*(i/4), element i%4
// so a lane holds 8 consecutive k of one row (A) or column (B): in NHWC
// with C % 8 == 0 those are 8 channels of one pixel and one tap - a single
// 16-byte load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "fdsurr.h"

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define FDS_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(FDS_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

__device__ __forceinline__ v8bf bf8(u4v v) { return __builtin_bit_cast(v8bf, v); }

__device__ __forceinline__ uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// The bf16 halves of v where the matching halves of m (a ReLU output) are
// > 0 - for bf16 that is the bit pattern read as a positive int16 - and 0
// elsewhere: the ReLU's backward mask applied on load.
__device__ __forceinline__ u4v relu_keep(u4v v, u4v m) {
    u4v r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const unsigned lo = ((int16_t)(m[i] & 0xFFFFu) > 0) ? 0x0000FFFFu : 0u;
        const unsigned hi = ((int16_t)(m[i] >> 16) > 0) ? 0xFFFF0000u : 0u;
        r[i] = v[i] & (lo | hi);
    }
    return r;
}

// ---------------------------------------------------------------- forward
// y[p][co] = act(bias[co] + sum_k A[p][k] w[co][k]), p over N*H*W (NHWC),
// k = tap*Cin + ci with tap = 3*(dh+1) + (dw+1) and A[p][k] = x at pixel
// (n, h+dh, w+dw), channel ci (0 outside the image).  One wave computes
// 16*MP consecutive pixels x 16*NT output channels (MP x NT MFMA tiles, so
// every A fragment is used NT times and every B fragment MP times); the
// four waves of a workgroup take consecutive pixel ranges with the same
// channels, so they share the weight fragments through L1.  The K loop
// takes 32 k per step (one tap and 8 channels per lane, a 16-byte load);
// k beyond 9*Cin and channels beyond Cout load zeros.  With MASK, x is
// multiplied on load by the ReLU mask of xm (the input gradient pass).
// Workgroups of a 1-D grid are dispatched to the 8 XCDs round robin
// (workgroup i runs on XCD i mod 8), and each XCD has its own L2.  The
// logical index returned gives every XCD a contiguous band of logical
// workgroups, so workgroups adjacent in logical order, which read the same
// or overlapping data (the channel blocks of one tile, the taps of one
// pixel slice, neighbouring tiles' halo rows), share an L2.
__device__ __forceinline__ int xcd_band() {
    const int n = (int)gridDim.x, q = n >> 3, rem = n & 7, xcd = (int)blockIdx.x & 7;
    return xcd * q + min(xcd, rem) + ((int)blockIdx.x >> 3);
}

template <int MP, int NT, bool MASK>
__global__ __launch_bounds__(256) void conv_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ xm,
                                                const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                                float* __restrict__ y32, uint16_t* __restrict__ y16, int N, int H,
                                                int W, int Cin, int Cout, int relu) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)N * H * W;
    // logical workgroup = pixel block x channel block, channel blocks adjacent
    const int ncb = (Cout + 16 * NT - 1) / (16 * NT);
    const int L = xcd_band();
    const int pb = L / ncb;
    const int64_t pw = ((int64_t)pb * 4 + wave) * (16 * MP);
    const int co0 = (L - pb * ncb) * (16 * NT);
    const int K = 9 * Cin;
    int hq[MP], wq[MP];
    int64_t img[MP];
#pragma unroll
    for (int g = 0; g < MP; ++g) {
        const int64_t pa = pw + 16 * g + (lane & 15);
        hq[g] = -4;  // out of range for every tap: pixels past the end load zeros
        wq[g] = 0;
        img[g] = 0;
        if (pa < P) {
            wq[g] = (int)(pa % W);
            const int64_t t = pa / W;
            hq[g] = (int)(t % H);
            img[g] = (t / H) * H;
        }
    }
    v4f acc[MP][NT];
#pragma unroll
    for (int g = 0; g < MP; ++g)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[g][t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < K; kc += 32) {
        const int k0 = kc + 8 * (lane >> 4);
        const bool kin = k0 < K;
        const int tap = k0 / Cin, ci = k0 - tap * Cin;
        const int dh = tap / 3 - 1, dw = tap % 3 - 1;
        // every load of the step is issued before any is used, so the
        // waits cover all of them at once
        u4v av[MP], mv[MP], bv[NT];
#pragma unroll
        for (int g = 0; g < MP; ++g) {
            av[g] = u4v{0u, 0u, 0u, 0u};
            mv[g] = u4v{0u, 0u, 0u, 0u};
            const int hh = hq[g] + dh, ww = wq[g] + dw;
            if (kin && hh >= 0 && hh < H && ww >= 0 && ww < W) {
                const int64_t o = ((img[g] + hh) * W + ww) * Cin + ci;
                av[g] = *reinterpret_cast<const u4v*>(x + o);
                if (MASK) mv[g] = *reinterpret_cast<const u4v*>(xm + o);
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = co0 + 16 * t + (lane & 15);
            bv[t] = u4v{0u, 0u, 0u, 0u};
            if (co < Cout && kin) bv[t] = *reinterpret_cast<const u4v*>(w + (int64_t)co * K + k0);
        }
        if (MASK) {
#pragma unroll
            for (int g = 0; g < MP; ++g) av[g] = relu_keep(av[g], mv[g]);
        }
#pragma unroll
        for (int g = 0; g < MP; ++g)
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf8(av[g]), bf8(bv[t]), acc[g][t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = co0 + 16 * t + (lane & 15);
        if (co >= Cout) continue;
        const float b = bias ? bias[co] : 0.f;
#pragma unroll
        for (int g = 0; g < MP; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t p = pw + 16 * g + 4 * (lane >> 4) + r;
                if (p >= P) continue;
                float v = acc[g][t][r] + b;
                if (relu) v = fmaxf(v, 0.f);
                if (y32) y32[p * Cout + co] = v;
                if (y16) y16[p * Cout + co] = to_bf16(v);
            }
    }
}

// Forward convolution with the input staged in LDS (Cin a multiple of 32).
// A workgroup owns an output block of WR*2 image rows x 32 columns of one
// image and WN*WC*16 output channels (WN = 4 / WR).  Per chunk of 32 input
// channels it stages the block's input halo - (WR*2 + 2) rows x 34 columns
// x 32 channels, zeros outside the image, the ReLU mask of xm applied on
// load - in LDS, each pixel padded to 40 channels (80 bytes: the 16 lanes
// of one MFMA row read 16 different bank groups), and every tap's A
// fragments are 16-byte LDS reads at the shifted pixel.  Wave (wr, wn)
// computes rows 2wr, 2wr+1 (four 16-pixel tiles) x WC channel tiles; its B
// fragments (weights) are 16-byte global loads that the L1 shares between
// the waves and workgroups using the same channels.  The next chunk's halo
// is loaded into registers while the current one is consumed (two LDS
// buffers, one barrier per chunk).
constexpr int kHaloC = 40;   // padded channels per staged pixel
constexpr int kHaloW = 34;   // 32 columns + one either side

template <int WR, int WC, bool MASK>
__global__ __launch_bounds__(256) void conv_fwd_lds(const uint16_t* __restrict__ x, const uint16_t* __restrict__ xm,
                                                    const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                                    float* __restrict__ y32, uint16_t* __restrict__ y16, int N, int H,
                                                    int W, int Cin, int Cout, int relu) {
    constexpr int WN = 4 / WR;
    constexpr int BR = 2 * WR;                    // output rows of the block
    constexpr int HR = BR + 2;                    // halo rows
    constexpr int HP = HR * kHaloW;               // halo pixels
    constexpr int PIECES = HP * 4;                // 16-byte pieces (8 channels) per chunk
    constexpr int PPT = (PIECES + 255) / 256;     // pieces per thread
    __shared__ __align__(16) uint16_t halo[2][HP * kHaloC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave % WR, wn = wave / WR;
    const int wblocks = (W + 31) / 32, hblocks = (H + BR - 1) / BR;
    // logical workgroup = tile x channel block, channel blocks adjacent
    const int ncb = (Cout + WN * WC * 16 - 1) / (WN * WC * 16);
    const int L = xcd_band();
    int b = L / ncb;
    const int cb = L - b * ncb;
    const int bw = b % wblocks;
    b /= wblocks;
    const int bh = b % hblocks;
    const int n = b / hblocks;
    const int h0 = bh * BR, w0 = bw * 32;
    const int co0 = cb * (WN * WC * 16) + wn * (WC * 16);
    const int K = 9 * Cin;
    const int64_t img = (int64_t)n * H;

    // staging: piece q = pixel q / 4 of the halo, channels 8 * (q % 4)
    auto fetch = [&](int ci0, u4v v[PPT]) {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int q = threadIdx.x + 256 * j;
            v[j] = u4v{0u, 0u, 0u, 0u};
            if (q < PIECES) {
                const int px = q >> 2, g = q & 3;
                const int hh = h0 - 1 + px / kHaloW, ww = w0 - 1 + px % kHaloW;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W) {
                    const int64_t o = ((img + hh) * W + ww) * Cin + ci0 + 8 * g;
                    v[j] = *reinterpret_cast<const u4v*>(x + o);
                    if (MASK) v[j] = relu_keep(v[j], *reinterpret_cast<const u4v*>(xm + o));
                }
            }
        }
    };
    auto stage = [&](int buf, const u4v v[PPT]) {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int q = threadIdx.x + 256 * j;
            if (q < PIECES)
                *reinterpret_cast<u4v*>(&halo[buf][(q >> 2) * kHaloC + 8 * (q & 3)]) = v[j];
        }
    };

    v4f acc[4][WC];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < WC; ++t) acc[m][t] = v4f{0.f, 0.f, 0.f, 0.f};

    u4v nxt[PPT];
    fetch(0, nxt);
    stage(0, nxt);
    __syncthreads();
    const int nchunks = Cin / 32;
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) fetch(32 * (c + 1), nxt);
        const uint16_t* hb = halo[buf];
        // fragments of tap t + 1 are loaded while tap t's MFMAs run
        auto load_b = [&](int tap, u4v bv[WC]) {
            const int k0 = tap * Cin + 32 * c + 8 * (lane >> 4);
#pragma unroll
            for (int t = 0; t < WC; ++t) {
                const int co = co0 + 16 * t + (lane & 15);
                bv[t] = co < Cout ? *reinterpret_cast<const u4v*>(w + (int64_t)co * K + k0) : u4v{0u, 0u, 0u, 0u};
            }
        };
        auto load_a = [&](int tap, u4v av[4]) {
            const int dh = tap / 3, dw = tap % 3;     // halo offsets (0..2)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                // tile m: block row 2 wr + m / 2, columns 16 (m % 2) .. + 15
                const int hr = 2 * wr + (m >> 1) + dh;
                const int hc = 16 * (m & 1) + (lane & 15) + dw;
                av[m] = *reinterpret_cast<const u4v*>(&hb[(hr * kHaloW + hc) * kHaloC + 8 * (lane >> 4)]);
            }
        };
        u4v bv[WC], av[4], bn[WC], an[4];
        load_b(0, bv);
        load_a(0, av);
#pragma unroll 1
        for (int tap = 0; tap < 9; ++tap) {
            if (tap < 8) {
                load_b(tap + 1, bn);
                load_a(tap + 1, an);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < WC; ++t)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf8(av[m]), bf8(bv[t]), acc[m][t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < WC; ++t) bv[t] = bn[t];
#pragma unroll
            for (int m = 0; m < 4; ++m) av[m] = an[m];
        }
        if (c + 1 < nchunks) {
            stage(buf ^ 1, nxt);
            __syncthreads();
        }
    }
#pragma unroll
    for (int t = 0; t < WC; ++t) {
        const int co = co0 + 16 * t + (lane & 15);
        if (co >= Cout) continue;
        const float bb = bias ? bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int h = h0 + 2 * wr + (m >> 1);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int wc = w0 + 16 * (m & 1) + 4 * (lane >> 4) + r;
                if (h >= H || wc >= W) continue;
                const int64_t p = (img + h) * W + wc;
                float v = acc[m][t][r] + bb;
                if (relu) v = fmaxf(v, 0.f);
                if (y32) y32[p * Cout + co] = v;
                if (y16) y16[p * Cout + co] = to_bf16(v);
            }
        }
    }
}

// wt[ci][tap][co] = w[co][8 - tap][ci]
__global__ __launch_bounds__(256) void flip_weights(const uint16_t* __restrict__ w, uint16_t* __restrict__ wt, int Cin,
                                                    int Cout) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)Cout * 9 * Cin;
    if (i >= total) return;
    const int ci = (int)(i % Cin);
    const int tap = (int)((i / Cin) % 9);
    const int co = (int)(i / (9 * (int64_t)Cin));
    wt[((int64_t)ci * 9 + (8 - tap)) * Cout + co] = w[i];
}

// The bf16 weights of every layer of a network and their flipped copies
// (flip_weights) from the FP32 master weights in one launch: a step's 15
// layers would otherwise take 30 small launches.  Layer l holds elements
// [start[l], start[l + 1]) of the grid.
constexpr int kCastLayers = 16;
struct CastFlipArgs {
    const float* w[kCastLayers];
    uint16_t* w16[kCastLayers];
    uint16_t* wt[kCastLayers];
    int cin[kCastLayers], cout[kCastLayers];
    int64_t start[kCastLayers + 1];
    int n;
};

__global__ __launch_bounds__(256) void cast_flip(CastFlipArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.start[a.n]) return;
    int l = 0;
    while (l + 1 < a.n && i >= a.start[l + 1]) ++l;
    const int64_t j = i - a.start[l];
    const int Cin = a.cin[l], Cout = a.cout[l];
    const uint16_t v = to_bf16(a.w[l][j]);
    a.w16[l][j] = v;
    const int ci = (int)(j % Cin);
    const int tap = (int)((j / Cin) % 9);
    const int co = (int)(j / (9 * (int64_t)Cin));
    a.wt[l][((int64_t)ci * 9 + (8 - tap)) * Cout + co] = v;
}

// ------------------------------------------------------- weight gradient
// dw[co][tap][ci] = sum_p dy[p][co] x[p + d(tap)][ci]: a GEMM whose
// reduction runs over pixels, the strided dimension of both operands.  Each
// 64-pixel step loads dy (64 px x 64 co) and the shifted x (64 px x 64 ci)
// with 16-byte loads (8 channels of one pixel) and stores them TRANSPOSED
// in LDS, [channel][pixel] with rows padded to 72 elements, so that every
// MFMA fragment (8 consecutive pixels of one channel) is one 16-byte LDS
// read.  Lane l of wave v stages the pixel pair 2(l & 31), 2(l & 31) + 1
// of channel group (l >> 5) + 2v, the pair's two values of a channel in
// one 32-bit write: for a fixed element the 64 lanes write two rows 8
// apart and 32 consecutive pixel pairs (16-bit writes of single pixels
// were 3-6 % slower).  dy is multiplied on load by the ReLU mask of dym
// when given.
// A workgroup owns one tap, 64 co x 64 ci and one slice of pixels; wave
// (wrow, wcol) = (v >> 1, v & 1) computes co 32 wrow .. 32 wrow + 31 x ci
// 32 wcol .. 32 wcol + 31 (2 x 2 tiles of 16 x 16, two 32-pixel MFMA steps per stage: four
// fragment reads per four MFMAs).  LDS holds two stages: the next step is
// written into the other buffer while this one's MFMAs run, one barrier per
// step.  The workgroups of tap 0 and the first ci block also form the bias
// gradient's slice sums, sum_p dy[p][co], on the waves with wcol = 0, with
// one more MFMA per co tile against a ones operand (every column of the
// tile then holds the row sum).  Slices write partial sums that reduce_slices
// adds in a fixed order.
constexpr int kStep = 64;
constexpr int kLds = kStep + 8;

__global__ __launch_bounds__(256) void conv_bwd_w(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                                                  const uint16_t* __restrict__ dym, float* __restrict__ part,
                                                  float* __restrict__ bpart, int N, int H, int W, int Cin, int Cout,
                                                  int64_t chunk) {
    __shared__ __align__(16) uint16_t sdy[2][64][kLds];
    __shared__ __align__(16) uint16_t sx[2][64][kLds];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wrow = wave >> 1, wcol = wave & 1;
    const int ncb = (Cout + 63) / 64;
    // logical workgroup = slice x (ci block, tap, co block), the 9 taps and
    // channel blocks of one slice adjacent (they read the same dy and x)
    const int G = ncb * 9 * ((Cin + 63) / 64);
    const int L = xcd_band();
    const int slice = L / G;
    int gy = L - slice * G;
    const int cb = gy % ncb;
    gy /= ncb;
    const int tap = gy % 9;
    const int ib = gy / 9;
    const int co0 = cb * 64, ci0 = ib * 64, dh = tap / 3 - 1, dw = tap % 3 - 1;
    const int64_t P = (int64_t)N * H * W;
    const int64_t pbeg = (int64_t)slice * chunk;
    const int64_t pend = std::min<int64_t>(P, pbeg + chunk);
    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    v4f accb[2] = {v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}};
    const bool with_bias = bpart != nullptr && tap == 0 && ib == 0 && wcol == 0;
    const u4v ones = {0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u};  // eight bf16 1.0
    // Thread t loads channels g8 .. g8 + 7 (g8 = 8 (t & 7)) of pixels
    // pl = t >> 3 and pl + 32 of the step: eight consecutive lanes read one
    // pixel's 64 channels, 128 contiguous bytes.  Lanes t and t ^ 8 hold
    // adjacent pixels 2m, 2m + 1 (m = pl >> 1); one DPP row rotation by 8
    // swaps half of each one's channels, after which the even lane holds both
    // pixels of channels g8 .. g8 + 3 and the odd lane those of g8 + 4 ..
    // g8 + 7, written as 32-bit pixel pairs.
    const int g8 = (threadIdx.x & 7) * 8, pl = threadIdx.x >> 3;
    const bool odd = pl & 1;
    const bool co_in = co0 + g8 < Cout, ci_in = ci0 + g8 < Cin;
    // the next step's operands are loaded once this step's are in LDS, so
    // the loads run during the barrier and the MFMAs
    u4v ldy[2], lm[2], lx[2];
    // (n, h, wc) of the thread's first pixel of the next load, divided out
    // once and then advanced by kStep pixels per step (64-bit divisions per
    // pixel and step cost more than the step's MFMAs)
    int pn, ph, pwc;
    {
        const int64_t p0 = pbeg + pl, t = p0 / W;
        pwc = (int)(p0 - t * W);
        pn = (int)(t / H);
        ph = (int)(t - (int64_t)pn * H);
    }
    auto advance = [&](int& n, int& h, int& wc, int d) {
        wc += d;
        while (wc >= W) {
            wc -= W;
            if (++h == H) {
                h = 0;
                ++n;
            }
        }
    };
    auto load = [&](int64_t pc) {
        int n = pn, h = ph, wc = pwc;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int64_t p = pc + pl + 32 * k;
            if (k) advance(n, h, wc, 32);
            ldy[k] = u4v{0u, 0u, 0u, 0u};
            lx[k] = u4v{0u, 0u, 0u, 0u};
            lm[k] = u4v{~0u, ~0u, ~0u, ~0u};
            if (p < pend) {
                if (co_in) {
                    ldy[k] = *reinterpret_cast<const u4v*>(dy + p * Cout + co0 + g8);
                    if (dym) lm[k] = *reinterpret_cast<const u4v*>(dym + p * Cout + co0 + g8);
                }
                const int hh = h + dh, ww = wc + dw;
                if (ci_in && hh >= 0 && hh < H && ww >= 0 && ww < W)
                    lx[k] = *reinterpret_cast<const u4v*>(x + (((int64_t)n * H + hh) * W + ww) * Cin + ci0 + g8);
            }
        }
        advance(pn, ph, pwc, kStep);
    };
    // LDS layout [channel R][pixel], the 16-byte chunk c (8 pixels) of row R
    // stored at chunk c ^ (R >> 3): the pair writes of the eight channel
    // groups and the fragment reads of 16 consecutive rows spread over the
    // banks
    auto pair_col = [](int R, int m) { return ((((m >> 2) ^ (R >> 3)) << 2) | (m & 3)) * 2; };
    auto swap8 = [](unsigned v) { return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false); };
    auto put = [&](uint16_t (*sb)[kLds], const u4v& v, int k) {
        const unsigned r0 = swap8(odd ? v[0] : v[2]), r1 = swap8(odd ? v[1] : v[3]);
        const unsigned o0 = odd ? v[2] : v[0], o1 = odd ? v[3] : v[1];
        const unsigned e0 = odd ? r0 : o0, e1 = odd ? r1 : o1;  // pixel 2m
        const unsigned d0 = odd ? o0 : r0, d1 = odd ? o1 : r1;  // pixel 2m + 1
        const int m = (pl >> 1) + 16 * k;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int R = g8 + 4 * odd + j, sh = 16 * (j & 1);
            const unsigned ev = j < 2 ? e0 : e1, ov = j < 2 ? d0 : d1;
            *reinterpret_cast<unsigned*>(&sb[R][pair_col(R, m)]) = ((ev >> sh) & 0xFFFFu) | (((ov >> sh) & 0xFFFFu) << 16);
        }
    };
    // LDS stage from the load registers (the ReLU mask applied here)
    auto stage = [&](int buf) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            put(sdy[buf], dym ? relu_keep(ldy[k], lm[k]) : ldy[k], k);
            put(sx[buf], lx[k], k);
        }
    };
    if (pbeg < pend) {
        load(pbeg);
        stage(0);
        if (pbeg + kStep < pend) load(pbeg + kStep);
    }
    __syncthreads();
    int buf = 0;
    for (int64_t pc = pbeg; pc < pend; pc += kStep) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const int kc = 4 * kk + (lane >> 4);  // 8-pixel chunk of the fragment
            v8bf a[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ra = 32 * wrow + 16 * i + (lane & 15), rb = 32 * wcol + 16 * i + (lane & 15);
                a[i] = bf8(*reinterpret_cast<const u4v*>(&sdy[buf][ra][8 * (kc ^ (ra >> 3))]));
                bv[i] = bf8(*reinterpret_cast<const u4v*>(&sx[buf][rb][8 * (kc ^ (rb >> 3))]));
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bv[j], acc[i][j], 0, 0, 0);
            if (with_bias) {
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    accb[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bf8(ones), accb[i], 0, 0, 0);
            }
        }
        // the other buffer was last read before the previous barrier
        if (pc + kStep < pend) {
            stage(buf ^ 1);
            if (pc + 2 * kStep < pend) load(pc + 2 * kStep);
        }
        __syncthreads();
        buf ^= 1;
    }
    const int64_t K = 9 * (int64_t)Cin;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int ci = ci0 + 32 * wcol + 16 * j + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + 32 * wrow + 16 * i + 4 * (lane >> 4) + r;
                if (co < Cout && ci < Cin) part[((int64_t)slice * Cout + co) * K + tap * Cin + ci] = acc[i][j][r];
            }
        }
    if (with_bias && (lane & 15) == 0) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + 32 * wrow + 16 * i + 4 * (lane >> 4) + r;
                if (co < Cout) bpart[(int64_t)slice * Cout + co] = accb[i][r];
            }
    }
}

// Weight gradient of thin layers (Cout <= 32): one workgroup owns a block
// of up to 32 input channels, all output channels and all 9 taps, and a
// slice of row segments (64 consecutive pixels of one image row).  Per
// segment it stages dy (transposed, [co][pixel]) and the three x rows the
// taps reach, each in three copies shifted by dw = -1, 0, +1
// ([dh][dw][ci][pixel]), so every tap's fragments are aligned 16-byte LDS
// reads and x is read once for all taps.  Staging writes pixel pairs as
// 32-bit words, 32 consecutive pairs of two channel rows 8 apart per wave
// instruction.  Wave v computes taps v, v+4, v+8
// (< 9), 16 x 16 tiles skipped where the channels end; wave 3 of the first
// channel block also forms the bias sums (ones operand).  Partial sums per
// slice, added by reduce_slices.
constexpr int kSmallCo = 32, kSmallCi = 32;

__global__ __launch_bounds__(256) void conv_bwd_w_small(const uint16_t* __restrict__ x,
                                                        const uint16_t* __restrict__ dy,
                                                        const uint16_t* __restrict__ dym, float* __restrict__ part,
                                                        float* __restrict__ bpart, int N, int H, int W, int Cin,
                                                        int Cout, int64_t segs_per_slice) {
    __shared__ __align__(16) uint16_t sdy[kSmallCo][kLds];
    __shared__ __align__(16) uint16_t sx[9][kSmallCi][kLds];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // logical workgroup = slice x ci block, the ci blocks of one slice adjacent
    const int nb = (Cin + kSmallCi - 1) / kSmallCi;
    const int L = xcd_band();
    const int slice = L / nb, ib = L - slice * nb;
    const int ci0 = ib * kSmallCi;
    const int wsegs = (W + 63) / 64;
    const int64_t nseg = (int64_t)N * H * wsegs;
    const int64_t sbeg = (int64_t)slice * segs_per_slice;
    const int64_t send = std::min<int64_t>(nseg, sbeg + segs_per_slice);
    const bool co2 = Cout > 16, ci2 = ci0 + 16 < Cin;
    const bool with_bias = bpart != nullptr && ib == 0 && wave == 3;
    const u4v ones = {0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u};
    v4f acc[3][2][2];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[j][a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    v4f accb[2] = {v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}};
    // The next segment's operands are loaded once this segment's are in LDS,
    // so the loads run during the barrier and the MFMAs.
    // dy: 32 pixel pairs x 4 groups of 8 output channels on threads 0..127,
    // lane t of group g taking pixels 2t, 2t + 1.
    // x: 3 rows x 32 pixel pairs x 4 groups of 8 input channels, items
    // i = (r, g, t) = thread and thread + 256 (threads 0..127); item i loads
    // columns w0 + 2t - 1 .. w0 + 2t + 2 and writes pixels 2t, 2t + 1 of the
    // three copies (copy dw holds column pp + dw at pixel pp), each
    // channel's pair as one 32-bit write.
    const bool low = threadIdx.x < 128;  // stages dy and takes a second x item
    u4v ldy[2], lm[2], lx[2][4];
    auto load = [&](int64_t sg) {
        const int w0 = (int)(sg % wsegs) * 64;
        const int64_t nh = sg / wsegs;
        const int h = (int)(nh % H);
        const int64_t n = nh / H;
        if (low) {
            const int t = threadIdx.x & 31, g = threadIdx.x >> 5;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                ldy[k] = u4v{0u, 0u, 0u, 0u};
                lm[k] = u4v{~0u, ~0u, ~0u, ~0u};
                if (w0 + 2 * t + k < W && 8 * g < Cout) {
                    const int64_t o = ((n * H + h) * W + w0 + 2 * t + k) * Cout + 8 * g;
                    ldy[k] = *reinterpret_cast<const u4v*>(dy + o);
                    if (dym) lm[k] = *reinterpret_cast<const u4v*>(dym + o);
                }
            }
        }
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            if (it == 1 && !low) break;
            const int i = threadIdx.x + 256 * it;
            const int t = i & 31, g = (i >> 5) & 3, r = i >> 7;
            const int hh = h + r - 1;
            const bool rowok = hh >= 0 && hh < H && ci0 + 8 * g < Cin;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int ww = w0 + 2 * t - 1 + k;
                lx[it][k] = u4v{0u, 0u, 0u, 0u};
                if (rowok && ww >= 0 && ww < W)
                    lx[it][k] = *reinterpret_cast<const u4v*>(x + ((n * H + hh) * W + ww) * Cin + ci0 + 8 * g);
            }
        }
    };
    if (sbeg < send) load(sbeg);
    for (int64_t sg = sbeg; sg < send; ++sg) {
        u4v vdy[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) vdy[k] = dym ? relu_keep(ldy[k], lm[k]) : ldy[k];
        if (low) {
            const int t = threadIdx.x & 31, g = threadIdx.x >> 5;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int sh = 16 * (e & 1);
                *reinterpret_cast<unsigned*>(&sdy[8 * g + e][2 * t]) =
                    ((vdy[0][e >> 1] >> sh) & 0xFFFFu) | (((vdy[1][e >> 1] >> sh) & 0xFFFFu) << 16);
            }
        }
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            if (it == 1 && !low) break;
            const int i = threadIdx.x + 256 * it;
            const int t = i & 31, g = (i >> 5) & 3, r = i >> 7;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int sh = 16 * (e & 1);
                    *reinterpret_cast<unsigned*>(&sx[3 * r + d][8 * g + e][2 * t]) =
                        ((lx[it][d][e >> 1] >> sh) & 0xFFFFu) | (((lx[it][d + 1][e >> 1] >> sh) & 0xFFFFu) << 16);
                }
        }
        if (sg + 1 < send) load(sg + 1);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const int kp = 32 * kk + 8 * (lane >> 4);
            v8bf a[2];
            a[0] = bf8(*reinterpret_cast<const u4v*>(&sdy[lane & 15][kp]));
            a[1] = bf8(*reinterpret_cast<const u4v*>(&sdy[16 + (lane & 15)][kp]));
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int tap = wave + 4 * j;
                if (tap >= 9) continue;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (b == 1 && !ci2) continue;
                    const v8bf bv = bf8(*reinterpret_cast<const u4v*>(&sx[tap][16 * b + (lane & 15)][kp]));
                    acc[j][0][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bv, acc[j][0][b], 0, 0, 0);
                    if (co2) acc[j][1][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bv, acc[j][1][b], 0, 0, 0);
                }
            }
            if (with_bias) {
                accb[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf8(ones), accb[0], 0, 0, 0);
                if (co2) accb[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bf8(ones), accb[1], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int64_t K = 9 * (int64_t)Cin;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int tap = wave + 4 * j;
        if (tap >= 9) continue;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int ci = ci0 + 16 * b + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int co = 16 * a + 4 * (lane >> 4) + r;
                    if (co < Cout && ci < Cin && ci < ci0 + kSmallCi)
                        part[((int64_t)slice * Cout + co) * K + tap * Cin + ci] = acc[j][a][b][r];
                }
            }
    }
    if (with_bias && (lane & 15) == 0) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = 16 * a + 4 * (lane >> 4) + r;
                if (co < Cout) bpart[(int64_t)slice * Cout + co] = accb[a][r];
            }
    }
}

// out[i] = sum over s of part[s][i]: eight interleaved partial sums a_j
// (slices s = j mod 8 in ascending order, from 0), then a fixed tree of the
// eight; the order is the same on every call.  A workgroup takes 32
// consecutive outputs, thread t the partial sum j = t >> 5 of output t & 31
// (eight times the threads of one per output: the sums are long, the
// outputs few), and the tree runs from LDS.
__global__ __launch_bounds__(256) void reduce_slices(const float* __restrict__ part, float* __restrict__ out,
                                                     int64_t n, int S) {
    __shared__ float sa[8][33];
    const int j = threadIdx.x >> 5, ii = threadIdx.x & 31;
    const int64_t i = (int64_t)blockIdx.x * 32 + ii;
    float a = 0.f;
    if (i < n) {
#pragma unroll 4
        for (int s = j; s < S; s += 8) a += part[(int64_t)s * n + i];
    }
    sa[j][ii] = a;
    __syncthreads();
    if (j == 0 && i < n)
        out[i] = ((sa[0][ii] + sa[1][ii]) + (sa[2][ii] + sa[3][ii])) + ((sa[4][ii] + sa[5][ii]) + (sa[6][ii] + sa[7][ii]));
}

// ------------------------------------------------------ U-Net plumbing
// The decoder input: out[n][h][w][0:C1] = lo[n][h/2][w/2][:] (nearest-
// neighbour upsampling of the coarser level) and out[n][h][w][C1:C1+C2] =
// skip[n][h][w][:] (the skip connection), one pass of 16-byte pieces of 8
// channels (G1 = C1 / 8, G2 = C2 / 8 pieces per pixel); n pieces in all.
__global__ __launch_bounds__(256) void up_cat(const uint16_t* __restrict__ lo, const uint16_t* __restrict__ skip,
                                              uint16_t* __restrict__ out, int H, int W, int G1, int G2, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int G = G1 + G2;
    const int64_t p = i / G;
    const int g = (int)(i - p * G);
    u4v v;
    if (g < G1) {
        const int w = (int)(p % W);
        const int64_t t = p / W;
        const int h = (int)(t % H);
        const int64_t nn = t / H;
        v = reinterpret_cast<const u4v*>(lo)[((nn * (H / 2) + h / 2) * (W / 2) + w / 2) * G1 + g];
    } else {
        v = reinterpret_cast<const u4v*>(skip)[p * G2 + (g - G1)];
    }
    reinterpret_cast<u4v*>(out)[i] = v;
}

__device__ __forceinline__ float bf_lo(unsigned v) { return __builtin_bit_cast(float, v << 16); }
__device__ __forceinline__ float bf_hi(unsigned v) { return __builtin_bit_cast(float, v & 0xFFFF0000u); }

// Backward of up_cat: dlo[n][h][w][c] = the sum of dout[n][2h+a][2w+b][c]
// over the 2 x 2 block, taken in the order of the gradients of the two
// repeats (along w, then along h): the two columns of each row summed and
// rounded to bf16, then the two rows summed and rounded; dskip =
// dout[..., C1:].  With lo_y (lo, the output of a ReLU), dlo is zeroed where
// lo_y is not > 0: that ReLU's backward.  Pieces [0, n_lo) are dlo's,
// [n_lo, n_lo + n_skip) dskip's.
__global__ __launch_bounds__(256) void up_cat_bwd(const uint16_t* __restrict__ dout,
                                                  const uint16_t* __restrict__ lo_y, uint16_t* __restrict__ dlo,
                                                  uint16_t* __restrict__ dskip, int H, int W, int G1, int G2,
                                                  int64_t n_lo, int64_t n_skip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int G = G1 + G2;
    const u4v* d = reinterpret_cast<const u4v*>(dout);
    if (i < n_lo) {
        const int64_t q = i / G1;
        const int g = (int)(i - q * G1);
        const int w2 = W / 2, h2 = H / 2;
        const int w = (int)(q % w2);
        const int64_t t = q / w2;
        const int h = (int)(t % h2);
        const int64_t nn = t / h2;
        const int64_t p00 = (nn * H + 2 * h) * W + 2 * w;
        const u4v a = d[p00 * G + g], b = d[(p00 + 1) * G + g];
        const u4v c = d[(p00 + W) * G + g], e = d[(p00 + W + 1) * G + g];
        u4v r;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float t0 = bf_lo(to_bf16(bf_lo(a[k]) + bf_lo(b[k]))), t1 = bf_lo(to_bf16(bf_lo(c[k]) + bf_lo(e[k])));
            const float u0 = bf_lo(to_bf16(bf_hi(a[k]) + bf_hi(b[k]))), u1 = bf_lo(to_bf16(bf_hi(c[k]) + bf_hi(e[k])));
            r[k] = (unsigned)to_bf16(t0 + t1) | ((unsigned)to_bf16(u0 + u1) << 16);
        }
        if (lo_y) {
            const u4v m = reinterpret_cast<const u4v*>(lo_y)[i];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                r[k] &= (bf_lo(m[k]) > 0.f ? 0x0000FFFFu : 0u) | (bf_hi(m[k]) > 0.f ? 0xFFFF0000u : 0u);
        }
        reinterpret_cast<u4v*>(dlo)[i] = r;
    } else if (i < n_lo + n_skip) {
        const int64_t j = i - n_lo;
        const int64_t p = j / G2;
        const int g = (int)(j - p * G2);
        reinterpret_cast<u4v*>(dskip)[j] = d[p * G + G1 + g];
    }
}

// out = dy where y > 0 (as a float: -0 and NaN are not), else 0: the
// backward mask of a ReLU whose bf16 output is y; n8 pieces of 8 values.
__global__ __launch_bounds__(256) void relu_mask(const uint16_t* __restrict__ dy, const uint16_t* __restrict__ y,
                                                 uint16_t* __restrict__ out, int64_t n8) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const u4v v = reinterpret_cast<const u4v*>(dy)[i], m = reinterpret_cast<const u4v*>(y)[i];
    u4v r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned lo = bf_lo(m[k]) > 0.f ? 0x0000FFFFu : 0u;
        const unsigned hi = bf_hi(m[k]) > 0.f ? 0xFFFF0000u : 0u;
        r[k] = v[k] & (lo | hi);
    }
    reinterpret_cast<u4v*>(out)[i] = r;
}

// Gradient of a ReLU output s [N][H][W][C] that feeds both a 2 x 2 average
// pool and a skip connection: out = where(s > 0, g, 0) with g = bf16(pool
// gradient + dskip), the pool gradient dp[n][h/2][w/2][c] / 4 rounded to
// bf16 first (the mean's backward), then added to dskip and rounded (the
// sum of a tensor's two gradients), then masked (the ReLU's backward).
// dskip's pixels are Gs pieces of 8 channels apart (a channel slice of a
// wider gradient may be read in place).  One 8-channel piece per thread.
__global__ __launch_bounds__(256) void pool_skip_bwd(const uint16_t* __restrict__ dp,
                                                     const uint16_t* __restrict__ dskip, int Gs,
                                                     const uint16_t* __restrict__ s, uint16_t* __restrict__ out,
                                                     int H, int W, int G, int64_t n8) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const int64_t p = i / G;
    const int g = (int)(i - p * G);
    const int w = (int)(p % W);
    const int64_t t = p / W;
    const int h = (int)(t % H);
    const int64_t nn = t / H;
    const int64_t q = (nn * (H / 2) + h / 2) * (W / 2) + w / 2;
    const u4v a = reinterpret_cast<const u4v*>(dp)[q * G + g];
    const u4v b = reinterpret_cast<const u4v*>(dskip)[p * Gs + g], m = reinterpret_cast<const u4v*>(s)[i];
    u4v r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float plo = bf_lo(to_bf16(bf_lo(a[k]) * 0.25f)), phi = bf_lo(to_bf16(bf_hi(a[k]) * 0.25f));
        const unsigned v = (unsigned)to_bf16(plo + bf_lo(b[k])) | ((unsigned)to_bf16(phi + bf_hi(b[k])) << 16);
        r[k] = v & ((bf_lo(m[k]) > 0.f ? 0x0000FFFFu : 0u) | (bf_hi(m[k]) > 0.f ? 0xFFFF0000u : 0u));
    }
    reinterpret_cast<u4v*>(out)[i] = r;
}

int check(int N, int H, int W, int Cin, int Cout) {
    if (N < 1 || H < 1 || W < 1 || Cin < 8 || Cout < 1) return fail(FDS_E_ARG, "bad shape");
    if (Cin % 8) return fail(FDS_E_ARG, "Cin must be a multiple of 8");
    if ((int64_t)N * H * W >= ((int64_t)1 << 40)) return fail(FDS_E_ARG, "too many pixels");
    return FDS_OK;
}

// Work split of the weight gradient.  Thin layers (Cout <= 32) go to
// conv_bwd_w_small with slices of row segments, about 768 workgroups in all
// (three per CU fit its LDS); the others to conv_bwd_w with pixel slices of
// whole 64-pixel steps, about 2048 workgroups (1..256 slices), at least 512
// pixels each.  chunk counts segments or pixels.
struct BwdWPlan {
    bool small;
    int64_t chunk;
    int S;
};

BwdWPlan bwd_w_plan(int N, int H, int W, int Cin, int Cout) {
    BwdWPlan pl{};
    if (Cout <= kSmallCo) {
        const int64_t nseg = (int64_t)N * H * ((W + 63) / 64);
        const int64_t nb = (Cin + kSmallCi - 1) / kSmallCi;
        const int64_t want = std::min<int64_t>(nseg, std::max<int64_t>(1, (768 + nb - 1) / nb));
        pl.small = true;
        pl.chunk = (nseg + want - 1) / want;
        pl.S = (int)((nseg + pl.chunk - 1) / pl.chunk);
        return pl;
    }
    const int64_t P = (int64_t)N * H * W;
    const int64_t per_slice = (int64_t)((Cout + 63) / 64) * 9 * ((Cin + 63) / 64);
    const int64_t want = std::min<int64_t>(256, std::max<int64_t>(1, (2048 + per_slice - 1) / per_slice));
    int64_t c = (P + want - 1) / want;
    c = std::max<int64_t>((c + kStep - 1) / kStep * kStep, 512);
    pl.small = false;
    pl.chunk = c;
    pl.S = (int)((P + c - 1) / c);
    return pl;
}

}  // namespace

namespace {
template <bool MASK>
void launch_fwd(const uint16_t* x, const uint16_t* xm, const uint16_t* w, const float* bias, float* y32, uint16_t* y16,
                int N, int H, int W, int Cin, int Cout, int relu, hipStream_t st) {
    const int64_t P = (int64_t)N * H * W;
    if (Cin % 32 == 0 && Cin >= 64 && Cout >= 32) {
        // input staged in LDS: blocks of 4 rows x 32 columns x 128 channels
        // (WR 2, WC 4) when Cout is a multiple of 128, 8 rows x 32 (WR 4,
        // WC 2) for Cout < 64, else 8 rows x 64 (WR 4, WC 4); the measured
        // per-layer times favour the direct kernel below when Cin is 32
        const unsigned wb = (unsigned)((W + 31) / 32);
        if (Cout < 64) {
            const dim3 grid(wb * (unsigned)((H + 7) / 8) * (unsigned)N * (unsigned)((Cout + 31) / 32));
            hipLaunchKernelGGL((conv_fwd_lds<4, 2, MASK>), grid, dim3(256), 0, st, x, xm, w, bias, y32, y16, N, H, W,
                               Cin, Cout, relu);
        } else if (Cout % 128 == 0) {
            const dim3 grid(wb * (unsigned)((H + 3) / 4) * (unsigned)N * (unsigned)((Cout + 127) / 128));
            hipLaunchKernelGGL((conv_fwd_lds<2, 4, MASK>), grid, dim3(256), 0, st, x, xm, w, bias, y32, y16, N, H, W,
                               Cin, Cout, relu);
        } else {
            const dim3 grid(wb * (unsigned)((H + 7) / 8) * (unsigned)N * (unsigned)((Cout + 63) / 64));
            hipLaunchKernelGGL((conv_fwd_lds<4, 4, MASK>), grid, dim3(256), 0, st, x, xm, w, bias, y32, y16, N, H, W,
                               Cin, Cout, relu);
        }
        return;
    }
    // wave tiles: 64 px x 64 co, 128 px x 32 co, or 128 px x 16 co
    if (Cout >= 64) {
        const dim3 grid((unsigned)((P + 255) / 256) * (unsigned)((Cout + 63) / 64));
        hipLaunchKernelGGL((conv_fwd<4, 4, MASK>), grid, dim3(256), 0, st, x, xm, w, bias, y32, y16, N, H, W, Cin,
                           Cout, relu);
    } else if (Cout >= 32) {
        const dim3 grid((unsigned)((P + 511) / 512) * (unsigned)((Cout + 31) / 32));
        hipLaunchKernelGGL((conv_fwd<8, 2, MASK>), grid, dim3(256), 0, st, x, xm, w, bias, y32, y16, N, H, W, Cin,
                           Cout, relu);
    } else {
        const dim3 grid((unsigned)((P + 511) / 512) * (unsigned)((Cout + 15) / 16));
        hipLaunchKernelGGL((conv_fwd<8, 1, MASK>), grid, dim3(256), 0, st, x, xm, w, bias, y32, y16, N, H, W, Cin,
                           Cout, relu);
    }
}

int conv_fwd_checked(const uint16_t* x, const uint16_t* xm, const uint16_t* w, const float* bias, float* y32,
                     uint16_t* y16, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t relu,
                     void* stream) {
    int rc = check(N, H, W, Cin, Cout);
    if (rc) return rc;
    if (Cout % 8) return fail(FDS_E_ARG, "Cout must be a multiple of 8");
    if (!x || !w || (!y32 && !y16)) return fail(FDS_E_ARG, "NULL pointer");
    if (xm)
        launch_fwd<true>(x, xm, w, bias, y32, y16, N, H, W, Cin, Cout, relu, (hipStream_t)stream);
    else
        launch_fwd<false>(x, nullptr, w, bias, y32, y16, N, H, W, Cin, Cout, relu, (hipStream_t)stream);
    FDS_HIP(hipGetLastError());
    return FDS_OK;
}
}  // namespace

extern "C" {

const char* fds_last_error(void) { return g_err.c_str(); }


int fds_conv3x3_fwd(const uint16_t* x, const uint16_t* w, const float* bias, float* y32, uint16_t* y16, int32_t N,
                    int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t relu, void* stream) {
    return conv_fwd_checked(x, nullptr, w, bias, y32, y16, N, H, W, Cin, Cout, relu, stream);
}

int fds_flip_weights(const uint16_t* w, uint16_t* wt, int32_t Cin, int32_t Cout, void* stream) {
    if (!w || !wt || Cin < 1 || Cout < 1) return fail(FDS_E_ARG, "bad arguments");
    const int64_t total = (int64_t)Cout * 9 * Cin;
    hipLaunchKernelGGL(flip_weights, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, w, wt,
                       Cin, Cout);
    FDS_HIP(hipGetLastError());
    return FDS_OK;
}

int fds_cast_flip(int32_t n, const float* const* w, uint16_t* const* w16, uint16_t* const* wt, const int32_t* cin,
                  const int32_t* cout, void* stream) {
    if (n < 1 || n > kCastLayers) return fail(FDS_E_ARG, "1 to 16 layers expected");
    if (!w || !w16 || !wt || !cin || !cout) return fail(FDS_E_ARG, "NULL pointer");
    CastFlipArgs a{};
    a.n = n;
    a.start[0] = 0;
    for (int l = 0; l < n; ++l) {
        if (!w[l] || !w16[l] || !wt[l]) return fail(FDS_E_ARG, "NULL pointer");
        if (cin[l] < 1 || cout[l] < 1) return fail(FDS_E_ARG, "bad arguments");
        a.w[l] = w[l];
        a.w16[l] = w16[l];
        a.wt[l] = wt[l];
        a.cin[l] = cin[l];
        a.cout[l] = cout[l];
        a.start[l + 1] = a.start[l] + (int64_t)cout[l] * 9 * cin[l];
    }
    hipLaunchKernelGGL(cast_flip, dim3((unsigned)((a.start[n] + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
    FDS_HIP(hipGetLastError());
    return FDS_OK;
}

int fds_conv3x3_bwd_data(const uint16_t* dy, const uint16_t* wt, const uint16_t* relu_y, float* dx32,
                         uint16_t* dx16, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout, void* stream) {
    // the transposed convolution is a forward convolution of dy (Cout
    // channels) with the flipped weights (Cin outputs)
    return conv_fwd_checked(dy, relu_y, wt, nullptr, dx32, dx16, N, H, W, Cout, Cin, 0, stream);
}

int64_t fds_bwd_weight_scratch(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout) {
    const BwdWPlan pl = bwd_w_plan(N, H, W, Cin, Cout);
    return (int64_t)pl.S * Cout * (9 * (int64_t)Cin + 1);
}

int fds_conv3x3_bwd_weight(const uint16_t* x, const uint16_t* dy, const uint16_t* relu_y, float* dw, float* dbias,
                           float* scratch, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout, void* stream) {
    int rc = check(N, H, W, Cin, Cout);
    if (rc) return rc;
    if (Cout % 8) return fail(FDS_E_ARG, "Cout must be a multiple of 8");
    if (!x || !dy || !dw || !scratch) return fail(FDS_E_ARG, "NULL pointer");
    const BwdWPlan pl = bwd_w_plan(N, H, W, Cin, Cout);
    const int S = pl.S;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nw = (int64_t)Cout * 9 * Cin;
    float* bp = dbias ? scratch + (int64_t)S * nw : nullptr;
    if (pl.small) {
        const dim3 grid((unsigned)S * (unsigned)((Cin + kSmallCi - 1) / kSmallCi));
        hipLaunchKernelGGL(conv_bwd_w_small, grid, dim3(256), 0, st, x, dy, relu_y, scratch, bp, N, H, W, Cin, Cout,
                           pl.chunk);
    } else {
        const dim3 grid((unsigned)S * (unsigned)(((Cout + 63) / 64) * 9 * ((Cin + 63) / 64)));
        hipLaunchKernelGGL(conv_bwd_w, grid, dim3(256), 0, st, x, dy, relu_y, scratch, bp, N, H, W, Cin, Cout,
                           pl.chunk);
    }
    FDS_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_slices, dim3((unsigned)((nw + 31) / 32)), dim3(256), 0, st, scratch, dw, nw, S);
    FDS_HIP(hipGetLastError());
    if (dbias) {
        hipLaunchKernelGGL(reduce_slices, dim3((unsigned)((Cout + 31) / 32)), dim3(256), 0, st, bp, dbias,
                           (int64_t)Cout, S);
        FDS_HIP(hipGetLastError());
    }
    return FDS_OK;
}

int fds_up_cat(const uint16_t* lo, const uint16_t* skip, uint16_t* out, int32_t N, int32_t H, int32_t W,
               int32_t C1, int32_t C2, void* stream) {
    if (!lo || !skip || !out) return fail(FDS_E_ARG, "NULL pointer");
    if (N < 1 || H < 2 || W < 2 || (H % 2) || (W % 2)) return fail(FDS_E_ARG, "N >= 1 and even H, W >= 2 expected");
    if (C1 < 8 || C2 < 8 || (C1 % 8) || (C2 % 8)) return fail(FDS_E_ARG, "C1, C2 must be positive multiples of 8");
    const int64_t n = (int64_t)N * H * W * ((C1 + C2) / 8);
    hipLaunchKernelGGL(up_cat, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, lo, skip, out, H,
                       W, C1 / 8, C2 / 8, n);
    FDS_HIP(hipGetLastError());
    return FDS_OK;
}

int fds_up_cat_bwd(const uint16_t* dout, uint16_t* dlo, uint16_t* dskip, int32_t N, int32_t H, int32_t W, int32_t C1,
                   int32_t C2, void* stream) {
    if (!dskip) return fail(FDS_E_ARG, "NULL pointer");
    return fds_up_cat_bwd_masked(dout, nullptr, dlo, dskip, N, H, W, C1, C2, stream);
}

int fds_up_cat_bwd_masked(const uint16_t* dout, const uint16_t* lo, uint16_t* dlo, uint16_t* dskip, int32_t N,
                          int32_t H, int32_t W, int32_t C1, int32_t C2, void* stream) {
    if (!dout || !dlo) return fail(FDS_E_ARG, "NULL pointer");
    if (N < 1 || H < 2 || W < 2 || (H % 2) || (W % 2)) return fail(FDS_E_ARG, "N >= 1 and even H, W >= 2 expected");
    if (C1 < 8 || C2 < 8 || (C1 % 8) || (C2 % 8)) return fail(FDS_E_ARG, "C1, C2 must be positive multiples of 8");
    const int64_t n_lo = (int64_t)N * (H / 2) * (W / 2) * (C1 / 8), n_skip = dskip ? (int64_t)N * H * W * (C2 / 8) : 0;
    hipLaunchKernelGGL(up_cat_bwd, dim3((unsigned)((n_lo + n_skip + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       dout, lo, dlo, dskip, H, W, C1 / 8, C2 / 8, n_lo, n_skip);
    FDS_HIP(hipGetLastError());
    return FDS_OK;
}

int fds_relu_mask(const uint16_t* dy, const uint16_t* y, uint16_t* out, int64_t n, void* stream) {
    if (!dy || !y || !out) return fail(FDS_E_ARG, "NULL pointer");
    if (n < 8 || (n % 8)) return fail(FDS_E_ARG, "n must be a positive multiple of 8");
    hipLaunchKernelGGL(relu_mask, dim3((unsigned)((n / 8 + 255) / 256)), dim3(256), 0, (hipStream_t)stream, dy, y,
                       out, n / 8);
    FDS_HIP(hipGetLastError());
    return FDS_OK;
}

int fds_pool_skip_bwd(const uint16_t* dp, const uint16_t* dskip, int32_t dskip_ld, const uint16_t* s, uint16_t* out,
                      int32_t N, int32_t H, int32_t W, int32_t C, void* stream) {
    if (!dp || !dskip || !s || !out) return fail(FDS_E_ARG, "NULL pointer");
    if (N < 1 || H < 2 || W < 2 || (H % 2) || (W % 2)) return fail(FDS_E_ARG, "N >= 1 and even H, W >= 2 expected");
    if (C < 8 || (C % 8)) return fail(FDS_E_ARG, "C must be a positive multiple of 8");
    if (dskip_ld < C || (dskip_ld % 8)) return fail(FDS_E_ARG, "dskip_ld must be a multiple of 8, at least C");
    const int64_t n8 = (int64_t)N * H * W * (C / 8);
    hipLaunchKernelGGL(pool_skip_bwd, dim3((unsigned)((n8 + 255) / 256)), dim3(256), 0, (hipStream_t)stream, dp,
                       dskip, dskip_ld / 8, s, out, H, W, C / 8, n8);
    FDS_HIP(hipGetLastError());
    return FDS_OK;
}

}  // extern "C"
