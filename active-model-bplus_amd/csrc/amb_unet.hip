// amb_unet.hip -- learned flux-form surrogate: a small periodic U-Net whose
// convolutions run on gfx950 BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulate), with hand-written forward, backward and Adam kernels.  No reference
// counterpart.  The network (3 levels, 32/64/128 channels, 3x3 kernels, GELU)
// predicts a current J so that phi <- phi - div J: mass is conserved exactly by the
// backward-difference divergence, whatever the network outputs.
//
// Layout: activations NHWC bfloat16 (pixel-major, channels contiguous), weights
// [tap][cout][cin] bfloat16 (forward) and [tap][cin][cout] (backward data), FP32
// master weights and Adam moments in one flat vector, tap = 3(dy+1) + (dx+1)
// reading in[y+dy][x+dx] with periodic wrap.
//
// Convolution as implicit GEMM: M = pixels, N = output channels, K = taps x cin.
// Each wave owns MB x NB tiles of 16 x 16 outputs; per K-step of 32 a lane loads one
// 16-byte chunk (8 channels of one tap) per A tile and per B tile straight from
// global memory (L1/L2 resident) into the MFMA operand layout:
//   A: lane l holds A[l%16][8(l/16) .. +7];  B: B[8(l/16) .. +7][l%16];
//   D: lane l holds D[4(l/16) + r][l%16], r = 0..3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "amb.h"
#include "amb_internal.h"
#include "amb_math.h"

namespace amb {
namespace unet {

typedef uint16_t bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Input channels: phi rounded to bfloat16, lamb/2, zeta/5, phi0/0.6, and the rounding
// remainder 256 (phi - bf16(phi)) so that the first convolution sees phi to ~16 bits;
// 11 zero pads.
constexpr int kCin0 = 16, kCinReal = 5;
constexpr float kLoScale = 256.f;
// The head's output is scaled by kJScale: coarse-step currents are 1e-5 .. 1e-2 while
// Adam moves weights by ~lr per step, so unscaled head weights could not settle.
constexpr float kJScale = 1e-3f;
constexpr float kParamScale[3] = {2.0f, 5.0f, 0.6f};
constexpr float kGeluC0 = 0.7978845608028654f, kGeluC1 = 0.044715f;

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bfloat16, round to nearest even: gfx950's v_cvt_pk_bf16_f32 (one instruction
// for two values where the compiler pairs them).
__device__ __forceinline__ bf16 f2bf(float f) { return __builtin_bit_cast(bf16, (__bf16)f); }

// tanh(u) = 1 - 2 / (exp(2u) + 1) with the hardware 2^x (v_exp_f32) and reciprocal
// (v_rcp_f32): saturates to +-1 for large |u|, absolute error ~1e-7 near 0, far below
// the bfloat16 rounding the activations get.
__device__ __forceinline__ float fast_tanh(float u)
{
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * u);  // 2^(2 u log2 e)
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ __forceinline__ float gelu(float x)
{
    return 0.5f * x * (1.0f + fast_tanh(kGeluC0 * (x + kGeluC1 * x * x * x)));
}

__device__ __forceinline__ float gelu_grad(float x)
{
    const float t = fast_tanh(kGeluC0 * (x + kGeluC1 * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluC0 * (1.0f + 3.0f * kGeluC1 * x * x);
}

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 load8u(const bf16* p) { return *reinterpret_cast<const u32x4*>(p); }
__device__ __forceinline__ float lo_f(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi_f(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// ---- convolution forward / backward-data --------------------------------------------

enum ConvMode { kOutBf16 = 0, kOutBf16Gelu = 1, kOutF32 = 2 };
// Where the A operand comes from: a plain NHWC tensor; the 2x2 average of a tensor of
// twice the resolution (fused avgpool); up2(coarse) + skip (fused up-sampling add);
// or the network input assembled from phi and the parameters (fused input layer).
// The fused forms round exactly as the separate kernels do, so results are identical.
enum ConvSrc { kSrcPlain = 0, kSrcPool = 1, kSrcUpAdd = 2, kSrcInput = 3 };

struct ConvIn {
    const bf16* a;       // plain / pool (fine tensor) / up-add (coarse tensor)
    const bf16* skip;    // up-add: fine skip tensor
    const float* phi;    // input layer
    const float* params; // input layer: 3 per sample
};

template <int SRC>
__device__ __forceinline__ bf16x8 load_a(const ConvIn& src, int cin, int n, int y, int x, int H, int W, int ci)
{
    if constexpr (SRC == kSrcPlain) {
        return load8(src.a + (((size_t)n * H + y) * W + x) * cin + ci);
    } else if constexpr (SRC == kSrcPool) {
        // 8 channels as 4 words of two bfloat16 each (element 2j in the low half)
        const int W2 = 2 * W;
        const bf16* b = src.a + (((size_t)n * 2 * H + 2 * y) * W2 + 2 * x) * cin + ci;
        const u32x4 v00 = load8u(b), v01 = load8u(b + cin), v10 = load8u(b + (size_t)W2 * cin),
                    v11 = load8u(b + (size_t)W2 * cin + cin);
        u32x4 r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float lo = lo_f(v00[j]) + lo_f(v01[j]) + lo_f(v10[j]) + lo_f(v11[j]);
            const float hi = hi_f(v00[j]) + hi_f(v01[j]) + hi_f(v10[j]) + hi_f(v11[j]);
            r[j] = (uint32_t)f2bf(0.25f * lo) | ((uint32_t)f2bf(0.25f * hi) << 16);
        }
        return __builtin_bit_cast(bf16x8, r);
    } else if constexpr (SRC == kSrcUpAdd) {
        const u32x4 u = load8u(src.a + (((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * cin + ci);
        const u32x4 k = load8u(src.skip + (((size_t)n * H + y) * W + x) * cin + ci);
        u32x4 r;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            r[j] = (uint32_t)f2bf(lo_f(u[j]) + lo_f(k[j])) | ((uint32_t)f2bf(hi_f(u[j]) + hi_f(k[j])) << 16);
        return __builtin_bit_cast(bf16x8, r);
    } else {
        bf16x8 r = {};
        if (ci == 0) {
            r[0] = __builtin_bit_cast(__bf16, f2bf(src.phi[((size_t)n * H + y) * W + x]));
            r[1] = __builtin_bit_cast(__bf16, f2bf(src.params[3 * n + 0] / kParamScale[0]));
            r[2] = __builtin_bit_cast(__bf16, f2bf(src.params[3 * n + 1] / kParamScale[1]));
            r[3] = __builtin_bit_cast(__bf16, f2bf(src.params[3 * n + 2] / kParamScale[2]));
            const float f = src.phi[((size_t)n * H + y) * W + x];
            r[4] = __builtin_bit_cast(__bf16, f2bf((f - bf2f(f2bf(f))) * kLoScale));
        }
        return r;
    }
}

// out[p][co] = bias[co] + sum_{tap, ci} A[p + shift(tap)][ci] * w[tap][co][ci]
// (flip: shift negated, used for the data gradient with the [tap][cin][cout] layout).
// The next K-step's operands are loaded before the current step's MFMAs issue.
template <int MB, int NB, int MODE, int SRC>
__global__ __launch_bounds__(256) void conv_kernel(const ConvIn src, int cin, const bf16* __restrict__ w,
                                                   const float* __restrict__ bias, int taps, int flip, int H, int W,
                                                   void* __restrict__ out, bf16* __restrict__ pre)
{
    constexpr int COUT = 16 * NB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int pix0 = (blockIdx.x * 4 + wave) * (16 * MB);
    int pn[MB], py[MB], px[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int p = pix0 + 16 * m + l16;
        px[m] = p % W;
        const int t = p / W;
        py[m] = t % H;
        pn[m] = t / H;
    }
    f32x4 acc[MB][NB];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = taps * cin;
    const bf16x8 zero = {};
    bf16x8 a[MB], b[NB];
    auto fetch = [&](int kb, bf16x8* av, bf16x8* bv) {
        const int k0 = kb + 8 * q;
        if (k0 < K) {
            const int tap = k0 / cin, ci = k0 - tap * cin;
            int dy = 0, dx = 0;
            if (taps == 9) {
                dy = tap / 3 - 1;
                dx = tap % 3 - 1;
                if (flip) {
                    dy = -dy;
                    dx = -dx;
                }
            }
#pragma unroll
            for (int m = 0; m < MB; ++m)
                av[m] = load_a<SRC>(src, cin, pn[m], wrapi(py[m] + dy, H), wrapi(px[m] + dx, W), H, W, ci);
#pragma unroll
            for (int n = 0; n < NB; ++n) bv[n] = load8(w + ((size_t)tap * COUT + 16 * n + l16) * cin + ci);
        } else {
#pragma unroll
            for (int m = 0; m < MB; ++m) av[m] = zero;
#pragma unroll
            for (int n = 0; n < NB; ++n) bv[n] = zero;
        }
    };
    fetch(0, a, b);
    for (int kb = 0; kb < K; kb += 32) {
        bf16x8 an[MB], bn[NB];
        if (kb + 32 < K) fetch(kb + 32, an, bn);
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
        if (kb + 32 < K) {
#pragma unroll
            for (int m = 0; m < MB; ++m) a[m] = an[m];
#pragma unroll
            for (int n = 0; n < NB; ++n) b[n] = bn[n];
        }
    }
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            const int co = 16 * n + l16;
            const float bco = MODE == kOutF32 ? 0.f : bias[co];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const size_t idx = (size_t)(pix0 + 16 * m + 4 * q + r) * COUT + co;
                const float v = acc[m][n][r] + bco;
                if constexpr (MODE == kOutF32) {
                    static_cast<float*>(out)[idx] = v;
                } else if constexpr (MODE == kOutBf16Gelu) {
                    if (pre) pre[idx] = f2bf(v);
                    static_cast<bf16*>(out)[idx] = f2bf(gelu(v));
                } else {
                    static_cast<bf16*>(out)[idx] = f2bf(v);
                }
            }
        }
}

// Inference convolution (3x3, periodic) on a TY x 16 pixel tile per block.  The block
// first stages its (TY + 2) x 18 input pixels (halo included) in LDS, applying the
// operand transform of SRC (input assembly, 2x2 average pooling or up-sampling add)
// once per pixel; the 3x3 taps then read A fragments from LDS.  The block's waves form
// WM row groups x WN channel groups x KW K groups: each accumulates TY / WM row tiles
// of 16 pixels against 16 NB / WN output channels over every KW-th 32-wide K-step of
// K = 9 CIN, with weight fragments of up to five of its K-steps ahead in flight; K
// groups are summed through LDS in a fixed order.  Output is GELU(conv + bias) in
// bfloat16; NB2 > 0 instead applies a fused 1x1 convolution to 16 NB2 channels (no
// activation) to the bfloat16 GELU tile kept in LDS.  Small tiles give many blocks
// per CU, whose memory latencies then overlap.
template <int SRC, int CIN, int NB, int TY, int WM, int WN, int NB2 = 0, int KW = 1>
__global__ __launch_bounds__(64 * WM * WN * KW) void conv_tile_kernel(const ConvIn src, const bf16* __restrict__ w,
                                                                      const float* __restrict__ bias, int H, int W,
                                                                      bf16* __restrict__ out,
                                                                      const bf16* __restrict__ w2,
                                                                      const float* __restrict__ b2)
{
    constexpr int COUT = 16 * NB, LDP = CIN + 8, PW = 18, K = 9 * CIN, NS = (K + 31) / 32;
    constexpr int MB = TY / WM, NBW = NB / WN, TLD = COUT + 8, NT = 64 * WM * WN * KW, NWAVE = NT / 64;
    constexpr int SPW = (NS + KW - 1) / KW;  // K-steps of the wave with kw = 0
    static_assert(TY % WM == 0 && NB % WN == 0 && CIN % 8 == 0 && NT <= 1024, "tile shape");
    static_assert(NB2 == 0 || COUT % 32 == 0, "fused 1x1 needs COUT % 32 == 0");
    __shared__ __align__(16) bf16 sin[(TY + 2) * PW * LDP];
    __shared__ __align__(16) bf16 tl[NB2 > 0 ? TY * 16 * TLD : 8];
    __shared__ float red[KW > 1 ? KW * TY * 16 * COUT : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int wm = wave % WM, wn = (wave / WM) % WN, kw = wave / (WM * WN);
    const int tiles_x = W / 16, tiles_y = H / TY;
    const int bx = blockIdx.x % tiles_x, t = blockIdx.x / tiles_x, by = t % tiles_y, n = t / tiles_y;
    const int x0 = 16 * bx, y0 = TY * by;

    // stage the padded input tile: every gather of the thread first, then the stores
    constexpr int ITEMS = (TY + 2) * PW * (CIN / 8), PER = (ITEMS + NT - 1) / NT;
    bf16x8 sv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + NT * k;
        if (e < ITEMS) {
            const int c8 = e % (CIN / 8), pp = e / (CIN / 8);
            const int ry = pp / PW, rx = pp - ry * PW;
            sv[k] = load_a<SRC>(src, CIN, n, wrapi(y0 + ry - 1, H), wrapi(x0 + rx - 1, W), H, W, 8 * c8);
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + NT * k;
        if (e < ITEMS) {
            const int c8 = e % (CIN / 8), pp = e / (CIN / 8);
            *reinterpret_cast<bf16x8*>(sin + pp * LDP + 8 * c8) = sv[k];
        }
    }

    const bf16x8 zero = {};
    auto load_b = [&](int st, bf16x8* bb) {
        const int k0 = 32 * st + 8 * q;
#pragma unroll
        for (int j = 0; j < NBW; ++j) bb[j] = zero;
        if (st < NS && k0 < K) {
            const int tap = k0 / CIN, ci = k0 - tap * CIN;
#pragma unroll
            for (int j = 0; j < NBW; ++j) bb[j] = load8(w + ((size_t)tap * COUT + 16 * (wn * NBW + j) + l16) * CIN + ci);
        }
    };
    f32x4 acc[MB][NBW];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int j = 0; j < NBW; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // ring of PF of the wave's K-steps of weight fragments in flight
    constexpr int PF = SPW < 6 ? SPW : 6;
    bf16x8 bq[PF][NBW];
#pragma unroll
    for (int i = 0; i < PF - 1; ++i) load_b(kw + KW * i, bq[i]);
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < SPW; ++jj) {
        if (jj + PF - 1 < SPW) load_b(kw + KW * (jj + PF - 1), bq[(jj + PF - 1) % PF]);
        const int st = kw + KW * jj;
        const int k0 = 32 * st + 8 * q;
        const int tap = k0 / CIN, ci = k0 - tap * CIN;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const bool live = st < NS && k0 < K;
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const int r = wm + WM * m;
            bf16x8 a = zero;
            if (live) a = *reinterpret_cast<const bf16x8*>(sin + ((r + 1 + dy) * PW + l16 + 1 + dx) * LDP + ci);
#pragma unroll
            for (int j = 0; j < NBW; ++j)
                acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq[jj % PF][j], acc[m][j], 0, 0, 0);
        }
    }

    // epilogue: D fragment element e of lane (q, l16) is pixel 4q + e, channel l16
    auto emit = [&](int r, int px, int co, float v) {
        const bf16 g = f2bf(gelu(v + bias[co]));
        if constexpr (NB2 > 0) {
            tl[(r * 16 + px) * TLD + co] = g;
        } else {
            out[(((size_t)n * H + y0 + r) * W + x0 + px) * COUT + co] = g;
        }
    };
    if constexpr (KW == 1) {
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int j = 0; j < NBW; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) emit(wm + WM * m, 4 * q + e, 16 * (wn * NBW + j) + l16, acc[m][j][e]);
    } else {
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int j = 0; j < NBW; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    red[(kw * TY * 16 + (wm + WM * m) * 16 + 4 * q + e) * COUT + 16 * (wn * NBW + j) + l16] =
                        acc[m][j][e];
        __syncthreads();
        for (int i = threadIdx.x; i < TY * 16 * COUT; i += NT) {
            float v = red[i];
#pragma unroll
            for (int k = 1; k < KW; ++k) v += red[k * TY * 16 * COUT + i];
            const int co = i % COUT, pix = i / COUT;
            emit(pix / 16, pix % 16, co, v);
        }
    }
    if constexpr (NB2 > 0) {
        __syncthreads();
        constexpr int C2 = 16 * NB2;
        // TY row tiles x NB2 channel blocks of the 1x1 output, spread over the waves
        for (int job = wave; job < TY * NB2; job += NWAVE) {
            const int r = job / NB2, j = job % NB2;
            bf16x8 a2[COUT / 32], bw[COUT / 32];
#pragma unroll
            for (int ks = 0; ks < COUT / 32; ++ks) {
                a2[ks] = *reinterpret_cast<const bf16x8*>(tl + (r * 16 + l16) * TLD + 32 * ks + 8 * q);
                bw[ks] = load8(w2 + (size_t)(16 * j + l16) * COUT + 32 * ks + 8 * q);
            }
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < COUT / 32; ++ks) d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[ks], bw[ks], d, 0, 0, 0);
            const float bb = b2[16 * j + l16];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                out[(((size_t)n * H + y0 + r) * W + x0 + 4 * q + e) * C2 + 16 * j + l16] = f2bf(d[e] + bb);
        }
    }
}

// Inference c1 and c2 in one kernel per TY x 16 tile of c2's output (half resolution).
// The block assembles the input on the (2 TY + 6) x 38 full-resolution pixels around the
// tile, computes c1 (3x3, 8 staged channels -> 32, GELU) with MFMA on the (2 TY + 4) x 36
// pixels that the tile's pooled 3x3 operand covers (16-pixel tiles of that region in
// row-major order), keeps them in LDS and writes its own 2 TY x 32 of them to e1 (c7 reads
// e1 as its skip), pools them 2x2 into c2's (TY + 2) x 18 operand tile and computes c2
// (3x3, 32 -> 64, GELU) into e2.  Every c1 and c2 output is formed from the same MFMA
// operands in the same K order, with the same bias, GELU and rounding, as
// conv_tile_kernel forms it, so e1 and e2 equal those of two separate launches bit for
// bit.  Neighbouring blocks recompute the c1 halo (1.69x the c1 work at TY = 4), which
// costs less than the launch it replaces.  TY waves per 16-channel group of c2, 4 groups.
template <int TY>
__global__ __launch_bounds__(256 * TY) void c12_kernel(const float* __restrict__ phi, const float* __restrict__ params,
                                                       const bf16* __restrict__ w1, const float* __restrict__ b1,
                                                       const bf16* __restrict__ w2, const float* __restrict__ b2, int H,
                                                       int W, bf16* __restrict__ e1, bf16* __restrict__ e2)
{
    constexpr int NT = 256 * TY, NWAVE = NT / 64;
    constexpr int IR = 2 * TY + 6, IC = 38, LDP1 = 16;  // input tile: 8 channels + 8 padding per pixel
    constexpr int ER = 2 * TY + 4, EC = 36, LDE = 40;   // c1 tile: 32 channels + 8 padding
    constexpr int NP1 = ER * EC, MT1 = (NP1 + 15) / 16, MB1 = (MT1 + NWAVE - 1) / NWAVE;
    constexpr int PR = TY + 2, PW = 18, LDP2 = 40;      // c2's pooled operand tile
    constexpr int NS2 = 9;                              // c2 K-steps (K = 9 x 32)
    __shared__ __align__(16) bf16 s_in[IR * IC * LDP1];
    __shared__ __align__(16) bf16 s_e1[NP1 * LDE];
    __shared__ __align__(16) bf16 s_p[PR * PW * LDP2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int H2 = H / 2, W2 = W / 2, tiles_x = W2 / 16, tiles_y = H2 / TY;
    const int bx = blockIdx.x % tiles_x, t = blockIdx.x / tiles_x, by = t % tiles_y, n = t / tiles_y;
    const int x0 = 16 * bx, y0 = TY * by;          // c2 tile origin (half resolution)
    const int fy0 = 2 * y0 - 3, fx0 = 2 * x0 - 3;  // input tile origin (full resolution)

    // input assembly, every gather of the thread first
    constexpr int PER1 = (IR * IC + NT - 1) / NT;
    const ConvIn src{nullptr, nullptr, phi, params};
    bf16x8 sv[PER1];
#pragma unroll
    for (int k = 0; k < PER1; ++k) {
        const int i = tid + NT * k;
        if (i < IR * IC) {
            const int ry = i / IC, rx = i - ry * IC;
            sv[k] = load_a<kSrcInput>(src, 8, n, wrapi(fy0 + ry, H), wrapi(fx0 + rx, W), H, W, 0);
        }
    }
    // weight fragments: c1's three K-steps (tap k0 / 8 of lane quarter q, zero past the
    // 72 real K) for both 16-channel tiles, and c2's nine K-steps for this wave's group
    const bf16x8 zero = {};
    bf16x8 bw1[3][2], bw2[NS2];
#pragma unroll
    for (int st = 0; st < 3; ++st) {
        const int k0 = 32 * st + 8 * q;
#pragma unroll
        for (int j = 0; j < 2; ++j) bw1[st][j] = k0 < 72 ? load8(w1 + ((size_t)(k0 / 8) * 32 + 16 * j + l16) * 8) : zero;
    }
    const int wm = wave % TY, wn = wave / TY;
#pragma unroll
    for (int st = 0; st < NS2; ++st) bw2[st] = load8(w2 + ((size_t)st * 64 + 16 * wn + l16) * 32 + 8 * q);
#pragma unroll
    for (int k = 0; k < PER1; ++k) {
        const int i = tid + NT * k;
        if (i < IR * IC) *reinterpret_cast<bf16x8*>(s_in + i * LDP1) = sv[k];
    }

    // c1 on the covered region; lanes past its last pixel read a valid pixel and are not stored
    f32x4 acc1[MB1][2];
    int abase[MB1];
#pragma unroll
    for (int m = 0; m < MB1; ++m) {
        acc1[m][0] = acc1[m][1] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int p = min(16 * (wave + NWAVE * m) + l16, NP1 - 1), er = p / EC, ec = p - er * EC;
        abase[m] = ((er + 1) * IC + ec + 1) * LDP1;  // input coordinates are c1 coordinates + 1
    }
    __syncthreads();
#pragma unroll
    for (int st = 0; st < 3; ++st) {
        const int k0 = 32 * st + 8 * q, tap = k0 / 8;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
        for (int m = 0; m < MB1; ++m) {
            if (wave + NWAVE * m < MT1) {
                bf16x8 a = zero;
                if (k0 < 72) a = *reinterpret_cast<const bf16x8*>(s_in + abase[m] + (dy * IC + dx) * LDP1);
                acc1[m][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw1[st][0], acc1[m][0], 0, 0, 0);
                acc1[m][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw1[st][1], acc1[m][1], 0, 0, 0);
            }
        }
    }
    // D fragment element e of lane (q, l16) is pixel 16 mt + 4q + e, channel 16 j + l16
#pragma unroll
    for (int m = 0; m < MB1; ++m) {
        const int mt = wave + NWAVE * m;
        if (mt < MT1) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int co = 16 * j + l16;
                const float bco = b1[co];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int p = 16 * mt + 4 * q + e;
                    if (p < NP1) s_e1[p * LDE + co] = f2bf(gelu(acc1[m][j][e] + bco));
                }
            }
        }
    }
    __syncthreads();

    // the block's own 2 TY x 32 pixels of c1 to e1, in 16-byte chunks from the LDS tile
    for (int i = tid; i < 2 * TY * 32 * 4; i += NT) {
        const int c4 = i & 3, pix = i >> 2, orow = pix >> 5, ocol = pix & 31;
        *reinterpret_cast<u32x4*>(e1 + (((size_t)n * H + 2 * y0 + orow) * W + 2 * x0 + ocol) * 32 + 8 * c4) =
            *reinterpret_cast<const u32x4*>(s_e1 + ((orow + 2) * EC + ocol + 2) * LDE + 8 * c4);
    }

    // 2x2 average pooling into c2's operand tile, summed and rounded as load_a<kSrcPool>
    constexpr int ITEMS2 = PR * PW * 4, PER2 = (ITEMS2 + NT - 1) / NT;
#pragma unroll
    for (int k = 0; k < PER2; ++k) {
        const int i = tid + NT * k;
        if (i < ITEMS2) {
            const int c4 = i & 3, pp = i >> 2, pr = pp / PW, pc = pp - pr * PW;
            const bf16* b = s_e1 + (2 * pr * EC + 2 * pc) * LDE + 8 * c4;
            const u32x4 v00 = *reinterpret_cast<const u32x4*>(b), v01 = *reinterpret_cast<const u32x4*>(b + LDE),
                        v10 = *reinterpret_cast<const u32x4*>(b + EC * LDE),
                        v11 = *reinterpret_cast<const u32x4*>(b + EC * LDE + LDE);
            u32x4 r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float lo = lo_f(v00[j]) + lo_f(v01[j]) + lo_f(v10[j]) + lo_f(v11[j]);
                const float hi = hi_f(v00[j]) + hi_f(v01[j]) + hi_f(v10[j]) + hi_f(v11[j]);
                r[j] = (uint32_t)f2bf(0.25f * lo) | ((uint32_t)f2bf(0.25f * hi) << 16);
            }
            *reinterpret_cast<u32x4*>(s_p + pp * LDP2 + 8 * c4) = r;
        }
    }
    __syncthreads();

    // c2: row wm of the tile, channels 16 wn .. 16 wn + 15
    f32x4 acc2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < NS2; ++st) {
        const int dy = st / 3 - 1, dx = st % 3 - 1;
        const bf16x8 a =
            *reinterpret_cast<const bf16x8*>(s_p + ((wm + 1 + dy) * PW + l16 + 1 + dx) * LDP2 + 8 * q);
        acc2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw2[st], acc2, 0, 0, 0);
    }
    const int co = 16 * wn + l16;
    const float bco = b2[co];
#pragma unroll
    for (int e = 0; e < 4; ++e)
        e2[(((size_t)n * H2 + y0 + wm) * W2 + x0 + 4 * q + e) * 64 + co] = f2bf(gelu(acc2[e] + bco));
}

// ---- weight gradient ----------------------------------------------------------------
// part[chunk][tap][co][ci] = sum_{p in chunk} dpre[p][co] * in[p + shift(tap)][ci].
// 32-pixel K-steps are staged transposed in LDS so each lane reads its 8 pixels as
// one 16-byte vector.
template <int NBO, int NBI>
__global__ __launch_bounds__(256) void wgrad_kernel(const bf16* __restrict__ dpre, const bf16* __restrict__ in,
                                                    int taps, int H, int W, int chunk, float* __restrict__ part)
{
    constexpr int COUT = 16 * NBO, CIN = 16 * NBI, PS = 128, LD = PS + 8;  // 128 pixels per stage
    constexpr int TILES = NBO * NBI, TPW = (TILES + 3) / 4;
    __shared__ __align__(16) bf16 sA[COUT * LD];
    __shared__ __align__(16) bf16 sB[CIN * LD];
    const int c = blockIdx.x, tap = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int dy = taps == 9 ? tap / 3 - 1 : 0, dx = taps == 9 ? tap % 3 - 1 : 0;
    f32x4 acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int pbeg = c * chunk;
    for (int p0 = pbeg; p0 < pbeg + chunk; p0 += PS) {
        // 16-byte loads (8 channels of one pixel), written transposed into LDS
        for (int e = threadIdx.x; e < PS * (COUT / 8); e += 256) {
            const int pl = e / (COUT / 8), c8 = (e - pl * (COUT / 8)) * 8;
            const u32x4 v = load8u(dpre + (size_t)(p0 + pl) * COUT + c8);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sA[(c8 + 2 * j) * LD + pl] = (bf16)(v[j] & 0xffffu);
                sA[(c8 + 2 * j + 1) * LD + pl] = (bf16)(v[j] >> 16);
            }
        }
        for (int e = threadIdx.x; e < PS * (CIN / 8); e += 256) {
            const int pl = e / (CIN / 8), c8 = (e - pl * (CIN / 8)) * 8;
            const int p = p0 + pl;
            const int x = p % W, t = p / W, y = t % H, n = t / H;
            const u32x4 v = load8u(in + (((size_t)n * H + wrapi(y + dy, H)) * W + wrapi(x + dx, W)) * CIN + c8);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sB[(c8 + 2 * j) * LD + pl] = (bf16)(v[j] & 0xffffu);
                sB[(c8 + 2 * j + 1) * LD + pl] = (bf16)(v[j] >> 16);
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < PS; ks += 32) {
#pragma unroll
            for (int i = 0; i < TPW; ++i) {
                const int tile = wave + 4 * i;
                if (tile < TILES) {
                    const int ib = tile / NBI, jb = tile - ib * NBI;
                    const bf16x8 a = load8(sA + (16 * ib + l16) * LD + ks + 8 * q);
                    const bf16x8 b = load8(sB + (16 * jb + l16) * LD + ks + 8 * q);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float* dst = part + ((size_t)c * taps + tap) * COUT * CIN;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const int tile = wave + 4 * i;
        if (tile < TILES) {
            const int ib = tile / NBI, jb = tile - ib * NBI;
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[(16 * ib + 4 * q + r) * CIN + 16 * jb + l16] = acc[i][r];
        }
    }
}

// Bias gradient partials: part[chunk][co] = sum_{p in chunk} dpre[p][co]; threads
// read consecutive channels of consecutive pixels (coalesced), then reduce in LDS.
__global__ __launch_bounds__(256) void bgrad_kernel(const bf16* __restrict__ dpre, int cout, int chunk,
                                                    float* __restrict__ part)
{
    __shared__ float red[256];
    const int c = blockIdx.x, co = threadIdx.x % cout, row = threadIdx.x / cout, rows = 256 / cout;
    float s = 0.f;
    for (int p = c * chunk + row; p < (c + 1) * chunk; p += rows) s += bf2f(dpre[(size_t)p * cout + co]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < cout) {
        float t = 0.f;
        for (int r = 0; r < rows; ++r) t += red[r * cout + threadIdx.x];
        part[(size_t)c * cout + threadIdx.x] = t;
    }
}

// out[g * len + i] = sum of part[c * len + i] over chunks c of group g (gsize chunks
// per group, blockIdx.y = g); fixed order, deterministic.
__global__ void reduce_chunks_kernel(const float* __restrict__ part, int nchunks, int len, int gsize,
                                     float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const int c0 = blockIdx.y * gsize, c1 = min(nchunks, c0 + gsize);
    float s = 0.f;
#pragma unroll 8
    for (int c = c0; c < c1; ++c) s += part[(size_t)c * len + i];
    out[(size_t)blockIdx.y * len + i] = s;
}

// ---- elementwise and resampling ------------------------------------------------------

__global__ void input_kernel(const float* __restrict__ phi, const float* __restrict__ params, int hw, int npix,
                             bf16* __restrict__ x0)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int n = p / hw;
    bf16* dst = x0 + (size_t)p * kCin0;
    dst[0] = f2bf(phi[p]);
    dst[1] = f2bf(params[3 * n + 0] / kParamScale[0]);
    dst[2] = f2bf(params[3 * n + 1] / kParamScale[1]);
    dst[3] = f2bf(params[3 * n + 2] / kParamScale[2]);
    dst[4] = f2bf((phi[p] - bf2f(f2bf(phi[p]))) * kLoScale);
    for (int c = kCinReal; c < kCin0; ++c) dst[c] = 0;
}

// out (N, H/2, W/2, C) = mean of 2x2 blocks of in (N, H, W, C).
__global__ void avgpool_kernel(const bf16* __restrict__ in, int H, int W, int C, int nout, bf16* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    const int c = i % C, t = i / C, Wo = W / 2, Ho = H / 2;
    const int x = t % Wo, t2 = t / Wo, y = t2 % Ho, n = t2 / Ho;
    const size_t base = (((size_t)n * H + 2 * y) * W + 2 * x) * C + c;
    const float s = bf2f(in[base]) + bf2f(in[base + C]) + bf2f(in[base + (size_t)W * C]) + bf2f(in[base + (size_t)W * C + C]);
    out[i] = f2bf(0.25f * s);
}

// s (N, H, W, C) = up2(u) + skip, u of size (N, H/2, W/2, C).
__global__ void upadd_kernel(const bf16* __restrict__ u, const bf16* __restrict__ skip, int H, int W, int C, int n,
                             bf16* __restrict__ s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = i % C, t = i / C, x = t % W, t2 = t / W, y = t2 % H, b = t2 / H;
    const float v = bf2f(u[(((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]) + bf2f(skip[i]);
    s[i] = f2bf(v);
}

// Up-sampling backward: out (N, H/2, W/2, C) bf16 = sum of the 2x2 block of g.
__global__ void sumpool_kernel(const float* __restrict__ g, int H, int W, int C, int nout, bf16* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    const int c = i % C, t = i / C, Wo = W / 2, Ho = H / 2;
    const int x = t % Wo, t2 = t / Wo, y = t2 % Ho, b = t2 / Ho;
    const size_t base = (((size_t)b * H + 2 * y) * W + 2 * x) * C + c;
    out[i] = f2bf(g[base] + g[base + C] + g[base + (size_t)W * C] + g[base + (size_t)W * C + C]);
}

// dpre = bf16((g + pool_grad) * gelu'(pre)); pool_grad(i) = gc[coarse(i)] / 4 when gc is given.
__global__ void gelu_bwd_kernel(const float* __restrict__ g, const float* __restrict__ gc, const bf16* __restrict__ pre,
                                int H, int W, int C, int n, bf16* __restrict__ dpre)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = g[i];
    if (gc) {
        const int c = i % C, t = i / C, x = t % W, t2 = t / W, y = t2 % H, b = t2 / H;
        v += 0.25f * gc[(((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    }
    dpre[i] = f2bf(v * gelu_grad(bf2f(pre[i])));
}

// ---- head (32 -> 2 channels, FP32 weights and output) ----------------------------------

__global__ void head_fwd_kernel(const bf16* __restrict__ d1, const float* __restrict__ wh, const float* __restrict__ bh,
                                int H, int W, int npix, float* __restrict__ J)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int x = p % W, t = p / W, y = t % H, n = t / H;
    // one accumulator per tap column, summed as in c7_head_flux_kernel (same rounding)
    float j0[3] = {bh[0], 0.f, 0.f}, j1[3] = {bh[1], 0.f, 0.f};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = wrapi(y + tap / 3 - 1, H), xx = wrapi(x + tap % 3 - 1, W);
        const bf16* src = d1 + (((size_t)n * H + yy) * W + xx) * 32;
        const float* w0 = wh + tap * 64;
#pragma unroll
        for (int ci = 0; ci < 32; ++ci) {
            const float v = bf2f(src[ci]);
            j0[tap % 3] = fmaf(v, w0[ci], j0[tap % 3]);
            j1[tap % 3] = fmaf(v, w0[32 + ci], j1[tap % 3]);
        }
    }
    J[2 * (size_t)p] = kJScale * ((j0[0] + j0[1]) + j0[2]);
    J[2 * (size_t)p + 1] = kJScale * ((j1[0] + j1[1]) + j1[2]);
}

// Inference tail in one kernel per TYO x 16 tile: c7 (3x3, 32 -> 32, GELU) on the
// up-sampled u1 plus the skip e1, the head (3x3, 32 -> 2, FP32 weights) and the flux
// update phi_out = phi - div J.  The block stages s1 = up(u1) + e1 over rows
// y0-3 .. y0+TYO+2 and cols x0-3 .. x0+17 in LDS, computes d1 on rows y0-2 .. y0+TYO
// and cols x0-2 .. x0+16 with MFMA (16-pixel tiles of that region in row-major order) into
// LDS, J on rows y0-1 .. y0+TYO-1 and cols x0-1 .. x0+15, then the update of its
// TYO x 16 pixels.  J_out (nullable) receives the tile's J.  Eight waves.
template <int TYO>
__global__ __launch_bounds__(512) void c7_head_flux_kernel(const bf16* __restrict__ u1, const bf16* __restrict__ e1,
                                                           const bf16* __restrict__ w7, const float* __restrict__ b7,
                                                           const float* __restrict__ wh, const float* __restrict__ bh,
                                                           const float* __restrict__ phi, int H, int W,
                                                           float* __restrict__ phi_out, float* __restrict__ J_out)
{
    constexpr int SP = 40;                                 // LDS pixel stride (32 channels + 8 padding)
    constexpr int SW = 21, DW = 19, JW = 17;                // staged, d1 and J region widths
    constexpr int SH = TYO + 6, DH = TYO + 3, JH = TYO + 1;  // their heights
    constexpr int NS = 9;                                  // K-steps of c7 (K = 9 x 32)
    constexpr int NP = DH * DW, MT = (NP + 15) / 16;        // d1 pixels, their 16-pixel tiles
    constexpr int MBW = (MT + 7) / 8;                      // tiles per wave
    static_assert(JH * JW * 2 * 4 <= SH * SW * SP * 2, "J must fit in the staging buffer");
    static_assert(JH * JW <= 256, "one head pixel per thread of each half block");
    constexpr int SPF = 36;                            // d1 pixel stride in floats (32 channels + 4 padding)
    __shared__ __align__(16) bf16 sin[SH * SW * SP];  // later reused for J
    __shared__ __align__(16) float sd[DH * DW * SPF];  // d1, bfloat16 values widened to float for the head
    float* sj = reinterpret_cast<float*>(sin);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    // XCD-aware order (gfx950 dispatches block b to XCD b % 8): the two 16-column tiles of
    // each 32-column strip get blocks B and B + half, which share an XCD.  c12_kernel's block
    // B covers the same strip and rows (TYO = 2 x its 4 half-resolution rows) and has
    // written its e1 there, and the phi written here is what that c12 block reads next step,
    // so both stay in one XCD's L2.  half is a multiple of 8 for grids in multiples of 64.
    const int half = gridDim.x / 2, s32 = blockIdx.x >= half ? 1 : 0, B = blockIdx.x - s32 * half;
    const int strips = W / 32, tilesy = H / TYO;
    const int t2 = B / strips, tx = 2 * (B - t2 * strips) + s32, ty = t2 % tilesy, n = t2 / tilesy;
    const int x0 = tx * 16, y0 = ty * TYO;

    // stage s1 = up(u1) + e1: every gather of the thread first, then the stores
    constexpr int ITEMS = SH * SW * 4, PER = (ITEMS + 511) / 512;
    bf16x8 sv[PER];
    ConvIn src{u1, e1, nullptr, nullptr};
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 512 * k;
        if (i < ITEMS) {
            const int c4 = i & 3, pix = i >> 2, r = pix / SW, c = pix % SW;
            sv[k] = load_a<kSrcUpAdd>(src, 32, n, wrapi(y0 - 3 + r, H), wrapi(x0 - 3 + c, W), H, W, 8 * c4);
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 512 * k;
        if (i < ITEMS) *reinterpret_cast<bf16x8*>(sin + (i >> 2) * SP + 8 * (i & 3)) = sv[k];
    }

    // c7 over the d1 region (rows y0-2 .. y0+TYO, columns x0-2 .. x0+16) taken row-major as
    // NP pixels: tile mt holds pixels 16 mt .. 16 mt + 15, so no tile straddles computed
    // pixels that are not needed.  K-step st is tap st (32 channels), lane quarter q reads
    // channels 8q .. 8q+7; lanes past the last pixel read a valid pixel and are not stored.
    f32x4 acc[MBW][2];
    int abase[MBW];
#pragma unroll
    for (int m = 0; m < MBW; ++m) {
        acc[m][0] = acc[m][1] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int p = min(16 * (wave + 8 * m) + l16, NP - 1), i = p / DW, c = p - i * DW;
        abase[m] = ((i + 1) * SW + c + 1) * SP + 8 * q;  // staged coordinates are d1 coordinates + 1
    }
    auto load_b = [&](int st, bf16x8* bb) {
        const int k0 = 32 * st + 8 * q, tap = k0 / 32, ci = k0 - 32 * tap;
        bb[0] = load8(w7 + ((size_t)tap * 32 + l16) * 32 + ci);
        bb[1] = load8(w7 + ((size_t)tap * 32 + 16 + l16) * 32 + ci);
    };
    constexpr int PF = 6;
    bf16x8 bq[PF][2];
#pragma unroll
    for (int i = 0; i < PF - 1; ++i) load_b(i, bq[i]);
    __syncthreads();
#pragma unroll
    for (int st = 0; st < NS; ++st) {
        if (st + PF - 1 < NS) load_b(st + PF - 1, bq[(st + PF - 1) % PF]);
        const int dy = st / 3 - 1, dx = st % 3 - 1;
#pragma unroll
        for (int m = 0; m < MBW; ++m) {
            if (wave + 8 * m < MT) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(sin + abase[m] + (dy * SW + dx) * SP);
                acc[m][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq[st % PF][0], acc[m][0], 0, 0, 0);
                acc[m][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq[st % PF][1], acc[m][1], 0, 0, 0);
            }
        }
    }
    // D fragment element e of lane (q, l16) is pixel 16 mt + 4q + e, channel 16 j + l16
#pragma unroll
    for (int m = 0; m < MBW; ++m) {
        const int mt = wave + 8 * m;
        if (mt < MT) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int co = 16 * j + l16;
                const float bco = b7[co];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int p = 16 * mt + 4 * q + e;
                    if (p < NP) sd[p * SPF + co] = bf2f(f2bf(gelu(acc[m][j][e] + bco)));
                }
            }
        }
    }
    __syncthreads();

    // head: J at d1 coordinates 1 .. JH (rows y0-1 .. y0+TYO-1), cols 1 .. 17; the first
    // half of the block computes J's component 0, the second half component 1, one pixel
    // per thread.  The component is uniform across each wave, so the weights are read as
    // scalars; summation order as in head_fwd_kernel (one chain per tap column).
    {
        const int ch = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), i = threadIdx.x & 255;
        if (i < JH * JW) {
            const int r = i / JW + 1, c = i % JW + 1;
            float jc[3] = {bh[ch], 0.f, 0.f};
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const float* sp = sd + ((r + tap / 3 - 1) * DW + (c + tap % 3 - 1)) * SPF;
                const float* w0 = wh + tap * 64 + 32 * ch;
#pragma unroll
                for (int c4 = 0; c4 < 8; ++c4) {
                    const f32x4 v = *reinterpret_cast<const f32x4*>(sp + 4 * c4);
#pragma unroll
                    for (int e = 0; e < 4; ++e) jc[tap % 3] = fmaf(v[e], w0[4 * c4 + e], jc[tap % 3]);
                }
            }
            sj[2 * i + ch] = kJScale * ((jc[0] + jc[1]) + jc[2]);
        }
    }
    __syncthreads();
    if (threadIdx.x < TYO * 16) {
        const int r = threadIdx.x / 16, c = threadIdx.x % 16;  // tile pixel
        const int i = (r + 1) * JW + (c + 1), up = r * JW + (c + 1), left = (r + 1) * JW + c;
        const float div = (sj[2 * i] - sj[2 * up]) + (sj[2 * i + 1] - sj[2 * left + 1]);
        const size_t p = ((size_t)n * H + y0 + r) * W + x0 + c;
        phi_out[p] = phi[p] - div;
        if (J_out) {
            J_out[2 * p] = sj[2 * i];
            J_out[2 * p + 1] = sj[2 * i + 1];
        }
    }
}

// dd1[p][ci] = sum_{tap, c} dJ[p - shift(tap)][c] * wh[tap][c][ci]
__global__ void head_bwd_data_kernel(const float* __restrict__ dJ, const float* __restrict__ wh, int H, int W, int npix,
                                     float* __restrict__ dd1)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int x = p % W, t = p / W, y = t % H, n = t / H;
    float acc[32];
#pragma unroll
    for (int ci = 0; ci < 32; ++ci) acc[ci] = 0.f;
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = wrapi(y - (tap / 3 - 1), H), xx = wrapi(x - (tap % 3 - 1), W);
        const size_t q = ((size_t)n * H + yy) * W + xx;
        const float g0 = dJ[2 * q], g1 = dJ[2 * q + 1];
        const float* w0 = wh + tap * 64;
#pragma unroll
        for (int ci = 0; ci < 32; ++ci) acc[ci] = fmaf(g0, w0[ci], fmaf(g1, w0[32 + ci], acc[ci]));
    }
#pragma unroll
    for (int ci = 0; ci < 32; ++ci) dd1[(size_t)p * 32 + ci] = acc[ci];
}

// part[chunk][tap*64 + c*32 + ci] = sum_p dJ[p][c] d1[p + shift(tap)][ci]; part[chunk][576 + c] = sum_p dJ[p][c]
__global__ __launch_bounds__(320) void head_bwd_weight_kernel(const float* __restrict__ dJ, const bf16* __restrict__ d1,
                                                              int H, int W, int chunk, float* __restrict__ part)
{
    const int c = blockIdx.x, t = threadIdx.x;
    if (t >= 290) return;
    float s0 = 0.f, s1 = 0.f;
    if (t < 288) {
        const int tap = t / 32, ci = t % 32, dy = tap / 3 - 1, dx = tap % 3 - 1;
        float a0[4] = {0.f, 0.f, 0.f, 0.f}, a1[4] = {0.f, 0.f, 0.f, 0.f};
        for (int p = c * chunk; p < (c + 1) * chunk; p += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int pp = p + u;
                const int x = pp % W, tt = pp / W, y = tt % H, n = tt / H;
                const float v = bf2f(d1[(((size_t)n * H + wrapi(y + dy, H)) * W + wrapi(x + dx, W)) * 32 + ci]);
                a0[u] = fmaf(dJ[2 * (size_t)pp], v, a0[u]);
                a1[u] = fmaf(dJ[2 * (size_t)pp + 1], v, a1[u]);
            }
        }
        s0 = (a0[0] + a0[1]) + (a0[2] + a0[3]);
        s1 = (a1[0] + a1[1]) + (a1[2] + a1[3]);
        float* dst = part + (size_t)c * 578;
        dst[tap * 64 + ci] = s0;
        dst[tap * 64 + 32 + ci] = s1;
    } else {
        const int comp = t - 288;
        for (int p = c * chunk; p < (c + 1) * chunk; ++p) s0 += dJ[2 * (size_t)p + comp];
        part[(size_t)c * 578 + 576 + comp] = s0;
    }
}

// ---- flux-form update, loss ------------------------------------------------------------

// out = phi - div J with the backward-difference divergence (exactly conservative).
__global__ void flux_update_kernel(const float* __restrict__ phi, const float* __restrict__ J, int H, int W, int npix,
                                   float* __restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int x = p % W, t = p / W, y = t % H, n = t / H;
    const size_t up = ((size_t)n * H + wrapi(y - 1, H)) * W + x;
    const size_t left = ((size_t)n * H + y) * W + wrapi(x - 1, W);
    const float div = (J[2 * (size_t)p] - J[2 * up]) + (J[2 * (size_t)p + 1] - J[2 * left + 1]);
    out[p] = phi[p] - div;
}

constexpr int kAuditBlocks = 256;

// part[2 b] = sum (a - e)^2, part[2 b + 1] = sum e^2 over block b's grid-stride share,
// in FP64 (non-finite values propagate into the sums).
__global__ __launch_bounds__(256) void audit_kernel(const float* __restrict__ a, const float* __restrict__ e, int n,
                                                    double* __restrict__ part)
{
    __shared__ double r0[256], r1[256];
    double s0 = 0.0, s1 = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kAuditBlocks * 256) {
        const double d = (double)a[i] - (double)e[i];
        s0 += d * d;
        s1 += (double)e[i] * e[i];
    }
    r0[threadIdx.x] = s0;
    r1[threadIdx.x] = s1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r0[0];
        part[2 * blockIdx.x + 1] = r1[0];
    }
}

// g = 2 (pred - target) / (norm_n N); part[block] = sum of (pred - target)^2 / norm_n.
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ pred, const float* __restrict__ target,
                                                   const float* __restrict__ norm, int hw, int npix, int batch,
                                                   float* __restrict__ g, float* __restrict__ part)
{
    __shared__ float s[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float v = 0.f;
    if (p < npix) {
        const int n = p / hw;
        const float r = pred[p] - target[p];
        v = r * r / norm[n];
        g[p] = 2.f * r / (norm[n] * batch);
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Gradient with respect to the unscaled head output: kJScale (-div^T g), where
// (-div^T g)_y(i,j) = g(i+1,j) - g(i,j) and (-div^T g)_x(i,j) = g(i,j+1) - g(i,j).
__global__ void dj_kernel(const float* __restrict__ g, int H, int W, int npix, float* __restrict__ dJ)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int x = p % W, t = p / W, y = t % H, n = t / H;
    const float gp = g[p];
    dJ[2 * (size_t)p] = kJScale * (g[((size_t)n * H + wrapi(y + 1, H)) * W + x] - gp);
    dJ[2 * (size_t)p + 1] = kJScale * (g[((size_t)n * H + y) * W + wrapi(x + 1, W)] - gp);
}

// Training-input noise: phi += div_back(eta), eta ~ N(0, sig_n^2) per pixel and
// component with sig_n = rel * sqrt(norm_n / hw) / 2 (norm_n: the sample's loss
// normaliser), so the perturbation is a divergence (sum(phi) unchanged) of about rel
// times the sample's RMS increment.  eta comes from Philox4x32-10 with counter
// {pixel, iteration low, iteration high, 0x6e6f6973} and key seed, Box-Muller.
__device__ __forceinline__ float2 train_eta(uint32_t pix, uint64_t iter, uint32_t k0, uint32_t k1)
{
    U4 c;
    c.x = pix;
    c.y = (uint32_t)iter;
    c.z = (uint32_t)(iter >> 32);
    c.w = 0x6e6f6973u;
    const U4 r = philox4x32_10(c, k0, k1);
    const float u1 = ((float)r.x + 0.5f) * 2.3283064e-10f, u2 = (float)r.y * 2.3283064e-10f;
    const float m = sqrtf(-2.f * logf(u1));
    float sn, cs;
    sincosf(6.2831853f * u2, &sn, &cs);
    return make_float2(m * cs, m * sn);
}

__global__ void train_noise_kernel(float* __restrict__ phi, const float* __restrict__ norm, int H, int W, int npix,
                                   float rel, uint64_t iter, uint32_t k0, uint32_t k1)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int x = p % W, t = p / W, y = t % H, n = t / H;
    const int hw = H * W;
    const float sig = 0.5f * rel * sqrtf(norm[n] / (float)hw);
    const int up = (n * H + wrapi(y - 1, H)) * W + x, left = (n * H + y) * W + wrapi(x - 1, W);
    const float2 e = train_eta((uint32_t)p, iter, k0, k1);
    const float ey_up = train_eta((uint32_t)up, iter, k0, k1).x, ex_left = train_eta((uint32_t)left, iter, k0, k1).y;
    phi[p] += sig * ((e.x - ey_up) + (e.y - ex_left));
}

// ---- optimiser -------------------------------------------------------------------------

__global__ void sumsq_kernel(const float* __restrict__ g, int n, float* __restrict__ part)
{
    __shared__ float s[256];
    float v = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) v += g[i] * g[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Adam with bias correction; the gradient is first scaled by min(1, clip / ||g||).
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, const float* __restrict__ sumsq_part, int nparts, float lr,
                            float b1, float b2, float eps, float bc1, float bc2, float clip)
{
    __shared__ float scale;
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int i = 0; i < nparts; ++i) s += sumsq_part[i];
        const float norm = sqrtf(s);
        scale = (clip > 0.f && norm > clip) ? clip / norm : 1.f;
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * scale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 forward [tap][co][ci] and backward-data [tap][ci][co] copies of a layer's weights.
__global__ void pack_kernel(const float* __restrict__ master, int taps, int cout, int cin, bf16* __restrict__ wf,
                            bf16* __restrict__ wb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= taps * cout * cin) return;
    const int ci = i % cin, t = i / cin, co = t % cout, tap = t / cout;
    const bf16 v = f2bf(master[i]);
    wf[i] = v;
    wb[((size_t)tap * cin + ci) * cout + co] = v;
}

// c1 inference weights: the first 8 of the 16 input channels (the data channels and
// zero pads), [tap][co][8].
__global__ void pack_c1_kernel(const float* __restrict__ master, bf16* __restrict__ w8)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * 32 * 8) return;
    const int ci = i % 8, t = i / 8;  // t = tap * 32 + co
    w8[i] = f2bf(master[(size_t)t * kCin0 + ci]);
}

__global__ void gather_kernel(const float* __restrict__ src, const int* __restrict__ idx, int per, int count,
                              float* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per * count) return;
    dst[i] = src[(size_t)idx[i / per] * per + i % per];
}

__global__ void f64_to_f32_kernel(const double* __restrict__ a, int n, float* __restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

}  // namespace unet
}  // namespace amb

using namespace amb;
using namespace amb::unet;

namespace {

struct LayerDef {
    int cin, cout, taps;
    bool act;
};

// c1 .. c7 on matrix cores, then the FP32 head.
const LayerDef kLayers[8] = {{kCin0, 32, 9, true}, {32, 64, 9, true},  {64, 128, 9, true}, {128, 64, 1, false},
                             {64, 64, 9, true},    {64, 32, 1, false}, {32, 32, 9, true},  {32, 2, 9, false}};
constexpr int kEw = 256;  // threads of elementwise kernels

unsigned blocks_for(size_t n, int t = kEw) { return (unsigned)((n + t - 1) / t); }

constexpr int kReduceGroup = 64;  // chunks summed per thread in each level of the reduction

// Pixels per weight-gradient partial: the largest of 4096 .. 256 dividing npix
// (npix is always a multiple of 256 since grid sides are multiples of 64).
int chunk_for(int npix)
{
    for (int c : {4096, 2048, 1024, 512, 256, 128})
        if (npix % c == 0) return c;
    return 0;
}

}  // namespace

struct amb_surrogate {
    int H = 0, W = 0, max_batch = 0;
    hipStream_t stream = nullptr;
    size_t w_off[8], b_off[8], nparams = 0;
    float* master = nullptr;  // FP32 weights and biases (flat)
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int64_t adam_t = 0;
    bf16* wf[7] = {};
    bf16* wf0_8 = nullptr;  // c1 weights [tap][co][8] over the first 8 input channels (inference)
    bf16* wb[7] = {};
    // activations (bf16 NHWC) and gradients
    bf16 *x0 = nullptr, *e1 = nullptr, *e1p = nullptr, *p1 = nullptr, *e2 = nullptr, *e2p = nullptr, *p2 = nullptr;
    bf16 *bt = nullptr, *btp = nullptr, *u2 = nullptr, *s2 = nullptr, *d2 = nullptr, *d2p = nullptr, *u1 = nullptr;
    bf16 *s1 = nullptr, *d1 = nullptr, *d1p = nullptr, *dpre = nullptr;
    float *J = nullptr, *phi = nullptr, *target = nullptr, *params = nullptr, *norm = nullptr, *pred = nullptr;
    float *g = nullptr, *dJ = nullptr, *dd1 = nullptr, *ds1 = nullptr, *dd2 = nullptr, *ds2 = nullptr, *dbt = nullptr;
    float *dp2 = nullptr, *dp1 = nullptr, *part = nullptr, *part2 = nullptr, *loss_part = nullptr,
          *sumsq_part = nullptr;
    size_t part_len = 0, part2_len = 0;
    double* audit_part = nullptr;  // kAuditBlocks x (|a - b|^2, |b|^2)
    float* audit_buf = nullptr;    // surrogate field of a pending overlapped audit
    bool audit_pending = false;
    int64_t audit_t0 = 0, audit_steps = 0;
    amb_handle* audit_main = nullptr;
    amb_handle* audit_judge = nullptr;
    float* p3 = nullptr;           // parameter triple of advance / audit (device)
    float last_p3[3] = {NAN, NAN, NAN};
    int* d_idx = nullptr;
    // dataset (device): phi_in, phi_out [S][H][W], params [S][3], norms [S]
    float *ds_in = nullptr, *ds_out = nullptr, *ds_par = nullptr, *ds_norm = nullptr;
    int64_t ds_count = 0;
    std::mt19937_64 rng;
    float train_noise = 0.f;  // relative size of the training-input noise (0: off)
    uint64_t noise_seed = 0, noise_iter = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // two-step graphs (field -> other -> field), one per buffer parity, batch 1
    hipGraphExec_t step_graph[2] = {};
    void* graph_field[2] = {};
    void* graph_out[2] = {};
};

namespace {

std::vector<void*> all_buffers(amb_surrogate* s)
{
    std::vector<void*> v = {s->master, s->grad, s->adam_m, s->adam_v, s->x0, s->e1, s->e1p, s->p1, s->e2, s->e2p,
                            s->p2, s->bt, s->btp, s->u2, s->s2, s->d2, s->d2p, s->u1, s->s1, s->d1, s->d1p, s->dpre,
                            s->J, s->phi, s->target, s->params, s->norm, s->pred, s->g, s->dJ, s->dd1, s->ds1,
                            s->dd2, s->ds2, s->dbt, s->dp2, s->dp1, s->part, s->part2, s->audit_part, s->audit_buf, s->p3,
                            s->loss_part, s->sumsq_part, s->d_idx,
                            s->ds_in, s->ds_out, s->ds_par, s->ds_norm, s->wf0_8};
    for (int l = 0; l < 7; ++l) {
        v.push_back(s->wf[l]);
        v.push_back(s->wb[l]);
    }
    return v;
}

void release(amb_surrogate* s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (hipGraphExec_t g : s->step_graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (void* p : all_buffers(s))
        if (p) (void)hipFree(p);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

template <int MB, int NB, int MODE, int SRC>
hipError_t conv_launch(const ConvIn& in, int cin, const bf16* w, const float* bias, int taps, int flip, int npix, int H,
                       int W, void* out, bf16* pre, hipStream_t st)
{
    hipLaunchKernelGGL((conv_kernel<MB, NB, MODE, SRC>), dim3(npix / (64 * MB)), dim3(256), 0, st, in, cin, w, bias,
                       taps, flip, H, W, out, pre);
    return hipGetLastError();
}

template <int SRC, int CIN, int NB, int TY, int WM, int WN, int NB2 = 0, int KW = 1>
hipError_t conv_tile_launch(const ConvIn& in, const bf16* w, const float* bias, int n, int H, int W, bf16* out,
                            hipStream_t st, const bf16* w2 = nullptr, const float* b2 = nullptr)
{
    if (H % TY || W % 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL((conv_tile_kernel<SRC, CIN, NB, TY, WM, WN, NB2, KW>), dim3(n * (H / TY) * (W / 16)),
                       dim3(64 * WM * WN * KW), 0, st, in, w, bias, H, W, out, w2, b2);
    return hipGetLastError();
}

// Dispatch on output channels (16 NB); MB (16-pixel rows per wave) is 2 for large
// launches and 1 for small.
template <int MODE, int SRC = kSrcPlain>
hipError_t conv(const ConvIn& in, int cin, const bf16* w, const float* bias, int cout, int taps, int flip, int npix,
                int H, int W, void* out, bf16* pre, hipStream_t st)
{
    const bool big = npix >= 16 * 2048;
    switch (cout) {
    case 32:
        return big ? conv_launch<2, 2, MODE, SRC>(in, cin, w, bias, taps, flip, npix, H, W, out, pre, st)
                   : conv_launch<1, 2, MODE, SRC>(in, cin, w, bias, taps, flip, npix, H, W, out, pre, st);
    case 64:
        return big ? conv_launch<2, 4, MODE, SRC>(in, cin, w, bias, taps, flip, npix, H, W, out, pre, st)
                   : conv_launch<1, 4, MODE, SRC>(in, cin, w, bias, taps, flip, npix, H, W, out, pre, st);
    case 128:
        return big ? conv_launch<2, 8, MODE, SRC>(in, cin, w, bias, taps, flip, npix, H, W, out, pre, st)
                   : conv_launch<1, 8, MODE, SRC>(in, cin, w, bias, taps, flip, npix, H, W, out, pre, st);
    default: return hipErrorInvalidValue;
    }
}

ConvIn plain(const bf16* a) { return ConvIn{a, nullptr, nullptr, nullptr}; }

template <int NBO, int NBI>
hipError_t wgrad_launch(const bf16* dpre, const bf16* in, int taps, int H, int W, int npix, float* part, hipStream_t st)
{
    const int chunk = chunk_for(npix);
    hipLaunchKernelGGL((wgrad_kernel<NBO, NBI>), dim3(npix / chunk, taps), dim3(256), 0, st, dpre, in, taps, H, W,
                       chunk, part);
    return hipGetLastError();
}

hipError_t wgrad(const bf16* dpre, int cout, const bf16* in, int cin, int taps, int H, int W, int npix, float* part,
                 hipStream_t st)
{
#define AMB_WG(co, ci) \
    if (cout == 16 * co && cin == 16 * ci) return wgrad_launch<co, ci>(dpre, in, taps, H, W, npix, part, st);
    AMB_WG(2, 1)  // c1: 32 <- 16
    AMB_WG(4, 2)  // c2: 64 <- 32
    AMB_WG(8, 4)  // c3: 128 <- 64
    AMB_WG(4, 8)  // c4: 64 <- 128
    AMB_WG(4, 4)  // c5: 64 <- 64
    AMB_WG(2, 4)  // c6: 32 <- 64
    AMB_WG(2, 2)  // c7: 32 <- 32
#undef AMB_WG
    return hipErrorInvalidValue;
}

// out[i] = sum over nch chunks of s->part[c * len + i]: one pass for up to 64 chunks,
// else groups of 64 into s->part2 and then the groups (a tree of fixed shape).
hipError_t reduce_parts(amb_surrogate* s, int nch, int len, float* out)
{
    if (nch <= kReduceGroup) {
        hipLaunchKernelGGL(reduce_chunks_kernel, dim3(blocks_for(len), 1), dim3(kEw), 0, s->stream, s->part, nch, len,
                           nch, out);
        return hipGetLastError();
    }
    const int groups = (nch + kReduceGroup - 1) / kReduceGroup;
    if ((size_t)groups * len > s->part2_len) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3(blocks_for(len), groups), dim3(kEw), 0, s->stream, s->part, nch, len,
                       kReduceGroup, s->part2);
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3(blocks_for(len), 1), dim3(kEw), 0, s->stream, s->part2, groups, len,
                       groups, out);
    return hipGetLastError();
}

// Weight and bias gradient of conv layer l into s->grad.
int layer_grads(amb_surrogate* s, int l, const bf16* dpre, const bf16* in, int H, int W, int npix)
{
    const LayerDef& L = kLayers[l];
    const int chunk = chunk_for(npix), nch = npix / chunk;
    const int len = L.taps * L.cout * L.cin;
    hipError_t e = wgrad(dpre, L.cout, in, L.cin, L.taps, H, W, npix, s->part, s->stream);
    if (e == hipSuccess) e = reduce_parts(s, nch, len, s->grad + s->w_off[l]);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(bgrad_kernel, dim3(npix / 256), dim3(256), 0, s->stream, dpre, L.cout, 256, s->part);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = reduce_parts(s, npix / 256, L.cout, s->grad + s->b_off[l]);
    return e == hipSuccess ? AMB_OK : fail_hip(e, "surrogate weight gradient");
}

int pack_weights(amb_surrogate* s)
{
    for (int l = 0; l < 7; ++l) {
        const LayerDef& L = kLayers[l];
        const int n = L.taps * L.cout * L.cin;
        hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(n)), dim3(kEw), 0, s->stream, s->master + s->w_off[l], L.taps,
                           L.cout, L.cin, s->wf[l], s->wb[l]);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail_hip(e, "weight pack");
    }
    hipLaunchKernelGGL(pack_c1_kernel, dim3(blocks_for(9 * 32 * 8)), dim3(kEw), 0, s->stream, s->master + s->w_off[0],
                       s->wf0_8);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AMB_OK : fail_hip(e, "weight pack");
}

#define SUR_CHECK(expr)                                   \
    do {                                                  \
        const hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return fail_hip(e_, #expr); \
    } while (0)

// Forward pass on s->phi / s->params (batch n) into s->J; with `train` the
// pre-activations are kept for the backward pass.
int forward(amb_surrogate* s, int n, const float* phi, const float* params, float* J)
{
    const int H = s->H, W = s->W, H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
    const int P1 = n * H * W, P2 = n * H2 * W2, P3 = n * H4 * W4;
    hipStream_t st = s->stream;
    const float* M = s->master;
    hipLaunchKernelGGL(input_kernel, dim3(blocks_for(P1)), dim3(kEw), 0, st, phi, params, H * W, P1, s->x0);
    SUR_CHECK(hipGetLastError());
    SUR_CHECK(conv<kOutBf16Gelu>(plain(s->x0), kCin0, s->wf[0], M + s->b_off[0], 32, 9, 0, P1, H, W, s->e1, s->e1p, st));
    hipLaunchKernelGGL(avgpool_kernel, dim3(blocks_for((size_t)P2 * 32)), dim3(kEw), 0, st, s->e1, H, W, 32, P2 * 32, s->p1);
    SUR_CHECK(hipGetLastError());
    SUR_CHECK(conv<kOutBf16Gelu>(plain(s->p1), 32, s->wf[1], M + s->b_off[1], 64, 9, 0, P2, H2, W2, s->e2, s->e2p, st));
    hipLaunchKernelGGL(avgpool_kernel, dim3(blocks_for((size_t)P3 * 64)), dim3(kEw), 0, st, s->e2, H2, W2, 64, P3 * 64, s->p2);
    SUR_CHECK(hipGetLastError());
    SUR_CHECK(conv<kOutBf16Gelu>(plain(s->p2), 64, s->wf[2], M + s->b_off[2], 128, 9, 0, P3, H4, W4, s->bt, s->btp, st));
    SUR_CHECK(conv<kOutBf16>(plain(s->bt), 128, s->wf[3], M + s->b_off[3], 64, 1, 0, P3, H4, W4, s->u2, nullptr, st));
    hipLaunchKernelGGL(upadd_kernel, dim3(blocks_for((size_t)P2 * 64)), dim3(kEw), 0, st, s->u2, s->e2, H2, W2, 64, P2 * 64, s->s2);
    SUR_CHECK(hipGetLastError());
    SUR_CHECK(conv<kOutBf16Gelu>(plain(s->s2), 64, s->wf[4], M + s->b_off[4], 64, 9, 0, P2, H2, W2, s->d2, s->d2p, st));
    SUR_CHECK(conv<kOutBf16>(plain(s->d2), 64, s->wf[5], M + s->b_off[5], 32, 1, 0, P2, H2, W2, s->u1, nullptr, st));
    hipLaunchKernelGGL(upadd_kernel, dim3(blocks_for((size_t)P1 * 32)), dim3(kEw), 0, st, s->u1, s->e1, H, W, 32, P1 * 32, s->s1);
    SUR_CHECK(hipGetLastError());
    SUR_CHECK(conv<kOutBf16Gelu>(plain(s->s1), 32, s->wf[6], M + s->b_off[6], 32, 9, 0, P1, H, W, s->d1, s->d1p, st));
    hipLaunchKernelGGL(head_fwd_kernel, dim3(blocks_for(P1)), dim3(kEw), 0, st, s->d1, M + s->w_off[7], M + s->b_off[7],
                       H, W, P1, J);
    SUR_CHECK(hipGetLastError());
    return AMB_OK;
}

// Inference: phi_out = phi - div J in four launches: c1 (input assembly) with c2 (pooled
// operand); c3 (pooled operand) with c4; c5 (up-sampled operand) with c6; c7
// (up-sampled operand) with the head and the flux update.  Pooling, up-sampling adds and
// the input layer are applied once per pixel while staging a tile in LDS; the 1x1 layers
// run in the epilogue of the 3x3 layer before them.  J_out (nullable) receives the current.
int forward_infer(amb_surrogate* s, int n, const float* phi, const float* params, float* phi_out,
                  float* J_out = nullptr)
{
    const int H = s->H, W = s->W, H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
    const int P1 = n * H * W, P2 = n * H2 * W2, P3 = n * H4 * W4;
    hipStream_t st = s->stream;
    const float* M = s->master;
    // <SRC, CIN, NB, TY, WM, WN, NB2>: small tiles, so that several blocks share a CU
    // and hide each other's memory latency
    static_assert(kCinReal <= 8, "c1 inference packs the data channels into 8");
    constexpr int kC12Rows = 4;  // c7_head_flux_kernel<8> pairs its tiles with these (8 full-resolution rows)
    if (H2 % kC12Rows || W2 % 16) return fail(AMB_E_ARG, "surrogate grid does not fit the inference tiles");
    hipLaunchKernelGGL(c12_kernel<kC12Rows>, dim3(n * (H2 / kC12Rows) * (W2 / 16)), dim3(256 * kC12Rows), 0, st, phi,
                       params, s->wf0_8, M + s->b_off[0], s->wf[1], M + s->b_off[1], H, W, s->e1, s->e2);
    SUR_CHECK(hipGetLastError());
    SUR_CHECK((conv_tile_launch<kSrcPool, 64, 8, 1, 1, 4, 4, 2>(plain(s->e2), s->wf[2], M + s->b_off[2], n, H4, W4,
                                                                 s->u2, st, s->wf[3], M + s->b_off[3])));
    SUR_CHECK((conv_tile_launch<kSrcUpAdd, 64, 4, 2, 1, 4, 2>(ConvIn{s->u2, s->e2, nullptr, nullptr}, s->wf[4],
                                                               M + s->b_off[4], n, H2, W2, s->u1, st, s->wf[5],
                                                               M + s->b_off[5])));
    hipLaunchKernelGGL(c7_head_flux_kernel<8>, dim3(n * (H / 8) * (W / 16)), dim3(512), 0, st, s->u1, s->e1,
                       s->wf[6], M + s->b_off[6], M + s->w_off[7], M + s->b_off[7], phi, H, W, phi_out, J_out);
    SUR_CHECK(hipGetLastError());
    return AMB_OK;
}

// Loss and full gradient into s->grad for the batch in s->phi, s->params, s->target, s->norm.
int loss_and_grad(amb_surrogate* s, int n, double* loss_out)
{
    const int H = s->H, W = s->W, H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
    const int P1 = n * H * W, P2 = n * H2 * W2, P3 = n * H4 * W4;
    hipStream_t st = s->stream;
    const float* M = s->master;
    int rc = forward(s, n, s->phi, s->params, s->J);
    if (rc != AMB_OK) return rc;
    hipLaunchKernelGGL(flux_update_kernel, dim3(blocks_for(P1)), dim3(kEw), 0, st, s->phi, s->J, H, W, P1, s->pred);
    const unsigned lb = blocks_for(P1);
    hipLaunchKernelGGL(loss_kernel, dim3(lb), dim3(256), 0, st, s->pred, s->target, s->norm, H * W, P1, n, s->g,
                       s->loss_part);
    hipLaunchKernelGGL(dj_kernel, dim3(blocks_for(P1)), dim3(kEw), 0, st, s->g, H, W, P1, s->dJ);
    SUR_CHECK(hipGetLastError());
    // head
    hipLaunchKernelGGL(head_bwd_data_kernel, dim3(blocks_for(P1)), dim3(kEw), 0, st, s->dJ, M + s->w_off[7], H, W, P1,
                       s->dd1);
    const int hchunk = 256, hch = P1 / hchunk;
    hipLaunchKernelGGL(head_bwd_weight_kernel, dim3(hch), dim3(320), 0, st, s->dJ, s->d1, H, W, hchunk, s->part);
    SUR_CHECK(hipGetLastError());
    SUR_CHECK(reduce_parts(s, hch, 578, s->grad + s->w_off[7]));
    // c7 (in s1), then s1 = up(u1) + e1
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks_for((size_t)P1 * 32)), dim3(kEw), 0, st, s->dd1, nullptr, s->d1p,
                       H, W, 32, P1 * 32, s->dpre);
    if ((rc = layer_grads(s, 6, s->dpre, s->s1, H, W, P1)) != AMB_OK) return rc;
    SUR_CHECK(conv<kOutF32>(plain(s->dpre), 32, s->wb[6], nullptr, 32, 9, 1, P1, H, W, s->ds1, nullptr, st));
    hipLaunchKernelGGL(sumpool_kernel, dim3(blocks_for((size_t)P2 * 32)), dim3(kEw), 0, st, s->ds1, H, W, 32, P2 * 32,
                       s->dpre);
    // c6 (1x1, in d2)
    if ((rc = layer_grads(s, 5, s->dpre, s->d2, H2, W2, P2)) != AMB_OK) return rc;
    SUR_CHECK(conv<kOutF32>(plain(s->dpre), 32, s->wb[5], nullptr, 64, 1, 1, P2, H2, W2, s->dd2, nullptr, st));
    // c5 (in s2), then s2 = up(u2) + e2
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks_for((size_t)P2 * 64)), dim3(kEw), 0, st, s->dd2, nullptr, s->d2p,
                       H2, W2, 64, P2 * 64, s->dpre);
    if ((rc = layer_grads(s, 4, s->dpre, s->s2, H2, W2, P2)) != AMB_OK) return rc;
    SUR_CHECK(conv<kOutF32>(plain(s->dpre), 64, s->wb[4], nullptr, 64, 9, 1, P2, H2, W2, s->ds2, nullptr, st));
    hipLaunchKernelGGL(sumpool_kernel, dim3(blocks_for((size_t)P3 * 64)), dim3(kEw), 0, st, s->ds2, H2, W2, 64, P3 * 64,
                       s->dpre);
    // c4 (1x1, in bt)
    if ((rc = layer_grads(s, 3, s->dpre, s->bt, H4, W4, P3)) != AMB_OK) return rc;
    SUR_CHECK(conv<kOutF32>(plain(s->dpre), 64, s->wb[3], nullptr, 128, 1, 1, P3, H4, W4, s->dbt, nullptr, st));
    // c3 (in p2)
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks_for((size_t)P3 * 128)), dim3(kEw), 0, st, s->dbt, nullptr, s->btp,
                       H4, W4, 128, P3 * 128, s->dpre);
    if ((rc = layer_grads(s, 2, s->dpre, s->p2, H4, W4, P3)) != AMB_OK) return rc;
    SUR_CHECK(conv<kOutF32>(plain(s->dpre), 128, s->wb[2], nullptr, 64, 9, 1, P3, H4, W4, s->dp2, nullptr, st));
    // c2 (in p1): de2 = ds2 + pool_grad(dp2)
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks_for((size_t)P2 * 64)), dim3(kEw), 0, st, s->ds2, s->dp2, s->e2p,
                       H2, W2, 64, P2 * 64, s->dpre);
    if ((rc = layer_grads(s, 1, s->dpre, s->p1, H2, W2, P2)) != AMB_OK) return rc;
    SUR_CHECK(conv<kOutF32>(plain(s->dpre), 64, s->wb[1], nullptr, 32, 9, 1, P2, H2, W2, s->dp1, nullptr, st));
    // c1 (in x0): de1 = ds1 + pool_grad(dp1)
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks_for((size_t)P1 * 32)), dim3(kEw), 0, st, s->ds1, s->dp1, s->e1p,
                       H, W, 32, P1 * 32, s->dpre);
    if ((rc = layer_grads(s, 0, s->dpre, s->x0, H, W, P1)) != AMB_OK) return rc;
    SUR_CHECK(hipGetLastError());
    if (loss_out) {
        std::vector<float> lp(lb);
        SUR_CHECK(hipMemcpyAsync(lp.data(), s->loss_part, lb * sizeof(float), hipMemcpyDeviceToHost, st));
        SUR_CHECK(hipStreamSynchronize(st));
        double sum = 0;
        for (float v : lp) sum += v;
        *loss_out = sum / n;
    }
    return AMB_OK;
}

int adam_step(amb_surrogate* s, float lr, float clip)
{
    const int n = (int)s->nparams;
    const int nparts = 64;
    hipLaunchKernelGGL(sumsq_kernel, dim3(nparts), dim3(256), 0, s->stream, s->grad, n, s->sumsq_part);
    s->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)s->adam_t), bc2 = 1.f - std::pow(b2, (float)s->adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(kEw), 0, s->stream, s->master, s->grad, s->adam_m,
                       s->adam_v, n, s->sumsq_part, nparts, lr, b1, b2, 1e-8f, bc1, bc2, clip);
    SUR_CHECK(hipGetLastError());
    return pack_weights(s);
}

int check_dims(int64_t ny, int64_t nx)
{
    if (ny < 64 || nx < 64 || ny % 64 || nx % 64 || ny > 4096 || nx > 4096)
        return fail(AMB_E_ARG, "surrogate grids must be multiples of 64 between 64 and 4096");
    return AMB_OK;
}

// Per-sample loss normaliser |target - phi|^2 + kLossFloor * (grid sites): the floor
// keeps samples whose field barely changes over a coarse step (uniform fields outside
// the spinodal region) from dominating the loss.
constexpr double kLossFloor = 1e-8;
float loss_norm(double sq, size_t hw) { return (float)(sq + kLossFloor * (double)hw); }

// Upload host doubles for a batch into phi/params/target/norm.
int upload_batch(amb_surrogate* s, int n, const double* phi, const double* params, const double* target)
{
    const size_t hw = (size_t)s->H * s->W;
    std::vector<float> f(hw * n), p(3 * (size_t)n), t, nm;
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)phi[i];
    for (size_t i = 0; i < p.size(); ++i) p[i] = (float)params[i];
    SUR_CHECK(hipMemcpyAsync(s->phi, f.data(), f.size() * 4, hipMemcpyHostToDevice, s->stream));
    SUR_CHECK(hipMemcpyAsync(s->params, p.data(), p.size() * 4, hipMemcpyHostToDevice, s->stream));
    if (target) {
        t.resize(hw * n);
        nm.assign(n, 0.f);
        for (int b = 0; b < n; ++b) {
            double acc = 0;
            for (size_t i = 0; i < hw; ++i) {
                t[b * hw + i] = (float)target[b * hw + i];
                const double d = target[b * hw + i] - phi[b * hw + i];
                acc += d * d;
            }
            nm[b] = loss_norm(acc, hw);
        }
        SUR_CHECK(hipMemcpyAsync(s->target, t.data(), t.size() * 4, hipMemcpyHostToDevice, s->stream));
        SUR_CHECK(hipMemcpyAsync(s->norm, nm.data(), nm.size() * 4, hipMemcpyHostToDevice, s->stream));
    }
    SUR_CHECK(hipStreamSynchronize(s->stream));
    return AMB_OK;
}

}  // namespace

extern "C" {

int amb_surrogate_create(int64_t ny, int64_t nx, int max_batch, uint64_t seed, amb_surrogate** out)
{
    clear_error();
    if (!out) return fail(AMB_E_ARG, "out must not be null");
    *out = nullptr;
    int rc = check_dims(ny, nx);
    if (rc != AMB_OK) return rc;
    if (max_batch < 1 || max_batch > 256) return fail(AMB_E_ARG, "max_batch must be in [1, 256]");
    // element counts of the 32-channel full-resolution tensors are 32-bit in the kernels
    if ((int64_t)max_batch * ny * nx > ((int64_t)1 << 26))
        return fail(AMB_E_ARG, "max_batch * ny * nx must not exceed 2^26");
    if ((rc = check_device()) != AMB_OK) return rc;
    amb_surrogate* s = new (std::nothrow) amb_surrogate();
    if (!s) return fail(AMB_E_ARG, "out of host memory");
    s->H = (int)ny;
    s->W = (int)nx;
    s->max_batch = max_batch;
    s->rng.seed(seed ^ 0x9E3779B97F4A7C15ull);
    size_t off = 0;
    for (int l = 0; l < 8; ++l) {
        s->w_off[l] = off;
        off += (size_t)kLayers[l].taps * kLayers[l].cout * kLayers[l].cin;
        s->b_off[l] = off;
        off += kLayers[l].cout;
    }
    s->nparams = off;
    const size_t B = max_batch, P1 = B * ny * nx, P2 = P1 / 4, P3 = P1 / 16;
    // Partial-sum buffer: the most chunks x entries any layer needs at any batch size.
    s->part_len = 0;
    for (int b = 1; b <= max_batch; ++b) {
        const int lv[3] = {(int)(b * ny * nx), (int)(b * ny * nx / 4), (int)(b * ny * nx / 16)};
        const int level_of[8] = {0, 1, 2, 2, 1, 1, 0, 0};
        for (int l = 0; l < 8; ++l) {
            const int npix = lv[level_of[l]];
            const size_t len = l == 7 ? 578 : (size_t)kLayers[l].taps * kLayers[l].cout * kLayers[l].cin;
            s->part_len = std::max(s->part_len, (size_t)(npix / chunk_for(npix)) * len);
            s->part_len = std::max(s->part_len, (size_t)(npix / 256) * (l == 7 ? 578 : kLayers[l].cout));
        }
    }
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc((void**)p, std::max<size_t>(bytes, 256));
    };
    A(&s->master, s->nparams * 4);
    A(&s->grad, s->nparams * 4);
    A(&s->adam_m, s->nparams * 4);
    A(&s->adam_v, s->nparams * 4);
    for (int l = 0; l < 7; ++l) {
        const size_t n = (size_t)kLayers[l].taps * kLayers[l].cout * kLayers[l].cin;
        A(&s->wf[l], n * 2);
        A(&s->wb[l], n * 2);
        if (l == 0) A(&s->wf0_8, 9 * 32 * 8 * 2);
    }
    A(&s->x0, P1 * kCin0 * 2);
    A(&s->e1, P1 * 32 * 2);
    A(&s->e1p, P1 * 32 * 2);
    A(&s->p1, P2 * 32 * 2);
    A(&s->e2, P2 * 64 * 2);
    A(&s->e2p, P2 * 64 * 2);
    A(&s->p2, P3 * 64 * 2);
    A(&s->bt, P3 * 128 * 2);
    A(&s->btp, P3 * 128 * 2);
    A(&s->u2, P3 * 64 * 2);
    A(&s->s2, P2 * 64 * 2);
    A(&s->d2, P2 * 64 * 2);
    A(&s->d2p, P2 * 64 * 2);
    A(&s->u1, P2 * 32 * 2);
    A(&s->s1, P1 * 32 * 2);
    A(&s->d1, P1 * 32 * 2);
    A(&s->d1p, P1 * 32 * 2);
    A(&s->dpre, P1 * 32 * 2);
    A(&s->J, P1 * 2 * 4);
    A(&s->phi, P1 * 4);
    A(&s->target, P1 * 4);
    A(&s->params, B * 3 * 4);
    A(&s->norm, B * 4);
    A(&s->pred, P1 * 4);
    A(&s->g, P1 * 4);
    A(&s->dJ, P1 * 2 * 4);
    A(&s->dd1, P1 * 32 * 4);
    A(&s->ds1, P1 * 32 * 4);
    A(&s->dd2, P2 * 64 * 4);
    A(&s->ds2, P2 * 64 * 4);
    A(&s->dbt, P3 * 128 * 4);
    A(&s->dp2, P3 * 64 * 4);
    A(&s->dp1, P2 * 32 * 4);
    A(&s->part, s->part_len * 4);
    s->part2_len = s->part_len / kReduceGroup + 9 * 128 * 64;
    A(&s->part2, s->part2_len * 4);
    A(&s->audit_part, kAuditBlocks * 2 * sizeof(double));
    A(&s->audit_buf, (size_t)ny * nx * sizeof(float));
    A(&s->p3, 4 * sizeof(float));
    A(&s->loss_part, (P1 / 256 + 1) * 4);
    A(&s->sumsq_part, 64 * 4);
    A(&s->d_idx, B * 4);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev1);
    if (e != hipSuccess) {
        rc = fail_hip(e, "surrogate allocation");
        release(s);
        return rc;
    }
    // He-uniform initialisation (zero on c1's padding channels), zero biases and moments.
    std::vector<float> init(s->nparams, 0.f);
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int l = 0; l < 8; ++l) {
        const LayerDef& L = kLayers[l];
        const int real_cin = l == 0 ? kCinReal : L.cin;
        double bound = std::sqrt(6.0 / (L.taps * real_cin));
        if (l == 7) bound *= 0.1;
        for (int t = 0; t < L.taps; ++t)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci)
                    init[s->w_off[l] + ((size_t)t * L.cout + co) * L.cin + ci] =
                        (l == 0 && ci >= kCinReal) ? 0.f : (float)(bound * U(g));
    }
    std::vector<float> zeros(s->nparams, 0.f);
    e = hipMemcpy(s->master, init.data(), s->nparams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->adam_m, zeros.data(), s->nparams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->adam_v, zeros.data(), s->nparams * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rc = fail_hip(e, "surrogate initialisation");
        release(s);
        return rc;
    }
    if ((rc = pack_weights(s)) != AMB_OK || hipStreamSynchronize(s->stream) != hipSuccess) {
        release(s);
        return rc != AMB_OK ? rc : fail(AMB_E_HIP, "surrogate initialisation");
    }
    *out = s;
    return AMB_OK;
}

int amb_surrogate_destroy(amb_surrogate* s)
{
    clear_error();
    release(s);
    return AMB_OK;
}

int amb_surrogate_param_count(const amb_surrogate* s, int64_t* n)
{
    if (!s || !n) return fail(AMB_E_ARG, "surrogate and n must not be null");
    *n = (int64_t)s->nparams;
    return AMB_OK;
}

int amb_surrogate_get_params(amb_surrogate* s, float* out)
{
    clear_error();
    if (!s || !out) return fail(AMB_E_ARG, "surrogate and out must not be null");
    SUR_CHECK(hipStreamSynchronize(s->stream));
    SUR_CHECK(hipMemcpy(out, s->master, s->nparams * 4, hipMemcpyDeviceToHost));
    return AMB_OK;
}

int amb_surrogate_set_params(amb_surrogate* s, const float* in)
{
    clear_error();
    if (!s || !in) return fail(AMB_E_ARG, "surrogate and in must not be null");
    SUR_CHECK(hipStreamSynchronize(s->stream));
    SUR_CHECK(hipMemcpy(s->master, in, s->nparams * 4, hipMemcpyHostToDevice));
    const int rc = pack_weights(s);
    if (rc != AMB_OK) return rc;
    SUR_CHECK(hipStreamSynchronize(s->stream));
    return AMB_OK;
}

int amb_surrogate_loss_grad(amb_surrogate* s, const double* phi, const double* params, const double* target, int batch,
                            double* loss, float* grad)
{
    clear_error();
    if (!s || !phi || !params || !target) return fail(AMB_E_ARG, "null argument");
    if (batch < 1 || batch > s->max_batch) return fail(AMB_E_ARG, "batch out of range");
    int rc = upload_batch(s, batch, phi, params, target);
    if (rc == AMB_OK) rc = loss_and_grad(s, batch, loss);
    if (rc != AMB_OK) return rc;
    if (grad) {
        SUR_CHECK(hipStreamSynchronize(s->stream));
        SUR_CHECK(hipMemcpy(grad, s->grad, s->nparams * 4, hipMemcpyDeviceToHost));
    }
    return AMB_OK;
}

int amb_surrogate_set_dataset(amb_surrogate* s, const float* phi_in, const float* phi_out, const float* params,
                              int64_t count)
{
    clear_error();
    if (!s || !phi_in || !phi_out || !params || count < 1) return fail(AMB_E_ARG, "need a non-empty dataset");
    SUR_CHECK(hipStreamSynchronize(s->stream));
    for (float* p : {s->ds_in, s->ds_out, s->ds_par, s->ds_norm})
        if (p) (void)hipFree(p);
    s->ds_in = s->ds_out = s->ds_par = s->ds_norm = nullptr;
    s->ds_count = 0;
    const size_t hw = (size_t)s->H * s->W;
    std::vector<float> nm(count);
    for (int64_t b = 0; b < count; ++b) {
        double acc = 0;
        for (size_t i = 0; i < hw; ++i) {
            const double d = (double)phi_out[b * hw + i] - phi_in[b * hw + i];
            acc += d * d;
        }
        nm[b] = loss_norm(acc, hw);
    }
    SUR_CHECK(hipMalloc(&s->ds_in, count * hw * 4));
    SUR_CHECK(hipMalloc(&s->ds_out, count * hw * 4));
    SUR_CHECK(hipMalloc(&s->ds_par, count * 3 * 4));
    SUR_CHECK(hipMalloc(&s->ds_norm, count * 4));
    SUR_CHECK(hipMemcpy(s->ds_in, phi_in, count * hw * 4, hipMemcpyHostToDevice));
    SUR_CHECK(hipMemcpy(s->ds_out, phi_out, count * hw * 4, hipMemcpyHostToDevice));
    SUR_CHECK(hipMemcpy(s->ds_par, params, count * 3 * 4, hipMemcpyHostToDevice));
    SUR_CHECK(hipMemcpy(s->ds_norm, nm.data(), count * 4, hipMemcpyHostToDevice));
    s->ds_count = count;
    return AMB_OK;
}

// `iterations` Adam steps on random mini-batches of the dataset (uniform, with
// replacement across steps, without within a step).  mean_loss: average of the
// per-step losses (nullable).
int amb_surrogate_train(amb_surrogate* s, int64_t iterations, int batch, double lr, double clip, double* mean_loss)
{
    clear_error();
    if (!s) return fail(AMB_E_ARG, "surrogate must not be null");
    if (s->ds_count < batch || batch < 1 || batch > s->max_batch) return fail(AMB_E_ARG, "bad batch or empty dataset");
    const int hw = s->H * s->W;
    double total = 0;
    std::vector<int> idx(batch);
    for (int64_t it = 0; it < iterations; ++it) {
        for (int b = 0; b < batch; ++b) {
            bool dup;
            do {
                idx[b] = (int)(s->rng() % (uint64_t)s->ds_count);
                dup = false;
                for (int c = 0; c < b; ++c) dup |= idx[c] == idx[b];
            } while (dup);
        }
        SUR_CHECK(hipMemcpyAsync(s->d_idx, idx.data(), batch * sizeof(int), hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(gather_kernel, dim3(blocks_for((size_t)hw * batch)), dim3(kEw), 0, s->stream, s->ds_in,
                           s->d_idx, hw, batch, s->phi);
        hipLaunchKernelGGL(gather_kernel, dim3(blocks_for((size_t)hw * batch)), dim3(kEw), 0, s->stream, s->ds_out,
                           s->d_idx, hw, batch, s->target);
        hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(3 * batch)), dim3(kEw), 0, s->stream, s->ds_par, s->d_idx, 3,
                           batch, s->params);
        hipLaunchKernelGGL(gather_kernel, dim3(1), dim3(kEw), 0, s->stream, s->ds_norm, s->d_idx, 1, batch, s->norm);
        if (s->train_noise > 0.f) {
            hipLaunchKernelGGL(train_noise_kernel, dim3(blocks_for((size_t)hw * batch)), dim3(kEw), 0, s->stream, s->phi,
                               s->norm, s->H, s->W, hw * batch, s->train_noise, s->noise_iter++,
                               (uint32_t)s->noise_seed, (uint32_t)(s->noise_seed >> 32));
        }
        SUR_CHECK(hipGetLastError());
        double loss = 0;
        int rc = loss_and_grad(s, batch, mean_loss ? &loss : nullptr);
        if (rc == AMB_OK) rc = adam_step(s, (float)lr, (float)clip);
        if (rc != AMB_OK) return rc;
        total += loss;
    }
    SUR_CHECK(hipStreamSynchronize(s->stream));
    if (mean_loss) *mean_loss = iterations > 0 ? total / iterations : 0.0;
    return AMB_OK;
}

int amb_surrogate_set_train_noise(amb_surrogate* s, double rel, uint64_t seed)
{
    clear_error();
    if (!s) return fail(AMB_E_ARG, "surrogate must not be null");
    if (!(rel >= 0.0) || rel > 10.0) return fail(AMB_E_ARG, "noise must be in [0, 10]");
    s->train_noise = (float)rel;
    s->noise_seed = seed;
    s->noise_iter = 0;
    return AMB_OK;
}

// One surrogate step on host fields: phi_next = phi - div J (batch fields of ny*nx).
int amb_surrogate_predict(amb_surrogate* s, const double* phi, const double* params, int batch, double* phi_next,
                          double* current)
{
    clear_error();
    if (!s || !phi || !params || !phi_next) return fail(AMB_E_ARG, "null argument");
    if (batch < 1 || batch > s->max_batch) return fail(AMB_E_ARG, "batch out of range");
    // the inference kernels of advance(), so both give identical fields
    int rc = upload_batch(s, batch, phi, params, nullptr);
    if (rc == AMB_OK) rc = forward_infer(s, batch, s->phi, s->params, s->pred, current ? s->J : nullptr);
    if (rc != AMB_OK) return rc;
    const int P1 = batch * s->H * s->W;
    std::vector<float> out(P1), jj;
    SUR_CHECK(hipMemcpyAsync(out.data(), s->pred, (size_t)P1 * 4, hipMemcpyDeviceToHost, s->stream));
    if (current) {
        jj.resize(2 * (size_t)P1);
        SUR_CHECK(hipMemcpyAsync(jj.data(), s->J, (size_t)P1 * 8, hipMemcpyDeviceToHost, s->stream));
    }
    SUR_CHECK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < P1; ++i) phi_next[i] = out[i];
    if (current)
        for (size_t i = 0; i < jj.size(); ++i) current[i] = jj[i];
    return AMB_OK;
}

}  // extern "C"

namespace {

// Parameter triple of advance / audit into s->p3 (skipped when unchanged).
int upload_params3(amb_surrogate* s, const double* params3)
{
    const float p3[3] = {(float)params3[0], (float)params3[1], (float)params3[2]};
    if (p3[0] == s->last_p3[0] && p3[1] == s->last_p3[1] && p3[2] == s->last_p3[2]) return AMB_OK;
    SUR_CHECK(hipMemcpyAsync(s->p3, p3, sizeof p3, hipMemcpyHostToDevice, s->stream));
    SUR_CHECK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < 3; ++i) s->last_p3[i] = p3[i];
    return AMB_OK;
}

int check_surrogate_view(const amb_surrogate* s, const FieldView& v)
{
    if (v.precision != AMB_F32 || v.ny != s->H || v.nx != s->W)
        return fail(AMB_E_ARG, "surrogate needs an FP32 integrator of the grid it was built for");
    return AMB_OK;
}

}  // namespace

extern "C" {

int amb_surrogate_audit(amb_surrogate* s, amb_handle* main, amb_handle* judge, const double* params3,
                        int64_t euler_steps, double tol, double* rel_err, int* accepted)
{
    clear_error();
    if (!s || !main || !judge || !params3 || !rel_err || !accepted) return fail(AMB_E_ARG, "null argument");
    if (main == judge) return fail(AMB_E_ARG, "main and judge must be different integrators");
    if (euler_steps < 1) return fail(AMB_E_ARG, "euler_steps must be positive");
    FieldView mv, jv;
    int rc = handle_view(main, &mv);
    if (rc == AMB_OK) rc = handle_view(judge, &jv);
    if (rc == AMB_OK) rc = check_surrogate_view(s, mv);
    if (rc == AMB_OK) rc = check_surrogate_view(s, jv);
    if (rc != AMB_OK) return rc;
    const size_t bytes = (size_t)s->H * s->W * sizeof(float);
    SUR_CHECK(hipStreamSynchronize(mv.stream));
    SUR_CHECK(hipStreamSynchronize(jv.stream));
    // judge: the same start, euler_steps explicit-Euler steps
    SUR_CHECK(hipMemcpyAsync(jv.cur, mv.cur, bytes, hipMemcpyDeviceToDevice, jv.stream));
    if ((rc = amb_run(judge, euler_steps)) != AMB_OK) return rc;
    if ((rc = handle_view(judge, &jv)) != AMB_OK) return rc;
    // surrogate: one step into main's other buffer
    if ((rc = upload_params3(s, params3)) != AMB_OK) return rc;
    if ((rc = forward_infer(s, 1, static_cast<float*>(mv.cur), s->p3, static_cast<float*>(mv.other))) != AMB_OK)
        return rc;
    hipLaunchKernelGGL(audit_kernel, dim3(kAuditBlocks), dim3(256), 0, s->stream, static_cast<const float*>(mv.other),
                       static_cast<const float*>(jv.cur), s->H * s->W, s->audit_part);
    SUR_CHECK(hipGetLastError());
    double part[2 * kAuditBlocks];
    SUR_CHECK(hipMemcpyAsync(part, s->audit_part, sizeof part, hipMemcpyDeviceToHost, s->stream));
    SUR_CHECK(hipStreamSynchronize(s->stream));
    double num = 0.0, den = 0.0;
    for (int b = 0; b < kAuditBlocks; ++b) {
        num += part[2 * b];
        den += part[2 * b + 1];
    }
    const double err = std::sqrt(num) / std::sqrt(den);
    *rel_err = err;
    *accepted = err <= tol ? 1 : 0;  // false for NaN
    if (*accepted) return handle_commit(main, euler_steps, true);
    SUR_CHECK(hipMemcpyAsync(mv.cur, jv.cur, bytes, hipMemcpyDeviceToDevice, s->stream));
    SUR_CHECK(hipStreamSynchronize(s->stream));
    return handle_commit(main, euler_steps, false);
}

int amb_surrogate_audit_begin(amb_surrogate* s, amb_handle* main, amb_handle* judge, const double* params3,
                              int64_t euler_steps)
{
    clear_error();
    if (!s || !main || !judge || !params3) return fail(AMB_E_ARG, "null argument");
    if (main == judge) return fail(AMB_E_ARG, "main and judge must be different integrators");
    if (euler_steps < 1) return fail(AMB_E_ARG, "euler_steps must be positive");
    if (s->audit_pending) return fail(AMB_E_ARG, "an audit is already pending");
    FieldView mv, jv;
    int rc = handle_view(main, &mv);
    if (rc == AMB_OK) rc = handle_view(judge, &jv);
    if (rc == AMB_OK) rc = check_surrogate_view(s, mv);
    if (rc == AMB_OK) rc = check_surrogate_view(s, jv);
    if (rc != AMB_OK) return rc;
    const size_t bytes = (size_t)s->H * s->W * sizeof(float);
    SUR_CHECK(hipStreamSynchronize(mv.stream));
    SUR_CHECK(hipStreamSynchronize(jv.stream));
    // judge: the same start, euler_steps explicit-Euler steps on its own stream.  The
    // copy completes before returning: the caller's next steps overwrite main's buffers.
    SUR_CHECK(hipMemcpyAsync(jv.cur, mv.cur, bytes, hipMemcpyDeviceToDevice, jv.stream));
    SUR_CHECK(hipStreamSynchronize(jv.stream));
    if ((rc = handle_enqueue(judge, euler_steps)) != AMB_OK) return rc;
    // surrogate: one step into main's other buffer, kept for the comparison
    if ((rc = upload_params3(s, params3)) != AMB_OK) return rc;
    if ((rc = forward_infer(s, 1, static_cast<float*>(mv.cur), s->p3, static_cast<float*>(mv.other))) != AMB_OK)
        return rc;
    SUR_CHECK(hipMemcpyAsync(s->audit_buf, mv.other, bytes, hipMemcpyDeviceToDevice, s->stream));
    SUR_CHECK(hipStreamSynchronize(s->stream));
    s->audit_t0 = handle_timestep(main);
    s->audit_steps = euler_steps;
    s->audit_main = main;
    s->audit_judge = judge;
    s->audit_pending = true;
    return handle_commit(main, euler_steps, true);
}

int amb_surrogate_audit_end(amb_surrogate* s, double tol, double* rel_err, int* accepted)
{
    clear_error();
    if (!s || !rel_err || !accepted) return fail(AMB_E_ARG, "null argument");
    if (!s->audit_pending) return fail(AMB_E_ARG, "no audit is pending");
    s->audit_pending = false;
    amb_handle* main = s->audit_main;
    amb_handle* judge = s->audit_judge;
    FieldView mv, jv;
    int rc = handle_view(main, &mv);
    if (rc == AMB_OK) rc = handle_view(judge, &jv);
    if (rc != AMB_OK) return rc;
    SUR_CHECK(hipStreamSynchronize(jv.stream));
    hipLaunchKernelGGL(audit_kernel, dim3(kAuditBlocks), dim3(256), 0, s->stream, s->audit_buf,
                       static_cast<const float*>(jv.cur), s->H * s->W, s->audit_part);
    SUR_CHECK(hipGetLastError());
    double part[2 * kAuditBlocks];
    SUR_CHECK(hipMemcpyAsync(part, s->audit_part, sizeof part, hipMemcpyDeviceToHost, s->stream));
    SUR_CHECK(hipStreamSynchronize(s->stream));
    double num = 0.0, den = 0.0;
    for (int b = 0; b < kAuditBlocks; ++b) {
        num += part[2 * b];
        den += part[2 * b + 1];
    }
    const double err = std::sqrt(num) / std::sqrt(den);
    *rel_err = err;
    *accepted = err <= tol ? 1 : 0;  // false for NaN
    if (*accepted) return AMB_OK;
    // roll main back to the Euler field at the audited step, discarding later steps
    SUR_CHECK(hipStreamSynchronize(mv.stream));
    SUR_CHECK(hipMemcpyAsync(mv.cur, jv.cur, (size_t)s->H * s->W * sizeof(float), hipMemcpyDeviceToDevice,
                             mv.stream));
    return handle_set_timestep(main, s->audit_t0 + s->audit_steps);
}

// nsteps surrogate steps applied in place to the field of an FP32 integrator whose
// grid matches; its step counter advances by nsteps * euler_steps_per_step.  The
// per-step launches are captured once in a hipGraph.  device_ms (nullable): HIP-event
// time of the steps on the integrator's stream.
int amb_surrogate_advance(amb_surrogate* s, amb_handle* h, int64_t nsteps, const double* params3,
                          int64_t euler_steps_per_step, double* device_ms)
{
    clear_error();
    if (!s || !h || !params3) return fail(AMB_E_ARG, "null argument");
    if (nsteps < 0 || euler_steps_per_step < 1) return fail(AMB_E_ARG, "bad step counts");
    FieldView v;
    int rc = handle_view(h, &v);
    if (rc == AMB_OK) rc = check_surrogate_view(s, v);
    if (rc != AMB_OK) return rc;
    // the surrogate works on its own stream; order it after the integrator's work
    SUR_CHECK(hipStreamSynchronize(v.stream));
    if ((rc = upload_params3(s, params3)) != AMB_OK) return rc;
    const int P1 = s->H * s->W;
    // Graph of two steps (field -> other -> field) so replays keep the buffer parity;
    // one cached graph per parity of the integrator's buffers.
    int slot = -1;
    for (int i = 0; i < 2; ++i)
        if (s->step_graph[i] && s->graph_field[i] == v.cur && s->graph_out[i] == v.other) slot = i;
    if (slot < 0) {
        slot = (s->step_graph[0] && s->graph_field[0] == v.other && s->graph_out[0] == v.cur) ? 1 : 0;
        if (s->step_graph[slot]) (void)hipGraphExecDestroy(s->step_graph[slot]);
        s->step_graph[slot] = nullptr;
        hipGraph_t g = nullptr;
        SUR_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int crc = forward_infer(s, 1, static_cast<float*>(v.cur), s->p3, static_cast<float*>(v.other));
        if (crc == AMB_OK) crc = forward_infer(s, 1, static_cast<float*>(v.other), s->p3, static_cast<float*>(v.cur));
        const hipError_t ee = hipStreamEndCapture(s->stream, &g);
        if (crc != AMB_OK) {
            if (g) (void)hipGraphDestroy(g);
            return crc;
        }
        if (ee != hipSuccess) return fail_hip(ee, "hipStreamEndCapture");
        const hipError_t ie = hipGraphInstantiate(&s->step_graph[slot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return fail_hip(ie, "hipGraphInstantiate");
        s->graph_field[slot] = v.cur;
        s->graph_out[slot] = v.other;
    }
    SUR_CHECK(hipEventRecord(s->ev0, s->stream));
    for (int64_t k = 0; k + 1 < nsteps; k += 2) SUR_CHECK(hipGraphLaunch(s->step_graph[slot], s->stream));
    if (nsteps & 1) {
        rc = forward_infer(s, 1, static_cast<float*>(v.cur), s->p3, static_cast<float*>(v.other));
        if (rc != AMB_OK) return rc;
    }
    SUR_CHECK(hipEventRecord(s->ev1, s->stream));
    SUR_CHECK(hipEventSynchronize(s->ev1));
    if (device_ms) {
        float ms = 0.f;
        SUR_CHECK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        *device_ms = ms;
    }
    // An odd number of steps leaves the field in the other buffer.
    return handle_commit(h, nsteps * euler_steps_per_step, (nsteps & 1) != 0);
}

}  // extern "C"
