// amb_tile.h -- device code of one fused step over one tile (shared by the single-field
// and the batched-ensemble kernels).  See amb_step.hip for the algorithm.
#pragma once

#include <hip/hip_runtime.h>

#include "amb_kernels.h"
#include "amb_math.h"

namespace amb {

constexpr int kThreads = 256;

// Noisy FP32 steps on fields of at least this many sites form the active current in the
// noise stage (step_tile); smaller fields and FP64 apply the 5x5 stencil of mu_a in stage B.
constexpr long long kActiveCurrentSites = 2048LL * 2048;

__device__ __forceinline__ int wrap_index(int g, int n)
{
    // Exact for any g; one iteration in the common case, several on tiny grids.
    while (g < 0) g += n;
    while (g >= n) g -= n;
    return g;
}

template <typename T, int TW, int TH, int NT>
struct Lds {
    static constexpr int PW = TW + 18, PH = TH + 18;  // phi: rows [-9,TH+9), cols [-9,TW+9)
    static constexpr int MW = TW + 16, MH = TH + 16;  // mu_p: rows [-8,TH+8), cols [-8,TW+8)
    static constexpr int AW = TW + 4, AH = TH + 4;    // mu_a: [-2, +2)
    static constexpr int QW = TW + 4, QH = TH + 4;    // circulating current: the box [-2, +2)
    static constexpr int XW = TW + 8, XH = TH;        // xi_x: rows [0,TH), cols [-4,TW+4)
    static constexpr int YW = TW, YH = TH + 8;        // xi_y: rows [-4,TH+4), cols [0,TW)
    static constexpr int n_phi = PW * PH;
    static constexpr int n_xi = XW * XH + YW * YH;
    static constexpr int n_union = n_phi > n_xi ? n_phi : n_xi;
    static constexpr int off_mup = n_union;
    static constexpr int off_mua = off_mup + MW * MH;
    static constexpr int off_qx = off_mua + AW * AH;
    static constexpr int off_qy = off_qx + QW * QH;
    static constexpr int total = off_qy + QW * QH;
    // per-wave FP64 partials of the fused mass / finite check (16 bytes per wave)
    static constexpr size_t red_off = (sizeof(T) * (size_t)total + 15) / 16 * 16;
    static constexpr size_t bytes = red_off + 16 * (size_t)(NT / 64);
    static_assert(NT % TW == 0, "tile width must divide the workgroup");
    static_assert((TW * TH) % NT == 0, "tile must be a multiple of the workgroup");
    static_assert(NT >= TW + 4, "workgroup must cover one stage-A row of the box");
    static_assert(TH % 8 == 0 && NT >= 2 * TW + 12 * (TH / 8), "workgroup must cover the four stage-A arms");
    static_assert(TW % 4 == 0, "noise column groups need TW % 4 == 0");
};

// Waves per SIMD the register allocation must leave room for: as many as the tile's LDS
// lets share a CU (160 KB per CU, 4 SIMDs), so that registers never limit occupancy below
// what LDS allows (e.g. 3 workgroups of 512 threads = 6 waves per SIMD for 32x48 FP32 tiles,
// 2 of 512 = 4 for 32x32 FP64 tiles).
template <typename T, int TW, int TH, int NT>
constexpr int min_waves_per_eu()
{
    constexpr int wg = (int)((160 * 1024) / Lds<T, TW, TH, NT>::bytes);
    constexpr int w = wg * (NT / 64) / 4;
    return w < 1 ? 1 : w > 8 ? 8 : w;
}

// ---- Box-Muller ---------------------------------------------------------------

__device__ __forceinline__ void box_muller(uint32_t wa, uint32_t wb, float& z0, float& z1)
{
    // u1 = (wa + 1/2) 2^-32 in (0, 1], u2 = wb 2^-32 in [0, 1) (revolutions).
    const float u1 = fmaf((float)wa, 0x1p-32f, 0x1p-33f);
    const float u2 = (float)wb * 0x1p-32f;
    // v_log_f32 is log2; -2 ln u1 = -2 ln2 log2 u1.  v_sin/v_cos take revolutions.
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    z0 = r * __builtin_amdgcn_cosf(u2);
    z1 = r * __builtin_amdgcn_sinf(u2);
}

template <typename T>
__device__ __forceinline__ void normals4(uint32_t group, uint32_t row, uint64_t step, int comp,
                                         uint32_t k0, uint32_t k1, T z[4])
{
    const U4 w = philox4x32_10(noise_counter(group, row, step, comp), k0, k1);
    box_muller(w.x, w.y, z[0], z[1]);
    box_muller(w.z, w.w, z[2], z[3]);
}

// Normal of one (row, column, component) when its column group straddles the seam.
template <typename T>
__device__ __forceinline__ T normal1(int row, int col, uint64_t step, int comp, uint32_t k0, uint32_t k1)
{
    T z[4];
    normals4<T>((uint32_t)(col >> 2), (uint32_t)row, step, comp, k0, k1, z);
    const int l = col & 3;
    return l == 0 ? z[0] : l == 1 ? z[1] : l == 2 ? z[2] : z[3];
}

// ---- local terms -------------------------------------------------------------

// L3 Laplacian from the centre m0, the sum of its x-neighbours xs, of its
// y-neighbours ys and of its four corners cs; and mu_p = m0 (a + m0 (b + m0 c)) -
// kappa lap.  Explicit fma fixes the rounding, so every stage-A path that evaluates
// them (box strips, arms) gives the same bits for the same site.
template <typename T>
__device__ __forceinline__ T laplacian3(const StepCoef<T>& K, T m0, T xs, T ys, T cs)
{
    return fma(K.lck, cs, fma(K.ley, ys, fma(K.lex, xs, K.lcc * m0)));
}

template <typename T>
__device__ __forceinline__ T mu_passive(const StepCoef<T>& K, T kb, T m0, T lap)
{
    return fma(-K.kappa, lap, m0 * fma(m0, fma(m0, K.c, kb), K.a));
}

// A wave-uniform coefficient held in a vector register: an instruction may read only one
// scalar register on CDNA, so fma(x, c, b) with both c and b scalar costs a move every time
// it is evaluated; one move per thread up front instead.  (Empty asm: no instruction.)
template <typename T>
__device__ __forceinline__ T in_vgpr(T x)
{
    asm volatile("" : "=v"(x) : "0"(x));
    return x;
}

// An LDS address as a base register of its own.  The arrays of a tile lie tens of KB into
// the workgroup's LDS; folded into one constant, such offsets overflow the 8-bit dword
// offsets of ds_read2 and the compiler spends an add per pair of reads on a fresh base.
// Held opaque in a register, the base serves every read within 1 KB of it by offset alone.
// (Empty asm on the 32-bit LDS address: no instruction, no memory access.)
template <typename T>
using lds_ptr = __attribute__((address_space(3))) T*;

template <typename T>
__device__ __forceinline__ lds_ptr<T> lds_base(T* p)
{
    uint32_t a = (uint32_t)(uintptr_t)(lds_ptr<T>)p;
    asm volatile("" : "+v"(a));
    return (lds_ptr<T>)(uintptr_t)a;
}

// mu_p at rows r0 .. r0+N-1 of tile column c (tile-relative; LDS phi holds rows and
// columns from -9, mu_p from -8), marched down: each phi row is read once and its
// horizontal sum serves the three rows that need it.
template <typename T, int N, int PW, int MW>
__device__ __forceinline__ void mu_column(const T* s_phi, T* s_mup, const StepCoef<T>& K, const T kb, const int r0,
                                          const int c)
{
    const T* P = s_phi + (r0 + 8) * PW + c + 8;  // phi at (r0 - 1, c - 1)
    T t0 = P[1], th2 = P[2] + P[0];
    T m0 = P[PW + 1], mh2 = P[PW + 2] + P[PW];
    T* M = s_mup + (r0 + 8) * MW + c + 8;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const T* Q = P + (k + 2) * PW;
        const T b0 = Q[1], bh2 = Q[2] + Q[0];
        M[k * MW] = mu_passive(K, kb, m0, laplacian3(K, m0, mh2, t0 + b0, th2 + bh2));
        t0 = m0; th2 = mh2;
        m0 = b0; mh2 = bh2;
    }
}

// Four consecutive LDS words at a 16-byte aligned address (32 bytes for double) as one
// or two 128-bit accesses.
__device__ __forceinline__ void lds_load4(const float* p, float v[4])
{
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void lds_load4(const double* p, double v[4])
{
    const double2 a = *reinterpret_cast<const double2*>(p), b = *reinterpret_cast<const double2*>(p + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void lds_store4(float* p, const float v[4])
{
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void lds_store4(double* p, const double v[4])
{
    *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
}

// ---- fused step --------------------------------------------------------------

template <typename T, int TW, int TH, int NT, int NOISE>
__device__ __forceinline__ void step_tile(T* const sm, const T* __restrict__ in, T* __restrict__ out, const int ny,
                                          const int nx, const int nx4, const int x0, const int y0,
                                          const StepCoef<T>& K, const uint32_t key0, const uint32_t key1,
                                          const uint64_t step, const double* noise, double* const red_out)
{
    using L = Lds<T, TW, TH, NT>;
    T* const s_phi = sm;
    T* const s_mup = sm + L::off_mup;
    T* const s_mua = sm + L::off_mua;
    T* const s_qx = sm + L::off_qx;
    T* const s_qy = sm + L::off_qy;

    const int tid = threadIdx.x;
    constexpr int RG = NT / TW;        // row groups of stage B
    constexpr int R = TH / RG;         // output rows per thread in stage B
    static_assert(TH % RG == 0, "tile height must split evenly over the row groups");
    const int cx = tid % TW;
    const int rg = tid / TW;

    // Interior tiles (halo inside the field) use affine addressing; tiles on the
    // periodic seam or past the edge of the field wrap every index.
    const bool interior = x0 >= 9 && x0 + TW + 9 <= nx && y0 >= 9 && y0 + TH + 9 <= ny;

    // Wave priorities: a workgroup's global loads (stage 0) are issued ahead of the other
    // resident workgroups' arithmetic, so that their memory latency overlaps it; noisy steps
    // keep the raised priority through stage A (whose LDS results the long Philox stage
    // waits for), and FP64 noisy steps raise it again for stage B.  Measured on MI355X at
    // 4096^2 against no priorities: FP32 noisy 107.4 -> 104.6 us, noiseless 90.7 -> 87.2 us,
    // FP64 noisy 266.0 -> 252.2 us; 1024^2 FP32 noisy 11.05 -> 10.2 us.  (Priorities change
    // only the order in which a SIMD picks ready waves, not what any wave computes.)
    constexpr bool kNoisyKernel = NOISE == kNoisePhilox || NOISE == kNoiseBuffer;
    __builtin_amdgcn_s_setprio(1);
    // ---- stage 0: phi with a 9-site halo -> LDS.  Each thread owns fixed region
    // columns and walks rows; loads are independent and unrolled.
    {
        constexpr int NCJ = (L::PW + TW - 1) / TW;
        if (interior) {
            const T* src = in + (size_t)(y0 - 9 + rg) * nx + (x0 - 9 + cx);
#pragma unroll
            for (int r = rg; r < L::PH; r += RG) {
#pragma unroll
                for (int j = 0; j < NCJ; ++j)
                    if (j < NCJ - 1 || cx + j * TW < L::PW) s_phi[r * L::PW + cx + j * TW] = src[j * TW];
                src += (size_t)RG * nx;
            }
        } else {
            int gcol[NCJ];
#pragma unroll
            for (int j = 0; j < NCJ; ++j) gcol[j] = wrap_index(x0 - 9 + cx + j * TW, nx);
#pragma unroll
            for (int r = rg; r < L::PH; r += RG) {
                const T* row = in + (size_t)wrap_index(y0 - 9 + r, ny) * nx;
#pragma unroll
                for (int j = 0; j < NCJ; ++j)
                    if (j < NCJ - 1 || cx + j * TW < L::PW) s_phi[r * L::PW + cx + j * TW] = row[gcol[j]];
            }
        }
    }
    __syncthreads();
    if constexpr (!kNoisyKernel) __builtin_amdgcn_s_setprio(0);

    T phi_c[R];
#pragma unroll
    for (int k = 0; k < R; ++k) phi_c[k] = s_phi[(rg * R + k + 9) * L::PW + cx + 9];

    // ---- stage A: local terms.  Stage B reads mu_a on the box rows/cols [-2, T+2), the
    // circulating current on [-1, T+1), and mu_p on a cross: that box plus four arms six
    // sites deep (rows [-8,-2) and [TH+2,TH+8) over the tile's columns, columns [-8,-2)
    // and [TW+2,TW+8) over its rows); the corners of the halo region are never read.
    // The box is marched down column strips with every term, so each phi row is read
    // once and its horizontal difference and sum serve the three rows that need them;
    // the arms get mu_p only, marched down columns so that neighbouring lanes touch
    // neighbouring words: the top and bottom arms one 6-row column per thread, the left
    // and right arms (6 columns each) in 8-row segments, 12 threads per segment.
    // Two layouts, chosen per tile shape by the busiest wave's work (a box row costs
    // about two arm rows): every thread on the box with the arms added to the first
    // waves, or the last wave on the arms and longer box strips on the others.
    {
        constexpr int BW = TW + 4, BH = TH + 4;  // box: br = r + 2, bc = c + 2
        constexpr int NARM_TB = 2 * TW, NARM_LR = 12 * (TH / 8);
        constexpr int SEG_ALL = (BH + NT / BW - 1) / (NT / BW);
        constexpr int NSEG_RES = (NT - 64) / BW;
        constexpr int SEG_RES = NSEG_RES > 0 ? (BH + NSEG_RES - 1) / NSEG_RES : BH;
        constexpr int ARM_ROWS_RES = 6 * ((NARM_TB + 63) / 64) + 8 * ((NARM_LR + 63) / 64);
        constexpr int COST_ALL = 2 * SEG_ALL + 8;
        constexpr int COST_RES = 2 * SEG_RES > ARM_ROWS_RES ? 2 * SEG_RES : ARM_ROWS_RES;
        constexpr bool ARM_WAVE = NSEG_RES > 0 && COST_RES < COST_ALL;
        constexpr int SEG = ARM_WAVE ? SEG_RES : SEG_ALL;
        // Box columns are dealt 32 to a half-wave, so that each LDS access of a strip row
        // touches 32 consecutive words, one per bank; the columns past the last multiple of
        // 32 follow.  (Dealt 36 to a segment, a half-wave straddled two strips whose rows
        // share banks: 43% of the step's LDS bank-conflict cycles, 1.2% of its time.)
        constexpr int NSEGS = (BH + SEG - 1) / SEG;
        constexpr int BWM = BW >= 32 ? BW / 32 * 32 : BW;
        int seg = NSEGS, bc = 0;
        if (tid < NSEGS * BWM) {
            seg = tid / BWM;
            bc = tid - seg * BWM;
        } else if (tid < NSEGS * BW) {
            const int j = tid - NSEGS * BWM;
            seg = j / (BW - BWM);
            bc = BWM + j - seg * (BW - BWM);
        }
        const T kb = in_vgpr(K.b);
        if (seg * SEG < BH) {
            const int br0 = seg * SEG;
            const auto P = lds_base(s_phi + (br0 + 6) * L::PW + bc + 6);  // phi at (r0 - 1, c - 1)
            const auto Mw = lds_base(s_mup + (br0 + 6) * L::MW + bc + 6);
            const auto Aw = lds_base(s_mua + br0 * L::AW + bc);
            const auto Qxw = lds_base(s_qx + br0 * L::QW + bc);
            const auto Qyw = lds_base(s_qy + br0 * L::QW + bc);
            T tm = P[0], t0 = P[1], tp = P[2];
            T mm = P[L::PW], m0 = P[L::PW + 1], mp = P[L::PW + 2];
            T th1 = tp - tm, th2 = tp + tm;
            T mh1 = mp - mm, mh2 = mp + mm;
#pragma unroll
            for (int k = 0; k < SEG; ++k) {
                const int br = br0 + k;
                if (BH % SEG == 0 || br < BH) {  // every strip is whole when SEG divides BH
                    const auto Q = P + (k + 2) * L::PW;
                    const T bm = Q[0], b0 = Q[1], bp = Q[2];
                    const T bh1 = bp - bm, bh2 = bp + bm;
                    const T gx = K.gxe * (th1 + bh1) + K.gxc * mh1;
                    const T gy = K.gye * (bh2 - th2) + K.gyc * (b0 - t0);
                    const T lap = laplacian3(K, m0, mh2, t0 + b0, th2 + bh2);
                    Mw[k * L::MW] = mu_passive(K, kb, m0, lap);
                    Aw[k * L::AW] = K.lam * (gy * gy + gx * gx);
                    // the circulating current on the whole box (stage B reads its [-1, T+1)
                    // core; storing the rim too spares a condition per row)
                    const T zl = K.zeta * lap;
                    Qxw[k * L::QW] = zl * gx;
                    Qyw[k * L::QW] = zl * gy;
                    tm = mm; t0 = m0; tp = mp; th1 = mh1; th2 = mh2;
                    mm = bm; m0 = b0; mp = bp; mh1 = bh1; mh2 = bh2;
                }
            }
        }
        if constexpr (ARM_WAVE) {
            if (tid >= NT - 64) {
                const int lane = tid - (NT - 64);
#pragma unroll
                for (int a = lane; a < NARM_TB; a += 64)
                    mu_column<T, 6, L::PW, L::MW>(s_phi, s_mup, K, kb, a < TW ? -8 : TH + 2, a % TW);
#pragma unroll
                for (int a = lane; a < NARM_LR; a += 64) {
                    const int q = a % 12;  // 0..5 left arm, 6..11 right arm
                    mu_column<T, 8, L::PW, L::MW>(s_phi, s_mup, K, kb, (a / 12) * 8, q < 6 ? q - 8 : TW + 2 + (q - 6));
                }
            }
        } else if (tid < NARM_TB) {
            mu_column<T, 6, L::PW, L::MW>(s_phi, s_mup, K, kb, tid < TW ? -8 : TH + 2, tid % TW);
        } else if (tid < NARM_TB + NARM_LR) {
            const int t = tid - NARM_TB;
            const int q = t % 12;
            mu_column<T, 8, L::PW, L::MW>(s_phi, s_mup, K, kb, (t / 12) * 8, q < 6 ? q - 8 : TW + 2 + (q - 6));
        }
    }
    __syncthreads();
    if constexpr (kNoisyKernel) __builtin_amdgcn_s_setprio(0);

    // ---- stage N (noisy steps): the currents through D8, J = amp xi - (1/dx) D8 mu_p,
    // each formed as soon as its normal is drawn and written over the phi buffer: J_x on
    // rows [0,TH), cols [-4,TW+4); J_y on rows [-4,TH+4), cols [0,TW).  Stage B takes
    // their D8 divergence, the passive and the noise term in one.  Every path below
    // evaluates J in the same order (amp xi, then j = 1..4), so all give the same bits.
    // FP32 noisy steps also form the active current C here (active_current below).
    bool noisy = NOISE == kNoisePhilox || NOISE == kNoiseBuffer;
    if constexpr (NOISE == kNoiseRuntime) noisy = K.amp != T(0);
    T* const s_jx = sm;  // aliases s_phi after stage A
    T* const s_jy = sm + L::XW * L::XH;
    auto mup = [&](int r, int col) -> const T* { return s_mup + (r + 8) * L::MW + col + 8; };
    // x-currents at row r, columns c0 .. c0+3 (c0 a multiple of 4, so the mu_p row and the
    // J_x row are read and written in aligned fours)
    auto emit_x4 = [&](int r, int c0, const T* z) {
        T m[12], jv[4];
        lds_load4(mup(r, c0 - 4), m);
        lds_load4(mup(r, c0), m + 4);
        lds_load4(mup(r, c0 + 4), m + 8);
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            T acc = K.amp * z[l];
#pragma unroll
            for (int j = 1; j <= 4; ++j) acc = fma(-K.dxw[j], m[4 + l + j] - m[4 + l - j], acc);
            jv[l] = acc;
        }
        lds_store4(s_jx + r * L::XW + c0 + 4, jv);
    };
    // y-currents at row r (from -4), columns c0 .. c0+3
    auto emit_y4 = [&](int r, int c0, const T* z) {
        T acc[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) acc[l] = K.amp * z[l];
#pragma unroll
        for (int j = 1; j <= 4; ++j) {
            T up[4], dn[4];
            lds_load4(mup(r + j, c0), up);
            lds_load4(mup(r - j, c0), dn);
#pragma unroll
            for (int l = 0; l < 4; ++l) acc[l] = fma(-K.dyw[j], up[l] - dn[l], acc[l]);
        }
        lds_store4(s_jy + (r + 4) * L::YW + c0, acc);
    };
    auto jx1 = [&](int r, int col, T z) {
        const T* m = mup(r, col);
        T acc = K.amp * z;
#pragma unroll
        for (int j = 1; j <= 4; ++j) acc = fma(-K.dxw[j], m[j] - m[-j], acc);
        return acc;
    };
    auto jy1 = [&](int r, int col, T z) {
        T acc = K.amp * z;
#pragma unroll
        for (int j = 1; j <= 4; ++j) acc = fma(-K.dyw[j], *mup(r + j, col) - *mup(r - j, col), acc);
        return acc;
    };
    // The G3 current of the active terms, C = q - G3 mu_a (q the circulating current), on
    // the box [-1, T+1), written over q; stage B then takes the G3 divergence of C, which
    // equals -G3xoG3x mu_a - G3yoG3y mu_a + G3 . q, the 5x5 stencil of mu_a and the 3x3 of q
    // that it applies otherwise.  Items are column segments of CSEG rows (each mu_a row read
    // once, its x-difference and y-smoothed sum shared by the three rows that need them),
    // dealt from the first wave after the Philox items.  It pays on large FP32 fields with
    // the default 32x32 tile of 512 threads, whose 2-row stage-B strips form each row of
    // mu_a work three times over: about 5% fewer VALU instructions and, on MI355X, 103.7 ->
    // 101.8 us per noisy step at 4096^2 (3072^2 -1.6%, 2048^2 -1.1%, but 1536^2 +0.8%; the
    // other tiles lose 1-3% at 4096^2, but none of them is chosen there).  In FP64 it was
    // slower (242 -> 250.5 us), so FP64 and the smaller FP32 fields keep the 5x5 stencil.
    // The choice depends on the precision and the field's size only, so every tile shape
    // and the batched kernels give the same bits.
    constexpr bool kActiveCurrent = sizeof(T) == 4 && NOISE != kNoiseNone;
    const bool active = kActiveCurrent && noisy && (long long)ny * nx >= kActiveCurrentSites;
    auto active_current = [&]() {
        if constexpr (kActiveCurrent) {
            constexpr int CW = TW + 2, CH = TH + 2, CSEG = 4, NSEG = (CH + CSEG - 1) / CSEG;
            // the first CW32 columns of each segment go 32 to a half-wave, so that its 32
            // lanes touch 32 consecutive words (one per LDS bank); the last columns follow
            constexpr int CW32 = CW >= 32 ? CW / 32 * 32 : CW;
            constexpr int NMAIN = CW32 * NSEG, NC = CW * NSEG;
            constexpr int GXN = TW / 4 + 2, GYN = TW / 4;
            constexpr int NXY = GXN * TH + GYN * (L::YH / 2);  // xi_x items and xi_y pairs
            constexpr int C0 = ((NXY + 63) / 64 * 64) % NT;
            for (int i = (tid + NT - C0) % NT; i < NC; i += NT) {
                int seg, bc;
                if (i < NMAIN) {
                    seg = i / CW32;
                    bc = i - seg * CW32;
                } else {
                    seg = (i - NMAIN) / (CW - CW32);
                    bc = CW32 + (i - NMAIN) - seg * (CW - CW32);
                }
                const int br0 = seg * CSEG;  // box row br is tile row br - 1, box column bc column bc - 1
                const auto P = lds_base(s_mua + br0 * L::AW + bc);  // mu_a at (br0 - 2, bc - 2)
                const auto Qx = lds_base(s_qx + (br0 + 1) * L::QW + bc + 1);
                const auto Qy = lds_base(s_qy + (br0 + 1) * L::QW + bc + 1);
                // every load ahead of the stores (the compiler cannot move an LDS read across
                // a write it cannot prove disjoint, and would wait out each one in turn)
                constexpr int NR = CSEG + 2;
                T m[NR][3], qx[CSEG], qy[CSEG];
#pragma unroll
                for (int k = 0; k < NR; ++k)
                    if (CH % CSEG == 0 || k < 2 || br0 + k - 2 < CH)
#pragma unroll
                        for (int j = 0; j < 3; ++j) m[k][j] = P[k * L::AW + j];
#pragma unroll
                for (int k = 0; k < CSEG; ++k)
                    if (CH % CSEG == 0 || br0 + k < CH) {
                        qx[k] = Qx[k * L::QW];
                        qy[k] = Qy[k * L::QW];
                    }
                T d[NR], sum[NR];  // x-difference and y-smoothed sum of each mu_a row
#pragma unroll
                for (int k = 0; k < NR; ++k)
                    if (CH % CSEG == 0 || k < 2 || br0 + k - 2 < CH) {
                        d[k] = m[k][2] - m[k][0];
                        sum[k] = fma(K.gye, m[k][0] + m[k][2], K.gyc * m[k][1]);
                    }
#pragma unroll
                for (int k = 0; k < CSEG; ++k)
                    if (CH % CSEG == 0 || br0 + k < CH) {
                        Qx[k * L::QW] = fma(-K.gxc, d[k + 1], fma(-K.gxe, d[k] + d[k + 2], qx[k]));
                        Qy[k * L::QW] = qy[k] - (sum[k + 2] - sum[k]);
                    }
            }
        }
    };
    if (NOISE == kNoisePhilox || (NOISE == kNoiseRuntime && noisy)) {
        // Philox items: xi_x groups (rows [0,TH), tile-relative column groups [-1, TW/4])
        // and xi_y groups (rows [-4,TH+4), groups [0, TW/4)), four normals per call.
        constexpr int GXN = TW / 4 + 2, GYN = TW / 4;
        constexpr int NXI = GXN * TH, NYI = GYN * L::YH, NITEMS = NXI + NYI;
        constexpr int PERX = (NXI + NT - 1) / NT;
        if (nx4 && interior) {
            // Items past the end of a list are skipped, not recomputed: a wave whose lanes
            // are all past it branches over the call.  The xi_y items go in vertical pairs,
            // dealt from the thread after the last xi_x item (modulo the workgroup), so the
            // two lists' last, partial rounds fall on different waves.
            // xi_x: rows [0, TH), column groups x0/4 - 1 .. x0/4 + TW/4
#pragma unroll
            for (int u = 0; u < PERX; ++u) {
                const int i = tid + u * NT;
                if ((u + 1) * NT <= NXI || i < NXI) {
                    // a row's TW/4 inner groups eight to a 128-bit LDS pass, then the halo groups
                    constexpr int GIN = GXN - 2;
                    int r, gl;
                    if (i < GIN * TH) {
                        r = i / GIN;
                        gl = 1 + i - r * GIN;
                    } else {
                        r = (i - GIN * TH) >> 1;
                        gl = (i & 1) ? GXN - 1 : 0;
                    }
                    T z[4];
                    normals4<T>((uint32_t)((x0 >> 2) + gl - 1), (uint32_t)(y0 + r), step, 1, key0, key1, z);
                    emit_x4(r, 4 * gl - 4, z);
                }
            }
            // xi_y in vertical pairs of items (rows r, r + 1 of one column group): the pair
            // reads mu_p rows r-4 .. r+5 once (10 loads instead of 16)
            constexpr int NYP = GYN * (L::YH / 2), PERYP = (NYP + NT - 1) / NT;
            const int ty = (tid + NT - (NXI % NT)) % NT;
#pragma unroll
            for (int u = 0; u < PERYP; ++u) {
                const int i = ty + u * NT;
                if ((u + 1) * NT <= NYP || i < NYP) {
                    const int rp = i / GYN, gl = i - rp * GYN;
                    const int r = 2 * rp - 4, c0 = 4 * gl;
                    T zA[4], zB[4];
                    normals4<T>((uint32_t)((x0 >> 2) + gl), (uint32_t)(y0 + r), step, 0, key0, key1, zA);
                    normals4<T>((uint32_t)((x0 >> 2) + gl), (uint32_t)(y0 + r + 1), step, 0, key0, key1, zB);
                    T accA[4], accB[4], dnB[4], up[4], nxt[4], dn[4];
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        accA[l] = K.amp * zA[l];
                        accB[l] = K.amp * zB[l];
                    }
                    lds_load4(mup(r, c0), dnB);
                    lds_load4(mup(r + 1, c0), up);
#pragma unroll
                    for (int j = 1; j <= 4; ++j) {
                        lds_load4(mup(r + j + 1, c0), nxt);
                        lds_load4(mup(r - j, c0), dn);
#pragma unroll
                        for (int l = 0; l < 4; ++l) {
                            accA[l] = fma(-K.dyw[j], up[l] - dn[l], accA[l]);
                            accB[l] = fma(-K.dyw[j], nxt[l] - dnB[l], accB[l]);
                            up[l] = nxt[l];
                            dnB[l] = dn[l];
                        }
                    }
                    lds_store4(s_jy + (r + 4) * L::YW + c0, accA);
                    lds_store4(s_jy + (r + 5) * L::YW + c0, accB);
                }
            }
        } else if (nx4) {
            constexpr int PER = (NITEMS + NT - 1) / NT;
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i0 = tid + u * NT;
                const int i = i0 < NITEMS ? i0 : NITEMS - 1;
                uint32_t group, grow;
                int comp, r, c0;
                if (i < NXI) {
                    r = i / GXN;
                    const int gl = i - r * GXN - 1;
                    grow = (uint32_t)wrap_index(y0 + r, ny);
                    group = (uint32_t)(wrap_index(x0 + 4 * gl, nx) >> 2);
                    comp = 1;
                    c0 = 4 * gl;
                } else {
                    const int j = i - NXI;
                    const int rr = j / GYN, gl = j - rr * GYN;
                    grow = (uint32_t)wrap_index(y0 + rr - 4, ny);
                    group = (uint32_t)((x0 >> 2) + gl);
                    comp = 0;
                    r = rr - 4;
                    c0 = 4 * gl;
                }
                T z[4];
                normals4<T>(group, grow, step, comp, key0, key1, z);
                if (i0 < NITEMS) {
                    if (comp == 1) emit_x4(r, c0, z);
                    else emit_y4(r, c0, z);
                }
            }
        } else {
            // Column groups straddle the periodic seam: one call per site.
            for (int i = tid; i < NITEMS; i += NT) {
                if (i < NXI) {
                    const int r = i / GXN, gl = i - r * GXN - 1;
                    const int grow = wrap_index(y0 + r, ny);
                    T* dst = s_jx + r * L::XW + 4 * (gl + 1);
                    for (int l = 0; l < 4; ++l)
                        dst[l] = jx1(r, 4 * gl + l,
                                     normal1<T>(grow, wrap_index(x0 + 4 * gl + l, nx), step, 1, key0, key1));
                } else {
                    const int j = i - NXI;
                    const int rr = j / GYN, gl = j - rr * GYN;
                    T z[4];
                    normals4<T>((uint32_t)((x0 >> 2) + gl), (uint32_t)wrap_index(y0 + rr - 4, ny), step, 0, key0,
                                key1, z);
                    T* dst = s_jy + rr * L::YW + 4 * gl;
                    for (int l = 0; l < 4; ++l) dst[l] = jy1(rr - 4, 4 * gl + l, z[l]);
                }
            }
        }
        if (active) active_current();
        __syncthreads();
    } else if (NOISE == kNoiseBuffer) {
        const size_t n = (size_t)ny * nx;
        for (int i = tid; i < L::XW * L::XH; i += NT) {
            const int r = i / L::XW, col = i - r * L::XW - 4;
            s_jx[i] = jx1(r, col, (T)noise[n + (size_t)wrap_index(y0 + r, ny) * nx + wrap_index(x0 + col, nx)]);
        }
        for (int i = tid; i < L::YW * L::YH; i += NT) {
            const int rr = i / L::YW, col = i - rr * L::YW;
            s_jy[i] = jy1(rr - 4, col, (T)noise[(size_t)wrap_index(y0 + rr - 4, ny) * nx + wrap_index(x0 + col, nx)]);
        }
        if (active) active_current();
        __syncthreads();
    }

    if constexpr (kNoisyKernel && sizeof(T) == 8) __builtin_amdgcn_s_setprio(1);

    // ---- stage B: one column strip of R output rows per thread.  Row-wise partial
    // sums of the separable stencils are formed once per row and shared by the
    // output rows that need them.
    // With noise the D8 part is the divergence of J; without, D8oD8 of mu_p as one 17-tap
    // stencil per axis (cheaper when there is no noise current to form).
    constexpr bool kAlwaysNoisy = NOISE == kNoisePhilox || NOISE == kNoiseBuffer;
    constexpr bool kNeverNoisy = NOISE == kNoiseNone;
    const int c = cx, r0 = rg * R;
    T* orow = out + (size_t)(y0 + r0) * nx + (x0 + c);
    T mcol[kAlwaysNoisy ? 1 : R + 16];  // mu_p at rows r0-8 .. r0+R+7 of this column
    if constexpr (!kAlwaysNoisy) {
        if (!noisy) {
            const auto Mc = lds_base(s_mup + r0 * L::MW + c + 8);
#pragma unroll
            for (int k = 0; k < R + 16; ++k) mcol[k] = Mc[k * L::MW];
        }
    }
    // The strip's neighbouring rows, each formed just before the first output row that
    // needs it rather than all up front: the same values, fewer of them live at once (FP64
    // noisy 32x32 tiles 97 VGPRs instead of 109, FP32 noisy 64x32 58 instead of 100, and
    // the FP32 noiseless 32x32 tile no longer spills; on MI355X the FP64 noiseless step at
    // 4096^2 went from 177 to 173 us and the FP32 noisy one at 1024^2 from 10.3 to 10.1 us).
    const bool lambda_b = !active;  // stage B applies G3xoG3x, G3yoG3y to mu_a
    T d2x[R + 4], sx[R + 4];        // mu_a rows r0-2 .. r0+R+1
    T hx[R + 2], sy[R + 2];         // rows r0-1 .. r0+R of q, or of C (above)
    T ycol[kNeverNoisy ? 1 : R + 8];  // J_y at rows r0-4 .. r0+R+3 of this column
    const auto Ab = lds_base(s_mua + r0 * L::AW + c + 2);
    const auto Qxb = lds_base(s_qx + (r0 + 1) * L::QW + c + 2);
    const auto Qyb = lds_base(s_qy + (r0 + 1) * L::QW + c + 2);
    const auto Yc = lds_base(s_jy + r0 * L::YW + c);
    auto row_a = [&](int k) {
        if (lambda_b) {
            const auto A = Ab + k * L::AW;
            const T o2 = A[-2] + A[2];
            d2x[k] = o2 - A[0] - A[0];                                            // (1,0,-2,0,1) along x
            sx[k] = T(kG3G3S[0]) * A[0] + T(kG3G3S[1]) * (A[-1] + A[1]) + T(kG3G3S[2]) * o2;  // (.01,.06,.11,.06,.01)
        }
    };
    auto row_q = [&](int k) {
        const auto Qx = Qxb + k * L::QW;
        const auto Qy = Qyb + k * L::QW;
        hx[k] = Qx[1] - Qx[-1];
        sy[k] = K.gyc * Qy[0] + K.gye * (Qy[-1] + Qy[1]);
    };
    auto row_y = [&](int k) {
        if constexpr (!kNeverNoisy) {
            if (noisy) ycol[k] = Yc[k * L::YW];
        }
    };
#pragma unroll
    for (int k = 0; k < 4; ++k) row_a(k);
#pragma unroll
    for (int k = 0; k < 2; ++k) row_q(k);
#pragma unroll
    for (int k = 0; k < 8; ++k) row_y(k);
    const auto Xb = lds_base(s_jx + r0 * L::XW + c + 4);
    const auto Mb = lds_base(s_mup + (r0 + 8) * L::MW + c + 8);
    T vout[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = r0 + k;
        row_a(k + 4);
        row_q(k + 2);
        row_y(k + 8);
        T acc = T(0);
        if constexpr (!kNeverNoisy) {
            if (noisy) {
                // passive and noise: (1/dx) D8_x J_x + (1/dy) D8_y J_y
                const auto X = Xb + k * L::XW;
                acc = K.dxw[1] * (X[1] - X[-1]);
                acc = fma(K.dyw[1], ycol[k + 5] - ycol[k + 3], acc);
#pragma unroll
                for (int j = 2; j <= 4; ++j) {
                    acc = fma(K.dxw[j], X[j] - X[-j], acc);
                    acc = fma(K.dyw[j], ycol[k + 4 + j] - ycol[k + 4 - j], acc);
                }
            }
        }
        if constexpr (!kAlwaysNoisy) {
            if (!noisy) {
                // passive: -(1/dx^2) D8oD8_x mu_p - (1/dy^2) D8oD8_y mu_p
                const auto m = Mb + k * L::MW;
                acc = K.p0 * mcol[k + 8];
#pragma unroll
                for (int j = 1; j <= 8; ++j) {
                    acc = fma(K.px[j], m[j] + m[-j], acc);
                    acc = fma(K.py[j], mcol[k + 8 + j] + mcol[k + 8 - j], acc);
                }
            }
        }
        if (lambda_b) {
            // lambda: -(1/dx^2) G3xoG3x mu_a - (1/dy^2) G3yoG3y mu_a
            acc = fma(K.lxx[0], d2x[k + 2], acc);
            acc = fma(K.lxx[1], d2x[k + 1] + d2x[k + 3], acc);
            acc = fma(K.lxx[2], d2x[k] + d2x[k + 4], acc);
            acc = fma(K.lyy[0], (sx[k] + sx[k + 4]) - sx[k + 2] - sx[k + 2], acc);
        }
        // circulating, (1/dx) G3x q_x + (1/dy) G3y q_y, or with the active current the
        // whole active part, (1/dx) G3x C_x + (1/dy) G3y C_y
        acc = fma(K.gxc, hx[k + 1], acc);
        acc = fma(K.gxe, hx[k] + hx[k + 2], acc);
        acc += sy[k + 2] - sy[k];
        const T v = fma(-K.dt, acc, phi_c[k]);
        vout[k] = v;
        if (interior || (y0 + r < ny && x0 + c < nx)) *orow = v;
        orow += nx;
    }

    // ---- fused finite check and FP64 mass (the last step of a run): the tile's sum of
    // finite new values and its count of non-finite ones, reduced in a fixed order (lane
    // butterfly, then waves in index order), so the result does not depend on timing.
    if (red_out != nullptr) {
        double s = 0.0, bad = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (interior || (y0 + r0 + k < ny && x0 + c < nx)) {
                const double d = (double)vout[k];
                if (isfinite(d)) s += d;
                else bad += 1.0;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_xor(s, off, 64);
            bad += __shfl_xor(bad, off, 64);
        }
        double* const red = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(sm) + L::red_off);
        if ((tid & 63) == 0) {
            red[2 * (tid >> 6)] = s;
            red[2 * (tid >> 6) + 1] = bad;
        }
        __syncthreads();
        if (tid == 0) {
            double ts = 0.0, tb = 0.0;
#pragma unroll
            for (int w = 0; w < NT / 64; ++w) {
                ts += red[2 * w];
                tb += red[2 * w + 1];
            }
            red_out[0] = ts;
            red_out[1] = tb;
        }
    }
}

}  // namespace amb
