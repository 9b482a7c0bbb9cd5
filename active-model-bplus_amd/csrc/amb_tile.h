// amb_tile.h -- device code of one fused step over one tile (shared by the single-field
// and the batched-ensemble kernels).  See amb_step.hip for the algorithm.
#pragma once

#include <hip/hip_runtime.h>

#include "amb_kernels.h"
#include "amb_math.h"

namespace amb {

constexpr int kThreads = 256;

__device__ __forceinline__ int wrap_index(int g, int n)
{
    // Exact for any g; one iteration in the common case, several on tiny grids.
    while (g < 0) g += n;
    while (g >= n) g -= n;
    return g;
}

template <typename T, int TW, int TH>
struct Lds {
    static constexpr int PW = TW + 18, PH = TH + 18;  // phi: rows [-9,TH+9), cols [-9,TW+9)
    static constexpr int MW = TW + 16, MH = TH + 16;  // mu_p: rows [-8,TH+8), cols [-8,TW+8)
    static constexpr int AW = TW + 4, AH = TH + 4;    // mu_a: [-2, +2)
    static constexpr int QW = TW + 2, QH = TH + 2;    // circulating current: [-1, +1)
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
    static constexpr size_t bytes = sizeof(T) * (size_t)total;
    static constexpr int rows_per_thread = TW * TH / kThreads;
    static_assert(kThreads % TW == 0, "tile width must divide the workgroup");
    static_assert((TW * TH) % kThreads == 0, "tile must be a multiple of the workgroup");
    static_assert(TW % 4 == 0, "noise column groups need TW % 4 == 0");
};

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

__device__ __forceinline__ void box_muller(uint32_t wa, uint32_t wb, double& z0, double& z1)
{
    const double u1 = ((double)wa + 0.5) * 0x1p-32;
    const double u2 = (double)wb * 0x1p-32;
    const double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincos(2.0 * M_PI * u2, &s, &c);
    z0 = r * c;
    z1 = r * s;
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

// ---- fused step --------------------------------------------------------------

template <typename T, int TW, int TH, int NOISE>
__device__ __forceinline__ void step_tile(T* const sm, const T* __restrict__ in, T* __restrict__ out, const int ny,
                                          const int nx, const int nx4, const int x0, const int y0,
                                          const StepCoef<T>& K, const uint32_t key0, const uint32_t key1,
                                          const uint64_t step, const double* noise)
{
    using L = Lds<T, TW, TH>;
    T* const s_phi = sm;
    T* const s_mup = sm + L::off_mup;
    T* const s_mua = sm + L::off_mua;
    T* const s_qx = sm + L::off_qx;
    T* const s_qy = sm + L::off_qy;
    T* const s_xix = sm;                 // aliases s_phi after stage A
    T* const s_xiy = sm + L::XW * L::XH;

    const int tid = threadIdx.x;

    // stage 0: phi with a 9-site halo
    for (int i = tid; i < L::n_phi; i += kThreads) {
        const int r = i / L::PW, c = i - r * L::PW;
        const int gr = wrap_index(y0 - 9 + r, ny), gc = wrap_index(x0 - 9 + c, nx);
        s_phi[i] = in[(size_t)gr * nx + gc];
    }
    __syncthreads();

    constexpr int R = L::rows_per_thread;
    const int cx = tid % TW;
    const int ry = (tid / TW) * R;
    T phi_c[R];
#pragma unroll
    for (int k = 0; k < R; ++k) phi_c[k] = s_phi[(ry + k + 9) * L::PW + cx + 9];

    // stage A: local terms on the halo region (corners beyond the cross skipped)
    for (int i = tid; i < L::MW * L::MH; i += kThreads) {
        const int rr = i / L::MW, cc = i - rr * L::MW;
        const int r = rr - 8, c = cc - 8;
        const bool rin = (unsigned)r < (unsigned)TH, cin = (unsigned)c < (unsigned)TW;
        if (!(rin || cin)) continue;
        const T* q = s_phi + (r + 9) * L::PW + (c + 9);
        const T pmm = q[-L::PW - 1], pm0 = q[-L::PW], pmp = q[-L::PW + 1];
        const T p0m = q[-1], p00 = q[0], p0p = q[1];
        const T ppm = q[L::PW - 1], pp0 = q[L::PW], ppp = q[L::PW + 1];
        const T gx = K.gxe * ((pmp - pmm) + (ppp - ppm)) + K.gxc * (p0p - p0m);
        const T gy = K.gye * ((ppm - pmm) + (ppp - pmp)) + K.gyc * (pp0 - pm0);
        const T lap = K.lcc * p00 + K.lex * (p0m + p0p) + K.ley * (pm0 + pp0) +
                      K.lck * ((pmm + pmp) + (ppm + ppp));
        s_mup[i] = p00 * (K.a + p00 * (K.b + p00 * K.c)) - K.kappa * lap;
        if ((unsigned)(r + 2) < (unsigned)(TH + 4) && (unsigned)(c + 2) < (unsigned)(TW + 4))
            s_mua[(r + 2) * L::AW + (c + 2)] = K.lam * (gy * gy + gx * gx);
        if ((unsigned)(r + 1) < (unsigned)(TH + 2) && (unsigned)(c + 1) < (unsigned)(TW + 2)) {
            const T zl = K.zeta * lap;
            s_qx[(r + 1) * L::QW + (c + 1)] = zl * gx;
            s_qy[(r + 1) * L::QW + (c + 1)] = zl * gy;
        }
    }
    __syncthreads();

    // stage N: noise normals (overwrites the phi buffer)
    bool noisy = NOISE == kNoisePhilox || NOISE == kNoiseBuffer;
    if constexpr (NOISE == kNoiseRuntime) noisy = K.amp != T(0);
    if (NOISE == kNoisePhilox || (NOISE == kNoiseRuntime && noisy)) {
        constexpr int GXN = TW / 4 + 2;  // xi_x column groups: tile-relative [-1, TW/4]
        for (int i = tid; i < GXN * TH; i += kThreads) {
            const int r = i / GXN, gl = i - r * GXN - 1;
            const int grow = wrap_index(y0 + r, ny);
            T* dst = s_xix + r * L::XW + 4 * (gl + 1);
            if (nx4) {
                const int gc = wrap_index(x0 + 4 * gl, nx);
                T z[4];
                normals4<T>((uint32_t)(gc >> 2), (uint32_t)grow, step, 1, key0, key1, z);
                dst[0] = z[0]; dst[1] = z[1]; dst[2] = z[2]; dst[3] = z[3];
            } else {
                for (int l = 0; l < 4; ++l)
                    dst[l] = normal1<T>(grow, wrap_index(x0 + 4 * gl + l, nx), step, 1, key0, key1);
            }
        }
        constexpr int GYN = TW / 4;  // xi_y column groups [0, TW/4), rows [-4, TH+4)
        for (int i = tid; i < GYN * L::YH; i += kThreads) {
            const int rr = i / GYN, gl = i - rr * GYN;
            const int grow = wrap_index(y0 + rr - 4, ny);
            T z[4];
            normals4<T>((uint32_t)((x0 >> 2) + gl), (uint32_t)grow, step, 0, key0, key1, z);
            T* dst = s_xiy + rr * L::YW + 4 * gl;
            dst[0] = z[0]; dst[1] = z[1]; dst[2] = z[2]; dst[3] = z[3];
        }
        __syncthreads();
    } else if (NOISE == kNoiseBuffer) {
        const size_t n = (size_t)ny * nx;
        for (int i = tid; i < L::XW * L::XH; i += kThreads) {
            const int r = i / L::XW, c = i - r * L::XW;
            s_xix[i] = (T)noise[n + (size_t)wrap_index(y0 + r, ny) * nx + wrap_index(x0 + c - 4, nx)];
        }
        for (int i = tid; i < L::YW * L::YH; i += kThreads) {
            const int r = i / L::YW, c = i - r * L::YW;
            s_xiy[i] = (T)noise[(size_t)wrap_index(y0 + r - 4, ny) * nx + wrap_index(x0 + c, nx)];
        }
        __syncthreads();
    }

    // stage B: divergence and update
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = ry + k, c = cx;
        // passive: -(1/dy^2) D8oD8_y mu_p - (1/dx^2) D8oD8_x mu_p
        const T* m = s_mup + (r + 8) * L::MW + (c + 8);
        T acc = K.p0 * m[0];
#pragma unroll
        for (int j = 1; j <= 8; ++j)
            acc += K.px[j] * (m[j] + m[-j]) + K.py[j] * (m[j * L::MW] + m[-j * L::MW]);
        // lambda: -(1/dx^2) G3xoG3x mu_a - (1/dy^2) G3yoG3y mu_a
        const T* A = s_mua + (r + 2) * L::AW + (c + 2);
        T gxx = K.lxx[0] * (A[2] + A[-2] - A[0] - A[0]);
        T gyy = K.lyy[0] * (A[2 * L::AW] + A[-2 * L::AW] - A[0] - A[0]);
#pragma unroll
        for (int a = 1; a <= 2; ++a) {
            const int u = a * L::AW;
            gxx += K.lxx[a] * ((A[u + 2] + A[u - 2] - A[u] - A[u]) + (A[-u + 2] + A[-u - 2] - A[-u] - A[-u]));
            gyy += K.lyy[a] * ((A[2 * L::AW + a] + A[-2 * L::AW + a] - A[a] - A[a]) +
                               (A[2 * L::AW - a] + A[-2 * L::AW - a] - A[-a] - A[-a]));
        }
        acc += gxx + gyy;
        // circulating: (1/dx) G3x q_x + (1/dy) G3y q_y
        const T* Qx = s_qx + (r + 1) * L::QW + (c + 1);
        const T* Qy = s_qy + (r + 1) * L::QW + (c + 1);
        acc += K.gxc * (Qx[1] - Qx[-1]) + K.gxe * ((Qx[L::QW + 1] - Qx[L::QW - 1]) + (Qx[-L::QW + 1] - Qx[-L::QW - 1]));
        acc += K.gyc * (Qy[L::QW] - Qy[-L::QW]) + K.gye * ((Qy[L::QW + 1] - Qy[-L::QW + 1]) + (Qy[L::QW - 1] - Qy[-L::QW - 1]));
        if (noisy) {
            const T* X = s_xix + r * L::XW + (c + 4);
            const T* Y = s_xiy + (r + 4) * L::YW + c;
            T nsum = 0;
#pragma unroll
            for (int j = 1; j <= 4; ++j)
                nsum += K.nxw[j] * (X[j] - X[-j]) + K.nyw[j] * (Y[j * L::YW] - Y[-j * L::YW]);
            acc += nsum;
        }
        const int gr = y0 + r, gc = x0 + c;
        if (gr < ny && gc < nx) out[(size_t)gr * nx + gc] = phi_c[k] - K.dt * acc;
    }
}

}  // namespace amb
