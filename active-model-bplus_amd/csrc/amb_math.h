// amb_math.h -- stencil weights, Philox4x32-10 and Box-Muller shared by the HIP kernels.
//
// Discretisation of one explicit-Euler step of Active Model B+ (the reference's
// scheme, restated from its mathematics; reference citations below):
//
//   G3x = rows (0.1, 0.3, 0.1) (x) columns (-1, 0, 1), scaled 1/dx      src/finite_difference.h:379-387
//   G3y = transpose of G3x, scaled 1/dy                                    src/finite_difference.h:404-407
//   L3  = S2x/dx^2 + S2y/dy^2, S2x = rows (-1/4, 3/2, -1/4) (x) (1,-2,1)    src/finite_difference.h:389-412,
//                                                                          src/finite_difference.cuh:178-183
//   D8  = 8th-order central first derivative, taps (4/5, -1/5, 4/105, -1/280) for offsets 1..4
//                                                                          src/finite_difference.h:73
//   mu_p = a phi + b phi^2 + c phi^3 - kappa L3 phi                        src/integrator.cu:34-37,99-100
//   mu_a = lambda |G3 phi|^2                                               src/integrator.cu:101
//   J_zeta = zeta (L3 phi) G3 phi,  J_noise = sqrt(2T/(dt dx dy)) xi       src/integrator.cu:103-109
//   J_p = -D8 mu_p,  J_lambda = -G3 mu_a                                   src/integrator.cu:171-175
//   div J = G3 . (J_lambda + J_zeta) + D8 . (J_p + J_noise);  phi -= dt div J
//                                                                          src/integrator.cu:259-261
//
// The fused kernel never materialises J_p or J_lambda: D8(D8 mu_p) is applied as
// one 17-tap stencil (the exact convolution of D8 with itself) and G3x(G3x mu_a)
// as the separable 5x5 product rows (.01,.06,.11,.06,.01) (x) columns (1,0,-2,0,1).
#pragma once

#include <stdint.h>

#include <cmath>

namespace amb {

// D8 taps for offsets 1..4 (antisymmetric).
constexpr double kD8[5] = {0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0};

// D8 o D8 for offsets 0..8 (symmetric): sum_{k+l=m} d_k d_l, exact rationals.
constexpr double kD8D8[9] = {
    -480841.0 / 352800.0, 411.0 / 1225.0, 1213.0 / 2100.0, -11.0 / 35.0, 53.0 / 525.0,
    -11.0 / 525.0,        127.0 / 44100.0, -1.0 / 3675.0,  1.0 / 78400.0};

// Smoothing profile of G3 across its derivative axis, and of G3 o G3.
constexpr double kG3S[2] = {0.3, 0.1};                 // centre, +-1
constexpr double kG3G3S[3] = {0.11, 0.06, 0.01};       // centre, +-1, +-2

// Noise mapping (identical in the CPU oracle): one Philox4x32-10 call per
// (column group of 4, row, step, component); the 4 words give 4 normals by two
// Box-Muller transforms.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

struct U4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c: one v_bitop3_b32 (truth table 0x96) on gfx950.
__host__ __device__ inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
        U4 n;
        n.x = xor3((uint32_t)(p1 >> 32), c.y, k0);
        n.y = (uint32_t)p1;
        n.z = xor3((uint32_t)(p0 >> 32), c.w, k1);
        n.w = (uint32_t)p0;
        c = n;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c;
}

__host__ __device__ inline U4 noise_counter(uint32_t group, uint32_t row, uint64_t step, int comp)
{
    U4 c;
    c.x = group;
    c.y = row;
    c.z = (uint32_t)step;
    c.w = ((uint32_t)(step >> 32) & 0x7fffffffu) | ((uint32_t)comp << 31);
    return c;
}

// FP64 Box-Muller on the two uniforms a Philox word pair gives.  The generic math library
// handles every input (zeros, infinities, huge arguments, denormals) with special-case code
// that dominated the FP64 noisy step; here the inputs are known, so short kernels of our own
// suffice, all accurate to a few ulp:
//   -2 ln u for u = (w + 1/2) 2^-32 in [2^-33, 1): u = m 2^e with m in [sqrt(1/2), sqrt(2))
//     (e from the exponent of u sqrt(2)), f = m - 1 (exact), s = f / (m + 1), |s| <= 3 - 2
//     sqrt(2) = 0.1716, ln m = 2 atanh(s) = 2 s + 2 s^3 g(s^2) with g a 7-term polynomial;
//     e ln 2 with ln 2 split in a high part whose products with |e| <= 34 are exact and a
//     low part;
//   sin and cos of 2 pi u for u = w 2^-32 in [0, 1): quadrant j = rint(4u), reduced
//     argument x = (4u - j) pi/2 in [-pi/4, pi/4] (4u - j is exact), sin x = x + x^3 S(x^2),
//     cos x = 1 - x^2/2 + x^4 C(x^2) with 6-term polynomials S and C, then the quadrant's
//     swap and signs;
//   the quotient s and the square root of -2 ln u (a normal number in [2.3e-10, 46.1]) by
//     Newton steps from the gfx950 reciprocal and reciprocal-square-root instructions.
// g, S and C are Chebyshev fits at 60 digits (noise_poly_fit.py prints them and their
// worst relative errors over their intervals: 4.7e-16, 2.0e-16, 5.7e-17).
constexpr double kLnG[7] = {0x1.2b584aae78a57p-4, 0x1.39fe606542ddep-4, 0x1.7462b4ab2ef6bp-4, 0x1.c71c62e5800a1p-4,
                            0x1.2492492df148dp-3, 0x1.99999999952e2p-3, 0x1.5555555555558p-2};
constexpr double kSinS[6] = {0x1.5e0b19f8b1451p-33, -0x1.ae600b02b6262p-26, 0x1.71de37968a100p-19,
                             -0x1.a01a019e83aaep-13, 0x1.1111111110bb2p-7,  -0x1.5555555555555p-3};
constexpr double kCosC[6] = {-0x1.907da367a37cbp-37, 0x1.1eeb68e93b64cp-29, -0x1.27e4fa17da09ep-22,
                             0x1.a01a019f4eb01p-16,  -0x1.6c16c16c16967p-10, 0x1.5555555555555p-5};

// Exponent e of a finite non-zero double, x = m 2^e with m in [1/2, 1); x 2^k.
__host__ __device__ inline int exponent_of(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_frexp_exp(x);
#else
    int e;
    (void)std::frexp(x, &e);
    return e;
#endif
}

__host__ __device__ inline double scale2(double x, int k)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ldexp(x, k);
#else
    return std::ldexp(x, k);
#endif
}

// Seeds for the Newton steps: v_rcp_f64 and v_rsq_f64 on the device, the correctly
// rounded values on the host (the steps below bring either to the same few-ulp result).
__host__ __device__ inline double rcp_seed(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcp(d);
#else
    return 1.0 / d;
#endif
}

__host__ __device__ inline double rsq_seed(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsq(x);
#else
    return 1.0 / std::sqrt(x);
#endif
}

// -2 ln u for u in [2^-33, 1).
__host__ __device__ inline double neg2_ln_unit(double u)
{
    const int e = exponent_of(u * 1.41421356237309504880) - 1;
    const double m = scale2(u, -e);  // exact, in [sqrt(1/2), sqrt(2)) up to the rounding of u sqrt(2)
    const double f = m - 1.0;        // exact
    const double d = 2.0 + f;
    double r = rcp_seed(d);
    r = fma(r, fma(-d, r, 1.0), r);  // one Newton step on 1/d
    double s = f * r;
    s = fma(fma(-d, s, f), r, s);    // one residual correction of f/d
    const double t = s * s;
    double p = kLnG[0];
#pragma unroll
    for (int k = 1; k < 7; ++k) p = fma(p, t, kLnG[k]);
    const double q = -4.0 * s;
    const double lnm2 = fma(q * t, p, q);  // -2 ln m = -4 s (1 + s^2 g(s^2))
    const double de = (double)e;
    constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 32 significant bits: de * kLn2Hi is exact
    constexpr double kLn2Lo = 1.90821492927058770002e-10;
    return fma(de, -2.0 * kLn2Hi, fma(de, -2.0 * kLn2Lo, lnm2));
}

// ln u for u in [2^-33, 1): the value the kernels use, halved back (exact).
__host__ __device__ inline double ln_unit(double u) { return -0.5 * neg2_ln_unit(u); }

// sqrt x for a normal x well inside the double range.
__host__ __device__ inline double sqrt_normal(double x)
{
    const double y = rsq_seed(x);
    double g = x * y, h = 0.5 * y;
    const double e = fma(-g, h, 0.5);
    g = fma(g, e, g);  // sqrt x
    h = fma(h, e, h);  // 1 / (2 sqrt x)
    return fma(fma(-g, g, x), h, g);
}

__host__ __device__ inline void sincos_turn(double u, double& s, double& c)
{
    const double q = 4.0 * u;
    const double j = rint(q);
    const double x = (q - j) * 1.57079632679489661923;  // (4u - j) pi/2
    const double x2 = x * x;
    double sp = kSinS[0], cp = kCosC[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) {
        sp = fma(sp, x2, kSinS[k]);
        cp = fma(cp, x2, kCosC[k]);
    }
    const double sx = fma(x * x2, sp, x);
    const double cx = fma(x2 * x2, cp, fma(-0.5, x2, 1.0));
    const int k = (int)j & 3;
    const double a = (k & 1) ? cx : sx;  // sin(theta) up to sign
    const double b = (k & 1) ? sx : cx;  // cos(theta) up to sign
    s = (k & 2) ? -a : a;
    c = (k == 1 || k == 2) ? -b : b;
}

__host__ __device__ inline void box_muller(uint32_t wa, uint32_t wb, double& z0, double& z1)
{
    const double u1 = ((double)wa + 0.5) * 0x1p-32;
    const double u2 = (double)wb * 0x1p-32;
    const double r = sqrt_normal(neg2_ln_unit(u1));
    double s, c;
    sincos_turn(u2, s, c);  // sin and cos of 2 pi u2
    z0 = r * c;
    z1 = r * s;
}

}  // namespace amb
