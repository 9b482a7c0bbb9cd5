/*
 * amb_oracle.h -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY; see amb_oracle.c).
 *
 * Noise mapping shared with the HIP kernels.  For step n, component c (0 = y,
 * 1 = x), row i and column j, the normal xi_c(i, j, n) comes from one
 * Philox4x32-10 call with
 *     counter = { j / 4, i, low32(n), (high32(n) & 0x7fffffff) | (c << 31) }
 *     key     = { low32(seed), high32(seed) }
 * whose four output words (w0, w1, w2, w3) give, by Box-Muller with
 *     u1 = (w_even + 1/2) * 2^-32 in (0, 1),  u2 = w_odd * 2^-32 in [0, 1),
 *     r = sqrt(-2 ln u1),  (z0, z1) = r (cos 2 pi u2, sin 2 pi u2),
 * the normals of columns 4g, 4g+1 (pair (w0, w1)) and 4g+2, 4g+3 (pair (w2, w3)).
 * The noise current is J_noise = sqrt(2T) sqrt(1/(dt dx dy)) * xi
 * (src/parameters.h:50,86; src/integrator.cu:106-109).
 */
#ifndef AMB_ORACLE_H
#define AMB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMO_TWO_M32 2.3283064365386963e-10 /* 2^-32 */

typedef struct { double dt, dx, dy; } amo_stencil;
typedef struct { double a, b, c, kappa, lamb, zeta, T; } amo_model;

/* Philox4x32-10 block function. */
void amo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The four normals of column group `group` in row `row` (mapping above). */
void amo_noise_group(uint64_t seed, int64_t step, int comp, int64_t row, int64_t group, double z[4]);

/* All normals of one step: xi[0..N) = component y, xi[N..2N) = component x. */
void amo_noise(int64_t ny, int64_t nx, uint64_t seed, int64_t step, double* xi);

/* Every intermediate of one step at the current field.  xi (2N unscaled normals)
 * may be NULL for zero noise.  Outputs may be NULL; currents are 2N: [y | x].
 * jl = lambda current -G3 mu_a, jz = circulating current zeta lap(phi) G3 phi,
 * jp = passive current -D8 mu_p, jn = noise current, divj = div J. */
int amo_terms(const double* phi, int64_t ny, int64_t nx, const amo_stencil* st,
              const amo_model* m, const double* xi, double* mup, double* mua, double* jp,
              double* jl, double* jz, double* jn, double* divj);

/* nsteps explicit-Euler steps in place.  Noise (only when T != 0): from `noise`
 * (nsteps * 2N normals, step-major) if given, otherwise the Philox stream for
 * steps step0 .. step0+nsteps-1 under `seed`.  nthreads >= 1 (OpenMP over rows). */
int amo_run(double* phi, int64_t ny, int64_t nx, const amo_stencil* st, const amo_model* m,
            int64_t nsteps, uint64_t seed, int64_t step0, const double* noise, int nthreads);

int amo_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
