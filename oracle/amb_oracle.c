/*
 * amb_oracle.c -- CPU FP64 oracle for the Active Model B+ explicit-Euler step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * What it restates (every formula is written from the mathematics; no code is
 * taken from the reference):
 *
 *   pass structure            test/tjhung_ambplus.h:192-251 (calculate_dphi),
 *                             test/tjhung_ambplus.h:182-190 (update_phi)
 *   term assembly             src/integrator.cu:34-37   bulk mu = a phi + b phi^2 + c phi^3
 *                             src/integrator.cu:96-109  mu_p, mu_a, J_zeta, J_noise
 *                             src/integrator.cu:171-175 J_p = -D8 mu_p, J_lambda = -G3 mu_a
 *                             src/integrator.cu:203-204,259-261 divergence + Euler update
 *   stencil taps              src/finite_difference.h:73 (D8 first derivative),
 *                             src/finite_difference.h:379-412 (Tjhung 3x3 G3x, S2x, transposes)
 *   per-axis scaling          src/finite_difference.cuh:29-39,166-183 (1/dx, 1/dy, 1/dx^2, 1/dy^2)
 *   derived constants         src/parameters.h:43-51,79-87 (noise amplitude sqrt(2T/(dt dx dy)))
 *   component convention      component 0 = y = row direction (scaled by 1/dy),
 *                             component 1 = x = column direction (scaled by 1/dx);
 *                             src/integrator.cu:97-104,171-175
 *
 * The reference's own CPU scheme divides its Laplacian by dx*dy
 * (test/tjhung_ambplus.h:140); the GPU path it ships uses 1/dx^2 and 1/dy^2 per
 * axis.  This oracle follows the GPU convention (they agree when dx == dy).
 *
 * Noise: the reference draws cuRAND normals from a non-deterministic seed
 * (src/integrator.cu:106-109,320-323), which no test can pin.  Here the noise is
 * a counter-based Philox4x32-10 stream (Salmon, Moraes, Dror, Shaw, SC'11)
 * mapped to normals by Box-Muller; the mapping is defined in amb_oracle.h and is
 * the same one the HIP kernels implement.  A caller may instead pass explicit
 * normals (shared-noise-buffer parity).
 *
 * Field layout: row-major (ny, nx), linear index = col + row*nx, periodic in both
 * axes (src/math_primitives.h:7, src/integrator.cu:59,72-90).
 */
#include "amb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- stencil taps ------------------------------------------------------- */

/* D8: 8th-order central first derivative, antisymmetric taps for offsets 1..4
 * (src/finite_difference.h:73: [1/280, -4/105, 1/5, -4/5, 0, 4/5, -1/5, 4/105, -1/280]). */
static const double D8[5] = {0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0};

/* Tjhung 3x3 first derivative along columns (src/finite_difference.h:379-387):
 * rows (y-1, y, y+1) weights (0.1, 0.3, 0.1), columns (x-1, x, x+1) weights (-1, 0, +1). */
static const double G3_SMOOTH[3] = {0.1, 0.3, 0.1};
/* Tjhung 3x3 second derivative along columns (src/finite_difference.h:389-400):
 * rows weights (-0.25, 1.5, -0.25) times columns (1, -2, 1). */
static const double S2_SMOOTH[3] = {-0.25, 1.5, -0.25};

static inline int64_t wrap(int64_t i, int64_t n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

#define AT(f, r, c) (f)[wrap((r), ny) * nx + wrap((c), nx)]

/* Tjhung first derivative along x (columns), unscaled. */
static inline double g3x(const double* f, int64_t ny, int64_t nx, int64_t i, int64_t j)
{
    double s = 0.0;
    for (int a = -1; a <= 1; ++a)
        s += G3_SMOOTH[a + 1] * (AT(f, i + a, j + 1) - AT(f, i + a, j - 1));
    return s;
}

/* Tjhung first derivative along y (rows), unscaled: the transpose of g3x. */
static inline double g3y(const double* f, int64_t ny, int64_t nx, int64_t i, int64_t j)
{
    double s = 0.0;
    for (int b = -1; b <= 1; ++b)
        s += G3_SMOOTH[b + 1] * (AT(f, i + 1, j + b) - AT(f, i - 1, j + b));
    return s;
}

/* Tjhung Laplacian with per-axis scaling: S2x/dx^2 + S2y/dy^2. */
static inline double l3(const double* f, int64_t ny, int64_t nx, int64_t i, int64_t j,
                        double idx2, double idy2)
{
    double sx = 0.0, sy = 0.0;
    for (int a = -1; a <= 1; ++a)
        sx += S2_SMOOTH[a + 1] * (AT(f, i + a, j + 1) - 2.0 * AT(f, i + a, j) + AT(f, i + a, j - 1));
    for (int b = -1; b <= 1; ++b)
        sy += S2_SMOOTH[b + 1] * (AT(f, i + 1, j + b) - 2.0 * AT(f, i, j + b) + AT(f, i - 1, j + b));
    return idx2 * sx + idy2 * sy;
}

static inline double d8x(const double* f, int64_t ny, int64_t nx, int64_t i, int64_t j)
{
    double s = 0.0;
    for (int k = 1; k <= 4; ++k) s += D8[k] * (AT(f, i, j + k) - AT(f, i, j - k));
    return s;
}

static inline double d8y(const double* f, int64_t ny, int64_t nx, int64_t i, int64_t j)
{
    double s = 0.0;
    for (int k = 1; k <= 4; ++k) s += D8[k] * (AT(f, i + k, j) - AT(f, i - k, j));
    return s;
}

/* ---- Philox4x32-10 and the noise mapping --------------------------------- */

void amo_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    /* Salmon et al. (SC'11), Philox4x32 with R = 10 rounds: multipliers
     * 0xD2511F53, 0xCD9E8D57 and Weyl key increments 0x9E3779B9, 0xBB67AE85.
     * Round r > 0 uses the key bumped r times. */
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void box_muller(uint32_t wa, uint32_t wb, double* z0, double* z1)
{
    const double u1 = ((double)wa + 0.5) * AMO_TWO_M32;
    const double u2 = (double)wb * AMO_TWO_M32;
    const double r = sqrt(-2.0 * log(u1));
    const double th = 2.0 * M_PI * u2;
    *z0 = r * cos(th);
    *z1 = r * sin(th);
}

void amo_noise_group(uint64_t seed, int64_t step, int comp, int64_t row, int64_t group, double z[4])
{
    uint32_t ctr[4], key[2], w[4];
    ctr[0] = (uint32_t)group;
    ctr[1] = (uint32_t)row;
    ctr[2] = (uint32_t)((uint64_t)step);
    ctr[3] = ((uint32_t)((uint64_t)step >> 32) & 0x7fffffffu) | ((uint32_t)comp << 31);
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    amo_philox4x32_10(ctr, key, w);
    box_muller(w[0], w[1], &z[0], &z[1]);
    box_muller(w[2], w[3], &z[2], &z[3]);
}

void amo_noise(int64_t ny, int64_t nx, uint64_t seed, int64_t step, double* xi)
{
    const int64_t n = ny * nx;
    const int64_t ngroups = (nx + 3) / 4;
    for (int comp = 0; comp < 2; ++comp)
        for (int64_t i = 0; i < ny; ++i)
            for (int64_t g = 0; g < ngroups; ++g) {
                double z[4];
                amo_noise_group(seed, step, comp, i, g, z);
                for (int l = 0; l < 4; ++l) {
                    const int64_t j = 4 * g + l;
                    if (j < nx) xi[comp * n + i * nx + j] = z[l];
                }
            }
}

/* ---- the four passes ----------------------------------------------------- */

typedef struct {
    double* gy; double* gx; double* lap;        /* pass 1 */
    double* mup; double* mua;                   /* pass 2 */
    double* jp; double* jl; double* jz; double* jn; /* pass 3: 2N each, [y | x] */
} scratch_t;

static int scratch_alloc(scratch_t* s, int64_t n)
{
    memset(s, 0, sizeof(*s));
    double* block = (double*)malloc(sizeof(double) * (size_t)n * 13);
    if (!block) return -1;
    s->gy = block; s->gx = block + n; s->lap = block + 2 * n;
    s->mup = block + 3 * n; s->mua = block + 4 * n;
    s->jp = block + 5 * n; s->jl = block + 7 * n; s->jz = block + 9 * n; s->jn = block + 11 * n;
    return 0;
}

static void scratch_free(scratch_t* s) { free(s->gy); }

static void compute_terms(const double* phi, int64_t ny, int64_t nx, const amo_stencil* st,
                          const amo_model* m, const double* xi, scratch_t* s, double* divj,
                          int nthreads)
{
    const int64_t n = ny * nx;
    const double idx = 1.0 / st->dx, idy = 1.0 / st->dy;
    const double idx2 = idx * idx, idy2 = idy * idy;
    const double amp = sqrt(2.0 * m->T) * sqrt(1.0 / (st->dt * st->dx * st->dy));
    (void)nthreads;

    /* pass 1: Tjhung gradient and Laplacian of phi (test/tjhung_ambplus.h:199-211) */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < ny; ++i)
        for (int64_t j = 0; j < nx; ++j) {
            const int64_t p = i * nx + j;
            s->gy[p] = idy * g3y(phi, ny, nx, i, j);
            s->gx[p] = idx * g3x(phi, ny, nx, i, j);
            s->lap[p] = l3(phi, ny, nx, i, j, idx2, idy2);
        }

    /* pass 2: chemical potentials, circulating and noise currents
     * (test/tjhung_ambplus.h:207-224; src/integrator.cu:99-109) */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const double f = phi[p];
        s->mup[p] = m->a * f + m->b * f * f + m->c * f * f * f - m->kappa * s->lap[p];
        s->mua[p] = m->lamb * (s->gy[p] * s->gy[p] + s->gx[p] * s->gx[p]);
        s->jz[p] = m->zeta * s->lap[p] * s->gy[p];
        s->jz[n + p] = m->zeta * s->lap[p] * s->gx[p];
        s->jn[p] = xi ? amp * xi[p] : 0.0;
        s->jn[n + p] = xi ? amp * xi[n + p] : 0.0;
    }

    /* pass 3: local currents (test/tjhung_ambplus.h:227-240; src/integrator.cu:171-175) */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < ny; ++i)
        for (int64_t j = 0; j < nx; ++j) {
            const int64_t p = i * nx + j;
            s->jp[p] = -idy * d8y(s->mup, ny, nx, i, j);
            s->jp[n + p] = -idx * d8x(s->mup, ny, nx, i, j);
            s->jl[p] = -idy * g3y(s->mua, ny, nx, i, j);
            s->jl[n + p] = -idx * g3x(s->mua, ny, nx, i, j);
        }

    if (!divj) return;

    /* pass 4: divergence, grouping J1 = Jp + Jn (D8) and J2 = Jl + Jz (G3) as the
     * GPU step does (src/integrator.cu:203-204,259-260). */
    double* j1 = (double*)malloc(sizeof(double) * (size_t)n * 4);
    double* j2 = j1 + 2 * n;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t p = 0; p < 2 * n; ++p) {
        j1[p] = s->jp[p] + s->jn[p];
        j2[p] = s->jl[p] + s->jz[p];
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < ny; ++i)
        for (int64_t j = 0; j < nx; ++j) {
            const int64_t p = i * nx + j;
            divj[p] = idy * g3y(j2, ny, nx, i, j) + idx * g3x(j2 + n, ny, nx, i, j)
                    + idy * d8y(j1, ny, nx, i, j) + idx * d8x(j1 + n, ny, nx, i, j);
        }
    free(j1);
}

int amo_terms(const double* phi, int64_t ny, int64_t nx, const amo_stencil* st,
              const amo_model* m, const double* xi, double* mup, double* mua, double* jp,
              double* jl, double* jz, double* jn, double* divj)
{
    const int64_t n = ny * nx;
    scratch_t s;
    if (ny <= 0 || nx <= 0 || scratch_alloc(&s, n)) return -1;
    compute_terms(phi, ny, nx, st, m, xi, &s, divj, 1);
    if (mup) memcpy(mup, s.mup, sizeof(double) * n);
    if (mua) memcpy(mua, s.mua, sizeof(double) * n);
    if (jp) memcpy(jp, s.jp, sizeof(double) * 2 * n);
    if (jl) memcpy(jl, s.jl, sizeof(double) * 2 * n);
    if (jz) memcpy(jz, s.jz, sizeof(double) * 2 * n);
    if (jn) memcpy(jn, s.jn, sizeof(double) * 2 * n);
    scratch_free(&s);
    return 0;
}

int amo_run(double* phi, int64_t ny, int64_t nx, const amo_stencil* st, const amo_model* m,
            int64_t nsteps, uint64_t seed, int64_t step0, const double* noise, int nthreads)
{
    const int64_t n = ny * nx;
    scratch_t s;
    if (ny <= 0 || nx <= 0 || nsteps < 0 || nthreads < 1) return -1;
    if (scratch_alloc(&s, n)) return -1;
    double* divj = (double*)malloc(sizeof(double) * (size_t)n);
    double* xi_gen = NULL;
    const int noisy = m->T != 0.0;
    if (noisy && !noise) xi_gen = (double*)malloc(sizeof(double) * (size_t)n * 2);
    for (int64_t t = 0; t < nsteps; ++t) {
        const double* xi = NULL;
        if (noisy) {
            if (noise) {
                xi = noise + (size_t)t * 2 * n;
            } else {
                amo_noise(ny, nx, seed, step0 + t, xi_gen);
                xi = xi_gen;
            }
        }
        compute_terms(phi, ny, nx, st, m, xi, &s, divj, nthreads);
        /* update (test/tjhung_ambplus.h:182-190; src/integrator.cu:261) */
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t p = 0; p < n; ++p) phi[p] -= st->dt * divj[p];
    }
    free(xi_gen);
    free(divj);
    scratch_free(&s);
    return 0;
}

int amo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
