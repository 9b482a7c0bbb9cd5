// amb_step.hip -- fused explicit-Euler step of Active Model B+ for gfx950 (CDNA4).
//
// One launch per step reads phi^n and writes phi^{n+1}; nothing else touches HBM
// (8 B per site in FP32, 16 B in FP64).  The reference needs three launches per
// step and ~280 B per site (src/integrator.cu:46-262, host loop :507-518).
//
// Per tile of TW x TH outputs (one workgroup of NT threads; 32x32 tiles of 512 threads,
// eight wave64s, on large fields):
//   stage 0  phi on rows [-9, TH+9) x cols [-9, TW+9) -> LDS (periodic wrap)
//   stage A  on the cross-shaped halo region: G3x phi, G3y phi, L3 phi, then
//            mu_p (rows/cols +-8), mu_a = lambda|G3 phi|^2 and the circulating current
//            q = zeta L3phi G3phi (box +-2) -> LDS
//   stage N  (noisy steps) conserved-noise normals xi_x (cols +-4), xi_y (rows +-4) from
//            counter-based Philox4x32-10, each turned at once into the current
//            J = amp xi - D8 mu_p (aliases the phi buffer); on FP32 fields of 2048^2
//            sites or more also the active current C = q - G3 mu_a (box +-1, over q)
//   stage B  div J per output from LDS: the D8 divergence of J (noisy) or D8oD8 on mu_p
//            (17 taps per axis, noiseless), G3oG3 on mu_a (5x5 separable) plus G3 on q,
//            or G3 on C alone; phi^{n+1} = phi^n - dt div J.
// A site's noise depends only on (seed, step, row, column, component), so the
// halo copies that neighbouring tiles generate are bit-identical to the owner's.
#include <hip/hip_runtime.h>

#include "amb_kernels.h"
#include "amb_math.h"
#include "amb_tile.h"

namespace amb {

// ---- single-field kernel ----------------------------------------------------------

template <typename T, int TW, int TH, int NT, int NOISE>
__global__ __launch_bounds__(NT, (min_waves_per_eu<T, TW, TH, NT>())) void step_kernel(const StepArgs<T> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // Tile index; with xcd_chunk the round-robin XCD dispatch gets contiguous tiles.
    int t = blockIdx.x;
    if (p.xcd_chunk > 0) t = (t & 7) * p.xcd_chunk + (t >> 3);
    const int ty = t / p.ntx, tx = t - ty * p.ntx;
    const uint64_t step = NOISE == kNoisePhilox ? (uint64_t)(*p.step_base) + (uint64_t)p.step_off : 0;
    step_tile<T, TW, TH, NT, NOISE>(reinterpret_cast<T*>(smem_raw), p.in, p.out, p.ny, p.nx, p.nx4, tx * TW, ty * TH,
                                    p.k, p.key0, p.key1, step, p.noise, p.part ? p.part + 2 * (size_t)t : nullptr);
}

// ---- tile table ---------------------------------------------------------------

#define AMB_SHAPE(i, tw, th, nt) {tw, th, nt},
static const TileShape kTilesF32[] = {AMB_F32_TILES(AMB_SHAPE)};
static const TileShape kTilesF64[] = {AMB_F64_TILES(AMB_SHAPE)};
#undef AMB_SHAPE

int tile_count(int precision)
{
    return precision == 32 ? (int)(sizeof(kTilesF32) / sizeof(kTilesF32[0]))
                           : (int)(sizeof(kTilesF64) / sizeof(kTilesF64[0]));
}

TileShape tile_shape(int precision, int index)
{
    return precision == 32 ? kTilesF32[index] : kTilesF64[index];
}

template <typename T, int TW, int TH, int NT, int NOISE>
static hipError_t launch_one(const StepArgs<T>& a, hipStream_t s)
{
    using L = Lds<T, TW, TH, NT>;
    const int nblocks = a.ntx * a.nty;
    hipLaunchKernelGGL((step_kernel<T, TW, TH, NT, NOISE>), dim3(nblocks), dim3(NT), L::bytes, s, a);
    return hipGetLastError();
}

template <typename T, int TW, int TH, int NT>
static hipError_t launch_noise(const StepArgs<T>& a, int noise, hipStream_t s)
{
    switch (noise) {
    case kNoiseNone: return launch_one<T, TW, TH, NT, kNoiseNone>(a, s);
    case kNoisePhilox: return launch_one<T, TW, TH, NT, kNoisePhilox>(a, s);
    default: return launch_one<T, TW, TH, NT, kNoiseBuffer>(a, s);
    }
}

#define AMB_CASE_F32(i, tw, th, nt) case i: return launch_noise<float, tw, th, nt>(a, noise, s);
#define AMB_CASE_F64(i, tw, th, nt) case i: return launch_noise<double, tw, th, nt>(a, noise, s);

hipError_t launch_step_f32(const StepArgs<float>& a, int tile, int noise, hipStream_t s)
{
    switch (tile) {
        AMB_F32_TILES(AMB_CASE_F32)
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step_f64(const StepArgs<double>& a, int tile, int noise, hipStream_t s)
{
    switch (tile) {
        AMB_F64_TILES(AMB_CASE_F64)
    default: return hipErrorInvalidValue;
    }
}

// Resident workgroups per CU of one step kernel (registers and LDS), at least 1.
template <typename T, int TW, int TH, int NT>
static int occupancy_of(int noise)
{
    using L = Lds<T, TW, TH, NT>;
    const void* f = noise == kNoiseNone     ? (const void*)step_kernel<T, TW, TH, NT, kNoiseNone>
                    : noise == kNoisePhilox ? (const void*)step_kernel<T, TW, TH, NT, kNoisePhilox>
                                            : (const void*)step_kernel<T, TW, TH, NT, kNoiseBuffer>;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, NT, L::bytes) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

#define AMB_OCC_F32(i, tw, th, nt) case i: return occupancy_of<float, tw, th, nt>(noise);
#define AMB_OCC_F64(i, tw, th, nt) case i: return occupancy_of<double, tw, th, nt>(noise);

int tile_occupancy(int precision, int tile, int noise)
{
    if (precision == 32) {
        switch (tile) {
            AMB_F32_TILES(AMB_OCC_F32)
        default: return 1;
        }
    }
    switch (tile) {
        AMB_F64_TILES(AMB_OCC_F64)
    default: return 1;
    }
}

template <typename T, int TW, int TH, int NT>
static hipError_t raise_lds_limit()
{
    using L = Lds<T, TW, TH, NT>;
    hipError_t e = hipSuccess;
    if (L::bytes > 65536) {
        e = hipFuncSetAttribute((const void*)step_kernel<T, TW, TH, NT, kNoiseNone>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)L::bytes);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)step_kernel<T, TW, TH, NT, kNoisePhilox>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)L::bytes);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)step_kernel<T, TW, TH, NT, kNoiseBuffer>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)L::bytes);
    }
    return e;
}

#define AMB_RAISE_F32(i, tw, th, nt) if ((e = raise_lds_limit<float, tw, th, nt>()) != hipSuccess) return e;
#define AMB_RAISE_F64(i, tw, th, nt) if ((e = raise_lds_limit<double, tw, th, nt>()) != hipSuccess) return e;

hipError_t prepare_step_kernels()
{
    hipError_t e = hipSuccess;
    AMB_F32_TILES(AMB_RAISE_F32)
    AMB_F64_TILES(AMB_RAISE_F64)
    return e;
}

// ---- step counter -------------------------------------------------------------

__global__ void advance_kernel(long long* step, long long n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) step[0] += n;
}

hipError_t launch_advance(long long* step, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, s, step, n);
    return hipGetLastError();
}

// ---- observables (getters) ------------------------------------------------------
// A plain per-site formulation with wrapped global loads: a second, independent
// device implementation of the same terms (two passes, like the reference's
// calculate_forces / calculate_local_currents, src/integrator.cu:46-176).

template <typename T>
__device__ __forceinline__ T ld(const T* f, int ny, int nx, int r, int c)
{
    return f[(size_t)wrap_index(r, ny) * nx + wrap_index(c, nx)];
}

// Local terms at one site from its 3x3 neighbourhood: mu_p, mu_a and the circulating
// current (y, x components), in the arithmetic type of the field.
template <typename T>
__device__ __forceinline__ void local_terms(const StepArgs<T>& p, int r, int c, T& mup, T& mua, T& jzy, T& jzx)
{
    const int ny = p.ny, nx = p.nx;
    const StepCoef<T>& K = p.k;
    const T* f = p.in;
    const T pmm = ld(f, ny, nx, r - 1, c - 1), pm0 = ld(f, ny, nx, r - 1, c), pmp = ld(f, ny, nx, r - 1, c + 1);
    const T p0m = ld(f, ny, nx, r, c - 1), p00 = ld(f, ny, nx, r, c), p0p = ld(f, ny, nx, r, c + 1);
    const T ppm = ld(f, ny, nx, r + 1, c - 1), pp0 = ld(f, ny, nx, r + 1, c), ppp = ld(f, ny, nx, r + 1, c + 1);
    const T gx = K.gxe * ((pmp - pmm) + (ppp - ppm)) + K.gxc * (p0p - p0m);
    const T gy = K.gye * ((ppm - pmm) + (ppp - pmp)) + K.gyc * (pp0 - pm0);
    const T lap = K.lcc * p00 + K.lex * (p0m + p0p) + K.ley * (pm0 + pp0) + K.lck * ((pmm + pmp) + (ppm + ppp));
    mup = p00 * (K.a + p00 * (K.b + p00 * K.c)) - K.kappa * lap;
    mua = K.lam * (gy * gy + gx * gx);
    const T zl = K.zeta * lap;
    jzy = zl * gy;
    jzx = zl * gx;
}

// Observables of one kind (AMB_OBS_*), computed only as far as that kind needs:
// obs_local_kernel writes mu_p / mu_a into `tmp` ([mu_p | mu_a], n doubles each) when a
// current needs their derivatives, and the scalar kinds straight into `out`;
// obs_current_kernel forms the requested current into `out` ([y | x], 2n doubles) as
// ((J_zeta + J_lambda) + J_p) + J_noise over the terms the kind includes.
enum ObsTerm { kTermP = 1, kTermL = 2, kTermZ = 4, kTermN = 8 };

__host__ __device__ inline int obs_terms(int which)
{
    switch (which) {
    case 3: return kTermP | kTermL | kTermZ | kTermN;  // AMB_OBS_J
    case 4: return kTermP | kTermL | kTermZ;           // AMB_OBS_J_DETERMINISTIC
    case 5: return kTermP;                             // AMB_OBS_J_PASSIVE
    case 6: return kTermL | kTermZ;                    // AMB_OBS_J_ACTIVE
    case 7: return kTermL;                             // AMB_OBS_J_LAMBDA
    case 8: return kTermZ;                             // AMB_OBS_J_CIRCULATING
    case 9: return kTermN;                             // AMB_OBS_J_RANDOM
    default: return 0;                                 // scalars
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void obs_local_kernel(const StepArgs<T> p, int which, double* tmp,
                                                             double* out)
{
    const int ny = p.ny, nx = p.nx;
    const size_t n = (size_t)ny * nx;
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int r = (int)(i / nx), c = (int)(i - (size_t)r * nx);
    T mup, mua, jzy, jzx;
    local_terms(p, r, c, mup, mua, jzy, jzx);
    const int terms = obs_terms(which);
    if (terms & kTermP) tmp[i] = (double)mup;
    if (terms & kTermL) tmp[n + i] = (double)mua;
    if (which == 0) out[i] = (double)mup + (double)mua;  // AMB_OBS_MU
    else if (which == 1) out[i] = (double)mup;          // AMB_OBS_MU_PASSIVE
    else if (which == 2) out[i] = (double)mua;          // AMB_OBS_MU_ACTIVE
}

template <typename T>
__global__ __launch_bounds__(kThreads) void obs_current_kernel(const StepArgs<T> p, int which, const double* tmp,
                                                               double* out)
{
    const int ny = p.ny, nx = p.nx;
    const size_t n = (size_t)ny * nx;
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int r = (int)(i / nx), c = (int)(i - (size_t)r * nx);
    const StepCoef<T>& K = p.k;
    const int terms = obs_terms(which);
    double zy = 0.0, zx = 0.0, ly = 0.0, lx = 0.0, py = 0.0, px = 0.0, ny_ = 0.0, nx_ = 0.0;
    if (terms & kTermZ) {
        T mup, mua, jzy, jzx;
        local_terms(p, r, c, mup, mua, jzy, jzx);
        zy = (double)jzy;
        zx = (double)jzx;
    }
    if (terms & kTermP) {
        // J_p = -D8 mu_p (src/integrator.cu:171-172); K.zx = 1/dx, K.zy = 1/dy.
        const double* mp = tmp;
        T dy8 = 0, dx8 = 0;
#pragma unroll
        for (int k = 1; k <= 4; ++k) {
            dy8 += (T)kD8[k] * ((T)ld(mp, ny, nx, r + k, c) - (T)ld(mp, ny, nx, r - k, c));
            dx8 += (T)kD8[k] * ((T)ld(mp, ny, nx, r, c + k) - (T)ld(mp, ny, nx, r, c - k));
        }
        py = (double)(-K.zy * dy8);
        px = (double)(-K.zx * dx8);
    }
    if (terms & kTermL) {
        // J_lambda = -G3 mu_a (src/integrator.cu:174-175)
        const double* ma = tmp + n;
        const T amm = (T)ld(ma, ny, nx, r - 1, c - 1), am0 = (T)ld(ma, ny, nx, r - 1, c),
                amp_ = (T)ld(ma, ny, nx, r - 1, c + 1);
        const T a0m = (T)ld(ma, ny, nx, r, c - 1), a0p = (T)ld(ma, ny, nx, r, c + 1);
        const T apm = (T)ld(ma, ny, nx, r + 1, c - 1), ap0 = (T)ld(ma, ny, nx, r + 1, c),
                app = (T)ld(ma, ny, nx, r + 1, c + 1);
        const T gx = K.gxe * ((amp_ - amm) + (app - apm)) + K.gxc * (a0p - a0m);
        const T gy = K.gye * ((apm - amm) + (app - amp_)) + K.gyc * (ap0 - am0);
        ly = (double)(-gy);
        lx = (double)(-gx);
    }
    if (terms & kTermN) {
        T xy = 0, xx = 0;
        if (p.noise) {
            xy = (T)p.noise[i];
            xx = (T)p.noise[n + i];
        } else if (K.amp != T(0)) {
            const uint64_t step = (uint64_t)(*p.step_base);
            xy = normal1<T>(r, c, step, 0, p.key0, p.key1);
            xx = normal1<T>(r, c, step, 1, p.key0, p.key1);
        }
        ny_ = (double)(K.amp * xy);
        nx_ = (double)(K.amp * xx);
    }
    out[i] = ((zy + ly) + py) + ny_;
    out[n + i] = ((zx + lx) + px) + nx_;
}

template <typename T>
static hipError_t launch_obs(const StepArgs<T>& a, int which, double* tmp, double* out, hipStream_t s)
{
    const size_t n = (size_t)a.ny * a.nx;
    const unsigned blocks = (unsigned)((n + kThreads - 1) / kThreads);
    const int terms = obs_terms(which);
    if (which <= 2 || (terms & (kTermP | kTermL))) {
        hipLaunchKernelGGL(obs_local_kernel<T>, dim3(blocks), dim3(kThreads), 0, s, a, which, tmp, out);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess || which <= 2) return e;
    }
    hipLaunchKernelGGL(obs_current_kernel<T>, dim3(blocks), dim3(kThreads), 0, s, a, which,
                       static_cast<const double*>(tmp), out);
    return hipGetLastError();
}

hipError_t launch_observable_f32(const StepArgs<float>& a, int which, double* tmp, double* out, hipStream_t s)
{
    return launch_obs(a, which, tmp, out, s);
}
hipError_t launch_observable_f64(const StepArgs<double>& a, int which, double* tmp, double* out, hipStream_t s)
{
    return launch_obs(a, which, tmp, out, s);
}

// ---- FP64 mass and finite check ----------------------------------------------------

template <typename T>
__global__ __launch_bounds__(kThreads) void reduce_partial_kernel(const T* phi, int64_t n, double* part,
                                                                  unsigned long long* bad_part)
{
    __shared__ double ssum[kThreads];
    __shared__ unsigned long long sbad[kThreads];
    double s = 0.0;
    unsigned long long b = 0;
    // Each thread visits i0, i0 + G, i0 + 2G, ... (G = all threads of the grid) and adds in
    // that order; the values are fetched eight at a time first, so the loads overlap.
    constexpr int kBatch = 8;
    const int64_t G = (int64_t)gridDim.x * kThreads;
    for (int64_t i0 = (int64_t)blockIdx.x * kThreads + threadIdx.x; i0 < n; i0 += kBatch * G) {
        T e[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int64_t i = i0 + u * G;
            e[u] = i < n ? phi[i] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (i0 + u * G < n) {
                const double v = (double)e[u];
                if (isfinite(v)) s += v; else ++b;
            }
        }
    }
    ssum[threadIdx.x] = s;
    sbad[threadIdx.x] = b;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            sbad[threadIdx.x] += sbad[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ssum[0];
        bad_part[blockIdx.x] = sbad[0];
    }
}

__global__ __launch_bounds__(kThreads) void reduce_final_kernel(const double* part, const unsigned long long* bad_part,
                                                                int nparts, double* sum, unsigned long long* bad)
{
    __shared__ double ssum[kThreads];
    __shared__ unsigned long long sbad[kThreads];
    double s = 0.0;
    unsigned long long b = 0;
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
        s += part[i];
        b += bad_part[i];
    }
    ssum[threadIdx.x] = s;
    sbad[threadIdx.x] = b;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            sbad[threadIdx.x] += sbad[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sum[0] = ssum[0];
        bad[0] = sbad[0];
    }
}

// Partials of the fused finite check: entry i = (sum, count) of tile i.  Each thread
// sums a fixed stride of entries in index order, then a fixed tree.  The entries are
// fetched sixteen at a time before they are added, so the loads overlap while the
// additions keep their order (65536 partials of an 8192^2 field: four batches per thread).
constexpr int kTileReduceThreads = 1024;

__global__ __launch_bounds__(kTileReduceThreads) void reduce_tiles_kernel(const double* part, int64_t n, double* sum,
                                                                          unsigned long long* bad)
{
    __shared__ double ssum[kTileReduceThreads];
    __shared__ double sbad[kTileReduceThreads];
    constexpr int kBatch = 16;
    // entry i as one 16-byte load (every partials array starts 16-byte aligned: hipMalloc,
    // or 256 bytes into the no-allocation entry point's 256-byte aligned scratch array)
    const double2* p2 = reinterpret_cast<const double2*>(part);
    double s = 0.0, b = 0.0;
    for (int64_t i0 = threadIdx.x; i0 < n; i0 += (int64_t)kBatch * kTileReduceThreads) {
        double2 e[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int64_t i = i0 + (int64_t)u * kTileReduceThreads;
            if (i < n) e[u] = p2[i];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (i0 + (int64_t)u * kTileReduceThreads < n) {
                s += e[u].x;
                b += e[u].y;
            }
        }
    }
    ssum[threadIdx.x] = s;
    sbad[threadIdx.x] = b;
    __syncthreads();
    for (int w = kTileReduceThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            sbad[threadIdx.x] += sbad[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sum[0] = ssum[0];
        bad[0] = (unsigned long long)sbad[0];
    }
}

hipError_t launch_reduce_tiles(const double* part, int64_t ntiles, double* sum, unsigned long long* bad,
                               hipStream_t s)
{
    hipLaunchKernelGGL(reduce_tiles_kernel, dim3(1), dim3(kTileReduceThreads), 0, s, part, ntiles, sum, bad);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_reduce(const T* phi, int64_t n, double* part, unsigned long long* bad_part, double* sum,
                                unsigned long long* bad, hipStream_t s)
{
    // enough workgroups that every thread has at least 32 values (four batches) to add, up to
    // kReduceBlocks (a 4096^2 field uses all of them: 32 waves per CU keep the loads in flight)
    const int64_t want = (n + (int64_t)kThreads * 32 - 1) / ((int64_t)kThreads * 32);
    const int blocks = (int)(want < 1 ? 1 : want > kReduceBlocks ? kReduceBlocks : want);
    hipLaunchKernelGGL(reduce_partial_kernel<T>, dim3(blocks), dim3(kThreads), 0, s, phi, n, part, bad_part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kThreads), 0, s, part, bad_part, blocks, sum, bad);
    return hipGetLastError();
}

hipError_t launch_reduce_f32(const float* phi, int64_t n, double* part, unsigned long long* bad_part, double* sum,
                             unsigned long long* bad, hipStream_t s)
{
    return launch_reduce(phi, n, part, bad_part, sum, bad, s);
}

hipError_t launch_reduce_f64(const double* phi, int64_t n, double* part, unsigned long long* bad_part, double* sum,
                             unsigned long long* bad, hipStream_t s)
{
    return launch_reduce(phi, n, part, bad_part, sum, bad, s);
}

}  // namespace amb
