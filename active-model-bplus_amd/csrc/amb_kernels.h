// amb_kernels.h -- host-side declarations of the HIP kernels' launchers (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amb {

// Coefficients of one step with dx, dy, dt and the model folded in (host-computed
// in double, cast to the arithmetic type T).
template <typename T>
struct StepCoef {
    T a, b, c, kappa, lam, zeta;
    T gxc, gxe, gyc, gye;  // G3: centre/edge smoothing weights times 1/dx (1/dy)
    T lcc, lex, ley, lck;  // L3: centre, x-neighbours, y-neighbours, corners
    T p0;                  // D8oD8 centre weight of -(mu_p'' along x and y)
    T px[9], py[9];        // D8oD8 weights (index 1..8) times -1/dx^2, -1/dy^2
    T lxx[3], lyy[3];      // G3xoG3x row weights times -1/dx^2; lyy[0] = -1/dy^2
    T zx, zy;              // 1/dx, 1/dy (divergence of the circulating current)
    T dxw[5], dyw[5];      // D8 weights (index 1..4) times 1/dx, 1/dy (currents through D8)
    T dt;
    T amp;                 // sqrt(2T) sqrt(1/(dt dx dy))
};

template <typename T>
struct StepArgs {
    const T* in;
    T* out;
    int ny, nx;
    int nx4;          // nx % 4 == 0: column groups of 4 never straddle the periodic seam
    int ntx, nty;     // tiles along x and y
    int xcd_chunk;    // > 0: blockIdx remapped so each XCD gets a contiguous run of tiles
    StepCoef<T> k;
    uint32_t key0, key1;
    const long long* step_base;  // device step counter; noise uses *step_base + step_off
    int step_off;
    const double* noise;         // explicit normals [2][ny][nx] of this step (noise mode 2)
    double* part;                // nullable: per-tile (sum, non-finite count) of the new field
};

// kNoiseRuntime: Philox when the coefficient amp is non-zero (batched realisations).
enum NoiseMode { kNoiseNone = 0, kNoisePhilox = 1, kNoiseBuffer = 2, kNoiseRuntime = 3 };

struct TileShape {
    int tw, th, nt;  // tile columns, tile rows, workgroup size
};

// Compiled tile configurations (index = autotuner choice): X(index, TW, TH, NT).  The
// default tile follows an order of preference per kernel (default_tile in amb_capi.hip).
#define AMB_F32_TILES(X) X(0, 32, 48, 512) X(1, 32, 24, 256) X(2, 32, 32, 256) X(3, 64, 32, 512) X(4, 64, 16, 256) X(5, 32, 16, 256) X(6, 16, 16, 128) X(7, 16, 8, 128) X(8, 8, 8, 64) X(9, 32, 32, 512)
#define AMB_F64_TILES(X) X(0, 32, 32, 256) X(1, 64, 16, 256) X(2, 32, 16, 256) X(3, 32, 32, 512) X(4, 16, 16, 128) X(5, 16, 8, 128) X(6, 8, 8, 64)
// FP32 orders of preference, measured on MI355X (the times below predate the step's wave
// priorities, which took 2-4% off every shape).  Noisy: 32x32 tiles of 512 threads are
// fastest at 4096^2 (108.1 us per step; 32x48 111.5 us, 32x16 116.9 us, the others 120-139
// us): their 35 KB of LDS and 61 VGPRs let four workgroups, 8 waves per SIMD, share a CU,
// against three for 32x48, and the extra halo work costs less than the extra waves hide;
// at 1024^2 they tie with 64x32 (11.02 / 11.04 us).  Noiseless, measured with the current
// step (stage B's rows formed just before use, so its 17-tap windows fit 58 VGPRs without
// spilling): 32x32 tiles of 512 threads are fastest at 4096^2 (84.9 us; 32x48 86.3 us, 32x16
// 93.4 us), 2048^2 (25.0 / 25.6 us) and 1024^2 (8.92 us, 64x32 8.94 us, 32x48 9.79 us).
// 16x8 serves the smallest fields (256^2: 5.1 us noiseless, 6.2 us noisy, 8x8 6.0 us).
#define AMB_F32_NOISY_ORDER 9, 0, 3, 2, 1, 4, 5, 6, 7, 8
#define AMB_F32_NOISELESS_ORDER 9, 0, 3, 1, 5, 2, 4, 6, 7, 8
// FP64, noisy and noiseless alike: 32x32 tiles of 512 threads (held to 128 VGPRs, so two
// workgroups, 16 waves, share a CU) are fastest at 4096^2 (291.0 / 178.6 us per step against
// 321.0 / 213.2 us for 256 threads, both before the fitted noise polynomials and the priorities); 16x8 serves the smallest fields (256^2: 5.56 us, 16x16
// 5.77 us, 8x8 5.88 us).
#define AMB_F64_ORDER 3, 0, 2, 1, 4, 5, 6

// Tile shapes compiled per precision (index = autotuner choice).
int tile_count(int precision);
TileShape tile_shape(int precision, int index);

// Launches one fused step.  Returns a hipError_t.
hipError_t launch_step_f32(const StepArgs<float>& a, int tile, int noise, hipStream_t s);
hipError_t launch_step_f64(const StepArgs<double>& a, int tile, int noise, hipStream_t s);
hipError_t prepare_step_kernels();  // raises the dynamic-LDS limit of large tiles
int tile_occupancy(int precision, int tile, int noise);  // resident workgroups per CU (>= 1)

// Adds `n` to the device step counter (a one-lane kernel, graph-capturable).
hipError_t launch_advance(long long* step, long long n, hipStream_t s);

// One observable (AMB_OBS_*) of the current step into `out` (n = ny*nx doubles for the
// chemical potentials, 2n for currents); `tmp` is 2n doubles of workspace.
hipError_t launch_observable_f32(const StepArgs<float>& a, int which, double* tmp, double* out, hipStream_t s);
hipError_t launch_observable_f64(const StepArgs<double>& a, int which, double* tmp, double* out, hipStream_t s);

// Final reduction of the per-tile partials (sum, non-finite count) that the last step of
// a run writes: one block, fixed order.
hipError_t launch_reduce_tiles(const double* part, int64_t ntiles, double* sum, unsigned long long* bad,
                               hipStream_t s);

// FP64 sum and non-finite count of a field: deterministic two-pass reduction (at most
// kReduceBlocks partials; their number depends only on the field's size).
constexpr int kReduceBlocks = 2048;
hipError_t launch_reduce_f32(const float* phi, int64_t n, double* part, unsigned long long* bad_part,
                             double* sum, unsigned long long* bad, hipStream_t s);
hipError_t launch_reduce_f64(const double* phi, int64_t n, double* part, unsigned long long* bad_part,
                             double* sum, unsigned long long* bad, hipStream_t s);

}  // namespace amb
