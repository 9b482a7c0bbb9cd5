// amb_noalloc.hip -- amb_run_noalloc: the explicit-Euler steps of amb_run on caller-owned
// device buffers, for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The step kernels are amb_run's own (same tile choice, same arithmetic, same noise);
// only the handle's scratch memory moves: the step counter that keys the noise, the
// per-tile partials of the fused finite check and the final (sum, count) live in one
// file-scope device array, `scratch`, of kScratchBytes.  Each GPU holds its own copy
// until the process ends; its address on the current device is looked up once per
// device with hipGetSymbolAddress and used for every kernel argument.  Every call on a
// GPU reuses that GPU's array, so calls that overlap on one GPU (two host threads, or
// two streams in flight) race; a call finishes with a stream synchronisation, so calls
// made one after another from one thread do not.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "amb.h"
#include "amb_internal.h"
#include "amb_kernels.h"

namespace amb {
namespace {

// Layout of `scratch`: [0, 8) step counter; [16, 24) sum; [24, 32) non-finite count;
// [256, 256 + 16 * kNoAllocMaxTiles) the per-tile (sum, count) partials.
constexpr int64_t kNoAllocMaxTiles = 65536;
constexpr size_t kPartOffset = 256;
constexpr size_t kScratchBytes = kPartOffset + 16 * (size_t)kNoAllocMaxTiles;  // 1 MiB + 256 B

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_scratch[kMaxDevices] = {};  // address of `scratch` on each device

// Address of `scratch` on the current device (looked up once per device).
int scratch_on_device(unsigned char** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(AMB_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail_hip(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress");
        e = prepare_step_kernels();  // the dynamic-LDS limit of the large tiles (no allocation)
        if (e != hipSuccess) return fail_hip(e, "hipFuncSetAttribute");
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return AMB_OK;
}

static_assert(16 * (size_t)kReduceBlocks <= 16 * (size_t)kNoAllocMaxTiles, "full-read partials fit the scratch");

hipError_t launch_step_t(const StepArgs<float>& a, int tile, int noise, hipStream_t s)
{
    return launch_step_f32(a, tile, noise, s);
}
hipError_t launch_step_t(const StepArgs<double>& a, int tile, int noise, hipStream_t s)
{
    return launch_step_f64(a, tile, noise, s);
}
hipError_t launch_reduce_t(const float* f, int64_t n, double* part, unsigned long long* bad_part, double* sum,
                           unsigned long long* bad, hipStream_t s)
{
    return launch_reduce_f32(f, n, part, bad_part, sum, bad, s);
}
hipError_t launch_reduce_t(const double* f, int64_t n, double* part, unsigned long long* bad_part, double* sum,
                           unsigned long long* bad, hipStream_t s)
{
    return launch_reduce_f64(f, n, part, bad_part, sum, bad, s);
}

__global__ void set_counter_kernel(long long* step, long long value)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) step[0] = value;
}

template <typename T>
int run_noalloc(T* field, T* work, int64_t ny, int64_t nx, const amb_stencil& st, const amb_model& md, uint64_t seed,
                int64_t timestep, int64_t nsteps, int tile, int noise, unsigned char* base, hipStream_t s,
                double* mass, int64_t* nonfinite)
{
    long long* step = reinterpret_cast<long long*>(base);
    double* sum = reinterpret_cast<double*>(base + 16);
    unsigned long long* bad = reinterpret_cast<unsigned long long*>(base + 24);
    double* part = reinterpret_cast<double*>(base + kPartOffset);
    const TileShape ts = tile_shape(sizeof(T) == 4 ? AMB_F32 : AMB_F64, tile);
    StepArgs<T> a{};
    a.ny = (int)ny;
    a.nx = (int)nx;
    a.nx4 = (nx % 4) == 0;
    a.ntx = grid_tiles(nx, ts.tw);
    a.nty = grid_tiles(ny, ts.th);
    a.xcd_chunk = xcd_chunk(a.ntx * a.nty);
    a.k = make_coef<T>(st, md);
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.step_base = step;
    a.noise = nullptr;
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(64), 0, s, step, (long long)timestep);
    hipError_t e = hipGetLastError();
    T* buf[2] = {field, work};
    for (int64_t k = 0; k < nsteps && e == hipSuccess; ++k) {
        a.in = buf[k & 1];
        a.out = buf[(k + 1) & 1];
        a.step_off = (int)k;
        a.part = k == nsteps - 1 ? part : nullptr;
        e = launch_step_t(a, tile, noise, s);
    }
    if (e != hipSuccess) return fail_hip(e, "step launch");
    if (nsteps > 0) {
        if ((e = launch_reduce_tiles(part, (int64_t)a.ntx * a.nty, sum, bad, s)) != hipSuccess)
            return fail_hip(e, "reduction launch");
        if (nsteps & 1) {  // the last step wrote `work`
            e = hipMemcpyAsync(field, work, (size_t)(ny * nx) * sizeof(T), hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return fail_hip(e, "hipMemcpyAsync (field)");
        }
    } else {
        // no step: the mass and count of the field as it is, through the same partials
        // reduction the handle's amb_mass uses (with the partials area as its workspace)
        e = launch_reduce_t(field, ny * nx, part, reinterpret_cast<unsigned long long*>(part + kReduceBlocks), sum,
                            bad, s);
        if (e != hipSuccess) return fail_hip(e, "reduction launch");
    }
    double h_sum = 0.0;
    unsigned long long h_bad = 0;
    if ((e = hipMemcpyAsync(&h_sum, sum, sizeof h_sum, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(&h_bad, bad, sizeof h_bad, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return fail_hip(e, "hipMemcpyAsync (mass)");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    if (mass) *mass = h_sum;
    if (nonfinite) *nonfinite = (int64_t)h_bad;
    if (h_bad != 0) return fail(AMB_E_NONFINITE, "an unknown numerical error occurred during simulation");
    return AMB_OK;
}

}  // namespace
}  // namespace amb

using namespace amb;

extern "C" int amb_run_noalloc(void* field, void* work, int64_t ny, int64_t nx, int precision,
                               const amb_stencil* stencil, const amb_model* model, uint64_t seed, int64_t timestep,
                               int64_t nsteps, void* stream, double* mass, int64_t* nonfinite)
{
    clear_error();
    if (!field || !work || field == work) return fail(AMB_E_ARG, "field and work must be two distinct device buffers");
    if (precision != AMB_F32 && precision != AMB_F64) return fail(AMB_E_ARG, "precision must be AMB_F32 or AMB_F64");
    if (nsteps < 0) return fail(AMB_E_ARG, "nsteps must be non-negative");
    int rc = validate_params(ny, nx, stencil, model);
    if (rc != AMB_OK) return rc;
    if ((rc = check_device()) != AMB_OK) return rc;
    unsigned char* base = nullptr;
    if ((rc = scratch_on_device(&base)) != AMB_OK) return rc;
    const int noise = model->T > 0.0 ? kNoisePhilox : kNoiseNone;
    const int tile = default_tile(precision, ny, nx, noise);
    const TileShape ts = tile_shape(precision, tile);
    const int64_t tiles = (int64_t)grid_tiles(nx, ts.tw) * grid_tiles(ny, ts.th);
    if (tiles > kNoAllocMaxTiles)
        return fail(AMB_E_ARG, "field too large for amb_run_noalloc: its " + std::to_string(tiles) +
                                   " tiles exceed the scratch area's 65536 (use amb_create / amb_run)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return precision == AMB_F32
               ? run_noalloc<float>(static_cast<float*>(field), static_cast<float*>(work), ny, nx, *stencil, *model,
                                    seed, timestep, nsteps, tile, noise, base, s, mass, nonfinite)
               : run_noalloc<double>(static_cast<double*>(field), static_cast<double*>(work), ny, nx, *stencil,
                                     *model, seed, timestep, nsteps, tile, noise, base, s, mass, nonfinite);
}
