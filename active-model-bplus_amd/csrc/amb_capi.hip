// amb_capi.hip -- host runtime behind the C ABI declared in include/amb.h.
//
// One handle owns: two device field buffers (ping-pong; the fused step cannot
// update in place because each tile reads its neighbours' halo), a device step
// counter that keys the Philox noise, a private stream, cached hipGraph
// executables of kGraphSteps steps, and small reduction buffers.  The reference's
// Integrator (src/integrator.cu:286-538) instead owns 11 FP64 arrays plus cuRAND
// state and launches three kernels per step from a host loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "amb.h"
#include "amb_internal.h"
#include "amb_kernels.h"
#include "amb_math.h"

namespace {

thread_local std::string g_error;

}  // namespace

namespace amb {

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

int fail_hip(hipError_t e, const char* what)
{
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return AMB_E_HIP;
}

void clear_error() { g_error.clear(); }

int check_device()
{
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(AMB_E_NODEVICE, "no HIP device is visible (the integrator runs on an AMD GPU only)");
    }
    return AMB_OK;
}

template <typename T>
StepCoef<T> make_coef(const amb_stencil& st, const amb_model& m)
{
    StepCoef<T> k{};
    const double idx = 1.0 / st.dx, idy = 1.0 / st.dy;
    const double idx2 = idx * idx, idy2 = idy * idy;
    k.a = (T)m.a;
    k.b = (T)m.b;
    k.c = (T)m.c;
    k.kappa = (T)m.kappa;
    k.lam = (T)m.lamb;
    k.zeta = (T)m.zeta;
    k.gxc = (T)(kG3S[0] * idx);
    k.gxe = (T)(kG3S[1] * idx);
    k.gyc = (T)(kG3S[0] * idy);
    k.gye = (T)(kG3S[1] * idy);
    // L3 = S2x/dx^2 + S2y/dy^2 written per neighbour class.
    k.lcc = (T)(-3.0 * (idx2 + idy2));
    k.lex = (T)(1.5 * idx2 + 0.5 * idy2);
    k.ley = (T)(0.5 * idx2 + 1.5 * idy2);
    k.lck = (T)(-0.25 * (idx2 + idy2));
    k.p0 = (T)(-(idx2 + idy2) * kD8D8[0]);
    for (int j = 0; j < 9; ++j) {
        k.px[j] = (T)(-idx2 * kD8D8[j]);
        k.py[j] = (T)(-idy2 * kD8D8[j]);
    }
    // G3xoG3x: the smoothing weights across rows are folded in (lxx[|a|]); for
    // G3yoG3y the smoothing is applied to each row first, so lyy[0] = -1/dy^2.
    for (int a = 0; a < 3; ++a) {
        k.lxx[a] = (T)(-idx2 * kG3G3S[a]);
        k.lyy[a] = (T)(a == 0 ? -idy2 : 0.0);
    }
    k.zx = (T)idx;
    k.zy = (T)idy;
    // Noise amplitude sqrt(2T) * sqrt(1/(dt dx dy)) (src/parameters.h:50,86).
    const double amp = std::sqrt(2.0 * m.T) * std::sqrt(1.0 / (st.dt * st.dx * st.dy));
    for (int j = 0; j < 5; ++j) {
        k.dxw[j] = (T)(idx * kD8[j]);
        k.dyw[j] = (T)(idy * kD8[j]);
    }
    k.dt = (T)st.dt;
    k.amp = (T)amp;
    return k;
}

template StepCoef<float> make_coef<float>(const amb_stencil&, const amb_model&);
template StepCoef<double> make_coef<double>(const amb_stencil&, const amb_model&);

int validate_params(int64_t ny, int64_t nx, const amb_stencil* st, const amb_model* m)
{
    if (!st || !m) return fail(AMB_E_ARG, "stencil and model must not be null");
    if (ny <= 0 || nx <= 0) return fail(AMB_E_ARG, "field must have at least one row and one column");
    if (ny > (int64_t)0x7fffffff || nx > (int64_t)0x7fffffff || ny * nx > ((int64_t)1 << 40))
        return fail(AMB_E_ARG, "field is too large");
    const double v[] = {st->dt, st->dx, st->dy, m->a, m->b, m->c, m->kappa, m->lamb, m->zeta, m->T};
    for (double x : v)
        if (!std::isfinite(x)) return fail(AMB_E_ARG, "stencil and model parameters must be finite");
    if (!(st->dx > 0.0) || !(st->dy > 0.0)) return fail(AMB_E_ARG, "dx and dy must be positive");
    if (m->T < 0.0) return fail(AMB_E_ARG, "temperature must be non-negative");
    if (m->T > 0.0 && !(st->dt > 0.0)) return fail(AMB_E_ARG, "noise needs a positive dt");
    return AMB_OK;
}

int grid_tiles(int64_t n, int t) { return (int)((n + t - 1) / t); }

// Default tile: in order of preference (AMB_F32_NOISY_ORDER, AMB_F32_NOISELESS_ORDER or
// AMB_F64_ORDER), the first shape of 256 or more threads whose tiles fill the device's
// resident workgroups either for at least 8 rounds or in whole-or-mostly-full rounds (a
// last round under half full leaves most CUs idle for a round).  Fields too small for one
// round of such a shape take the shape that puts nearest 8 waves on each CU (in ratio; the
// earlier in the order on a tie): fewer waves leave each CU a single workgroup's latency to
// hide, while many small tiles pay their halo several times over.  The tiles of 64 and 128
// threads only serve that way: their boxes hold 4.5-10.6x their area of phi (32x32 tiles
// 2.4x), so where they fill a round (FP64 16x8 at 512^2) a half round of 32x32 tiles is
// still faster (9.0 against 10.2 us).  Measured from 128^2 to 4096^2 in both precisions, with and without noise, this
// picks the fastest shape or one within 2% of it.
int default_tile(int precision, int64_t ny, int64_t nx, int noise)
{
    int device = 0, cus = 256;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        cus = 256;
    }
    static constexpr int kNoisyF32[] = {AMB_F32_NOISY_ORDER};
    static constexpr int kNoiselessF32[] = {AMB_F32_NOISELESS_ORDER};
    static constexpr int kF64[] = {AMB_F64_ORDER};
#define AMB_COUNT_ONE(i, tw, th, nt) +1
    static_assert(sizeof(kNoisyF32) / sizeof(kNoisyF32[0]) == 0 AMB_F32_TILES(AMB_COUNT_ONE) &&
                      sizeof(kNoiselessF32) / sizeof(kNoiselessF32[0]) == 0 AMB_F32_TILES(AMB_COUNT_ONE),
                  "the FP32 orders must list every FP32 tile shape");
    static_assert(sizeof(kF64) / sizeof(kF64[0]) == 0 AMB_F64_TILES(AMB_COUNT_ONE),
                  "the FP64 order must list every FP64 tile shape");
#undef AMB_COUNT_ONE
    const int n = tile_count(precision);
    const int* order = precision != AMB_F32 ? kF64 : noise == kNoiseNone ? kNoiselessF32 : kNoisyF32;
    int best = 0;
    double best_gap = 0.0;
    for (int k = 0; k < n; ++k) {
        const int i = order ? order[k] : k;
        const TileShape t = tile_shape(precision, i);
        const int64_t tiles = (int64_t)grid_tiles(nx, t.tw) * grid_tiles(ny, t.th);
        const int64_t slots = (int64_t)cus * tile_occupancy(precision, i, noise);
        const int64_t full = tiles / slots, rest = tiles % slots;
        if (t.nt >= 256 && (full >= 8 || (full >= 1 && (rest == 0 || 2 * rest >= slots)))) return i;
        const double waves = (double)tiles * (t.nt / 64);
        const double gap = std::fabs(std::log(waves / (8.0 * cus)));
        if (k == 0 || gap < best_gap - 1e-9) {
            best = i;
            best_gap = gap;
        }
    }
    return best;
}

// XCD-contiguous tile order: gfx950 dispatches workgroup i to XCD i % 8.
int xcd_chunk(int nblocks) { return (nblocks % 8 == 0 && nblocks >= 64) ? nblocks / 8 : 0; }

// ---- tuning table: the autotuners' choices, per (precision, grid, noise) ------------
namespace {
std::mutex g_tune_mutex;
std::map<std::tuple<int, int64_t, int64_t, int>, int> g_tune;
}  // namespace

bool tuned_tile(int precision, int64_t ny, int64_t nx, int noise, int* tile)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    const auto it = g_tune.find({precision, ny, nx, noise != kNoiseNone});
    if (it == g_tune.end()) return false;
    *tile = it->second;
    return true;
}

void record_tile(int precision, int64_t ny, int64_t nx, int noise, int tile)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune[{precision, ny, nx, noise != kNoiseNone}] = tile;
}

// The tile a new integrator starts with: the tuning table's entry when there is one,
// otherwise the occupancy rule.
int initial_tile(int precision, int64_t ny, int64_t nx, int noise)
{
    int t = 0;
    if (tuned_tile(precision, ny, nx, noise, &t)) return t;
    return default_tile(precision, ny, nx, noise);
}

}  // namespace amb

using namespace amb;

struct amb_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t ny = 0, nx = 0;
    amb_stencil stencil{};
    amb_model model{};
    amb_options options{};
    int precision = AMB_F64;
    size_t esize = 8;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    long long* d_step = nullptr;
    int64_t timestep = 0;
    double* d_part = nullptr;
    unsigned long long* d_bad_part = nullptr;
    double* d_tile_part = nullptr;         // (sum, non-finite count) per tile, written by a run's last step
    double* d_sum = nullptr;
    unsigned long long* d_bad = nullptr;
    double last_sum = 0.0;                 // mass and non-finite count found by the last amb_run
    int64_t last_bad = 0;
    double* h_sum = nullptr;               // pinned
    unsigned long long* h_bad = nullptr;   // pinned
    double* d_obs = nullptr;               // 10 n doubles, allocated by the first getter
    double* d_noise = nullptr;             // explicit normals (test mode)
    int64_t noise_steps = 0, noise_used = 0;
    int tile = 0;                          // index into the precision's tile table
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_tile[2] = {-1, -1};
    int graph_noise[2] = {-1, -1};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    SkCache* sk = nullptr;                 // structure-factor plan and buffers (lazy)
};

namespace {

constexpr int kGraphSteps = 32;  // even: a graph returns the field to the buffer it started in

template <typename T>
StepArgs<T> step_args(const amb_handle* h, int parity_in, int step_off, int noise_mode, bool reduce = false)
{
    StepArgs<T> a{};
    const TileShape ts = tile_shape(h->precision, h->tile);
    a.in = static_cast<const T*>(h->buf[parity_in & 1]);
    a.out = static_cast<T*>(h->buf[(parity_in + 1) & 1]);
    a.ny = (int)h->ny;
    a.nx = (int)h->nx;
    a.nx4 = (h->nx % 4) == 0;
    a.ntx = grid_tiles(h->nx, ts.tw);
    a.nty = grid_tiles(h->ny, ts.th);
    a.xcd_chunk = xcd_chunk(a.ntx * a.nty);
    a.k = make_coef<T>(h->stencil, h->model);
    a.key0 = (uint32_t)h->options.seed;
    a.key1 = (uint32_t)(h->options.seed >> 32);
    a.step_base = h->d_step;
    a.step_off = step_off;
    a.noise = nullptr;
    if (noise_mode == kNoiseBuffer)
        a.noise = h->d_noise + (size_t)h->noise_used * 2 * (size_t)(h->ny * h->nx);
    a.part = reduce ? h->d_tile_part : nullptr;
    return a;
}

// One step; with `reduce` (the last step of a run) every tile also writes its partial
// of the fused finite check and FP64 mass.
hipError_t launch_step(const amb_handle* h, int parity_in, int step_off, int noise_mode, bool reduce = false)
{
    if (h->precision == AMB_F32)
        return launch_step_f32(step_args<float>(h, parity_in, step_off, noise_mode, reduce), h->tile, noise_mode,
                               h->stream);
    return launch_step_f64(step_args<double>(h, parity_in, step_off, noise_mode, reduce), h->tile, noise_mode,
                           h->stream);
}

int64_t tiles_of(const amb_handle* h, int tile)
{
    const TileShape ts = tile_shape(h->precision, tile);
    return (int64_t)grid_tiles(h->nx, ts.tw) * grid_tiles(h->ny, ts.th);
}

int noise_mode_of(const amb_handle* h)
{
    if (h->noise_used < h->noise_steps) return kNoiseBuffer;
    return h->model.T > 0.0 ? kNoisePhilox : kNoiseNone;
}

void drop_graphs(amb_handle* h)
{
    for (int p = 0; p < 2; ++p) {
        if (h->graph[p]) (void)hipGraphExecDestroy(h->graph[p]);
        h->graph[p] = nullptr;
        h->graph_tile[p] = -1;
        h->graph_noise[p] = -1;
    }
}

int build_graph(amb_handle* h, int parity, int noise_mode)
{
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail_hip(e, "hipStreamBeginCapture");
    for (int k = 0; k < kGraphSteps && e == hipSuccess; ++k)
        e = launch_step(h, parity + k, k, noise_mode, k == kGraphSteps - 1);
    if (e == hipSuccess) e = launch_advance(h->d_step, kGraphSteps, h->stream);
    const hipError_t e2 = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail_hip(e, "step launch during graph capture");
    }
    if (e2 != hipSuccess) return fail_hip(e2, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail_hip(e, "hipGraphInstantiate");
    if (h->graph[parity]) (void)hipGraphExecDestroy(h->graph[parity]);
    h->graph[parity] = exec;
    h->graph_tile[parity] = h->tile;
    h->graph_noise[parity] = noise_mode;
    return AMB_OK;
}

// Enqueues nsteps steps on the handle's stream (no synchronisation).
int enqueue_steps(amb_handle* h, int64_t nsteps)
{
    hipError_t e;
    int64_t left = nsteps;
    while (left > 0) {
        const int mode = noise_mode_of(h);
        if (mode == kNoiseBuffer) {
            if ((e = launch_step(h, h->cur, 0, mode, true)) != hipSuccess) return fail_hip(e, "step launch");
            if ((e = launch_advance(h->d_step, 1, h->stream)) != hipSuccess) return fail_hip(e, "advance launch");
            h->cur ^= 1;
            ++h->noise_used;
            --left;
            continue;
        }
        if (left >= kGraphSteps) {
            const int p = h->cur;
            if (!h->graph[p] || h->graph_tile[p] != h->tile || h->graph_noise[p] != mode) {
                const int rc = build_graph(h, p, mode);
                if (rc != AMB_OK) return rc;
            }
            const int64_t reps = left / kGraphSteps;
            for (int64_t r = 0; r < reps; ++r)
                if ((e = hipGraphLaunch(h->graph[p], h->stream)) != hipSuccess) return fail_hip(e, "hipGraphLaunch");
            left -= reps * kGraphSteps;
            continue;
        }
        for (int k = 0; k < left; ++k)
            if ((e = launch_step(h, h->cur + k, k, mode, k == left - 1)) != hipSuccess) return fail_hip(e, "step launch");
        if ((e = launch_advance(h->d_step, left, h->stream)) != hipSuccess) return fail_hip(e, "advance launch");
        h->cur = (int)((h->cur + left) & 1);
        left = 0;
    }
    return AMB_OK;
}

int reduce_now(amb_handle* h, double* sum, int64_t* bad)
{
    const int64_t n = h->ny * h->nx;
    hipError_t e = h->precision == AMB_F32
                       ? launch_reduce_f32(static_cast<const float*>(h->buf[h->cur]), n, h->d_part, h->d_bad_part,
                                           h->d_sum, h->d_bad, h->stream)
                       : launch_reduce_f64(static_cast<const double*>(h->buf[h->cur]), n, h->d_part,
                                           h->d_bad_part, h->d_sum, h->d_bad, h->stream);
    if (e != hipSuccess) return fail_hip(e, "reduction launch");
    if ((e = hipMemcpyAsync(h->h_sum, h->d_sum, sizeof(double), hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipMemcpyAsync(h->h_bad, h->d_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream)) !=
        hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    *sum = *h->h_sum;
    *bad = (int64_t)*h->h_bad;
    return AMB_OK;
}

// Mass and non-finite count after a run of nsteps: the last step of the run wrote one
// partial per tile, summed here by one small fixed-order reduction; for nsteps = 0 the
// field is read in full.
int finish_run(amb_handle* h, int64_t nsteps, double* sum, int64_t* bad)
{
    if (nsteps == 0) {
        const int rc = reduce_now(h, sum, bad);
        if (rc == AMB_OK) {
            h->last_sum = *sum;
            h->last_bad = *bad;
        }
        return rc;
    }
    hipError_t e = launch_reduce_tiles(h->d_tile_part, tiles_of(h, h->tile), h->d_sum, h->d_bad, h->stream);
    if (e != hipSuccess) return fail_hip(e, "reduction launch");
    if ((e = hipMemcpyAsync(h->h_sum, h->d_sum, sizeof(double), hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipMemcpyAsync(h->h_bad, h->d_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream)) !=
        hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    *sum = *h->h_sum;
    *bad = (int64_t)*h->h_bad;
    h->last_sum = *sum;
    h->last_bad = *bad;
    return AMB_OK;
}

void release(amb_handle* h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    sk_cache_free(h->sk);
    for (void* p : {h->buf[0], h->buf[1], (void*)h->d_step, (void*)h->d_part, (void*)h->d_bad_part,
                    (void*)h->d_tile_part, (void*)h->d_sum, (void*)h->d_bad, (void*)h->d_obs, (void*)h->d_noise})
        if (p) (void)hipFree(p);
    if (h->h_sum) (void)hipHostFree(h->h_sum);
    if (h->h_bad) (void)hipHostFree(h->h_bad);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int upload_field(amb_handle* h, const double* field, int which)
{
    const size_t n = (size_t)(h->ny * h->nx);
    hipError_t e;
    if (h->precision == AMB_F32) {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)field[i];
        e = hipMemcpy(h->buf[which], tmp.data(), n * sizeof(float), hipMemcpyHostToDevice);
    } else {
        e = hipMemcpy(h->buf[which], field, n * sizeof(double), hipMemcpyHostToDevice);
    }
    return e == hipSuccess ? AMB_OK : fail_hip(e, "hipMemcpy (field upload)");
}

// Computes observable `which` into d_obs + 2n (workspace d_obs[0, 2n)); allocated on
// the first getter call.
int compute_observable(amb_handle* h, int which)
{
    const size_t n = (size_t)(h->ny * h->nx);
    hipError_t e;
    if (!h->d_obs) {
        if ((e = hipMalloc(&h->d_obs, 4 * n * sizeof(double))) != hipSuccess) return fail_hip(e, "hipMalloc");
    }
    const int mode = noise_mode_of(h);
    if (h->precision == AMB_F32) {
        StepArgs<float> a = step_args<float>(h, h->cur, 0, mode);
        if (mode != kNoiseBuffer) a.noise = nullptr;
        e = launch_observable_f32(a, which, h->d_obs, h->d_obs + 2 * n, h->stream);
    } else {
        StepArgs<double> a = step_args<double>(h, h->cur, 0, mode);
        if (mode != kNoiseBuffer) a.noise = nullptr;
        e = launch_observable_f64(a, which, h->d_obs, h->d_obs + 2 * n, h->stream);
    }
    if (e != hipSuccess) return fail_hip(e, "observable launch");
    return AMB_OK;
}

}  // namespace

namespace amb {

int handle_view(amb_handle* h, FieldView* v)
{
    if (!h || !v) return fail(AMB_E_ARG, "handle must not be null");
    v->cur = h->buf[h->cur];
    v->other = h->buf[h->cur ^ 1];
    v->ny = h->ny;
    v->nx = h->nx;
    v->precision = h->precision;
    v->stream = h->stream;
    return AMB_OK;
}

int handle_commit(amb_handle* h, int64_t steps, bool swapped)
{
    if (swapped) h->cur ^= 1;
    h->timestep += steps;
    const hipError_t e = launch_advance(h->d_step, steps, h->stream);
    return e == hipSuccess ? AMB_OK : fail_hip(e, "advance launch");
}

int handle_enqueue(amb_handle* h, int64_t steps)
{
    const int rc = enqueue_steps(h, steps);
    if (rc == AMB_OK) h->timestep += steps;
    return rc;
}

int handle_set_timestep(amb_handle* h, int64_t timestep)
{
    const long long v = timestep;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h->d_step, &v, sizeof v, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "set timestep");
    h->timestep = timestep;
    return AMB_OK;
}

int64_t handle_timestep(const amb_handle* h) { return h->timestep; }

SkCache** handle_sk_cache(amb_handle* h) { return &h->sk; }

}  // namespace amb

extern "C" {

const char* amb_last_error(void) { return g_error.c_str(); }

int amb_set_field(amb_handle* h, const double* field)
{
    clear_error();
    if (!h || !field) return fail(AMB_E_ARG, "handle and field must not be null");
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    return upload_field(h, field, h->cur);
}

int amb_device_count(int* count)
{
    if (!count) return fail(AMB_E_ARG, "count must not be null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return AMB_OK;
}

int amb_tile_count(int precision, int* count)
{
    if (!count || (precision != AMB_F32 && precision != AMB_F64)) return fail(AMB_E_ARG, "bad precision");
    *count = tile_count(precision);
    return AMB_OK;
}

int amb_tile_shape(int precision, int index, int* tile_cols, int* tile_rows, int* threads)
{
    if (precision != AMB_F32 && precision != AMB_F64) return fail(AMB_E_ARG, "bad precision");
    if (index < 0 || index >= tile_count(precision)) return fail(AMB_E_ARG, "tile index out of range");
    const TileShape t = tile_shape(precision, index);
    if (tile_cols) *tile_cols = t.tw;
    if (tile_rows) *tile_rows = t.th;
    if (threads) *threads = t.nt;
    return AMB_OK;
}

int amb_create(const double* field, int64_t ny, int64_t nx, const amb_stencil* stencil, const amb_model* model,
               const amb_options* options, amb_handle** out)
{
    clear_error();
    if (!out) return fail(AMB_E_ARG, "out must not be null");
    *out = nullptr;
    if (!field) return fail(AMB_E_ARG, "field must not be null");
    int rc = validate_params(ny, nx, stencil, model);
    if (rc != AMB_OK) return rc;
    amb_options opt{AMB_F64, 0, 0, 0};
    if (options) opt = *options;
    if (opt.precision != AMB_F32 && opt.precision != AMB_F64)
        return fail(AMB_E_ARG, "precision must be AMB_F32 or AMB_F64");
    if (opt.tile < 0 || opt.tile > tile_count(opt.precision)) return fail(AMB_E_ARG, "tile index out of range");
    if ((rc = check_device()) != AMB_OK) return rc;

    amb_handle* h = new (std::nothrow) amb_handle();
    if (!h) return fail(AMB_E_ARG, "out of host memory");
    h->ny = ny;
    h->nx = nx;
    h->stencil = *stencil;
    h->model = *model;
    h->options = opt;
    h->precision = opt.precision;
    h->esize = opt.precision == AMB_F32 ? 4 : 8;
    h->timestep = opt.timestep;
    const size_t n = (size_t)(ny * nx);
    hipError_t e;
#define AMB_CHECK(call)                                  \
    do {                                                 \
        e = (call);                                      \
        if (e != hipSuccess) {                           \
            const int code_ = fail_hip(e, #call);        \
            release(h);                                  \
            return code_;                                \
        }                                                \
    } while (0)
    AMB_CHECK(hipGetDevice(&h->device));
    AMB_CHECK(prepare_step_kernels());
    h->tile = opt.tile > 0 ? opt.tile - 1
                           : initial_tile(opt.precision, ny, nx, model->T > 0.0 ? kNoisePhilox : kNoiseNone);
    AMB_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    AMB_CHECK(hipMalloc(&h->buf[0], n * h->esize));
    AMB_CHECK(hipMalloc(&h->buf[1], n * h->esize));
    AMB_CHECK(hipMalloc(&h->d_step, sizeof(long long)));
    AMB_CHECK(hipMalloc(&h->d_part, kReduceBlocks * sizeof(double)));
    AMB_CHECK(hipMalloc(&h->d_bad_part, kReduceBlocks * sizeof(unsigned long long)));
    {
        int64_t most = 1;  // partials for the tile shape with the most tiles (the tile may change)
        for (int t = 0; t < tile_count(h->precision); ++t) most = std::max(most, tiles_of(h, t));
        AMB_CHECK(hipMalloc(&h->d_tile_part, (size_t)most * 2 * sizeof(double)));
    }
    AMB_CHECK(hipMalloc(&h->d_sum, sizeof(double)));
    AMB_CHECK(hipMalloc(&h->d_bad, sizeof(unsigned long long)));
    AMB_CHECK(hipHostMalloc(&h->h_sum, sizeof(double), hipHostMallocDefault));
    AMB_CHECK(hipHostMalloc(&h->h_bad, sizeof(unsigned long long), hipHostMallocDefault));
    AMB_CHECK(hipEventCreate(&h->ev0));
    AMB_CHECK(hipEventCreate(&h->ev1));
    const long long step0 = (long long)h->timestep;
    AMB_CHECK(hipMemcpy(h->d_step, &step0, sizeof(long long), hipMemcpyHostToDevice));
#undef AMB_CHECK
    rc = upload_field(h, field, 0);
    if (rc != AMB_OK) {
        release(h);
        return rc;
    }
    *out = h;
    return AMB_OK;
}

int amb_destroy(amb_handle* h)
{
    clear_error();
    release(h);
    return AMB_OK;
}

int amb_run(amb_handle* h, int64_t nsteps)
{
    clear_error();
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    if (nsteps < 0) return fail(AMB_E_ARG, "nsteps must be non-negative");
    int rc = enqueue_steps(h, nsteps);
    if (rc != AMB_OK) return rc;
    h->timestep += nsteps;
    double sum;
    int64_t bad;
    if ((rc = finish_run(h, nsteps, &sum, &bad)) != AMB_OK) return rc;
    if (bad != 0) return fail(AMB_E_NONFINITE, "an unknown numerical error occurred during simulation");
    return AMB_OK;
}

int amb_run_timed(amb_handle* h, int64_t nsteps, double* device_ms)
{
    clear_error();
    if (!h || !device_ms) return fail(AMB_E_ARG, "handle and device_ms must not be null");
    if (nsteps < 0) return fail(AMB_E_ARG, "nsteps must be non-negative");
    hipError_t e;
    if ((e = hipEventRecord(h->ev0, h->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    int rc = enqueue_steps(h, nsteps);
    if (rc != AMB_OK) return rc;
    if ((e = hipEventRecord(h->ev1, h->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    if ((e = hipEventSynchronize(h->ev1)) != hipSuccess) return fail_hip(e, "hipEventSynchronize");
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, h->ev0, h->ev1)) != hipSuccess) return fail_hip(e, "hipEventElapsedTime");
    *device_ms = ms;
    h->timestep += nsteps;
    double sum;
    int64_t bad;
    if ((rc = finish_run(h, nsteps, &sum, &bad)) != AMB_OK) return rc;
    if (bad != 0) return fail(AMB_E_NONFINITE, "an unknown numerical error occurred during simulation");
    return AMB_OK;
}

int amb_get_field(amb_handle* h, double* out)
{
    clear_error();
    if (!h || !out) return fail(AMB_E_ARG, "handle and out must not be null");
    const size_t n = (size_t)(h->ny * h->nx);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    if (h->precision == AMB_F32) {
        std::vector<float> tmp(n);
        e = hipMemcpy(tmp.data(), h->buf[h->cur], n * sizeof(float), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    } else {
        e = hipMemcpy(out, h->buf[h->cur], n * sizeof(double), hipMemcpyDeviceToHost);
    }
    return e == hipSuccess ? AMB_OK : fail_hip(e, "hipMemcpy (field download)");
}

int amb_get_observable(amb_handle* h, int which, double* out)
{
    clear_error();
    if (!h || !out) return fail(AMB_E_ARG, "handle and out must not be null");
    if (which < AMB_OBS_MU || which > AMB_OBS_J_RANDOM) return fail(AMB_E_ARG, "unknown observable");
    const int rc = compute_observable(h, which);
    if (rc != AMB_OK) return rc;
    const size_t n = (size_t)(h->ny * h->nx);
    const size_t count = which <= AMB_OBS_MU_ACTIVE ? n : 2 * n;
    hipError_t e = hipMemcpyAsync(out, h->d_obs + 2 * n, count * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy (observable)");
    return AMB_OK;
}

int amb_get_timestep(const amb_handle* h, int64_t* timestep)
{
    if (!h || !timestep) return fail(AMB_E_ARG, "handle and timestep must not be null");
    *timestep = h->timestep;
    return AMB_OK;
}

int amb_get_params(const amb_handle* h, amb_stencil* stencil, amb_model* model)
{
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    if (stencil) *stencil = h->stencil;
    if (model) *model = h->model;
    return AMB_OK;
}

int amb_get_shape(const amb_handle* h, int64_t* ny, int64_t* nx)
{
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    if (ny) *ny = h->ny;
    if (nx) *nx = h->nx;
    return AMB_OK;
}

int amb_get_options(const amb_handle* h, amb_options* options)
{
    if (!h || !options) return fail(AMB_E_ARG, "handle and options must not be null");
    *options = h->options;
    options->timestep = h->timestep;
    options->tile = h->tile + 1;
    return AMB_OK;
}

int amb_mass(amb_handle* h, double* sum, int64_t* nonfinite)
{
    clear_error();
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    double s;
    int64_t b;
    const int rc = reduce_now(h, &s, &b);
    if (rc != AMB_OK) return rc;
    if (sum) *sum = s;
    if (nonfinite) *nonfinite = b;
    return AMB_OK;
}

int amb_last_mass(amb_handle* h, double* sum, int64_t* nonfinite)
{
    clear_error();
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    if (sum) *sum = h->last_sum;
    if (nonfinite) *nonfinite = h->last_bad;
    return AMB_OK;
}

int amb_set_noise(amb_handle* h, const double* noise, int64_t nsteps)
{
    clear_error();
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    if (nsteps < 0 || (nsteps > 0 && !noise)) return fail(AMB_E_ARG, "noise buffer must cover nsteps >= 0 steps");
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    if (h->d_noise) (void)hipFree(h->d_noise);
    h->d_noise = nullptr;
    h->noise_steps = h->noise_used = 0;
    if (nsteps == 0) return AMB_OK;
    const size_t bytes = (size_t)nsteps * 2 * (size_t)(h->ny * h->nx) * sizeof(double);
    if ((e = hipMalloc(&h->d_noise, bytes)) != hipSuccess) return fail_hip(e, "hipMalloc (noise)");
    if ((e = hipMemcpy(h->d_noise, noise, bytes, hipMemcpyHostToDevice)) != hipSuccess)
        return fail_hip(e, "hipMemcpy (noise)");
    h->noise_steps = nsteps;
    return AMB_OK;
}

int amb_autotune(amb_handle* h, int reps, int* chosen)
{
    clear_error();
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    const int mode0 = h->model.T > 0.0 ? kNoisePhilox : kNoiseNone;
    if (reps < 0) {  // the tuning table's choice, when it has one
        int t = 0;
        if (tuned_tile(h->precision, h->ny, h->nx, mode0, &t)) {
            h->tile = t;
            if (chosen) *chosen = t;
            return AMB_OK;
        }
        reps = -reps;
    }
    if (reps < 1) reps = 8;
    const size_t bytes = (size_t)(h->ny * h->nx) * h->esize;
    amb_handle trial = {};
    // A shallow trial handle on scratch buffers and a scratch step counter.
    hipError_t e = hipStreamSynchronize(h->stream);
    void* scratch[2] = {nullptr, nullptr};
    long long* step = nullptr;
    if (e == hipSuccess) e = hipMalloc(&scratch[0], bytes);
    if (e == hipSuccess) e = hipMalloc(&scratch[1], bytes);
    if (e == hipSuccess) e = hipMalloc(&step, sizeof(long long));
    if (e == hipSuccess) e = hipMemcpyAsync(scratch[0], h->buf[h->cur], bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(step, h->d_step, sizeof(long long), hipMemcpyDeviceToDevice, h->stream);
    float best = 1e30f;
    int best_t = h->tile;
    const int mode = h->model.T > 0.0 ? kNoisePhilox : kNoiseNone;
    for (int t = 0; t < tile_count(h->precision) && e == hipSuccess; ++t) {
        trial.ny = h->ny;
        trial.nx = h->nx;
        trial.stencil = h->stencil;
        trial.model = h->model;
        trial.options = h->options;
        trial.precision = h->precision;
        trial.buf[0] = scratch[0];
        trial.buf[1] = scratch[1];
        trial.d_step = step;
        trial.stream = h->stream;
        trial.tile = t;
        for (int k = 0; k <= reps && e == hipSuccess; ++k) {
            if (k == 1) e = hipEventRecord(h->ev0, h->stream);
            if (e == hipSuccess) e = launch_step(&trial, k, 0, mode);
        }
        if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
        if (e == hipSuccess && ms < best) {
            best = ms;
            best_t = t;
        }
    }
    (void)hipStreamSynchronize(h->stream);
    for (void* p : {scratch[0], scratch[1], (void*)step})
        if (p) (void)hipFree(p);
    trial.stream = nullptr;
    if (e != hipSuccess) return fail_hip(e, "autotune");
    h->tile = best_t;
    record_tile(h->precision, h->ny, h->nx, mode0, best_t);
    if (chosen) *chosen = best_t;
    return AMB_OK;
}

int amb_tuning_set(int precision, int64_t ny, int64_t nx, int noisy, int tile)
{
    clear_error();
    if (precision != AMB_F32 && precision != AMB_F64) return fail(AMB_E_ARG, "bad precision");
    if (ny <= 0 || nx <= 0) return fail(AMB_E_ARG, "bad grid");
    if (tile < 0 || tile >= tile_count(precision)) return fail(AMB_E_ARG, "tile index out of range");
    record_tile(precision, ny, nx, noisy ? kNoisePhilox : kNoiseNone, tile);
    return AMB_OK;
}

int amb_tuning_get(int precision, int64_t ny, int64_t nx, int noisy, int* tile)
{
    clear_error();
    if (!tile) return fail(AMB_E_ARG, "tile must not be null");
    if (!tuned_tile(precision, ny, nx, noisy ? kNoisePhilox : kNoiseNone, tile)) *tile = -1;
    return AMB_OK;
}

int amb_tuning_count(int64_t* count)
{
    if (!count) return fail(AMB_E_ARG, "count must not be null");
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    *count = (int64_t)g_tune.size();
    return AMB_OK;
}

int amb_tuning_entry(int64_t index, int* precision, int64_t* ny, int64_t* nx, int* noisy, int* tile)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (index < 0 || index >= (int64_t)g_tune.size()) return fail(AMB_E_ARG, "tuning entry out of range");
    auto it = g_tune.begin();
    std::advance(it, index);
    if (precision) *precision = std::get<0>(it->first);
    if (ny) *ny = std::get<1>(it->first);
    if (nx) *nx = std::get<2>(it->first);
    if (noisy) *noisy = std::get<3>(it->first);
    if (tile) *tile = it->second;
    return AMB_OK;
}

int amb_tuning_clear(void)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune.clear();
    return AMB_OK;
}

int amb_set_tile(amb_handle* h, int index)
{
    clear_error();
    if (!h) return fail(AMB_E_ARG, "handle must not be null");
    if (index < 0 || index >= tile_count(h->precision)) return fail(AMB_E_ARG, "tile index out of range");
    h->tile = index;
    return AMB_OK;
}

}  // extern "C"
