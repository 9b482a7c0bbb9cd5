// amb_internal.h -- helpers shared by the host runtime's translation units (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "amb.h"
#include "amb_kernels.h"

namespace amb {

int fail(int code, const std::string& msg);
int fail_hip(hipError_t e, const char* what);
void clear_error();
int check_device();
int validate_params(int64_t ny, int64_t nx, const amb_stencil* st, const amb_model* m);
int grid_tiles(int64_t n, int t);
int default_tile(int precision, int64_t ny, int64_t nx, int noise);
// The autotuners' table (process-wide): lookup, record, and the tile a new field starts with.
bool tuned_tile(int precision, int64_t ny, int64_t nx, int noise, int* tile);
void record_tile(int precision, int64_t ny, int64_t nx, int noise, int tile);
int initial_tile(int precision, int64_t ny, int64_t nx, int noise);
int xcd_chunk(int nblocks);

template <typename T>
StepCoef<T> make_coef(const amb_stencil& st, const amb_model& m);

// Access to an integrator's state for other parts of the library (the surrogate).
struct FieldView {
    void* cur;          // current field (ny*nx elements of the handle's precision)
    void* other;        // the second ping-pong buffer
    int64_t ny, nx;
    int precision;
    hipStream_t stream;
};
int handle_view(amb_handle* h, FieldView* v);
// Advances the step counter by `steps` (host and device); with `swapped` the field
// now lives in the buffer that was `other`.
int handle_commit(amb_handle* h, int64_t steps, bool swapped);
// Enqueues `steps` explicit-Euler steps on the handle's stream without waiting (no
// finite check) and advances the step counter.
int handle_enqueue(amb_handle* h, int64_t steps);
// Sets the step counter (host and device), after the handle's stream is idle.
int handle_set_timestep(amb_handle* h, int64_t timestep);
int64_t handle_timestep(const amb_handle* h);

// Per-handle state of the structure factor (FFT plan and buffers), created by
// amb_structure_factor on first use and freed with the handle.
struct SkCache;
SkCache** handle_sk_cache(amb_handle* h);
void sk_cache_free(SkCache* c);

}  // namespace amb
