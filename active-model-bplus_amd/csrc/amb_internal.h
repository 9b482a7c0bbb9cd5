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
int default_tile(int precision, int64_t ny, int64_t nx);
int xcd_chunk(int nblocks);

template <typename T>
StepCoef<T> make_coef(const amb_stencil& st, const amb_model& m);

}  // namespace amb
