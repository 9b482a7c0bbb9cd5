// Host build of the device noise arithmetic (active-model-bplus_amd/csrc/amb_math.h) for the
// CPU tests: Philox4x32-10 plus the FP64 Box-Muller (own logarithm, square root and sin/cos)
// as the kernels evaluate them, except for the seeds of the Newton steps (the correctly
// rounded reciprocal and reciprocal square root here, the gfx950 instructions there).
// Built by tests/test_noise_math.py with g++.
#define __host__
#define __device__
#include <cstdint>

#include "amb_math.h"

extern "C" void stub_normals(uint64_t seed, int64_t step, int comp, int64_t row, int64_t group, double z[4])
{
    const amb::U4 w = amb::philox4x32_10(amb::noise_counter((uint32_t)group, (uint32_t)row, (uint64_t)step, comp),
                                         (uint32_t)seed, (uint32_t)(seed >> 32));
    amb::box_muller(w.x, w.y, z[0], z[1]);
    amb::box_muller(w.z, w.w, z[2], z[3]);
}

extern "C" void stub_box_muller(const uint32_t* words, int64_t n, double* z)
{
    for (int64_t i = 0; i < n; ++i) amb::box_muller(words[2 * i], words[2 * i + 1], z[2 * i], z[2 * i + 1]);
}

extern "C" double stub_ln_unit(double u) { return amb::ln_unit(u); }

extern "C" void stub_sincos_turn(double u, double* s, double* c) { amb::sincos_turn(u, *s, *c); }
