// amb_batch.hip -- batched ensemble: many independent realisations (own grid size,
// model, stencil and Philox seed) advanced together.  No reference counterpart.
//
// All fields live back to back in two device buffers (ping-pong), addressed with
// 64-bit offsets.  A step is one launch per tile-shape group over a flattened
// work list of (realisation, tile) pairs, so hundreds of small grids fill the
// 256 CUs in one dispatch instead of one under-filled launch each; kGraphSteps
// steps (all groups, then the step-counter advance) are captured in one hipGraph
// and replayed.  The autotuner times every compiled tile shape on every distinct
// grid size and assigns each size class its fastest shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <new>
#include <utility>
#include <vector>

#include "amb.h"
#include "amb_internal.h"
#include "amb_kernels.h"
#include "amb_tile.h"

namespace amb {

template <typename T>
struct BatchDesc {
    int64_t off;  // element offset of the realisation in the concatenated buffers
    int ny, nx, nx4, ntx;
    uint32_t key0, key1;
    StepCoef<T> k;
};

// NOISE is kNoiseNone or kNoisePhilox: members with and without noise run in separate
// launches, so each kernel holds only its own stage-B windows in registers.
template <typename T, int TW, int TH, int NT, int NOISE>
__global__ __launch_bounds__(NT, (min_waves_per_eu<T, TW, TH, NT>())) void batch_step_kernel(const BatchDesc<T>* __restrict__ desc,
                                                              const int2* __restrict__ tiles, const T* in, T* out,
                                                              const long long* step_base, int step_off, double* part)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int2 t = tiles[blockIdx.x];
    double* const red = part ? part + 2 * (size_t)blockIdx.x : nullptr;
    if (t.x < 0) {  // padding entry of the XCD-ordered list
        if (red && threadIdx.x == 0) {
            red[0] = 0.0;
            red[1] = 0.0;
        }
        return;
    }
    const BatchDesc<T>& d = desc[t.x];
    const int ty = t.y / d.ntx, tx = t.y - ty * d.ntx;
    const uint64_t step = (uint64_t)(*step_base) + (uint64_t)step_off;
    step_tile<T, TW, TH, NT, NOISE>(reinterpret_cast<T*>(smem_raw), in + d.off, out + d.off, d.ny, d.nx, d.nx4,
                                    tx * TW, ty * TH, d.k, d.key0, d.key1, step, nullptr, red);
}

// ---- resident members: one workgroup keeps a whole small realisation in LDS ------------
//
// For members whose two field buffers and one tile's scratch fit in kResidentLds bytes,
// one launch advances `nsteps` steps with no global traffic in between: the workgroup
// loads its member's field into LDS, applies step_tile to each tile of the member in turn
// (reading one LDS buffer, writing the other; step_tile accepts any address space through
// generic pointers), swaps the buffers after each step and writes the field back once.
// No workgroup waits for another.  Same per-site arithmetic as every other path, so a
// resident member's trajectory equals a plain integrator's bit for bit.  The last act of a
// launch is the member's partial of the fused finite check (its FP64 sum and non-finite
// count, reduced in a fixed order) at part[slot].
constexpr size_t kResidentLds = 80 * 1024;

template <typename T>
__host__ __device__ constexpr size_t resident_field_bytes(int64_t n)
{
    return ((size_t)n * sizeof(T) + 15) / 16 * 16;
}

template <typename T, int TW, int TH, int NT, int NOISE>
__global__ __launch_bounds__(NT) void resident_kernel(const BatchDesc<T>* __restrict__ desc,
                                                      const int* __restrict__ members, const T* in, T* out,
                                                      const long long* step_base, int nsteps, double* part,
                                                      int64_t part_base)
{
    using L = Lds<T, TW, TH, NT>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const BatchDesc<T>& d = desc[members[blockIdx.x]];
    const int64_t n = (int64_t)d.ny * d.nx;
    const size_t fb = resident_field_bytes<T>(n);
    T* a = reinterpret_cast<T*>(smem_raw);
    T* b = reinterpret_cast<T*>(smem_raw + fb);
    T* const scratch = reinterpret_cast<T*>(smem_raw + 2 * fb);
    const int tid = threadIdx.x;
    for (int64_t i = tid; i < n; i += NT) a[i] = in[d.off + i];
    __syncthreads();
    const uint64_t step0 = (uint64_t)(*step_base);
    const int ntx = (d.nx + TW - 1) / TW, nty = (d.ny + TH - 1) / TH;
    for (int k = 0; k < nsteps; ++k) {
        for (int t = 0; t < ntx * nty; ++t) {
            const int ty = t / ntx, tx = t - ty * ntx;
            step_tile<T, TW, TH, NT, NOISE>(scratch, a, b, d.ny, d.nx, d.nx4, tx * TW, ty * TH, d.k, d.key0, d.key1,
                                            step0 + (uint64_t)k, nullptr, nullptr);
            __syncthreads();  // the next tile's stage 0 overwrites what this one's stage B reads
        }
        T* const c = a;
        a = b;
        b = c;
    }
    double s = 0.0, bad = 0.0;
    for (int64_t i = tid; i < n; i += NT) {
        const T v = a[i];
        out[d.off + i] = v;
        const double dv = (double)v;
        if (isfinite(dv)) s += dv;
        else bad += 1.0;
    }
    // fixed-order block reduction: lane butterfly, then the waves in index order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        bad += __shfl_xor(bad, off, 64);
    }
    double* const red = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(scratch) + L::red_off);
    __syncthreads();
    if ((tid & 63) == 0) {
        red[2 * (tid >> 6)] = s;
        red[2 * (tid >> 6) + 1] = bad;
    }
    __syncthreads();
    if (tid == 0) {
        double ts = 0.0, tb = 0.0;
        for (int w = 0; w < NT / 64; ++w) {
            ts += red[2 * w];
            tb += red[2 * w + 1];
        }
        part[2 * (part_base + blockIdx.x)] = ts;
        part[2 * (part_base + blockIdx.x) + 1] = tb;
    }
}

template <typename T>
struct ResidentShape;
template <>
struct ResidentShape<float> {
    static constexpr int tw = 32, th = 32, nt = 256;
};
template <>
struct ResidentShape<double> {
    static constexpr int tw = 32, th = 16, nt = 256;
};

template <typename T>
constexpr size_t resident_lds(int64_t n)
{
    using S = ResidentShape<T>;
    return 2 * resident_field_bytes<T>(n) + Lds<T, S::tw, S::th, S::nt>::bytes;
}

// A member runs resident when its two buffers and one tile's scratch fit kResidentLds.
static bool fits_resident(int precision, int64_t ny, int64_t nx)
{
    const int64_t n = ny * nx;
    return precision == AMB_F32 ? resident_lds<float>(n) <= kResidentLds : resident_lds<double>(n) <= kResidentLds;
}

template <typename T, int NOISE>
static hipError_t launch_resident_t(const void* desc, const int* members, int count, int64_t max_n, const void* in,
                                    void* out, const long long* step, int nsteps, double* part, int64_t part_base,
                                    hipStream_t s)
{
    using S = ResidentShape<T>;
    hipLaunchKernelGGL((resident_kernel<T, S::tw, S::th, S::nt, NOISE>), dim3(count), dim3(S::nt),
                       resident_lds<T>(max_n), s, static_cast<const BatchDesc<T>*>(desc), members,
                       static_cast<const T*>(in), static_cast<T*>(out), step, nsteps, part, part_base);
    return hipGetLastError();
}

static hipError_t launch_resident(int precision, int noise, const void* desc, const int* members, int count,
                                  int64_t max_n, const void* in, void* out, const long long* step, int nsteps,
                                  double* part, int64_t part_base, hipStream_t s)
{
    if (precision == AMB_F32)
        return noise == kNoisePhilox
                   ? launch_resident_t<float, kNoisePhilox>(desc, members, count, max_n, in, out, step, nsteps, part,
                                                            part_base, s)
                   : launch_resident_t<float, kNoiseNone>(desc, members, count, max_n, in, out, step, nsteps, part,
                                                          part_base, s);
    return noise == kNoisePhilox
               ? launch_resident_t<double, kNoisePhilox>(desc, members, count, max_n, in, out, step, nsteps, part,
                                                         part_base, s)
               : launch_resident_t<double, kNoiseNone>(desc, members, count, max_n, in, out, step, nsteps, part,
                                                       part_base, s);
}

template <typename T>
static hipError_t raise_resident_lds_t()
{
    using S = ResidentShape<T>;
    hipError_t e = hipFuncSetAttribute((const void*)resident_kernel<T, S::tw, S::th, S::nt, kNoiseNone>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResidentLds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)resident_kernel<T, S::tw, S::th, S::nt, kNoisePhilox>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResidentLds);
    return e;
}

template <typename T, int TW, int TH, int NT>
static hipError_t raise_batch_lds()
{
    using L = Lds<T, TW, TH, NT>;
    if (L::bytes <= 65536) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void*)batch_step_kernel<T, TW, TH, NT, kNoiseNone>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)L::bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)batch_step_kernel<T, TW, TH, NT, kNoisePhilox>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)L::bytes);
    return e;
}

#define AMB_BRAISE_F32(i, tw, th, nt) if ((e = raise_batch_lds<float, tw, th, nt>()) != hipSuccess) return e;
#define AMB_BRAISE_F64(i, tw, th, nt) if ((e = raise_batch_lds<double, tw, th, nt>()) != hipSuccess) return e;

// Raises the dynamic-LDS limit of every batched instantiation (before any capture).
static hipError_t prepare_batch_kernels()
{
    hipError_t e = hipSuccess;
    AMB_F32_TILES(AMB_BRAISE_F32)
    AMB_F64_TILES(AMB_BRAISE_F64)
    if ((e = raise_resident_lds_t<float>()) != hipSuccess) return e;
    return raise_resident_lds_t<double>();
}

template <typename T, int TW, int TH, int NT>
static hipError_t launch_batch_one(int noise, const void* desc, const int2* tiles, int ntiles, const void* in,
                                   void* out, const long long* step, int step_off, double* part, hipStream_t s)
{
    using L = Lds<T, TW, TH, NT>;
    const auto* d = static_cast<const BatchDesc<T>*>(desc);
    if (noise == kNoisePhilox)
        hipLaunchKernelGGL((batch_step_kernel<T, TW, TH, NT, kNoisePhilox>), dim3(ntiles), dim3(NT), L::bytes, s, d,
                           tiles, static_cast<const T*>(in), static_cast<T*>(out), step, step_off, part);
    else
        hipLaunchKernelGGL((batch_step_kernel<T, TW, TH, NT, kNoiseNone>), dim3(ntiles), dim3(NT), L::bytes, s, d,
                           tiles, static_cast<const T*>(in), static_cast<T*>(out), step, step_off, part);
    return hipGetLastError();
}

#define AMB_BCASE_F32(i, tw, th, nt) \
    case i: return launch_batch_one<float, tw, th, nt>(noise, desc, tiles, ntiles, in, out, step, step_off, part, s);
#define AMB_BCASE_F64(i, tw, th, nt) \
    case i: return launch_batch_one<double, tw, th, nt>(noise, desc, tiles, ntiles, in, out, step, step_off, part, s);

// One step of the members listed in `tiles` (noise: kNoiseNone or kNoisePhilox, the same
// for all of them).  part (nullable): the per-tile partials of the fused finite check,
// one entry per entry of `tiles` (padding entries write zeros).
static hipError_t launch_batch(int precision, int tile, int noise, const void* desc, const int2* tiles, int ntiles,
                               const void* in, void* out, const long long* step, int step_off, double* part,
                               hipStream_t s)
{
    if (precision == AMB_F32) {
        switch (tile) {
            AMB_F32_TILES(AMB_BCASE_F32)
        default: return hipErrorInvalidValue;
        }
    }
    switch (tile) {
        AMB_F64_TILES(AMB_BCASE_F64)
    default: return hipErrorInvalidValue;
    }
}

}  // namespace amb

using namespace amb;

namespace {

constexpr int kBatchGraphSteps = 32;

struct Group {
    int tile = 0;
    int noise = 0;  // kNoiseNone or kNoisePhilox
    std::vector<int> members;
    int2* d_tiles = nullptr;
    int ntiles = 0;
    int64_t part_base = 0;  // first entry of this group's tiles in the partials buffer
};

}  // namespace

struct amb_batch {
    int precision = AMB_F32;
    size_t esize = 4;
    hipStream_t stream = nullptr;
    std::vector<amb_realisation> desc;
    std::vector<int64_t> off;
    int64_t total = 0;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    long long* d_step = nullptr;
    long long* d_step_scratch = nullptr;
    void* d_desc = nullptr;
    std::vector<int> tile_of;  // tile index per realisation
    std::vector<Group> groups;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    double* d_tile_part = nullptr;  // (sum, non-finite count) per tile of the last step of a run
    int64_t ntiles_total = 0;
    int64_t tile_part_cap = 0;      // entries allocated in d_tile_part
    double last_sum = 0.0;          // total mass and non-finite count after the last run
    int64_t last_bad = 0;
    // members kept in LDS by the resident kernel (when enabled and they fit)
    bool resident_on = true;
    std::vector<int> resident;      // noiseless members first, then the noisy ones
    int64_t resident_quiet = 0;     // how many of them are noiseless
    int* d_resident = nullptr;
    int64_t resident_max_n = 0;
    hipStream_t side = nullptr;     // capture branch of the resident launch, beside the tiled steps
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<hipStream_t> branch;      // capture branch per tile-shape group
    std::vector<hipEvent_t> branch_done;
    double* d_part = nullptr;
    unsigned long long* d_bad_part = nullptr;
    double* d_sum = nullptr;
    unsigned long long* d_bad = nullptr;
    double* h_sum = nullptr;
    unsigned long long* h_bad = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void drop_graphs(amb_batch* b)
{
    for (auto& g : b->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void free_groups(amb_batch* b)
{
    for (auto& g : b->groups)
        if (g.d_tiles) (void)hipFree(g.d_tiles);
    b->groups.clear();
}

void release(amb_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    free_groups(b);
    if (b->side) (void)hipStreamSynchronize(b->side);
    for (void* p : {b->buf[0], b->buf[1], (void*)b->d_step, (void*)b->d_step_scratch, b->d_desc, (void*)b->d_part,
                    (void*)b->d_bad_part, (void*)b->d_sum, (void*)b->d_bad, (void*)b->d_tile_part,
                    (void*)b->d_resident})
        if (p) (void)hipFree(p);
    for (hipStream_t st : b->branch) (void)hipStreamSynchronize(st);
    for (hipEvent_t ev : b->branch_done) (void)hipEventDestroy(ev);
    for (hipStream_t st : b->branch) (void)hipStreamDestroy(st);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->h_sum) (void)hipHostFree(b->h_sum);
    if (b->h_bad) (void)hipHostFree(b->h_bad);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Tile list of `members` for tile shape `tile`, ordered so that the round-robin
// workgroup-to-XCD dispatch hands each XCD a contiguous run of tiles (neighbouring
// tiles share halo rows in that XCD's L2).  Padded to a multiple of 8 with (-1, 0).
std::vector<int2> tile_list(const amb_batch* b, const std::vector<int>& members, int tile)
{
    const TileShape ts = tile_shape(b->precision, tile);
    std::vector<int2> natural;
    for (int r : members) {
        const int ntx = grid_tiles(b->desc[r].nx, ts.tw), nty = grid_tiles(b->desc[r].ny, ts.th);
        for (int t = 0; t < ntx * nty; ++t) natural.push_back(make_int2(r, t));
    }
    const size_t n = natural.size();
    const size_t padded = (n + 7) / 8 * 8;
    const size_t chunk = padded / 8;
    std::vector<int2> out(padded, make_int2(-1, 0));
    for (size_t i = 0; i < n; ++i) out[(i % chunk) * 8 + i / chunk] = natural[i];
    return out;
}

int rebuild_groups(amb_batch* b)
{
    drop_graphs(b);
    free_groups(b);
    std::map<std::pair<int, int>, std::vector<int>> by_tile;  // (tile, noise mode) -> members
    std::vector<int> quiet, noisy;
    b->resident_max_n = 0;
    for (size_t r = 0; r < b->desc.size(); ++r) {
        const amb_realisation& m = b->desc[r];
        const int noise = m.model.T > 0.0 ? kNoisePhilox : kNoiseNone;
        if (b->resident_on && fits_resident(b->precision, m.ny, m.nx)) {
            (noise == kNoisePhilox ? noisy : quiet).push_back((int)r);
            b->resident_max_n = std::max(b->resident_max_n, m.ny * m.nx);
        } else {
            by_tile[{b->tile_of[r], noise}].push_back((int)r);
        }
    }
    b->resident = quiet;
    b->resident.insert(b->resident.end(), noisy.begin(), noisy.end());
    b->resident_quiet = (int64_t)quiet.size();
    if (b->d_resident) (void)hipFree(b->d_resident);
    b->d_resident = nullptr;
    if (!b->resident.empty()) {
        hipError_t e = hipMalloc(&b->d_resident, b->resident.size() * sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpy(b->d_resident, b->resident.data(), b->resident.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail_hip(e, "resident member list upload");
    }
    b->ntiles_total = 0;
    for (auto& kv : by_tile) {
        Group g;
        g.tile = kv.first.first;
        g.noise = kv.first.second;
        g.members = kv.second;
        const std::vector<int2> list = tile_list(b, g.members, g.tile);
        g.ntiles = (int)list.size();
        g.part_base = b->ntiles_total;
        b->ntiles_total += g.ntiles;
        hipError_t e = hipMalloc(&g.d_tiles, list.size() * sizeof(int2));
        if (e == hipSuccess) e = hipMemcpy(g.d_tiles, list.data(), list.size() * sizeof(int2), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (g.d_tiles) (void)hipFree(g.d_tiles);
            return fail_hip(e, "tile list upload");
        }
        b->groups.push_back(std::move(g));
    }
    const int64_t nparts = b->ntiles_total + (int64_t)b->resident.size();  // resident partials after the tiles'
    if (nparts > b->tile_part_cap) {
        if (b->d_tile_part) (void)hipFree(b->d_tile_part);
        b->d_tile_part = nullptr;
        b->tile_part_cap = 0;
        const hipError_t e = hipMalloc(&b->d_tile_part, (size_t)nparts * 2 * sizeof(double));
        if (e != hipSuccess) return fail_hip(e, "hipMalloc (tile partials)");
        b->tile_part_cap = nparts;
    }
    return AMB_OK;
}

template <typename T>
std::vector<BatchDesc<T>> make_descs(const amb_batch* b)
{
    std::vector<BatchDesc<T>> d(b->desc.size());
    for (size_t r = 0; r < b->desc.size(); ++r) {
        const amb_realisation& s = b->desc[r];
        d[r].off = b->off[r];
        d[r].ny = (int)s.ny;
        d[r].nx = (int)s.nx;
        d[r].nx4 = (s.nx % 4) == 0;
        d[r].ntx = 0;  // set per launch group below
        d[r].key0 = (uint32_t)s.seed;
        d[r].key1 = (uint32_t)(s.seed >> 32);
        d[r].k = make_coef<T>(s.stencil, s.model);
    }
    return d;
}

// Descriptors carry ntx of the realisation's own tile shape.
int upload_descs(amb_batch* b)
{
    hipError_t e;
    if (b->precision == AMB_F32) {
        auto d = make_descs<float>(b);
        for (size_t r = 0; r < d.size(); ++r) d[r].ntx = grid_tiles(d[r].nx, tile_shape(b->precision, b->tile_of[r]).tw);
        e = hipMemcpy(b->d_desc, d.data(), d.size() * sizeof(d[0]), hipMemcpyHostToDevice);
    } else {
        auto d = make_descs<double>(b);
        for (size_t r = 0; r < d.size(); ++r) d[r].ntx = grid_tiles(d[r].nx, tile_shape(b->precision, b->tile_of[r]).tw);
        e = hipMemcpy(b->d_desc, d.data(), d.size() * sizeof(d[0]), hipMemcpyHostToDevice);
    }
    return e == hipSuccess ? AMB_OK : fail_hip(e, "descriptor upload");
}

// One step of every group; with `reduce` each tile also writes its partial of the fused
// finite check (the last step of a run).
hipError_t launch_groups(amb_batch* b, int parity_in, const long long* step, int step_off, bool reduce)
{
    for (const Group& g : b->groups) {
        const hipError_t e = launch_batch(b->precision, g.tile, g.noise, b->d_desc, g.d_tiles, g.ntiles,
                                          b->buf[parity_in & 1],
                                          b->buf[(parity_in + 1) & 1], step, step_off,
                                          reduce ? b->d_tile_part + 2 * g.part_base : nullptr, b->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// nsteps steps of every resident member in one launch on stream s, from buffer
// `parity_in` to buffer parity_in + nsteps.
hipError_t launch_resident_members(amb_batch* b, int parity_in, int nsteps, hipStream_t s)
{
    if (b->resident.empty() || nsteps <= 0) return hipSuccess;
    const int64_t nq = b->resident_quiet, nn = (int64_t)b->resident.size() - nq;
    hipError_t e = hipSuccess;
    if (nq > 0)
        e = launch_resident(b->precision, kNoiseNone, b->d_desc, b->d_resident, (int)nq, b->resident_max_n,
                            b->buf[parity_in & 1], b->buf[(parity_in + nsteps) & 1], b->d_step, nsteps,
                            b->d_tile_part, b->ntiles_total, s);
    if (e == hipSuccess && nn > 0)
        e = launch_resident(b->precision, kNoisePhilox, b->d_desc, b->d_resident + nq, (int)nn, b->resident_max_n,
                            b->buf[parity_in & 1], b->buf[(parity_in + nsteps) & 1], b->d_step, nsteps,
                            b->d_tile_part, b->ntiles_total + nq, s);
    return e;
}

// The graph of kBatchGraphSteps steps: the resident members' one launch on a branch of
// its own (fork from and join to the main stream), beside the tiled members' steps.
int build_graph(amb_batch* b, int parity)
{
    // one capture branch per tile-shape group (its kBatchGraphSteps steps in a chain) and
    // one for the resident members: the groups touch disjoint members, so their chains run
    // side by side and one group's last, partial round of workgroups overlaps the others'
    while (b->branch.size() < b->groups.size()) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        hipError_t ce = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (ce == hipSuccess) ce = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (ce != hipSuccess) {
            if (st) (void)hipStreamDestroy(st);
            return fail_hip(ce, "branch stream");
        }
        b->branch.push_back(st);
        b->branch_done.push_back(ev);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail_hip(e, "hipStreamBeginCapture");
    e = hipEventRecord(b->fork, b->stream);
    if (e == hipSuccess && !b->resident.empty()) {
        e = hipStreamWaitEvent(b->side, b->fork, 0);
        if (e == hipSuccess) e = launch_resident_members(b, parity, kBatchGraphSteps, b->side);
        if (e == hipSuccess) e = hipEventRecord(b->join, b->side);
    }
    for (size_t i = 0; i < b->groups.size() && e == hipSuccess; ++i) {
        const Group& gr = b->groups[i];
        hipStream_t st = b->branch[i];
        e = hipStreamWaitEvent(st, b->fork, 0);
        for (int k = 0; k < kBatchGraphSteps && e == hipSuccess; ++k) {
            const int pin = parity + k;
            e = launch_batch(b->precision, gr.tile, gr.noise, b->d_desc, gr.d_tiles, gr.ntiles, b->buf[pin & 1],
                             b->buf[(pin + 1) & 1], b->d_step, k,
                             k == kBatchGraphSteps - 1 ? b->d_tile_part + 2 * gr.part_base : nullptr, st);
        }
        if (e == hipSuccess) e = hipEventRecord(b->branch_done[i], st);
    }
    for (size_t i = 0; i < b->groups.size() && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->stream, b->branch_done[i], 0);
    if (e == hipSuccess && !b->resident.empty()) e = hipStreamWaitEvent(b->stream, b->join, 0);
    if (e == hipSuccess) e = launch_advance(b->d_step, kBatchGraphSteps, b->stream);
    const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail_hip(e, "batched launch during graph capture");
    }
    if (e2 != hipSuccess) return fail_hip(e2, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail_hip(e, "hipGraphInstantiate");
    b->graph[parity] = exec;
    return AMB_OK;
}

int enqueue(amb_batch* b, int64_t nsteps)
{
    hipError_t e;
    int64_t left = nsteps;
    if (left >= kBatchGraphSteps) {
        const int p = b->cur;
        if (!b->graph[p]) {
            const int rc = build_graph(b, p);
            if (rc != AMB_OK) return rc;
        }
        const int64_t reps = left / kBatchGraphSteps;
        for (int64_t r = 0; r < reps; ++r)
            if ((e = hipGraphLaunch(b->graph[p], b->stream)) != hipSuccess) return fail_hip(e, "hipGraphLaunch");
        left -= reps * kBatchGraphSteps;
    }
    if (left > 0 && (e = launch_resident_members(b, b->cur, (int)left, b->stream)) != hipSuccess)
        return fail_hip(e, "resident launch");
    for (int k = 0; k < left; ++k)
        if ((e = launch_groups(b, b->cur + k, b->d_step, k, k == left - 1)) != hipSuccess)
            return fail_hip(e, "batched launch");
    if (left > 0) {
        if ((e = launch_advance(b->d_step, left, b->stream)) != hipSuccess) return fail_hip(e, "advance launch");
        b->cur = (int)((b->cur + left) & 1);
    }
    return AMB_OK;
}

int reduce_range(amb_batch* b, int64_t off, int64_t n, double* sum, int64_t* bad)
{
    hipError_t e = b->precision == AMB_F32
                       ? launch_reduce_f32(static_cast<const float*>(b->buf[b->cur]) + off, n, b->d_part,
                                           b->d_bad_part, b->d_sum, b->d_bad, b->stream)
                       : launch_reduce_f64(static_cast<const double*>(b->buf[b->cur]) + off, n, b->d_part,
                                           b->d_bad_part, b->d_sum, b->d_bad, b->stream);
    if (e != hipSuccess) return fail_hip(e, "reduction launch");
    if ((e = hipMemcpyAsync(b->h_sum, b->d_sum, sizeof(double), hipMemcpyDeviceToHost, b->stream)) != hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipMemcpyAsync(b->h_bad, b->d_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost, b->stream)) !=
        hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    *sum = *b->h_sum;
    *bad = (int64_t)*b->h_bad;
    return AMB_OK;
}

// Mass and non-finite count after a run of nsteps: from the per-tile partials that its
// last step wrote (one small fixed-order reduction), or, for nsteps = 0, by reading
// every field.
int finish_run(amb_batch* b, int64_t nsteps, double* sum, int64_t* bad)
{
    if (nsteps == 0) return reduce_range(b, 0, b->total, sum, bad);
    hipError_t e = launch_reduce_tiles(b->d_tile_part, b->ntiles_total + (int64_t)b->resident.size(), b->d_sum,
                                       b->d_bad, b->stream);
    if (e != hipSuccess) return fail_hip(e, "reduction launch");
    if ((e = hipMemcpyAsync(b->h_sum, b->d_sum, sizeof(double), hipMemcpyDeviceToHost, b->stream)) != hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipMemcpyAsync(b->h_bad, b->d_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost, b->stream)) !=
        hipSuccess)
        return fail_hip(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    *sum = *b->h_sum;
    *bad = (int64_t)*b->h_bad;
    return AMB_OK;
}

}  // namespace

extern "C" {

int amb_batch_create(const amb_realisation* desc, int64_t count, const double* const* fields, int precision,
                     amb_batch** out)
{
    clear_error();
    if (!out) return fail(AMB_E_ARG, "out must not be null");
    *out = nullptr;
    if (!desc || !fields || count <= 0) return fail(AMB_E_ARG, "need at least one realisation");
    if (count > (1 << 24)) return fail(AMB_E_ARG, "too many realisations");
    if (precision != AMB_F32 && precision != AMB_F64) return fail(AMB_E_ARG, "precision must be AMB_F32 or AMB_F64");
    int64_t total = 0;
    for (int64_t r = 0; r < count; ++r) {
        const int rc = validate_params(desc[r].ny, desc[r].nx, &desc[r].stencil, &desc[r].model);
        if (rc != AMB_OK) return rc;
        if (!fields[r]) return fail(AMB_E_ARG, "field pointer must not be null");
        total += desc[r].ny * desc[r].nx;
    }
    int rc = check_device();
    if (rc != AMB_OK) return rc;
    amb_batch* b = new (std::nothrow) amb_batch();
    if (!b) return fail(AMB_E_ARG, "out of host memory");
    b->precision = precision;
    b->esize = precision == AMB_F32 ? 4 : 8;
    b->desc.assign(desc, desc + count);
    b->total = total;
    b->off.resize(count);
    int64_t o = 0;
    for (int64_t r = 0; r < count; ++r) {
        b->off[r] = o;
        o += desc[r].ny * desc[r].nx;
    }
    b->tile_of.assign(count, 0);
    for (int64_t r = 0; r < count; ++r)  // the tuning table's entries where it has them
        (void)tuned_tile(precision, desc[r].ny, desc[r].nx, desc[r].model.T > 0.0 ? kNoisePhilox : kNoiseNone,
                         &b->tile_of[r]);
    hipError_t e;
#define AMB_CHECK(call)                           \
    do {                                          \
        e = (call);                               \
        if (e != hipSuccess) {                    \
            const int code_ = fail_hip(e, #call); \
            release(b);                           \
            return code_;                         \
        }                                         \
    } while (0)
    const size_t desc_bytes = (size_t)count * (precision == AMB_F32 ? sizeof(BatchDesc<float>) : sizeof(BatchDesc<double>));
    AMB_CHECK(prepare_batch_kernels());
    AMB_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    AMB_CHECK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    AMB_CHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    AMB_CHECK(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
    AMB_CHECK(hipMalloc(&b->buf[0], (size_t)total * b->esize));
    AMB_CHECK(hipMalloc(&b->buf[1], (size_t)total * b->esize));
    AMB_CHECK(hipMalloc(&b->d_step, sizeof(long long)));
    AMB_CHECK(hipMalloc(&b->d_step_scratch, sizeof(long long)));
    AMB_CHECK(hipMalloc(&b->d_desc, desc_bytes));
    AMB_CHECK(hipMalloc(&b->d_part, kReduceBlocks * sizeof(double)));
    AMB_CHECK(hipMalloc(&b->d_bad_part, kReduceBlocks * sizeof(unsigned long long)));
    AMB_CHECK(hipMalloc(&b->d_sum, sizeof(double)));
    AMB_CHECK(hipMalloc(&b->d_bad, sizeof(unsigned long long)));
    AMB_CHECK(hipHostMalloc(&b->h_sum, sizeof(double), hipHostMallocDefault));
    AMB_CHECK(hipHostMalloc(&b->h_bad, sizeof(unsigned long long), hipHostMallocDefault));
    AMB_CHECK(hipEventCreate(&b->ev0));
    AMB_CHECK(hipEventCreate(&b->ev1));
    const long long zero = 0;
    AMB_CHECK(hipMemcpy(b->d_step, &zero, sizeof(long long), hipMemcpyHostToDevice));
    // Upload fields (converted to the storage precision on the host).
    for (int64_t r = 0; r < count; ++r) {
        const size_t n = (size_t)(desc[r].ny * desc[r].nx);
        if (precision == AMB_F32) {
            std::vector<float> tmp(n);
            for (size_t i = 0; i < n; ++i) tmp[i] = (float)fields[r][i];
            AMB_CHECK(hipMemcpy(static_cast<float*>(b->buf[0]) + b->off[r], tmp.data(), n * sizeof(float),
                                hipMemcpyHostToDevice));
        } else {
            AMB_CHECK(hipMemcpy(static_cast<double*>(b->buf[0]) + b->off[r], fields[r], n * sizeof(double),
                                hipMemcpyHostToDevice));
        }
    }
#undef AMB_CHECK
    if ((rc = upload_descs(b)) != AMB_OK || (rc = rebuild_groups(b)) != AMB_OK) {
        release(b);
        return rc;
    }
    *out = b;
    return AMB_OK;
}

int amb_batch_destroy(amb_batch* b)
{
    clear_error();
    release(b);
    return AMB_OK;
}

int amb_batch_run(amb_batch* b, int64_t nsteps)
{
    clear_error();
    if (!b) return fail(AMB_E_ARG, "batch must not be null");
    if (nsteps < 0) return fail(AMB_E_ARG, "nsteps must be non-negative");
    int rc = enqueue(b, nsteps);
    if (rc != AMB_OK) return rc;
    double s;
    int64_t bad;
    if ((rc = finish_run(b, nsteps, &s, &bad)) != AMB_OK) return rc;
    b->last_sum = s;
    b->last_bad = bad;
    if (bad != 0) return fail(AMB_E_NONFINITE, "an unknown numerical error occurred during simulation");
    return AMB_OK;
}

int amb_batch_run_timed(amb_batch* b, int64_t nsteps, double* device_ms)
{
    clear_error();
    if (!b || !device_ms) return fail(AMB_E_ARG, "batch and device_ms must not be null");
    if (nsteps < 0) return fail(AMB_E_ARG, "nsteps must be non-negative");
    hipError_t e;
    if ((e = hipEventRecord(b->ev0, b->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    int rc = enqueue(b, nsteps);
    if (rc != AMB_OK) return rc;
    if ((e = hipEventRecord(b->ev1, b->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    if ((e = hipEventSynchronize(b->ev1)) != hipSuccess) return fail_hip(e, "hipEventSynchronize");
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess) return fail_hip(e, "hipEventElapsedTime");
    *device_ms = ms;
    double s;
    int64_t bad;
    if ((rc = finish_run(b, nsteps, &s, &bad)) != AMB_OK) return rc;
    b->last_sum = s;
    b->last_bad = bad;
    if (bad != 0) return fail(AMB_E_NONFINITE, "an unknown numerical error occurred during simulation");
    return AMB_OK;
}

int amb_batch_get_field(amb_batch* b, int64_t index, double* out)
{
    clear_error();
    if (!b || !out) return fail(AMB_E_ARG, "batch and out must not be null");
    if (index < 0 || index >= (int64_t)b->desc.size()) return fail(AMB_E_ARG, "realisation index out of range");
    const size_t n = (size_t)(b->desc[index].ny * b->desc[index].nx);
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    if (b->precision == AMB_F32) {
        std::vector<float> tmp(n);
        e = hipMemcpy(tmp.data(), static_cast<const float*>(b->buf[b->cur]) + b->off[index], n * sizeof(float),
                      hipMemcpyDeviceToHost);
        for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    } else {
        e = hipMemcpy(out, static_cast<const double*>(b->buf[b->cur]) + b->off[index], n * sizeof(double),
                      hipMemcpyDeviceToHost);
    }
    return e == hipSuccess ? AMB_OK : fail_hip(e, "hipMemcpy (field download)");
}

int amb_batch_mass(amb_batch* b, int64_t index, double* sum, int64_t* nonfinite)
{
    clear_error();
    if (!b) return fail(AMB_E_ARG, "batch must not be null");
    if (index < 0 || index >= (int64_t)b->desc.size()) return fail(AMB_E_ARG, "realisation index out of range");
    double s;
    int64_t bad;
    const int rc = reduce_range(b, b->off[index], b->desc[index].ny * b->desc[index].nx, &s, &bad);
    if (rc != AMB_OK) return rc;
    if (sum) *sum = s;
    if (nonfinite) *nonfinite = bad;
    return AMB_OK;
}

int amb_batch_set_resident(amb_batch* b, int enable)
{
    clear_error();
    if (!b) return fail(AMB_E_ARG, "batch must not be null");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    b->resident_on = enable != 0;
    return rebuild_groups(b);
}

int amb_batch_resident_count(amb_batch* b, int64_t* count)
{
    if (!b || !count) return fail(AMB_E_ARG, "batch and count must not be null");
    *count = (int64_t)b->resident.size();
    return AMB_OK;
}

int amb_batch_last_mass(amb_batch* b, double* sum, int64_t* nonfinite)
{
    clear_error();
    if (!b) return fail(AMB_E_ARG, "batch must not be null");
    if (sum) *sum = b->last_sum;
    if (nonfinite) *nonfinite = b->last_bad;
    return AMB_OK;
}

int amb_batch_set_tiles(amb_batch* b, const int* tile_per_realisation)
{
    clear_error();
    if (!b || !tile_per_realisation) return fail(AMB_E_ARG, "batch and tiles must not be null");
    const int nt = tile_count(b->precision);
    for (size_t r = 0; r < b->desc.size(); ++r)
        if (tile_per_realisation[r] < 0 || tile_per_realisation[r] >= nt) return fail(AMB_E_ARG, "tile out of range");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    b->tile_of.assign(tile_per_realisation, tile_per_realisation + b->desc.size());
    int rc = upload_descs(b);
    if (rc == AMB_OK) rc = rebuild_groups(b);
    return rc;
}

int amb_batch_get_tiles(amb_batch* b, int* tile_per_realisation)
{
    if (!b || !tile_per_realisation) return fail(AMB_E_ARG, "batch and tiles must not be null");
    std::copy(b->tile_of.begin(), b->tile_of.end(), tile_per_realisation);
    return AMB_OK;
}

// Autotuner: for every distinct (ny, nx) class, time `reps` steps of each compiled
// tile shape on that class alone (on a saved copy of the state, with a scratch
// step counter) and keep the fastest.  The field and step counter are restored.
int amb_batch_autotune(amb_batch* b, int reps, int* chosen)
{
    clear_error();
    if (!b) return fail(AMB_E_ARG, "batch must not be null");
    const int reps_in = reps;
    if (reps < 0) reps = -reps;
    if (reps < 1) reps = 4;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    // classes of equal grid and noise mode (each noise mode has its own kernel)
    std::map<std::tuple<int64_t, int64_t, int>, std::vector<int>> classes;
    std::vector<bool> is_resident(b->desc.size(), false);
    for (int r : b->resident) is_resident[r] = true;
    for (size_t r = 0; r < b->desc.size(); ++r)
        if (!is_resident[r])
            classes[{b->desc[r].ny, b->desc[r].nx, b->desc[r].model.T > 0.0 ? kNoisePhilox : kNoiseNone}].push_back(
                (int)r);
    void* saved = nullptr;
    if ((e = hipMalloc(&saved, (size_t)b->total * b->esize)) != hipSuccess) return fail_hip(e, "hipMalloc");
    e = hipMemcpy(saved, b->buf[b->cur], (size_t)b->total * b->esize, hipMemcpyDeviceToDevice);
    const long long zero = 0;
    if (e == hipSuccess) e = hipMemcpy(b->d_step_scratch, &zero, sizeof(long long), hipMemcpyHostToDevice);
    const int nt = tile_count(b->precision);
    std::vector<int> best_tile = b->tile_of;  // resident members keep theirs (unused while resident)
    const bool prefer_table = reps_in < 0;
    for (auto& kv : classes) {
        if (e != hipSuccess) break;
        int known = 0;
        if (prefer_table && tuned_tile(b->precision, std::get<0>(kv.first), std::get<1>(kv.first),
                                       std::get<2>(kv.first), &known)) {
            for (int r : kv.second) best_tile[r] = known;
            continue;
        }
        float best = 1e30f;
        int best_t = 0;
        for (int t = 0; t < nt && e == hipSuccess; ++t) {
            // Descriptors must carry ntx for this tile shape while it is timed.
            std::vector<int> trial = b->tile_of;
            for (int r : kv.second) trial[r] = t;
            std::swap(trial, b->tile_of);
            int rc = upload_descs(b);
            std::swap(trial, b->tile_of);
            if (rc != AMB_OK) {
                (void)hipFree(saved);
                return rc;
            }
            const std::vector<int2> list = tile_list(b, kv.second, t);
            int2* d_tiles = nullptr;
            e = hipMalloc(&d_tiles, list.size() * sizeof(int2));
            if (e == hipSuccess)
                e = hipMemcpy(d_tiles, list.data(), list.size() * sizeof(int2), hipMemcpyHostToDevice);
            // one warm-up step, then `reps` timed steps
            for (int k = 0; k <= reps && e == hipSuccess; ++k) {
                if (k == 1) e = hipEventRecord(b->ev0, b->stream);
                if (e == hipSuccess)
                    e = launch_batch(b->precision, t, std::get<2>(kv.first), b->d_desc, d_tiles, (int)list.size(),
                                     b->buf[(b->cur + k) & 1],
                                     b->buf[(b->cur + k + 1) & 1], b->d_step_scratch, 0, nullptr, b->stream);
            }
            if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
            if (d_tiles) (void)hipFree(d_tiles);
            if (e == hipSuccess && ms < best) {
                best = ms;
                best_t = t;
            }
        }
        for (int r : kv.second) best_tile[r] = best_t;
        if (e == hipSuccess)
            record_tile(b->precision, std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), best_t);
    }
    if (e == hipSuccess) e = hipMemcpy(b->buf[b->cur], saved, (size_t)b->total * b->esize, hipMemcpyDeviceToDevice);
    (void)hipFree(saved);
    if (e != hipSuccess) return fail_hip(e, "autotune");
    if (chosen) std::copy(best_tile.begin(), best_tile.end(), chosen);
    return amb_batch_set_tiles(b, best_tile.data());
}

}  // extern "C"
