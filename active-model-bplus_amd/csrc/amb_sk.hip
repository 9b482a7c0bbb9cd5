// amb_sk.hip -- radially binned structure factor of an integrator's field, on the device.
//
// The reference computes it on the host in examples/structureFactor.ipynb (cell 7):
// phi_k = fftshift(fft2(phi - phi0)) dx dy, |k| on the fftfreq grid scaled by 2 pi,
// indices = digitize(|k|, bin_edges, right=True), S per bin = mean of |phi_k|^2.
//
// Here the transform is a real-to-complex hipFFT (FP64 for FP64 fields, FP32 for FP32
// fields): a real field's spectrum is Hermitian, so the half spectrum (columns 0..nx/2)
// holds every |phi_k|^2, the other columns counting twice.  Binning is one kernel in
// which every wave owns a private LDS histogram and its lanes add their modes to it one
// lane after another, so each bin's FP64 sum is formed in an order fixed by the grid
// alone; the waves' histograms are then added in wave order and the blocks' partials
// on the host in block order.  Two calls on the same field give the same bits.  The
// FFT plan and buffers belong to the integrator handle (created on first use, freed
// with the handle), so separate handles share nothing.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <new>
#include <vector>

#include "amb.h"
#include "amb_internal.h"

namespace amb {

struct SkCache {
    hipfftHandle plan = 0;
    bool has_plan = false;
    int precision = 0;
    void* real = nullptr;      // ny*nx field minus its mean
    void* spec = nullptr;      // ny*(nx/2+1) complex
    double* d_edges = nullptr; // kMaxEdges
    double* psum = nullptr;    // kSkBlocks*(kMaxEdges+1)
    unsigned long long* pcnt = nullptr;
};

void sk_cache_free(SkCache* c)
{
    if (!c) return;
    if (c->has_plan) (void)hipfftDestroy(c->plan);
    for (void* p : {c->real, c->spec, (void*)c->d_edges, (void*)c->psum, (void*)c->pcnt})
        if (p) (void)hipFree(p);
    delete c;
}

namespace {

constexpr int kSkBlocks = 240;
constexpr int kSkThreads = 256;
constexpr int kSkWaves = kSkThreads / 64;
constexpr int kMaxEdges = 1024;

template <typename T>
struct Fft;
template <>
struct Fft<float> {
    using complex = hipfftComplex;
    static constexpr hipfftType kType = HIPFFT_R2C;
    static hipfftResult exec(hipfftHandle p, void* in, void* out)
    {
        return hipfftExecR2C(p, static_cast<hipfftReal*>(in), static_cast<hipfftComplex*>(out));
    }
};
template <>
struct Fft<double> {
    using complex = hipfftDoubleComplex;
    static constexpr hipfftType kType = HIPFFT_D2Z;
    static hipfftResult exec(hipfftHandle p, void* in, void* out)
    {
        return hipfftExecD2Z(p, static_cast<hipfftDoubleReal*>(in), static_cast<hipfftDoubleComplex*>(out));
    }
};

template <typename T>
__global__ void sk_load_kernel(const T* __restrict__ phi, int64_t n, double mean, T* __restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (T)((double)phi[i] - mean);
}

// numpy.fft.fftfreq index -> signed frequency index.
__device__ __forceinline__ int freq_index(int j, int n) { return j < (n - 1) / 2 + 1 ? j : j - n; }

template <typename C>
__global__ __launch_bounds__(kSkThreads) void sk_bin_kernel(const C* __restrict__ f, int ny, int nx, double kx_unit,
                                                            double ky_unit, double scale2,
                                                            const double* __restrict__ edges, int ne,
                                                            double* __restrict__ psum,
                                                            unsigned long long* __restrict__ pcnt)
{
    __shared__ double se[kMaxEdges];
    __shared__ double hs[kSkWaves][kMaxEdges + 1];
    __shared__ unsigned int hc[kMaxEdges + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < ne; i += kSkThreads) se[i] = edges[i];
    for (int i = threadIdx.x; i <= ne; i += kSkThreads) {
        for (int w = 0; w < kSkWaves; ++w) hs[w][i] = 0.0;
        hc[i] = 0u;
    }
    __syncthreads();
    const int hx = nx / 2 + 1;  // columns of the half spectrum
    const int64_t n = (int64_t)ny * hx;
    double* const mine = hs[wave];
    // Modes in chunks of 64, chunk q to wave q of the grid in turn (the loop count is
    // the same on every lane of a wave).
    const int64_t nwaves = (int64_t)gridDim.x * kSkWaves;
    for (int64_t base = ((int64_t)blockIdx.x * kSkWaves + wave) * 64; base < n; base += nwaves * 64) {
        const int64_t i = base + lane;
        int bin = 0;
        double v = 0.0;
        if (i < n) {
            const int row = (int)(i / hx), col = (int)(i - (int64_t)row * hx);
            const double kx = kx_unit * col, ky = ky_unit * freq_index(row, ny);
            const double k = sqrt(kx * kx + ky * ky);
            // digitize(right=True): the number of edges strictly below k
            int lo = 0, hi = ne;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (se[mid] < k) lo = mid + 1;
                else hi = mid;
            }
            bin = lo;
            // column 0 and (nx even) the Nyquist column stand for themselves, the others
            // also for their mirror images nx - col
            const unsigned w = (col == 0 || 2 * col == nx) ? 1u : 2u;
            const double re = f[i].x, im = f[i].y;
            v = (re * re + im * im) * scale2 * (double)w;
            atomicAdd(&hc[bin], w);  // integer: the order does not matter
        }
        // lanes add to the wave's histogram one after another
        for (int l = 0; l < 64; ++l) {
            if (lane == l && i < n) mine[bin] += v;
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= ne; i += kSkThreads) {
        double s = 0.0;
        for (int w = 0; w < kSkWaves; ++w) s += hs[w][i];
        psum[(size_t)blockIdx.x * (ne + 1) + i] = s;
        pcnt[(size_t)blockIdx.x * (ne + 1) + i] = hc[i];
    }
}

template <typename T>
int prepare(SkCache* c, int64_t ny, int64_t nx)
{
    if (c->has_plan && c->precision == (int)sizeof(T)) return AMB_OK;
    if (c->has_plan) (void)hipfftDestroy(c->plan);
    c->has_plan = false;
    for (void** p : {&c->real, &c->spec})
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    const size_t nreal = (size_t)(ny * nx), nspec = (size_t)(ny * (nx / 2 + 1));
    hipError_t e = hipMalloc(&c->real, nreal * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&c->spec, nspec * sizeof(typename Fft<T>::complex));
    if (e == hipSuccess && !c->d_edges) e = hipMalloc(&c->d_edges, kMaxEdges * sizeof(double));
    if (e == hipSuccess && !c->psum) e = hipMalloc(&c->psum, (size_t)kSkBlocks * (kMaxEdges + 1) * sizeof(double));
    if (e == hipSuccess && !c->pcnt)
        e = hipMalloc(&c->pcnt, (size_t)kSkBlocks * (kMaxEdges + 1) * sizeof(unsigned long long));
    if (e != hipSuccess) return fail_hip(e, "structure factor buffers");
    if (hipfftPlan2d(&c->plan, (int)ny, (int)nx, Fft<T>::kType) != HIPFFT_SUCCESS)
        return fail(AMB_E_HIP, "hipfftPlan2d failed");
    c->has_plan = true;
    c->precision = (int)sizeof(T);
    return AMB_OK;
}

template <typename T>
int structure_factor(SkCache* c, const FieldView& v, double mean, double dx, double dy, const double* edges, int ne,
                     double* sums, int64_t* counts)
{
    using C = typename Fft<T>::complex;
    int rc = prepare<T>(c, v.ny, v.nx);
    if (rc != AMB_OK) return rc;
    const int64_t n = v.ny * v.nx;
    hipError_t e = hipMemcpyAsync(c->d_edges, edges, ne * sizeof(double), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL((sk_load_kernel<T>), dim3(1024), dim3(256), 0, v.stream, static_cast<const T*>(v.cur), n,
                           mean, static_cast<T*>(c->real));
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail_hip(e, "structure factor load");
    hipfftResult fr = hipfftSetStream(c->plan, v.stream);
    if (fr == HIPFFT_SUCCESS) fr = Fft<T>::exec(c->plan, c->real, c->spec);
    if (fr != HIPFFT_SUCCESS) return fail(AMB_E_HIP, "hipFFT execution failed");
    const double two_pi = 6.283185307179586;
    hipLaunchKernelGGL((sk_bin_kernel<C>), dim3(kSkBlocks), dim3(kSkThreads), 0, v.stream,
                       static_cast<const C*>(c->spec), (int)v.ny, (int)v.nx, two_pi / (v.nx * dx),
                       two_pi / (v.ny * dy), dx * dx * dy * dy, c->d_edges, ne, c->psum, c->pcnt);
    e = hipGetLastError();
    std::vector<double> hsum((size_t)kSkBlocks * (ne + 1));
    std::vector<unsigned long long> hcnt(hsum.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(hsum.data(), c->psum, hsum.size() * sizeof(double), hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hcnt.data(), c->pcnt, hcnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return fail_hip(e, "structure factor binning");
    for (int i = 0; i <= ne; ++i) {
        double s = 0.0;
        int64_t k = 0;
        for (int b = 0; b < kSkBlocks; ++b) {
            s += hsum[(size_t)b * (ne + 1) + i];
            k += (int64_t)hcnt[(size_t)b * (ne + 1) + i];
        }
        sums[i] = s;
        counts[i] = k;
    }
    return AMB_OK;
}

}  // namespace
}  // namespace amb

using namespace amb;

extern "C" int amb_structure_factor(amb_handle* h, const double* edges, int nedges, double* sums, int64_t* counts)
{
    clear_error();
    if (!h || !edges || !sums || !counts) return fail(AMB_E_ARG, "null argument");
    if (nedges < 1 || nedges > kMaxEdges) return fail(AMB_E_ARG, "nedges must be in [1, 1024]");
    for (int i = 1; i < nedges; ++i)
        if (!(edges[i] > edges[i - 1])) return fail(AMB_E_ARG, "edges must increase");
    FieldView v;
    int rc = handle_view(h, &v);
    if (rc != AMB_OK) return rc;
    if (v.ny > (1 << 15) || v.nx > (1 << 15)) return fail(AMB_E_ARG, "grid too large for the structure factor");
    amb_stencil st;
    amb_model md;
    if ((rc = amb_get_params(h, &st, &md)) != AMB_OK) return rc;
    double sum = 0.0;
    if ((rc = amb_mass(h, &sum, nullptr)) != AMB_OK) return rc;
    const double mean = sum / (double)(v.ny * v.nx);
    SkCache** slot = handle_sk_cache(h);
    if (!*slot) {
        *slot = new (std::nothrow) SkCache();
        if (!*slot) return fail(AMB_E_ARG, "out of host memory");
    }
    return v.precision == AMB_F64 ? structure_factor<double>(*slot, v, mean, st.dx, st.dy, edges, nedges, sums, counts)
                                  : structure_factor<float>(*slot, v, mean, st.dx, st.dy, edges, nedges, sums, counts);
}
