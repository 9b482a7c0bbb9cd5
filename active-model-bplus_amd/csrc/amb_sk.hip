// amb_sk.hip -- radially binned structure factor of an integrator's field, on the device.
//
// The reference computes it on the host in examples/structureFactor.ipynb (cell 7):
// phi_k = fftshift(fft2(phi - phi0)) dx dy, |k| on the fftfreq grid scaled by 2 pi,
// indices = digitize(|k|, bin_edges, right=True), S per bin = mean of |phi_k|^2.
// Here the transform is hipFFT (complex-to-complex, FP64 for FP64 fields and FP32 for
// FP32 fields) and the binning is a kernel: per-block LDS histograms (FP64 sums,
// integer counts) written as partials and summed over blocks in a fixed order.  The
// caller gets raw sums and counts per digitize index and forms S, so the notebook's
// own index convention can be reproduced exactly.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "amb.h"
#include "amb_internal.h"

namespace amb {
namespace {

constexpr int kSkBlocks = 240;
constexpr int kSkThreads = 256;
constexpr int kMaxEdges = 1024;

template <typename T>
struct Complex;
template <>
struct Complex<float> {
    using type = hipfftComplex;
    static constexpr hipfftType kType = HIPFFT_C2C;
};
template <>
struct Complex<double> {
    using type = hipfftDoubleComplex;
    static constexpr hipfftType kType = HIPFFT_Z2Z;
};

template <typename T, typename C>
__global__ void sk_load_kernel(const T* __restrict__ phi, int64_t n, double mean, C* __restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        out[i].x = (T)((double)phi[i] - mean);
        out[i].y = (T)0;
    }
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
    __shared__ double hs[kMaxEdges + 1];
    __shared__ unsigned int hc[kMaxEdges + 1];
    for (int i = threadIdx.x; i < ne; i += kSkThreads) se[i] = edges[i];
    for (int i = threadIdx.x; i <= ne; i += kSkThreads) {
        hs[i] = 0.0;
        hc[i] = 0u;
    }
    __syncthreads();
    const int64_t n = (int64_t)ny * nx;
    for (int64_t i = blockIdx.x * (int64_t)kSkThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kSkThreads) {
        const int row = (int)(i / nx), col = (int)(i - (int64_t)row * nx);
        const double kx = kx_unit * freq_index(col, nx), ky = ky_unit * freq_index(row, ny);
        const double k = sqrt(kx * kx + ky * ky);
        // digitize(right=True): the number of edges strictly below k
        int lo = 0, hi = ne;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (se[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        const double re = f[i].x, im = f[i].y;
        atomicAdd(&hs[lo], (re * re + im * im) * scale2);
        atomicAdd(&hc[lo], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= ne; i += kSkThreads) {
        psum[(size_t)blockIdx.x * (ne + 1) + i] = hs[i];
        pcnt[(size_t)blockIdx.x * (ne + 1) + i] = hc[i];
    }
}

// One hipFFT plan per (ny, nx, type), created on first use and kept.
std::mutex g_plan_mutex;
std::map<std::tuple<int64_t, int64_t, int>, hipfftHandle> g_plans;

int get_plan(int64_t ny, int64_t nx, hipfftType type, hipfftHandle* out)
{
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    const auto key = std::make_tuple(ny, nx, (int)type);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        *out = it->second;
        return AMB_OK;
    }
    hipfftHandle plan;
    if (hipfftPlan2d(&plan, (int)ny, (int)nx, type) != HIPFFT_SUCCESS) return fail(AMB_E_HIP, "hipfftPlan2d failed");
    g_plans[key] = plan;
    *out = plan;
    return AMB_OK;
}

template <typename T>
int structure_factor(const FieldView& v, double mean, double dx, double dy, const double* edges, int ne, double* sums,
                     int64_t* counts)
{
    using C = typename Complex<T>::type;
    const int64_t n = v.ny * v.nx;
    C* buf = nullptr;
    double* d_edges = nullptr;
    double* psum = nullptr;
    unsigned long long* pcnt = nullptr;
    hipError_t e = hipMalloc(&buf, n * sizeof(C));
    if (e == hipSuccess) e = hipMalloc(&d_edges, ne * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&psum, (size_t)kSkBlocks * (ne + 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&pcnt, (size_t)kSkBlocks * (ne + 1) * sizeof(unsigned long long));
    int rc = e == hipSuccess ? AMB_OK : fail_hip(e, "structure factor buffers");
    hipfftHandle plan;
    if (rc == AMB_OK) rc = get_plan(v.ny, v.nx, Complex<T>::kType, &plan);
    if (rc == AMB_OK) {
        e = hipMemcpyAsync(d_edges, edges, ne * sizeof(double), hipMemcpyHostToDevice, v.stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL((sk_load_kernel<T, C>), dim3(1024), dim3(256), 0, v.stream,
                               static_cast<const T*>(v.cur), n, mean, buf);
            e = hipGetLastError();
        }
        rc = e == hipSuccess ? AMB_OK : fail_hip(e, "structure factor load");
    }
    if (rc == AMB_OK) {
        hipfftResult fr = hipfftSetStream(plan, v.stream);
        if (fr == HIPFFT_SUCCESS) {
            if constexpr (sizeof(T) == 8) fr = hipfftExecZ2Z(plan, buf, buf, HIPFFT_FORWARD);
            else fr = hipfftExecC2C(plan, buf, buf, HIPFFT_FORWARD);
        }
        if (fr != HIPFFT_SUCCESS) rc = fail(AMB_E_HIP, "hipFFT execution failed");
    }
    if (rc == AMB_OK) {
        const double two_pi = 6.283185307179586;
        hipLaunchKernelGGL((sk_bin_kernel<C>), dim3(kSkBlocks), dim3(kSkThreads), 0, v.stream, buf, (int)v.ny,
                           (int)v.nx, two_pi / (v.nx * dx), two_pi / (v.ny * dy), dx * dx * dy * dy, d_edges, ne,
                           psum, pcnt);
        e = hipGetLastError();
        std::vector<double> hsum((size_t)kSkBlocks * (ne + 1));
        std::vector<unsigned long long> hcnt(hsum.size());
        if (e == hipSuccess)
            e = hipMemcpyAsync(hsum.data(), psum, hsum.size() * sizeof(double), hipMemcpyDeviceToHost, v.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hcnt.data(), pcnt, hcnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               v.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
        if (e == hipSuccess) {
            for (int i = 0; i <= ne; ++i) {
                double s = 0.0;
                int64_t c = 0;
                for (int b = 0; b < kSkBlocks; ++b) {
                    s += hsum[(size_t)b * (ne + 1) + i];
                    c += (int64_t)hcnt[(size_t)b * (ne + 1) + i];
                }
                sums[i] = s;
                counts[i] = c;
            }
        }
        if (e != hipSuccess) rc = fail_hip(e, "structure factor binning");
    }
    (void)hipFree(buf);
    (void)hipFree(d_edges);
    (void)hipFree(psum);
    (void)hipFree(pcnt);
    return rc;
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
    return v.precision == AMB_F64 ? structure_factor<double>(v, mean, st.dx, st.dy, edges, nedges, sums, counts)
                                  : structure_factor<float>(v, mean, st.dx, st.dy, edges, nedges, sums, counts);
}
