// Allocation-free form of the reference rebuild estimate (hdm_skin_heuristic,
// HiDEMLib::getMaxDistOwned, HiDEMLib.h:424-461) on Diff arrays the caller
// owns in device memory, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The per-workgroup partial maxima live in `scratch`, a file-scope device
// array that each GPU holds from the first call on it until the process
// ends.  Host code never passes the array's name to a kernel: its device
// address on the current GPU comes from hipGetSymbolAddress, taken once per
// device on the first call there.  Calls that overlap on one GPU share the
// array and race; every slot a call reads is written earlier in that call.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "hdm_context.h"
#include "kernel_common.h"

namespace {

constexpr int NA_TPB = 256;
constexpr int NA_WAVES = NA_TPB / 64;
constexpr int64_t NA_CHUNK = (int64_t)NA_TPB * 32;           // particles per first-pass workgroup
constexpr int64_t NA_MAX_PARTIALS = 16384;                   // first-pass workgroups
constexpr int64_t NA_MAX_N = NA_CHUNK * NA_MAX_PARTIALS;     // 134,217,728 particles
constexpr int NA_MAX_DEVICES = 64;
constexpr int64_t NA_RESULT = 3 * NA_MAX_PARTIALS;           // offset of the three maxima

}  // namespace

// [x partials | y partials | z partials] (NA_MAX_PARTIALS doubles each), then
// the three maxima: 393,240 B per GPU.
static __device__ __align__(256) double scratch[3 * NA_MAX_PARTIALS + 3];

namespace {

__device__ inline double block_max(double v, double *lds) {
    v = hdm::wave_max(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double m = lds[0];
#pragma unroll
    for (int k = 1; k < NA_WAVES; k++) m = lds[k] > m ? lds[k] : m;
    __syncthreads();
    return m;
}

// Pass 1: workgroup b reduces particles [b*NA_CHUNK, min(n, (b+1)*NA_CHUNK))
// to the signed maxima of the three Diff components, starting from 0.0 as
// the reference's transform_reduce does (negative motion is ignored).
__global__ void __launch_bounds__(NA_TPB) k_na_partials(const double *__restrict__ dx, const double *__restrict__ dy,
                                                        const double *__restrict__ dz, int64_t n,
                                                        double *__restrict__ part) {
    __shared__ double lds[NA_WAVES];
    const int64_t begin = (int64_t)blockIdx.x * NA_CHUNK;
    const int64_t end = begin + NA_CHUNK < n ? begin + NA_CHUNK : n;
    double m[3] = {0.0, 0.0, 0.0};
    for (int64_t i = begin + threadIdx.x; i < end; i += NA_TPB) {
        const double v[3] = {dx[i], dy[i], dz[i]};
#pragma unroll
        for (int c = 0; c < 3; c++) m[c] = v[c] > m[c] ? v[c] : m[c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double r = block_max(m[c], lds);
        if (threadIdx.x == 0) part[c * NA_MAX_PARTIALS + blockIdx.x] = r;
    }
}

// Pass 2 (one workgroup): the maxima of the nb partials of each component.
__global__ void __launch_bounds__(NA_TPB) k_na_final(const double *__restrict__ part, int64_t nb,
                                                     double *__restrict__ out) {
    __shared__ double lds[NA_WAVES];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double m = 0.0;
        for (int64_t k = threadIdx.x; k < nb; k += NA_TPB) {
            const double v = part[c * NA_MAX_PARTIALS + k];
            m = v > m ? v : m;
        }
        const double r = block_max(m, lds);
        if (threadIdx.x == 0) out[c] = r;
    }
}

std::mutex g_mu;
double *g_scratch[NA_MAX_DEVICES];   // device address of `scratch` on each GPU, once taken

int na_fail(int code, const std::string &msg) {
    hdm::set_free_error(msg);
    return code;
}

int na_hip_fail(hipError_t e, const char *where) {
    return na_fail(e == hipErrorOutOfMemory ? HDM_E_NOMEM : HDM_E_HIP,
                   std::string("hdm_skin_heuristic_noalloc: ") + where + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" int hdm_skin_heuristic_noalloc(const double *xDiff, const double *yDiff, const double *zDiff, int64_t n,
                                          int32_t stepSize, void *stream, double *dist) {
    if (!dist || n < 0 || (n > 0 && (!xDiff || !yDiff || !zDiff)))
        return na_fail(HDM_E_INVALID, "hdm_skin_heuristic_noalloc: null array or negative n");
    if (n > NA_MAX_N)
        return na_fail(HDM_E_CAPACITY, "hdm_skin_heuristic_noalloc: n = " + std::to_string(n) +
                                           " exceeds the maximum of " + std::to_string(NA_MAX_N) +
                                           " particles (its fixed device scratch holds " +
                                           std::to_string(NA_MAX_PARTIALS) + " partial maxima)");
    if (n == 0) {   // the reduction's initial value
        *dist = 0.0;
        return HDM_OK;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return na_hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= NA_MAX_DEVICES) return na_fail(HDM_E_NODEVICE, "hdm_skin_heuristic_noalloc: device index");
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_scratch[dev]) {   // first call on this GPU, which is current
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return na_hip_fail(e, "hipGetSymbolAddress(scratch)");
            g_scratch[dev] = static_cast<double *>(p);
        }
        base = g_scratch[dev];
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nb = (n + NA_CHUNK - 1) / NA_CHUNK;
    hipLaunchKernelGGL(k_na_partials, dim3((unsigned)nb), dim3(NA_TPB), 0, s, xDiff, yDiff, zDiff, n, base);
    if ((e = hipGetLastError()) != hipSuccess) return na_hip_fail(e, "k_na_partials");
    hipLaunchKernelGGL(k_na_final, dim3(1), dim3(NA_TPB), 0, s, base, nb, base + NA_RESULT);
    if ((e = hipGetLastError()) != hipSuccess) return na_hip_fail(e, "k_na_final");
    double m[3];
    if ((e = hipMemcpyAsync(m, base + NA_RESULT, sizeof(m), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return na_hip_fail(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return na_hip_fail(e, "hipStreamSynchronize");
    *dist = stepSize * std::sqrt((m[0] * m[0]) + (m[1] * m[1]) + (m[2] * m[2]));
    return HDM_OK;
}
