// HBM bandwidth probes: a streaming copy and a six-array read-modify-write
// (the CT side-state access pattern), used to calibrate the roofline and the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters against kernels of known size.
#include "hdm_context.h"
#include "kernel_common.h"

namespace hdm {
namespace {

constexpr int TPB = 256;

__global__ void __launch_bounds__(TPB) k_probe_copy(const double4 *__restrict__ src, double4 *__restrict__ dst,
                                                     int64_t n4) {
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n4; i += (int64_t)gridDim.x * TPB) dst[i] = src[i];
}

__global__ void __launch_bounds__(TPB) k_probe_rmw6(double *__restrict__ a0, double *__restrict__ a1,
                                                     double *__restrict__ a2, double *__restrict__ a3,
                                                     double *__restrict__ a4, double *__restrict__ a5, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        a0[i] = a0[i] + 1.0;
        a1[i] = a1[i] + 1.0;
        a2[i] = a2[i] + 1.0;
        a3[i] = a3[i] + 1.0;
        a4[i] = a4[i] + 1.0;
        a5[i] = a5[i] + 1.0;
    }
}

__global__ void __launch_bounds__(TPB) k_probe_zero(double *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) p[i] = 0.0;
}

}  // namespace

hipError_t launch_probe(Context &c, int kind, void *buf, int64_t bytes) {
    int grid = (c.numCU > 0 ? c.numCU : 256) * 8;
    if (kind == 0) {
        int64_t half = bytes / 2;
        int64_t n4 = half / 32;
        const double4 *src = static_cast<const double4 *>(buf);
        double4 *dst = reinterpret_cast<double4 *>(static_cast<char *>(buf) + n4 * 32);
        hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(TPB), 0, c.stream, src, dst, n4);
    } else {
        int64_t n = bytes / 6 / 8;
        double *b = static_cast<double *>(buf);
        hipLaunchKernelGGL(k_probe_rmw6, dim3(grid), dim3(TPB), 0, c.stream, b, b + n, b + 2 * n, b + 3 * n, b + 4 * n,
                           b + 5 * n, n);
    }
    return hipGetLastError();
}

hipError_t launch_probe_zero(Context &c, void *buf, int64_t bytes) {
    int grid = (c.numCU > 0 ? c.numCU : 256) * 8;
    hipLaunchKernelGGL(k_probe_zero, dim3(grid), dim3(TPB), 0, c.stream, static_cast<double *>(buf), bytes / 8);
    return hipGetLastError();
}

}  // namespace hdm
