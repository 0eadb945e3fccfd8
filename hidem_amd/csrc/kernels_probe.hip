// HBM bandwidth probes: a streaming copy and a six-array read-modify-write
// (the CT side-state access pattern), used to calibrate the roofline and the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters against kernels of known size.
#include "hdm_context.h"
#include "kernel_common.h"

namespace hdm {
namespace {

constexpr int TPB = 256;

// One pass over the buffer: each thread moves COPY_UNROLL 16 B elements,
// all loads issued before the stores, non-temporal both ways (the fastest of
// the copy forms measured on MI355X: 5.9-6.4 TB/s for 1-4 GiB, against
// 5.0-5.2 TB/s for a grid-stride loop and 5.2-5.3 TB/s for hipMemcpy).
constexpr int COPY_UNROLL = 4;
typedef float probe_f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(TPB) k_probe_copy(const probe_f4 *__restrict__ src, probe_f4 *__restrict__ dst,
                                                     int64_t n16) {
    const int64_t base = (int64_t)blockIdx.x * TPB * COPY_UNROLL + threadIdx.x;
    probe_f4 v[COPY_UNROLL];
#pragma unroll
    for (int k = 0; k < COPY_UNROLL; k++) {
        const int64_t i = base + (int64_t)k * TPB;
        if (i < n16) v[k] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int k = 0; k < COPY_UNROLL; k++) {
        const int64_t i = base + (int64_t)k * TPB;
        if (i < n16) __builtin_nontemporal_store(v[k], dst + i);
    }
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
        const int64_t n16 = bytes / 2 / 16;
        const probe_f4 *src = static_cast<const probe_f4 *>(buf);
        probe_f4 *dst = reinterpret_cast<probe_f4 *>(static_cast<char *>(buf) + n16 * 16);
        const int64_t per = (int64_t)TPB * COPY_UNROLL;
        hipLaunchKernelGGL(k_probe_copy, dim3((unsigned)((n16 + per - 1) / per)), dim3(TPB), 0, c.stream, src, dst, n16);
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
