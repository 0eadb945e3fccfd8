// C ABI implementation: host-side orchestration of the gfx950 kernels.
// Mirrors the accelerator calls HiDEMLib makes (HiDEMLib.h:256-382, 533-628)
// on device data owned by the context; see include/hidem/hdm_abi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "hdm_context.h"

using hdm::Context;

struct hdm_ctx {
    Context c;
};

namespace {

thread_local std::string g_free_error;

int fail(hdm_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->c.err = msg;
    return code;
}

int hip_fail(hdm_ctx *ctx, hipError_t e, const char *where) {
    return fail(ctx, e == hipErrorOutOfMemory ? HDM_E_NOMEM : HDM_E_HIP,
                std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_CHECK(ctx, expr)                                      \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return hip_fail((ctx), _e, #expr);  \
    } while (0)

// Event-pair timing of one kernel class on the context stream.
struct Timer {
    Context &c;
    int cls;
    hipEvent_t start = nullptr, stop = nullptr;
    Timer(Context &cc, int k) : c(cc), cls(k) {
        if (!c.timing || c.capturing) return;
        start = take();
        stop = take();
        if (start && stop) (void)hipEventRecord(start, c.stream);
    }
    ~Timer() {
        if (!c.timing || c.capturing || !start || !stop) return;
        (void)hipEventRecord(stop, c.stream);
        c.pending.push_back(hdm::TimedLaunch{cls, start, stop});
    }
    hipEvent_t take() {
        if (!c.eventPool.empty()) {
            hipEvent_t e = c.eventPool.back();
            c.eventPool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};

int collect_timing(hdm_ctx *ctx) {
    Context &c = ctx->c;
    if (c.pending.empty()) return HDM_OK;
    HIP_CHECK(ctx, hipStreamSynchronize(c.stream));
    for (auto &t : c.pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            c.ms[t.cls] += ms;
            c.launches[t.cls] += 1;
        }
        c.eventPool.push_back(t.start);
        c.eventPool.push_back(t.stop);
    }
    c.pending.clear();
    return HDM_OK;
}

// The caller's arrays may live in host memory or in device (or managed)
// memory of this GPU, as HiDEMLib's ManVec buffers do in its GPU builds:
// plain copies let the runtime see which (hipMemcpyDefault), conversions
// between FP64 and the FP32 form's fields run on the side that holds the
// array.
bool is_device_ptr(const void *p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable host memory is not known to the runtime
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
}

template <class S, class D>
__global__ void __launch_bounds__(256) k_convert(const S *__restrict__ src, D *__restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (D)src[i];
}

hipError_t copy_in(void *dst, const void *src, size_t bytes, hipStream_t s) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s);
}

template <class T>
hipError_t upload_as(hdm::DevBuf &dst, const double *src, int64_t n, hipStream_t s) {
    hipError_t e = dst.ensure(sizeof(T) * (size_t)(n > 0 ? n : 1));
    if (e != hipSuccess || n == 0) return e;
    if (src == nullptr) return hipMemsetAsync(dst.ptr, 0, sizeof(T) * (size_t)n, s);
    if (sizeof(T) == sizeof(double)) return copy_in(dst.ptr, src, sizeof(double) * (size_t)n, s);
    if (is_device_ptr(src)) {
        hipLaunchKernelGGL((k_convert<double, T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src,
                           dst.as<T>(), n);
        return hipGetLastError();
    }
    std::vector<T> tmp((size_t)n);
    for (int64_t i = 0; i < n; i++) tmp[(size_t)i] = (T)src[i];
    e = hipMemcpyAsync(dst.ptr, tmp.data(), sizeof(T) * (size_t)n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

template <class T>
hipError_t download_as(double *dst, const hdm::DevBuf &src, int64_t n, hipStream_t s) {
    if (dst == nullptr || n == 0) return hipSuccess;
    if (sizeof(T) == sizeof(double)) {
        hipError_t e = hipMemcpyAsync(dst, src.ptr, sizeof(double) * (size_t)n, hipMemcpyDefault, s);
        return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    if (is_device_ptr(dst)) {
        hipLaunchKernelGGL((k_convert<T, double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           src.as<T>(), dst, n);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    std::vector<T> tmp((size_t)n);
    hipError_t e = hipMemcpyAsync(tmp.data(), src.ptr, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    for (int64_t i = 0; i < n; i++) dst[i] = (double)tmp[(size_t)i];
    return hipSuccess;
}

// Device buffer -> the caller's array (host or device), waited for.
hipError_t d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

bool valid(const hdm_ctx *ctx) { return ctx != nullptr; }

}  // namespace

namespace hdm {
int xcd_grid(int64_t nBlocks) { return (int)nBlocks; }
void set_free_error(const std::string &msg) { g_free_error = msg; }
}  // namespace hdm

extern "C" {

int hdm_abi_version(void) { return HDM_ABI_VERSION; }

int hdm_device_count(int32_t *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (count) *count = (e == hipSuccess) ? n : 0;
    return e == hipSuccess ? HDM_OK : HDM_E_NODEVICE;
}

int hdm_ctx_create(int32_t precision, int32_t device, hdm_ctx **out) {
    if (!out) return HDM_E_INVALID;
    *out = nullptr;
    if (precision != HDM_PREC_FP64 && precision != HDM_PREC_MIXED) return HDM_E_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return HDM_E_NODEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return HDM_E_NODEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return HDM_E_NODEVICE;
    if (hipSetDevice(device) != hipSuccess) return HDM_E_NODEVICE;
    hdm_ctx *ctx = new (std::nothrow) hdm_ctx();
    if (!ctx) return HDM_E_NOMEM;
    Context &c = ctx->c;
    c.prec = precision;
    c.device = device;
    c.numCU = prop.multiProcessorCount;
    c.coopLaunch = prop.cooperativeLaunch != 0;
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return HDM_E_HIP;
    }
    if (c.red.ensure(sizeof(uint64_t) * 64) != hipSuccess) {
        (void)hipStreamDestroy(c.stream);
        delete ctx;
        return HDM_E_NOMEM;
    }
    *out = ctx;
    return HDM_OK;
}

int hdm_ctx_destroy(hdm_ctx *ctx) {
    if (!ctx) return HDM_OK;
    Context &c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (auto &t : c.pending) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : c.eventPool) (void)hipEventDestroy(e);
    if (c.graphExec) (void)hipGraphExecDestroy(c.graphExec);
    if (c.graphExecPair) (void)hipGraphExecDestroy(c.graphExecPair);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
    return HDM_OK;
}

const char *hdm_last_error(const hdm_ctx *ctx) {
    if (ctx) return ctx->c.err.c_str();
    return g_free_error.empty() ? "null context" : g_free_error.c_str();
}

int hdm_sync(hdm_ctx *ctx) {
    if (!valid(ctx)) return HDM_E_INVALID;
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->c.stream));
    return HDM_OK;
}

int hdm_set_timing(hdm_ctx *ctx, int32_t enabled) {
    if (!valid(ctx)) return HDM_E_INVALID;
    int r = collect_timing(ctx);
    ctx->c.timing = enabled != 0;
    return r;
}

int hdm_kernel_times(hdm_ctx *ctx, double ms[HDM_T_COUNT], int64_t launches[HDM_T_COUNT]) {
    if (!valid(ctx)) return HDM_E_INVALID;
    int r = collect_timing(ctx);
    if (r != HDM_OK) return r;
    for (int k = 0; k < HDM_T_COUNT; k++) {
        if (ms) ms[k] = ctx->c.ms[k];
        if (launches) launches[k] = ctx->c.launches[k];
    }
    return HDM_OK;
}

int hdm_reset_kernel_times(hdm_ctx *ctx) {
    if (!valid(ctx)) return HDM_E_INVALID;
    int r = collect_timing(ctx);
    for (int k = 0; k < HDM_T_COUNT; k++) {
        ctx->c.ms[k] = 0;
        ctx->c.launches[k] = 0;
    }
    return r;
}

int hdm_device_bytes(hdm_ctx *ctx, int64_t *bytes) {
    if (!valid(ctx) || !bytes) return HDM_E_INVALID;
    Context &c = ctx->c;
    size_t t = 0;
    const hdm::DevBuf *bufs[] = {&c.x, &c.y, &c.z, &c.tx, &c.ty, &c.tz, &c.radius, &c.type, &c.la, &c.lb,
                                 &c.state, &c.initialLen, &c.maxLoad, &c.linkCount, &c.linkCursor, &c.slotLink,
                                 &c.slotOther, &c.sidePos, &c.nbrCount, &c.nbrLinked, &c.nbr, &c.colour, &c.order,
                                 &c.cellStart, &c.cellCursor, &c.scanPartials, &c.red, &c.x0, &c.y0, &c.z0, &c.x0f, &c.y0f, &c.z0f,
                                 &c.skinPartials, &c.blockId, &c.blockWater, &c.probe, &c.rec, &c.recAlt,
                                 &c.drec, &c.drecAlt, &c.dtail, &c.dtailAlt, &c.slotThr};
    for (auto *b : bufs) t += b->bytes;
    t += c.ct.bytes;
    for (int k = 0; k < 6; k++) t += c.diff[k].bytes + c.force[k].bytes + c.ctStage[k].bytes;
    for (int k = 0; k < 12; k++) t += c.alt[k].bytes;
    *bytes = (int64_t)t;
    return HDM_OK;
}

// ---------------------------------------------------------------------------
int hdm_particles_upload(hdm_ctx *ctx, const hdm_particles_host *p) {
    if (!valid(ctx) || !p || p->n < 0) return HDM_E_INVALID;
    Context &c = ctx->c;
    c.stateEpoch++;
    if (p->n > 0 && (!p->xCord || !p->yCord || !p->zCord || !p->radius || !p->xTheta || !p->yTheta ||
                     !p->zTheta || !p->xDiff || !p->yDiff || !p->zDiff || !p->uDiff || !p->vDiff || !p->wDiff))
        return fail(ctx, HDM_E_INVALID, "hdm_particles_upload: missing required field");
    if (p->n >= (int64_t)INT32_MAX) return fail(ctx, HDM_E_CAPACITY, "more than int32 particles");
    if (c.L > 0 && p->n != c.n) return fail(ctx, HDM_E_STATE, "particle count changed while links are bound");
    (void)hipSetDevice(c.device);
    if (p->n != c.n) c.nBlocks = 0;
    c.n = p->n;
    hipStream_t s = c.stream;
    const double *dsrc[6] = {p->xDiff, p->yDiff, p->zDiff, p->uDiff, p->vDiff, p->wDiff};
    const double *fsrc[6] = {p->xForce, p->yForce, p->zForce, p->uForce, p->vForce, p->wForce};
    HIP_CHECK(ctx, upload_as<double>(c.x, p->xCord, c.n, s));
    HIP_CHECK(ctx, upload_as<double>(c.y, p->yCord, c.n, s));
    HIP_CHECK(ctx, upload_as<double>(c.z, p->zCord, c.n, s));
    HIP_CHECK(ctx, upload_as<double>(c.tx, p->xTheta, c.n, s));
    HIP_CHECK(ctx, upload_as<double>(c.ty, p->yTheta, c.n, s));
    HIP_CHECK(ctx, upload_as<double>(c.tz, p->zTheta, c.n, s));
    HIP_CHECK(ctx, upload_as<double>(c.radius, p->radius, c.n, s));
    for (int k = 0; k < 6; k++) HIP_CHECK(ctx, upload_as<double>(c.diff[k], dsrc[k], c.n, s));
    for (int k = 0; k < 6; k++) {
        if (c.prec == HDM_PREC_FP64) HIP_CHECK(ctx, upload_as<double>(c.force[k], fsrc[k], c.n, s));
        else HIP_CHECK(ctx, upload_as<float>(c.force[k], fsrc[k], c.n, s));
    }
    HIP_CHECK(ctx, c.type.ensure((size_t)(c.n > 0 ? c.n : 1)));
    if (c.n > 0) {
        if (p->type) HIP_CHECK(ctx, copy_in(c.type.ptr, p->type, (size_t)c.n, s));
        else HIP_CHECK(ctx, hipMemsetAsync(c.type.ptr, 0, (size_t)c.n, s));
    }
    for (hdm::DevBuf *b : {&c.x0, &c.y0, &c.z0}) HIP_CHECK(ctx, b->ensure(sizeof(double) * (size_t)(c.n > 0 ? c.n : 1)));
    for (hdm::DevBuf *b : {&c.x0f, &c.y0f, &c.z0f}) HIP_CHECK(ctx, b->ensure(sizeof(float) * (size_t)(c.n > 0 ? c.n : 1)));
    HIP_CHECK(ctx, hipStreamSynchronize(s));
    c.grid.built = false;
    c.neighBuilt = false;
    c.snapshotValid = false;
    c.skinFused = false;
    c.recCurrent = false;
    c.soaStale = false;
    c.altHoldsPrev = false;
    return HDM_OK;
}

int hdm_particles_download(hdm_ctx *ctx, hdm_particles_host *p) {
    if (!valid(ctx) || !p) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (p->n != c.n) return fail(ctx, HDM_E_INVALID, "hdm_particles_download: n mismatch");
    hipStream_t s = c.stream;
    HIP_CHECK(ctx, hdm::sync_state(c));
    HIP_CHECK(ctx, download_as<double>(p->xCord, c.x, c.n, s));
    HIP_CHECK(ctx, download_as<double>(p->yCord, c.y, c.n, s));
    HIP_CHECK(ctx, download_as<double>(p->zCord, c.z, c.n, s));
    HIP_CHECK(ctx, download_as<double>(p->xTheta, c.tx, c.n, s));
    HIP_CHECK(ctx, download_as<double>(p->yTheta, c.ty, c.n, s));
    HIP_CHECK(ctx, download_as<double>(p->zTheta, c.tz, c.n, s));
    HIP_CHECK(ctx, download_as<double>(p->radius, c.radius, c.n, s));
    double *dd[6] = {p->xDiff, p->yDiff, p->zDiff, p->uDiff, p->vDiff, p->wDiff};
    double *fd[6] = {p->xForce, p->yForce, p->zForce, p->uForce, p->vForce, p->wForce};
    for (int k = 0; k < 6; k++) HIP_CHECK(ctx, download_as<double>(dd[k], c.diff[k], c.n, s));
    for (int k = 0; k < 6; k++) {
        if (c.prec == HDM_PREC_FP64) HIP_CHECK(ctx, download_as<double>(fd[k], c.force[k], c.n, s));
        else HIP_CHECK(ctx, download_as<float>(fd[k], c.force[k], c.n, s));
    }
    if (p->type && c.n > 0) HIP_CHECK(ctx, d2h(p->type, c.type.ptr, (size_t)c.n, s));
    return HDM_OK;
}

int hdm_links_upload(hdm_ctx *ctx, const hdm_links_host *l) {
    if (!valid(ctx) || !l || l->n < 0) return HDM_E_INVALID;
    Context &c = ctx->c;
    c.stateEpoch++;
    c.brokenCountValid = false;
    if (l->n > 0 && (!l->localPartIndA || !l->localPartIndB || !l->state || !l->initialLen))
        return fail(ctx, HDM_E_INVALID, "hdm_links_upload: missing required field");
    if (2 * l->n >= (int64_t)INT32_MAX) return fail(ctx, HDM_E_CAPACITY, "more than int32 link sides");
    hipStream_t s = c.stream;
    c.L = l->n;
    size_t L = (size_t)(c.L > 0 ? c.L : 1);
    HIP_CHECK(ctx, c.brokenCount.ensure(sizeof(uint64_t)));
    HIP_CHECK(ctx, c.la.ensure(sizeof(int32_t) * L));
    HIP_CHECK(ctx, c.lb.ensure(sizeof(int32_t) * L));
    HIP_CHECK(ctx, c.state.ensure(L));
    if (c.L > 0) {
        HIP_CHECK(ctx, copy_in(c.la.ptr, l->localPartIndA, sizeof(int32_t) * c.L, s));
        HIP_CHECK(ctx, copy_in(c.lb.ptr, l->localPartIndB, sizeof(int32_t) * c.L, s));
        HIP_CHECK(ctx, copy_in(c.state.ptr, l->state, (size_t)c.L, s));
    }
    HIP_CHECK(ctx, upload_as<double>(c.initialLen, l->initialLen, c.L, s));
    c.hasMaxLoad = l->maxLoad != nullptr;
    if (c.hasMaxLoad) HIP_CHECK(ctx, upload_as<double>(c.maxLoad, l->maxLoad, c.L, s));
    for (int k = 0; k < 6; k++) HIP_CHECK(ctx, upload_as<double>(c.ctStage[k], l->CT[k], 2 * c.L, s));
    HIP_CHECK(ctx, hipStreamSynchronize(s));
    c.linksBuilt = false;
    return HDM_OK;
}

int hdm_links_download(hdm_ctx *ctx, hdm_links_host *l) {
    if (!valid(ctx) || !l) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (l->n != c.L) return fail(ctx, HDM_E_INVALID, "hdm_links_download: n mismatch");
    hipStream_t s = c.stream;
    if (c.L == 0) return HDM_OK;
    if (l->localPartIndA) HIP_CHECK(ctx, d2h(l->localPartIndA, c.la.ptr, sizeof(int32_t) * c.L, s));
    if (l->localPartIndB) HIP_CHECK(ctx, d2h(l->localPartIndB, c.lb.ptr, sizeof(int32_t) * c.L, s));
    if (l->state) HIP_CHECK(ctx, d2h(l->state, c.state.ptr, (size_t)c.L, s));
    HIP_CHECK(ctx, download_as<double>(l->initialLen, c.initialLen, c.L, s));
    if (l->maxLoad && c.hasMaxLoad) HIP_CHECK(ctx, download_as<double>(l->maxLoad, c.maxLoad, c.L, s));
    bool want = false;
    for (int k = 0; k < 6; k++) want |= l->CT[k] != nullptr;
    if (want) {
        if (c.linksBuilt) HIP_CHECK(ctx, hdm::launch_ct_export(c));
        for (int k = 0; k < 6; k++) HIP_CHECK(ctx, download_as<double>(l->CT[k], c.ctStage[k], 2 * c.L, s));
    }
    return HDM_OK;
}

// ---------------------------------------------------------------------------
int hdm_grid_build(hdm_ctx *ctx, double searchDistance) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!(searchDistance > 0)) return fail(ctx, HDM_E_INVALID, "searchDistance must be > 0");
    if (c.n <= 1) return fail(ctx, HDM_E_STATE, "Generating grid for empty data (ParticleData::getLimits)");
    Timer t(c, HDM_T_GRID);
    HIP_CHECK(ctx, hdm::launch_bbox(c));
    uint64_t keys[6];
    HIP_CHECK(ctx, d2h(keys, c.red.ptr, sizeof(keys), c.stream));
    hdm::GridState &g = c.grid;
    g.xMin = hdm::ordered_value(keys[0]);
    g.yMin = hdm::ordered_value(keys[1]);
    g.zMin = hdm::ordered_value(keys[2]);
    g.xMax = hdm::ordered_value(keys[3]);
    g.yMax = hdm::ordered_value(keys[4]);
    g.zMax = hdm::ordered_value(keys[5]);
    g.spacing = searchDistance;
    double ex = std::floor((g.xMax - g.xMin) / searchDistance);
    double ey = std::floor((g.yMax - g.yMin) / searchDistance);
    double ez = std::floor((g.zMax - g.zMin) / searchDistance);
    if (!(ex < 1e9 && ey < 1e9 && ez < 1e9)) return fail(ctx, HDM_E_CAPACITY, "grid extent not finite or too large");
    g.X = 1 + (int)ex;
    g.Y = 1 + (int)ey;
    g.Z = 1 + (int)ez;
    double cells = (double)g.X * (double)g.Y * (double)g.Z;
    const double maxCells = (double)(1ll << 30);
    if (cells > maxCells)
        return fail(ctx, HDM_E_CAPACITY, "grid has " + std::to_string(cells) +
                                             " cells (limit 2^30): particles spread too far, did something explode?");
    g.totalCells = (int64_t)g.X * g.Y * g.Z;
    HIP_CHECK(ctx, c.colour.ensure(sizeof(int32_t) * (size_t)c.n));
    HIP_CHECK(ctx, c.order.ensure(sizeof(int32_t) * (size_t)c.n));
    HIP_CHECK(ctx, c.cellStart.ensure(sizeof(int32_t) * (size_t)(g.totalCells + 1)));
    HIP_CHECK(ctx, c.cellCursor.ensure(sizeof(int32_t) * (size_t)(g.totalCells + 1)));
    HIP_CHECK(ctx, hdm::launch_colour_count(c));
    HIP_CHECK(ctx, hdm::launch_exclusive_scan_i32(c, c.cellCursor.as<int32_t>(), c.cellStart.as<int32_t>(), g.totalCells));
    HIP_CHECK(ctx, hdm::launch_scatter_sort(c));
    uint64_t st[3];
    HIP_CHECK(ctx, d2h(st, c.red.as<uint64_t>() + 8, sizeof(st), c.stream));
    if (st[0]) return fail(ctx, HDM_E_HIP, "grid colour out of range");
    g.occupied = (int64_t)st[1];
    g.maxPerCell = (int32_t)st[2];
    g.built = true;
    c.neighBuilt = false;
    return HDM_OK;
}

int hdm_grid_info_get(hdm_ctx *ctx, hdm_grid_info *info) {
    if (!valid(ctx) || !info) return HDM_E_INVALID;
    const hdm::GridState &g = ctx->c.grid;
    if (!g.built) return fail(ctx, HDM_E_STATE, "grid not built");
    info->spacing = g.spacing;
    info->xMin = g.xMin; info->yMin = g.yMin; info->zMin = g.zMin;
    info->xMax = g.xMax; info->yMax = g.yMax; info->zMax = g.zMax;
    info->X = g.X; info->Y = g.Y; info->Z = g.Z;
    info->totalCells = g.totalCells;
    info->occupiedCells = g.occupied;
    info->maxPerCell = g.maxPerCell;
    return HDM_OK;
}

int hdm_grid_download(hdm_ctx *ctx, int64_t *colours, int32_t *iterationOrder) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.grid.built) return fail(ctx, HDM_E_STATE, "grid not built");
    if (colours) {
        std::vector<int32_t> tmp((size_t)c.n);
        HIP_CHECK(ctx, d2h(tmp.data(), c.colour.ptr, sizeof(int32_t) * c.n, c.stream));
        for (int64_t i = 0; i < c.n; i++) colours[i] = tmp[(size_t)i];
    }
    if (iterationOrder) HIP_CHECK(ctx, d2h(iterationOrder, c.order.ptr, sizeof(int32_t) * c.n, c.stream));
    return HDM_OK;
}

int hdm_neighbours_build(hdm_ctx *ctx, double interactDist) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.grid.built) return fail(ctx, HDM_E_STATE, "grid not built");
    Timer t(c, HDM_T_NEIGH);
    HIP_CHECK(ctx, c.nbrCount.ensure(sizeof(int32_t) * (size_t)c.n));
    HIP_CHECK(ctx, c.nbrLinked.ensure(sizeof(int32_t) * (size_t)c.n));
    // the cell records must hold the current positions
    if (c.grid.recEpoch != c.stateEpoch || c.grid.recTs != c.ts) HIP_CHECK(ctx, hdm::launch_cell_records(c));
    uint64_t *mx = c.red.as<uint64_t>() + 12;
    // One fill pass into the previous build's width; a count pass first
    // only when there is none, and a second fill when some count exceeds it.
    const bool fillFirst = c.maxNeigh > 0 && c.nbr.bytes >= sizeof(int32_t) * (size_t)c.tiles() * 64 * (size_t)c.maxNeigh;
    if (fillFirst) HIP_CHECK(ctx, hdm::launch_neigh_fill(c, interactDist));
    else HIP_CHECK(ctx, hdm::launch_neigh_count(c, interactDist));
    HIP_CHECK(ctx, hdm::launch_max_i32(c, c.nbrCount.as<int32_t>(), c.n, mx));
    uint64_t m = 0;
    HIP_CHECK(ctx, d2h(&m, mx, sizeof(m), c.stream));
    c.maxNeighCount = (int32_t)m;
    if (!fillFirst || (int64_t)m > c.maxNeigh) {
        c.maxNeigh = (int32_t)(m > 0 ? m : 1);
        HIP_CHECK(ctx, c.nbr.ensure(sizeof(int32_t) * (size_t)c.tiles() * 64 * (size_t)c.maxNeigh));
        HIP_CHECK(ctx, hdm::launch_neigh_fill(c, interactDist));
    }
    c.neighBuilt = true;
    c.nbrExcludesLinked = hdm::merge_linked_contacts(c);
    return HDM_OK;
}

int hdm_neighbours_total(hdm_ctx *ctx, int64_t *total, int32_t *maxPerParticle) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.neighBuilt) return fail(ctx, HDM_E_STATE, "neighbour list not built");
    std::vector<int32_t> cnt((size_t)c.n);
    HIP_CHECK(ctx, d2h(cnt.data(), c.nbrCount.ptr, sizeof(int32_t) * c.n, c.stream));
    int64_t t = 0;
    for (auto v : cnt) t += v;
    if (total) *total = t;
    if (maxPerParticle) *maxPerParticle = c.maxNeighCount;
    return HDM_OK;
}

int hdm_neighbours_download(hdm_ctx *ctx, int32_t *counts, int64_t *starts, int32_t *list, int64_t listCap) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.neighBuilt) return fail(ctx, HDM_E_STATE, "neighbour list not built");
    std::vector<int32_t> cnt((size_t)c.n);
    HIP_CHECK(ctx, d2h(cnt.data(), c.nbrCount.ptr, sizeof(int32_t) * c.n, c.stream));
    int64_t total = 0;
    for (auto v : cnt) total += v;
    if (counts) std::memcpy(counts, cnt.data(), sizeof(int32_t) * c.n);
    if (starts) {
        starts[0] = 0;
        for (int64_t i = 0; i < c.n; i++) starts[i + 1] = starts[i] + cnt[(size_t)i];
    }
    if (list) {
        if (listCap < total) return fail(ctx, HDM_E_CAPACITY, "neighbour list buffer too small");
        size_t w = (size_t)c.maxNeigh;
        std::vector<int32_t> ell((size_t)c.tiles() * 64 * w);
        HIP_CHECK(ctx, d2h(ell.data(), c.nbr.ptr, sizeof(int32_t) * ell.size(), c.stream));
        int64_t pos = 0;
        for (int64_t p = 0; p < c.n; p++)
            for (int32_t k = 0; k < cnt[(size_t)p]; k++)
                list[pos++] = ell[(size_t)(p >> 6) * 64 * w + (size_t)k * 64 + (size_t)(p & 63)];
    }
    return HDM_OK;
}

int hdm_links_build(hdm_ctx *ctx) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (c.n == 0) return fail(ctx, HDM_E_STATE, "no particles");
    Timer t(c, HDM_T_LINKS);
    // Side data already in wave tiles goes back to checkpoint order first, so
    // a rebuild is a pure re-permutation (LinkDataHandler::reorderSideDataPadded).
    if (c.linksBuilt) HIP_CHECK(ctx, hdm::launch_ct_export(c));
    HIP_CHECK(ctx, c.linkCount.ensure(sizeof(uint32_t) * (size_t)c.n));
    HIP_CHECK(ctx, c.linkCursor.ensure(sizeof(uint32_t) * (size_t)c.n));
    HIP_CHECK(ctx, hdm::launch_link_count(c));
    uint64_t *mx = c.red.as<uint64_t>() + 12;
    HIP_CHECK(ctx, hdm::launch_max_i32(c, c.linkCount.as<int32_t>(), c.n, mx));
    uint64_t m[1], errf = 0;
    HIP_CHECK(ctx, d2h(m, mx, sizeof(m), c.stream));
    HIP_CHECK(ctx, d2h(&errf, c.red.as<uint64_t>() + 8, sizeof(errf), c.stream));
    if (errf) return fail(ctx, HDM_E_INVALID, "Dangling links found, connecting links failed (LinkData.h:105)");
    c.maxLinks = (int32_t)(m[0] > 0 ? m[0] : 1);
    size_t slots = (size_t)c.tiles() * 64 * (size_t)c.maxLinks;
    HIP_CHECK(ctx, c.slotLink.ensure(sizeof(int32_t) * slots));
    HIP_CHECK(ctx, c.slotOther.ensure(sizeof(int32_t) * slots));
    HIP_CHECK(ctx, c.sidePos.ensure(sizeof(int32_t) * (size_t)(2 * c.L > 0 ? 2 * c.L : 1)));
    HIP_CHECK(ctx, c.ct.ensure(6 * c.ctBytes() * slots));
    HIP_CHECK(ctx, c.slotThr.ensure(sizeof(uint16_t) * slots));
    HIP_CHECK(ctx, hdm::launch_link_fill(c));
    HIP_CHECK(ctx, hdm::launch_ct_import(c));
    uint64_t *words = c.red.as<uint64_t>() + 48;  // [0] partner check flag, [1..2] largest L0 and threshold
    HIP_CHECK(ctx, hdm::launch_partner_check(c, words));
    HIP_CHECK(ctx, hdm::launch_side_thresholds(c, words + 1, &c.l0Max, &c.thrScale));
    uint64_t w[1] = {0};
    HIP_CHECK(ctx, d2h(w, words, sizeof(w), c.stream));
    c.partnersUnique = w[0] == 0;
    c.linksBuilt = true;
    if (c.nbrExcludesLinked) c.neighBuilt = false;  // the partner set may have changed
    return HDM_OK;
}

int hdm_linklists_download(hdm_ctx *ctx, int32_t *counts, int64_t *starts, int32_t *linkId, int32_t *other,
                           int32_t *end, int64_t cap) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.linksBuilt) return fail(ctx, HDM_E_STATE, "link lists not built");
    std::vector<int32_t> cnt((size_t)c.n);
    HIP_CHECK(ctx, d2h(cnt.data(), c.linkCount.ptr, sizeof(int32_t) * c.n, c.stream));
    int64_t total = 0;
    for (auto v : cnt) total += v;
    if (counts) std::memcpy(counts, cnt.data(), sizeof(int32_t) * c.n);
    if (starts) {
        starts[0] = 0;
        for (int64_t i = 0; i < c.n; i++) starts[i + 1] = starts[i] + cnt[(size_t)i];
    }
    if (linkId || other || end) {
        if (cap < total) return fail(ctx, HDM_E_CAPACITY, "link list buffer too small");
        size_t w = (size_t)c.maxLinks;
        size_t slots = (size_t)c.tiles() * 64 * w;
        std::vector<int32_t> sl(slots), so(slots), A((size_t)c.L), B((size_t)c.L);
        HIP_CHECK(ctx, d2h(sl.data(), c.slotLink.ptr, sizeof(int32_t) * slots, c.stream));
        HIP_CHECK(ctx, d2h(so.data(), c.slotOther.ptr, sizeof(int32_t) * slots, c.stream));
        if (c.L) HIP_CHECK(ctx, d2h(B.data(), c.lb.ptr, sizeof(int32_t) * c.L, c.stream));
        std::vector<int32_t> sp((size_t)(2 * c.L));
        if (c.L) HIP_CHECK(ctx, d2h(sp.data(), c.sidePos.ptr, sizeof(int32_t) * 2 * c.L, c.stream));
        int64_t pos = 0;
        for (int64_t p = 0; p < c.n; p++)
            for (int32_t k = 0; k < cnt[(size_t)p]; k++) {
                size_t s = (size_t)(p >> 6) * 64 * w + (size_t)k * 64 + (size_t)(p & 63);
                int32_t l = sl[s] & hdm::SLOT_LINK_MASK;
                if (linkId) linkId[pos] = l;
                if (other) other[pos] = so[s] & hdm::SLOT_LINK_MASK;
                // the end is whichever of the link's two sides maps to this slot
                if (end) end[pos] = (sp[(size_t)(2 * l + 1)] == (int32_t)s) ? 1 : 0;
                pos++;
            }
    }
    return HDM_OK;
}

// ---------------------------------------------------------------------------
int hdm_force(hdm_ctx *ctx, const hdm_force_params *fp) {
    if (!valid(ctx) || !fp) return HDM_E_INVALID;
    Context &c = ctx->c;
    c.stateEpoch++;
    if (!c.neighBuilt || !c.linksBuilt) return fail(ctx, HDM_E_STATE, "neighbour list / link lists not built");
    Timer t(c, HDM_T_FORCE);
    HIP_CHECK(ctx, hdm::launch_force(c, *fp));
    return HDM_OK;
}

int hdm_integrate(hdm_ctx *ctx, const hdm_integ_params *ip) {
    if (!valid(ctx) || !ip) return HDM_E_INVALID;
    if (ip->variant != HDM_INTEG_ICEBLOCK && ip->variant != HDM_INTEG_GLACIER)
        return fail(ctx, HDM_E_INVALID, "unknown integrator variant");
    ctx->c.stateEpoch++;
    Timer t(ctx->c, HDM_T_INTEGRATE);
    HIP_CHECK(ctx, hdm::launch_integrate(ctx->c, *ip));
    ctx->c.skinFused = false;
    ctx->c.recCurrent = false;
    ctx->c.altHoldsPrev = false;
    return HDM_OK;
}

int hdm_break(hdm_ctx *ctx, const hdm_break_params *bp, int64_t *nBroken) {
    if (!valid(ctx) || !bp) return HDM_E_INVALID;
    Context &c = ctx->c;
    c.stateEpoch++;
    c.brokenCountValid = false;   // k_break counts this call's breaks only
    if (!c.linksBuilt) return fail(ctx, HDM_E_STATE, "link lists not built (link end indices unchecked)");
    uint64_t *counter = c.red.as<uint64_t>() + 16;
    if (nBroken) HIP_CHECK(ctx, hdm::launch_fill_u64(c.stream, counter, 0ull, 1));
    {
        Timer t(c, HDM_T_BREAK);
        HIP_CHECK(ctx, hdm::launch_break(c, *bp, nBroken ? counter : nullptr));
    }
    if (nBroken) {
        uint64_t v = 0;
        HIP_CHECK(ctx, d2h(&v, counter, sizeof(v), c.stream));
        *nBroken = (int64_t)v;
    }
    return HDM_OK;
}

int hdm_skin_heuristic(hdm_ctx *ctx, int32_t stepSize, double *dist) {
    if (!valid(ctx) || !dist) return HDM_E_INVALID;
    Context &c = ctx->c;
    uint64_t *out = c.red.as<uint64_t>() + 20;
    {
        Timer t(c, HDM_T_SKIN);
        HIP_CHECK(ctx, hdm::launch_skin_heuristic(c, out));
    }
    uint64_t k[3];
    HIP_CHECK(ctx, d2h(k, out, sizeof(k), c.stream));
    double xm = hdm::ordered_value(k[0]), ym = hdm::ordered_value(k[1]), zm = hdm::ordered_value(k[2]);
    *dist = stepSize * std::sqrt((xm * xm) + (ym * ym) + (zm * zm));
    return HDM_OK;
}

int hdm_skin_snapshot(hdm_ctx *ctx) {
    if (!valid(ctx)) return HDM_E_INVALID;
    HIP_CHECK(ctx, hdm::launch_snapshot(ctx->c));
    ctx->c.snapshotValid = true;
    ctx->c.skinFused = false;
    return HDM_OK;
}

int hdm_skin_exact(hdm_ctx *ctx, double *maxDisp) {
    if (!valid(ctx) || !maxDisp) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.snapshotValid) return fail(ctx, HDM_E_STATE, "no position snapshot");
    uint64_t *out = c.red.as<uint64_t>() + 24;
    {
        Timer t(c, HDM_T_SKIN);
        HIP_CHECK(ctx, hdm::launch_skin_exact(c, out));
    }
    uint64_t k = 0;
    HIP_CHECK(ctx, d2h(&k, out, sizeof(k), c.stream));
    *maxDisp = std::sqrt(hdm::ordered_value(k));
    return HDM_OK;
}

// The count every link-breaking kernel keeps (Context::brokenCount); a full
// count of the states refreshes it after anything else changed them.
int hdm_count_broken(hdm_ctx *ctx, int64_t *broken) {
    if (!valid(ctx) || !broken) return HDM_E_INVALID;
    Context &c = ctx->c;
    HIP_CHECK(ctx, c.brokenCount.ensure(sizeof(uint64_t)));
    uint64_t *counter = c.brokenCount.as<uint64_t>();
    if (!c.brokenCountValid) {
        HIP_CHECK(ctx, hdm::launch_count_broken(c, counter));
        c.brokenCountValid = true;
    }
    uint64_t v = 0;
    HIP_CHECK(ctx, d2h(&v, counter, sizeof(v), c.stream));
    *broken = (int64_t)v;
    return HDM_OK;
}

int hdm_particle_stats(hdm_ctx *ctx, int32_t *activeLinks, int32_t *iteratedNeighbours, int64_t *activeLinksSum,
                       int64_t *iteratedNeighboursSum) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.neighBuilt || !c.linksBuilt) return fail(ctx, HDM_E_STATE, "neighbour list / link lists not built");
    const bool perParticle = activeLinks || iteratedNeighbours;
    if (perParticle) HIP_CHECK(ctx, c.statsBuf.ensure(sizeof(int32_t) * 2 * (size_t)(c.n > 0 ? c.n : 1)));
    int32_t *dAct = perParticle ? c.statsBuf.as<int32_t>() : nullptr;
    int32_t *dNbr = perParticle ? c.statsBuf.as<int32_t>() + c.n : nullptr;
    uint64_t *sums = c.red.as<uint64_t>() + 56;
    HIP_CHECK(ctx, hdm::launch_particle_stats(c, dAct, dNbr, sums));
    uint64_t v[2] = {0, 0};
    HIP_CHECK(ctx, d2h(v, sums, sizeof(v), c.stream));
    if (activeLinks) HIP_CHECK(ctx, d2h(activeLinks, dAct, sizeof(int32_t) * (size_t)c.n, c.stream));
    if (iteratedNeighbours) HIP_CHECK(ctx, d2h(iteratedNeighbours, dNbr, sizeof(int32_t) * (size_t)c.n, c.stream));
    if (activeLinksSum) *activeLinksSum = (int64_t)v[0];
    if (iteratedNeighboursSum) *iteratedNeighboursSum = (int64_t)v[1];
    return HDM_OK;
}

int hdm_set_blocks(hdm_ctx *ctx, const int32_t *blockId, int32_t nBlocks, const double *waterLevel) {
    if (!valid(ctx) || nBlocks < 0) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (nBlocks == 0) {
        c.nBlocks = 0;
        return HDM_OK;
    }
    if (!blockId || !waterLevel || c.n == 0) return fail(ctx, HDM_E_INVALID, "hdm_set_blocks: missing arrays or particles");
    for (int64_t i = 0; i < c.n; i++)
        if (blockId[i] < 0 || blockId[i] >= nBlocks) return fail(ctx, HDM_E_INVALID, "hdm_set_blocks: block id out of range");
    HIP_CHECK(ctx, c.blockId.ensure(sizeof(int32_t) * (size_t)c.n));
    HIP_CHECK(ctx, c.blockWater.ensure(sizeof(double) * (size_t)nBlocks));
    HIP_CHECK(ctx, hipMemcpyAsync(c.blockId.ptr, blockId, sizeof(int32_t) * c.n, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(ctx, hipMemcpyAsync(c.blockWater.ptr, waterLevel, sizeof(double) * nBlocks, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(ctx, hipStreamSynchronize(c.stream));
    c.nBlocks = nBlocks;
    return HDM_OK;
}

int hdm_reorder(hdm_ctx *ctx, double searchDistance, int64_t *order, int64_t *linkOrder) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    c.stateEpoch++;
    if (!(searchDistance > 0)) return fail(ctx, HDM_E_INVALID, "hdm_reorder: searchDistance must be > 0");
    if (c.n <= 1) return fail(ctx, HDM_E_STATE, "hdm_reorder: fewer than two particles");
    // link endpoints are checked by the link build; the reorder indexes with them
    if (c.L > 0 && !c.linksBuilt) {
        int r = hdm_links_build(ctx);
        if (r != HDM_OK) return r;
    }
    const bool relink = c.linksBuilt;
    // CT back to checkpoint order first; the reorder permutes that copy
    if (c.linksBuilt) HIP_CHECK(ctx, hdm::launch_ct_export(c));
    hipError_t e = hdm::launch_reorder(c, searchDistance, order, linkOrder);
    if (e == hipErrorInvalidValue) return fail(ctx, HDM_E_CAPACITY, "hdm_reorder: grid extent not finite or too large");
    HIP_CHECK(ctx, e);
    c.linksBuilt = false;
    c.grid.built = false;
    c.neighBuilt = false;
    c.snapshotValid = false;
    c.skinFused = false;
    c.skinFusedDisp = false;
    c.recCurrent = false;
    c.soaStale = false;
    c.altHoldsPrev = false;
    c.schedAssigned = false;
    for (auto &lab : c.labels) lab.active = false;
    if (relink) return hdm_links_build(ctx);
    return HDM_OK;
}

int hdm_set_maps(hdm_ctx *ctx, const hdm_map2d *const *maps) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    hdm::MapDev desc[HDM_MAP_COUNT] = {};
    int32_t mask = 0;
    for (int k = 0; maps && k < HDM_MAP_COUNT; k++) {
        const hdm_map2d *m = maps[k];
        if (!m) continue;
        if (m->xPoints < 1 || m->yPoints < 1 || !m->data || (int64_t)m->xPoints * m->yPoints > (1ll << 30))
            return fail(ctx, HDM_E_INVALID, "hdm_set_maps: a map needs at least one point per axis and its data");
        const size_t n = (size_t)m->xPoints * (size_t)m->yPoints;
        HIP_CHECK(ctx, c.mapData[k].ensure(sizeof(double) * n));
        HIP_CHECK(ctx, hipMemcpyAsync(c.mapData[k].ptr, m->data, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        hdm::MapDev &d = desc[k];
        d.xPoints = m->xPoints;
        d.yPoints = m->yPoints;
        d.xMin = m->xMin;
        d.yMin = m->yMin;
        // Map2d constructor (Map2d.h:56-59), float arithmetic
        d.xRes = (m->xMax - m->xMin) / (float)(m->xPoints - 1);
        d.yRes = (m->yMax - m->yMin) / (float)(m->yPoints - 1);
        d.invXMaxMin = 1.0f / (m->xMax - m->xMin);
        d.invYMaxMin = 1.0f / (m->yMax - m->yMin);
        d.xOffset = m->xOffset;
        d.yOffset = m->yOffset;
        d.data = c.mapData[k].as<double>();
        mask |= 1 << k;
    }
    HIP_CHECK(ctx, c.mapDesc.ensure(sizeof(desc)));
    HIP_CHECK(ctx, hipMemcpyAsync(c.mapDesc.ptr, desc, sizeof(desc), hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(ctx, hipStreamSynchronize(c.stream));
    c.mapMask = mask;
    return HDM_OK;
}

int hdm_map2d_load(const char *path, int32_t fieldID, hdm_map2d *map) {
    if (!path || !map || fieldID < 0) return HDM_E_INVALID;
    *map = hdm_map2d{};
    std::ifstream f(path);
    if (!f) return HDM_E_INVALID;
    // Map2dLoader::readInput (Map2dLoader.cpp:8-86): whitespace-separated
    // floats, every row with the first row's column count; the extents from
    // the x and y columns; a grid row per distinct y, all of equal length
    std::vector<std::vector<float>> cols;
    std::vector<float> row;
    std::vector<std::pair<float, int64_t>> rowCounts;  // distinct y values in first-seen order
    std::string line;
    bool ok = true;
    while (ok && std::getline(f, line)) {
        row.clear();
        const char *q = line.c_str();
        while (*q) {
            while (*q == ' ' || *q == '\t' || *q == '\r') q++;
            if (!*q) break;
            char *end = nullptr;
            const float v = std::strtof(q, &end);
            if (end == q || (*end && *end != ' ' && *end != '\t' && *end != '\r')) {
                ok = false;
                break;
            }
            row.push_back(v);
            q = end;
        }
        if (!ok) break;
        if (row.empty()) continue;
        if (cols.empty()) cols.resize(row.size());
        if (row.size() != cols.size() || row.size() < 2) {
            ok = false;
            break;
        }
        for (size_t j = 0; j < row.size(); j++) cols[j].push_back(row[j]);
        bool seen = false;
        for (auto &rc : rowCounts)
            if (rc.first == row[1]) {
                rc.second++;
                seen = true;
            }
        if (!seen) rowCounts.emplace_back(row[1], 1);
    }
    if (!ok || cols.empty() || (size_t)fieldID >= cols.size()) return HDM_E_INVALID;
    const int64_t xPoints = rowCounts.front().second, yPoints = (int64_t)rowCounts.size();
    for (auto &rc : rowCounts)
        if (rc.second != xPoints) return HDM_E_INVALID;
    // the extents start where the reference's do: min from the largest
    // float, max from the smallest positive normal float (so the max of an
    // all-negative column reads 1.2e-38 there too)
    float xMin = FLT_MAX, xMax = FLT_MIN, yMin = FLT_MAX, yMax = FLT_MIN;
    for (size_t i = 0; i < cols[0].size(); i++) {
        xMin = std::min(xMin, cols[0][i]);
        xMax = std::max(xMax, cols[0][i]);
        yMin = std::min(yMin, cols[1][i]);
        yMax = std::max(yMax, cols[1][i]);
    }
    if (xPoints * yPoints != (int64_t)cols[0].size()) return HDM_E_INVALID;
    double *data = new (std::nothrow) double[(size_t)(xPoints * yPoints)];
    if (!data) return HDM_E_NOMEM;
    // Map2d constructor: every point -1 until set; Map2d::setPointCord
    // (Map2d.h:65-74) places each row at its rounded grid index
    for (int64_t i = 0; i < xPoints * yPoints; i++) data[i] = -1;
    const float xRes = (xMax - xMin) / (float)(xPoints - 1), yRes = (yMax - yMin) / (float)(yPoints - 1);
    for (size_t i = 0; i < cols[0].size(); i++) {
        float x = (cols[0][i] - xMin) / xRes, y = (cols[1][i] - yMin) / yRes;
        x = std::round(x);
        y = std::round(y);
        const int64_t ix = xPoints > 1 ? (int64_t)x : 0, iy = yPoints > 1 ? (int64_t)y : 0;
        if (ix < 0 || ix >= xPoints || iy < 0 || iy >= yPoints) {
            delete[] data;
            return HDM_E_INVALID;
        }
        data[ix + iy * xPoints] = (double)cols[(size_t)fieldID][i];
    }
    map->xPoints = (int32_t)xPoints;
    map->yPoints = (int32_t)yPoints;
    map->xMin = xMin;
    map->yMin = yMin;
    map->xMax = xMax;
    map->yMax = yMax;
    map->data = data;
    return HDM_OK;
}

void hdm_map2d_free(hdm_map2d *map) {
    if (!map) return;
    delete[] map->data;
    map->data = nullptr;
}

int hdm_bw_probe(hdm_ctx *ctx, int32_t kind, int64_t bytes, int32_t iters, double *gbps) {
    if (!valid(ctx) || (kind != HDM_PROBE_COPY && kind != HDM_PROBE_CT_RMW) || bytes < (1 << 20) || iters < 1)
        return HDM_E_INVALID;
    Context &c = ctx->c;
    HIP_CHECK(ctx, c.probe.ensure((size_t)bytes));
    HIP_CHECK(ctx, hdm::launch_probe_zero(c, c.probe.ptr, bytes));
    HIP_CHECK(ctx, hdm::launch_probe(c, kind, c.probe.ptr, bytes));  // warm-up
    hipEvent_t e0, e1;
    HIP_CHECK(ctx, hipEventCreate(&e0));
    HIP_CHECK(ctx, hipEventCreate(&e1));
    HIP_CHECK(ctx, hipEventRecord(e0, c.stream));
    for (int i = 0; i < iters; i++) HIP_CHECK(ctx, hdm::launch_probe(c, kind, c.probe.ptr, bytes));
    HIP_CHECK(ctx, hipEventRecord(e1, c.stream));
    HIP_CHECK(ctx, hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // copy: bytes/2 read + bytes/2 written; rmw6: bytes read + bytes written
    double moved = (kind == HDM_PROBE_COPY) ? (double)(bytes / 2 / 16 * 16) * 2.0 : (double)(bytes / 48 * 48) * 2.0;
    if (gbps) *gbps = moved * iters / ((double)ms * 1e-3) / 1e9;
    return HDM_OK;
}

int hdm_set_timestep(hdm_ctx *ctx, int64_t ts) {
    if (!valid(ctx) || ts < 0) return HDM_E_INVALID;
    ctx->c.stateEpoch++;
    ctx->c.ts = ts;
    ctx->c.accumulatedDist = 0;
    return HDM_OK;
}

// gridService (HiDEMLib.h:533-628) -> computePartForce -> applyForce ->
// linkUpdate (ts*dt > breakTime) -> finalizeTs, nSteps times.  The integrator
// of each step also reduces the rebuild-test quantities of the new state
// (max+ Diff and max |x - x0|^2 into red[32..36]), so the next step's
// decision costs one 40-byte read instead of a separate pass.
}  // extern "C"

namespace {

// gridService rebuild (HiDEMLib.h:589-627): grid, neighbour list, snapshot
// for the exact test; the scheduler's blocks follow the new grid.
int rebuild(hdm_ctx *ctx, const hdm_run_params *rp) {
    Context &c = ctx->c;
    c.stateEpoch++;
    c.accumulatedDist = 0;
    int r = hdm_grid_build(ctx, rp->searchDistance);
    if (r == HDM_OK) r = hdm_neighbours_build(ctx, rp->interactDist);
    if (r == HDM_OK) r = hdm_skin_snapshot(ctx);
    if (r != HDM_OK) return r;
    c.stepsSinceRebuild = 0;
    c.schedLastEval = 0;
    c.schedAssigned = false;
    return HDM_OK;
}

hdm::SchedGeometry sched_geometry(const Context &c) {
    hdm::SchedGeometry g;
    const int32_t bc = c.sched.blockCells > 0 ? c.sched.blockCells : 4;
    g.ox = c.grid.xMin;
    g.oy = c.grid.yMin;
    g.oz = c.grid.zMin;
    g.edge = c.grid.spacing * bc;
    g.BX = (c.grid.X + bc - 1) / bc;
    g.BY = (c.grid.Y + bc - 1) / bc;
    g.BZ = (c.grid.Z + bc - 1) / bc;
    return g;
}

// Features of the current state (assigning blocks first when the grid changed).
int sched_eval_features(hdm_ctx *ctx, const hdm_run_params *rp, int64_t stepsSince, int32_t *nBlocks) {
    Context &c = ctx->c;
    if (!c.grid.built) return fail(ctx, HDM_E_STATE, "scheduler features need a grid build");
    if (!c.snapshotValid) return fail(ctx, HDM_E_STATE, "scheduler features need a position snapshot");
    hdm::SchedGeometry g = sched_geometry(c);
    if (!c.schedAssigned) {
        HIP_CHECK(ctx, hdm::sched_assign_blocks(c, g));
        c.schedAssigned = true;
        c.schedLastEval = 0;
    }
    hdm::SchedFeatureInputs in;
    in.budget = 0.5 * (rp->interactDist - rp->force.attractDistance);
    in.waterLevel = rp->integ.waterLevel;
    in.stepsSince = stepsSince;
    in.stepsSincePrev = stepsSince - c.schedLastEval;
    HIP_CHECK(ctx, hdm::sched_features(c, g, in));
    c.schedLastEval = stepsSince;
    c.schedNb = g.BX * g.BY * g.BZ;
    if (nBlocks) *nBlocks = c.schedNb;
    return HDM_OK;
}

int sched_use_model(hdm_ctx *ctx) {
    Context &c = ctx->c;
    if (!c.mlpHoldsModel) {
        HIP_CHECK(ctx, hdm::mlp_upload(c, c.schedModel));
        c.mlpHoldsModel = true;
    }
    return HDM_OK;
}

// Weights for a standalone network call (no longer the scheduler's).
int mlp_use_weights(hdm_ctx *ctx, const hdm_mlp_weights *w) {
    Context &c = ctx->c;
    for (int k = 0; k < HDM_SCHED_FEATURES; k++)
        if (!(w->featStd[k] > 0.0f) || !std::isfinite(w->featMean[k]))
            return fail(ctx, HDM_E_INVALID, "hdm_mlp_weights: featStd must be > 0 and featMean finite");
    HIP_CHECK(ctx, hdm::mlp_upload(c, *w));
    c.mlpHoldsModel = false;
    return HDM_OK;
}

template <class T>
hipError_t h2d_buf(hdm::DevBuf &b, const T *src, int64_t count, hipStream_t s) {
    hipError_t e = b.ensure(sizeof(T) * (size_t)(count > 0 ? count : 1));
    if (e != hipSuccess || count == 0 || src == nullptr) return e;
    return hipMemcpyAsync(b.ptr, src, sizeof(T) * (size_t)count, hipMemcpyHostToDevice, s);
}

// Steps to enqueue before the next host check.
int sched_horizon(hdm_ctx *ctx, const hdm_run_params *rp, int64_t *H) {
    Context &c = ctx->c;
    const int64_t lo = c.sched.minBatch > 0 ? c.sched.minBatch : 1;
    const int64_t hi = c.sched.maxBatch >= lo ? c.sched.maxBatch : lo;
    if (!c.haveModel) {
        *H = hi;
        return HDM_OK;
    }
    int32_t nb = 0;
    int r = sched_eval_features(ctx, rp, c.stepsSinceRebuild, &nb);
    if (r == HDM_OK) r = sched_use_model(ctx);
    if (r != HDM_OK) return r;
    HIP_CHECK(ctx, hdm::mlp_forward(c, c.schedFeat.as<float>(), nb, c.schedPred.as<float>()));
    uint64_t *out = c.red.as<uint64_t>() + 44;
    HIP_CHECK(ctx, hdm::sched_min_prediction(c, nb, out));
    uint64_t k = 0;
    HIP_CHECK(ctx, d2h(&k, out, sizeof(k), c.stream));
    const double m = hdm::ordered_value(k);
    double h = (double)c.sched.safety * std::exp2(m < 30.0 ? m : 30.0);
    if (!(h >= (double)lo)) h = (double)lo;   // also NaN
    if (h > (double)hi) h = (double)hi;
    *H = (int64_t)h;
    return HDM_OK;
}

// Everything a captured step pair bakes into its kernel arguments.
std::vector<unsigned char> graph_key(const Context &c, const hdm_run_params *rp, const uint64_t *skin) {
    std::vector<unsigned char> k;
    auto put = [&](const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        k.insert(k.end(), b, b + n);
    };
    const hdm::DevBuf *bufs[] = {&c.x, &c.y, &c.z, &c.tx, &c.ty, &c.tz, &c.radius, &c.type, &c.rec, &c.recAlt,
                                 &c.drec, &c.drecAlt, &c.dtail, &c.dtailAlt, &c.nbrCount, &c.nbr, &c.linkCount, &c.slotLink, &c.slotOther,
                                 &c.state, &c.initialLen, &c.maxLoad, &c.slotThr, &c.skinPartials, &c.guardState,
                                 &c.x0, &c.y0, &c.z0, &c.x0f, &c.y0f, &c.z0f, &c.blockId, &c.blockWater};
    for (auto *b : bufs) put(&b->ptr, sizeof(void *));
    for (int q = 0; q < 6; q++) {
        put(&c.diff[q].ptr, sizeof(void *));
        put(&c.force[q].ptr, sizeof(void *));
    }
    put(&c.ct.ptr, sizeof(void *));
    put(&c.brokenCount.ptr, sizeof(void *));
    put(&c.guardCur, sizeof(c.guardCur));
    put(&c.nbrExcludesLinked, sizeof(c.nbrExcludesLinked));
    for (int q = 0; q < 12; q++) put(&c.alt[q].ptr, sizeof(void *));
    put(&c.n, sizeof(c.n));
    put(&c.L, sizeof(c.L));
    put(&c.maxNeigh, sizeof(c.maxNeigh));
    put(&c.maxLinks, sizeof(c.maxLinks));
    put(&c.nBlocks, sizeof(c.nBlocks));
    put(&c.mapMask, sizeof(c.mapMask));
    put(&c.mapDesc.ptr, sizeof(void *));
    put(&c.stepVariant, sizeof(c.stepVariant));
    put(&c.recTame, sizeof(c.recTame));
    put(&c.hasMaxLoad, sizeof(c.hasMaxLoad));
    put(&c.l0Max, sizeof(c.l0Max));
    put(&c.thrScale, sizeof(c.thrScale));
    put(&skin, sizeof(skin));
    put(&rp->force, sizeof(rp->force));
    put(&rp->integ, sizeof(rp->integ));
    put(&rp->brk, sizeof(rp->brk));
    put(&rp->interactDist, sizeof(rp->interactDist));
    return k;
}

// G fused steps (each with the previous step's break, no force store)
// replayed from a captured hipGraph, G = GRAPH_STEPS_LONG or 2 (one cached
// graph each); captured again whenever anything baked into the arguments
// changed (new parameters, a new launch shape, a neighbour list that needed
// a wider ELL).  G is even, so the steps swap the state buffers an even
// number of times and the host's view of the current state is unchanged by
// a replay.
constexpr int GRAPH_STEPS_LONG = 8;

int graph_steps(hdm_ctx *ctx, const hdm_run_params *rp, uint64_t *skin, const hdm::StepGuard *guard, int G) {
    Context &c = ctx->c;
    hipGraphExec_t &exec = G == 2 ? c.graphExecPair : c.graphExec;
    std::vector<unsigned char> &cached = G == 2 ? c.graphKeyPair : c.graphKey;
    std::vector<unsigned char> key = graph_key(c, rp, skin);
    if (!exec || key != cached) {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        hipGraph_t graph = nullptr;
        HIP_CHECK(ctx, hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        c.capturing = true;
        hipError_t e = hipSuccess;
        int launched = 0;
        for (; launched < G && e == hipSuccess; launched++)
            e = hdm::launch_step(c, rp->force, rp->integ, skin, true, &rp->brk, false, guard);
        if (e != hipSuccess) launched--;
        if (launched & 1) hdm::swap_step_state(c);   // keep the host's buffer view as before the capture
        hipError_t e3 = hipStreamEndCapture(c.stream, &graph);
        c.capturing = false;
        if (e != hipSuccess || e3 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_fail(ctx, e != hipSuccess ? e : e3, "timestep graph capture");
        }
        hipError_t e4 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e4 != hipSuccess) {
            exec = nullptr;
            return hip_fail(ctx, e4, "hipGraphInstantiate");
        }
        cached = std::move(key);
        c.graphCaptures++;
    }
    Timer t(c, HDM_T_STEP);
    HIP_CHECK(ctx, hipGraphLaunch(exec, c.stream));
    c.graphReplays++;
    // the replayed steps advanced the records, as launch_step notes for its
    // own launches: the SoA coordinates and Diff now lag them
    c.soaStale = true;
    return HDM_OK;
}

// HDM_REBUILD_LEARNED: the exact test runs on the device in every step's last
// workgroup (StepGuard); the host enqueues scheduler-sized batches and reads
// the guard once per batch.  A step launched after the test tripped does
// nothing, so the state advances exactly as under HDM_REBUILD_EXACT.
int run_learned(hdm_ctx *ctx, const hdm_run_params *rp, int64_t nSteps, uint8_t *rebuildLog, hdm_run_stats *stats) {
    Context &c = ctx->c;
    uint64_t *skin = c.red.as<uint64_t>() + hdm::SKIN_KEYS;
    const double thr = rp->interactDist - rp->force.attractDistance;
    hdm::StepGuard guard{thr};
    auto t0 = std::chrono::steady_clock::now();
    int64_t rebuilds = 0;
    if (rebuildLog) std::memset(rebuildLog, 0, (size_t)nSteps);
    // initial decision as HDM_REBUILD_EXACT takes it: known from the trip
    // words of the last learned run when nothing touched the state since
    bool regen = !c.neighBuilt || !c.snapshotValid;
    if (!regen && c.tripKnown && c.tripEpoch == c.stateEpoch && c.tripTs == c.ts) {
        regen = c.tripPending;
    } else if (!regen) {
        double d = 0;
        if (c.skinFused && c.skinFusedDisp) {
            uint64_t k = 0;
            HIP_CHECK(ctx, d2h(&k, skin + 4, sizeof(k), c.stream));
            d = std::sqrt(hdm::ordered_value(k));
        } else {
            int r = hdm_skin_exact(ctx, &d);
            if (r != HDM_OK) return r;
        }
        regen = 2.0 * d >= thr;
    }
    bool breakPending = false;
    int64_t step = 0;
    while (step < nSteps) {
        if (regen) {
            int r = rebuild(ctx, rp);
            if (r != HDM_OK) return r;
            rebuilds++;
            if (rebuildLog) rebuildLog[step] = 1;
        }
        int64_t H = 0;
        if (c.schedCarry > 0 && c.schedCarryEpoch == c.stateEpoch && c.schedCarryTs == c.ts) {
            H = c.schedCarry;   // the rest of the batch an earlier run scheduled
        } else {
            int r = sched_horizon(ctx, rp, &H);
            if (r != HDM_OK) return r;
        }
        c.schedCarry = 0;
        if (c.autotune && !c.tuned) H = 3 * hdm::STEP_VARIANTS;   // a tuning batch: every shape three times
        const int64_t Hplanned = H;
        if (H > nSteps - step) H = nSteps - step;
        // guard flag and executed-step count start at 0 (a fill kernel, capture-safe)
        HIP_CHECK(ctx, hdm::launch_guard_reset(c));
        // autotuning launches every shape in turn (same results) and times it
        const bool tuning = c.autotune && !c.tuned;
        std::vector<hipEvent_t> tev;
        std::vector<int> tvar;
        int64_t i = 0;
        while (i < H) {
            const int64_t s = step + i;
            // G replayed steps: inside the batch, not the run's last step
            // (which stores the force), each with the previous step's break
            auto graphOk = [&](int64_t G) {
                return !tuning && c.useGraphs && i + G <= H && breakPending && s + G < nSteps &&
                       (double)(c.ts + i) * rp->dt > rp->breakTime && c.recCurrent;
            };
            if (c.usePersistent && !tuning && breakPending && c.recCurrent &&
                (double)(c.ts + i) * rp->dt > rp->breakTime) {
                // the rest of the batch in one launch, but not the run's last
                // step (it stores the force)
                int64_t S = H - i;
                if (s + S >= nSteps) S = nSteps - 1 - s;
                if (S >= 2) {
                    bool failed = false;
                    {
                        Timer t(c, HDM_T_STEP);
                        HIP_CHECK(ctx, hdm::launch_step_loop(c, rp->force, rp->integ, skin, rp->brk, guard, (int32_t)S,
                                                             &failed));
                    }
                    if (failed)
                        return fail(ctx, HDM_E_HIP, "persistent step: a grid barrier timed out "
                                                    "(workgroups not co-resident)");
                    c.loopLaunches++;
                    i += S;
                    breakPending = (double)(c.ts + i - 1) * rp->dt > rp->breakTime;
                    continue;
                }
            }
            const int64_t G = graphOk(GRAPH_STEPS_LONG) ? GRAPH_STEPS_LONG : graphOk(2) ? 2 : 0;
            if (G) {
                int r2 = graph_steps(ctx, rp, skin, &guard, (int)G);
                if (r2 != HDM_OK) return r2;
                i += G;
                breakPending = (double)(c.ts + i - 1) * rp->dt > rp->breakTime;
                continue;
            }
            if (tuning) {
                c.stepVariant = (int32_t)(tvar.size() % hdm::STEP_VARIANTS);
                hipEvent_t a = nullptr, b = nullptr;
                HIP_CHECK(ctx, hipEventCreate(&a));
                HIP_CHECK(ctx, hipEventCreate(&b));
                // a tuning launch is a step like any other for the step timer
                Timer t(c, HDM_T_STEP);
                HIP_CHECK(ctx, hipEventRecord(a, c.stream));
                HIP_CHECK(ctx, hdm::launch_step(c, rp->force, rp->integ, skin, true, breakPending ? &rp->brk : nullptr,
                                                s == nSteps - 1, &guard));
                HIP_CHECK(ctx, hipEventRecord(b, c.stream));
                tev.push_back(a);
                tev.push_back(b);
                tvar.push_back(c.stepVariant);
            } else {
                Timer t(c, HDM_T_STEP);
                HIP_CHECK(ctx, hdm::launch_step(c, rp->force, rp->integ, skin, true, breakPending ? &rp->brk : nullptr,
                                                s == nSteps - 1, &guard));
            }
            breakPending = (double)(c.ts + i) * rp->dt > rp->breakTime;
            i++;
        }
        uint32_t words[hdm::GUARD_WORDS] = {0, 0, 0, 0};
        HIP_CHECK(ctx, d2h(words, c.guardState.ptr, sizeof(words), c.stream));
        const int64_t k = words[hdm::GUARD_STEPS];
        if (tuning) {
            // only launches that executed (the first k) time a real step
            for (size_t q = 0; q < tvar.size(); q++) {
                float ms = 0;
                if ((int64_t)q < k && hipEventElapsedTime(&ms, tev[2 * q], tev[2 * q + 1]) == hipSuccess && q > 0) {
                    const int v = tvar[q];
                    if (c.tuneCount[v] == 0 || ms < c.tuneMs[v]) c.tuneMs[v] = ms;   // best of the samples
                    c.tuneCount[v]++;
                }
                (void)hipEventDestroy(tev[2 * q]);
                (void)hipEventDestroy(tev[2 * q + 1]);
            }
            bool done = true;
            for (int v = 0; v < hdm::STEP_VARIANTS; v++) done &= c.tuneCount[v] >= 2;
            if (done) {
                c.stepVariant = hdm::pick_step_variant(c.tuneMs);
                c.tuned = true;
            } else {
                c.stepVariant = 0;
            }
        }
        regen = (words[hdm::GUARD_FLAG] | words[hdm::GUARD_FLAG + 1]) != 0;
        if ((H - k) & 1) {  // launch_step swapped once per launch; undo the skipped ones
            hdm::swap_step_state(c);
        }
        if (k > 0) {
            breakPending = (double)(c.ts + k - 1) * rp->dt > rp->breakTime;
            c.altHoldsPrev = true;
        }
        c.ts += k;
        c.stepsSinceRebuild += k;
        c.schedChecks++;
        c.schedLaunched += H;
        c.schedExecuted += k;
        step += k;
        // the guarded launches keep no maxima for the host-checked policies
        c.skinFused = false;
        c.skinFusedDisp = false;
        if (k == 0 && !regen) return fail(ctx, HDM_E_STATE, "guarded batch executed no step");
        if (!regen && k == H && H < Hplanned) {   // the run ended inside the batch
            c.schedCarry = Hplanned - H;
            c.schedCarryTs = c.ts;
            c.schedCarryEpoch = c.stateEpoch;
        }
    }
    // whether the state left trips the exact test (the link break below
    // does not move particles), for the next learned run's first decision
    c.tripKnown = true;
    c.tripPending = regen;
    c.tripEpoch = c.stateEpoch;
    c.tripTs = c.ts;
    if (breakPending) {
        Timer t(c, HDM_T_BREAK);
        HIP_CHECK(ctx, hdm::launch_break_records(c, rp->brk, nullptr));
    }
    HIP_CHECK(ctx, hipStreamSynchronize(c.stream));
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->timestep = c.ts;
        stats->steps = nSteps;
        stats->rebuilds = rebuilds;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        int64_t b = 0;
        int r = hdm_count_broken(ctx, &b);
        if (r != HDM_OK) return r;
        stats->brokenTotal = b;
    }
    return HDM_OK;
}

}  // namespace

extern "C" {

int hdm_run(hdm_ctx *ctx, const hdm_run_params *rp, int64_t nSteps, uint8_t *rebuildLog, hdm_run_stats *stats) {
    if (!valid(ctx) || !rp || nSteps < 0) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (rp->integ.variant != HDM_INTEG_ICEBLOCK && rp->integ.variant != HDM_INTEG_GLACIER)
        return fail(ctx, HDM_E_INVALID, "unknown integrator variant");
    if (rp->policy < HDM_REBUILD_HEURISTIC || rp->policy > HDM_REBUILD_LEARNED)
        return fail(ctx, HDM_E_INVALID, "unknown rebuild policy");
    if (!c.linksBuilt) {
        int r = hdm_links_build(ctx);
        if (r != HDM_OK) return r;
    }
    if (rp->policy == HDM_REBUILD_LEARNED) return run_learned(ctx, rp, nSteps, rebuildLog, stats);
    c.stateEpoch++;
    uint64_t *skin = c.red.as<uint64_t>() + hdm::SKIN_KEYS;
    const bool wantSkin = rp->policy != HDM_REBUILD_INTERVAL;
    const bool wantDisp = rp->policy == HDM_REBUILD_EXACT;
    auto t0 = std::chrono::steady_clock::now();
    int64_t rebuilds = 0;
    // The link break of step s runs inside step s+1's fused launch (on the
    // same state, before any force uses the links); the last step's break is
    // a standalone launch after the loop.
    bool breakPending = false;
    for (int64_t step = 0; step < nSteps; step++) {
        bool regen = !c.neighBuilt || !c.snapshotValid;
        if (rp->policy == HDM_REBUILD_INTERVAL) {
            if (rp->regenInterval > 0 && (c.ts % rp->regenInterval) == 0) regen = true;
        } else if (rp->policy == HDM_REBUILD_HEURISTIC) {
            if (c.ts % 10 == 0) {
                double d = 0;
                if (c.skinFused) {
                    uint64_t k[3];
                    HIP_CHECK(ctx, d2h(k, skin, sizeof(k), c.stream));
                    double xm = hdm::ordered_value(k[0]), ym = hdm::ordered_value(k[1]), zm = hdm::ordered_value(k[2]);
                    d = 10 * std::sqrt((xm * xm) + (ym * ym) + (zm * zm));
                } else {
                    int r = hdm_skin_heuristic(ctx, 10, &d);
                    if (r != HDM_OK) return r;
                }
                c.accumulatedDist += d;
                if (c.accumulatedDist >= rp->interactDist / 2.0) regen = true;
            }
        } else if (!regen) {
            double d = 0;
            if (c.skinFused && c.skinFusedDisp) {
                uint64_t k = 0;
                HIP_CHECK(ctx, d2h(&k, skin + 4, sizeof(k), c.stream));
                d = std::sqrt(hdm::ordered_value(k));
            } else {
                int r = hdm_skin_exact(ctx, &d);
                if (r != HDM_OK) return r;
            }
            if (2.0 * d >= rp->interactDist - rp->force.attractDistance) regen = true;
        }
        if (rebuildLog) rebuildLog[step] = regen ? 1 : 0;
        if (regen) {
            int r = rebuild(ctx, rp);
            if (r != HDM_OK) return r;
            rebuilds++;
        }
        {
            Timer t(c, HDM_T_STEP);
            HIP_CHECK(ctx, hdm::launch_step(c, rp->force, rp->integ, wantSkin ? skin : nullptr, wantDisp,
                                            breakPending ? &rp->brk : nullptr, step == nSteps - 1));
        }
        c.skinFused = wantSkin;
        c.skinFusedDisp = wantDisp;
        c.altHoldsPrev = true;
        breakPending = (double)c.ts * rp->dt > rp->breakTime;
        c.ts++;
        c.stepsSinceRebuild++;
    }
    if (breakPending) {
        Timer t(c, HDM_T_BREAK);
        HIP_CHECK(ctx, hdm::launch_break_records(c, rp->brk, nullptr));
    }
    HIP_CHECK(ctx, hipStreamSynchronize(c.stream));
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->timestep = c.ts;
        stats->steps = nSteps;
        stats->rebuilds = rebuilds;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        int64_t b = 0;
        int r = hdm_count_broken(ctx, &b);
        if (r != HDM_OK) return r;
        stats->brokenTotal = b;
    }
    return HDM_OK;
}


// ---- learned rebuild scheduler ---------------------------------------------------
int hdm_sched_configure(hdm_ctx *ctx, const hdm_sched_params *sp) {
    if (!valid(ctx) || !sp) return HDM_E_INVALID;
    if (sp->blockCells < 1 || sp->minBatch < 1 || sp->maxBatch < sp->minBatch || !(sp->safety > 0.0f))
        return fail(ctx, HDM_E_INVALID, "hdm_sched_params: need blockCells >= 1, 1 <= minBatch <= maxBatch, safety > 0");
    Context &c = ctx->c;
    c.stateEpoch++;
    if (sp->blockCells != c.sched.blockCells) c.schedAssigned = false;
    c.sched = *sp;
    return HDM_OK;
}

int hdm_sched_set_model(hdm_ctx *ctx, const hdm_mlp_weights *w) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    c.stateEpoch++;
    if (!w) {
        c.haveModel = false;
        return HDM_OK;
    }
    int r = mlp_use_weights(ctx, w);
    if (r != HDM_OK) return r;
    c.schedModel = *w;
    c.mlpHoldsModel = true;
    c.haveModel = true;
    return HDM_OK;
}

int hdm_sched_features(hdm_ctx *ctx, const hdm_run_params *rp, int64_t stepsSinceRebuild, int32_t capacity,
                       int32_t *nBlocks, float *features, float *blockDisp, uint8_t *validOut) {
    if (!valid(ctx) || !rp || stepsSinceRebuild < 0) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.linksBuilt) return fail(ctx, HDM_E_STATE, "link lists not built");
    if (c.grid.built) {
        const hdm::SchedGeometry g = sched_geometry(c);
        const int64_t need = (int64_t)g.BX * g.BY * g.BZ;
        if ((features || blockDisp || validOut) && need > capacity)
            return fail(ctx, HDM_E_CAPACITY, "hdm_sched_features: " + std::to_string(need) + " blocks exceed capacity");
    }
    int32_t nb = 0;
    int r = sched_eval_features(ctx, rp, stepsSinceRebuild, &nb);
    if (r != HDM_OK) return r;
    if (nBlocks) *nBlocks = nb;
    if (features) HIP_CHECK(ctx, d2h(features, c.schedFeat.ptr, sizeof(float) * HDM_SCHED_FEATURES * (size_t)nb, c.stream));
    if (blockDisp) HIP_CHECK(ctx, d2h(blockDisp, c.schedDisp.ptr, sizeof(float) * (size_t)nb, c.stream));
    if (validOut) HIP_CHECK(ctx, d2h(validOut, c.schedValid.ptr, (size_t)nb, c.stream));
    HIP_CHECK(ctx, hipStreamSynchronize(c.stream));
    return HDM_OK;
}

int hdm_sched_predict(hdm_ctx *ctx, int32_t capacity, float *pred, double *minPred) {
    if (!valid(ctx)) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (!c.haveModel) return fail(ctx, HDM_E_STATE, "no scheduler model set");
    if (c.schedNb == 0) return fail(ctx, HDM_E_STATE, "no features evaluated");
    if (pred && c.schedNb > capacity) return fail(ctx, HDM_E_CAPACITY, "hdm_sched_predict: capacity below nBlocks");
    int r = sched_use_model(ctx);
    if (r != HDM_OK) return r;
    HIP_CHECK(ctx, hdm::mlp_forward(c, c.schedFeat.as<float>(), c.schedNb, c.schedPred.as<float>()));
    if (minPred) {
        uint64_t *out = c.red.as<uint64_t>() + 44;
        HIP_CHECK(ctx, hdm::sched_min_prediction(c, c.schedNb, out));
        uint64_t k = 0;
        HIP_CHECK(ctx, d2h(&k, out, sizeof(k), c.stream));
        *minPred = hdm::ordered_value(k);
    }
    if (pred) HIP_CHECK(ctx, d2h(pred, c.schedPred.ptr, sizeof(float) * (size_t)c.schedNb, c.stream));
    HIP_CHECK(ctx, hipStreamSynchronize(c.stream));
    return HDM_OK;
}

int hdm_mlp_forward(hdm_ctx *ctx, const hdm_mlp_weights *w, const float *X, int64_t n, float *y) {
    if (!valid(ctx) || !w || n < 0 || (n > 0 && (!X || !y))) return HDM_E_INVALID;
    Context &c = ctx->c;
    (void)hipSetDevice(c.device);
    int r = mlp_use_weights(ctx, w);
    if (r != HDM_OK) return r;
    HIP_CHECK(ctx, h2d_buf(c.mlpX, X, n * HDM_SCHED_FEATURES, c.stream));
    HIP_CHECK(ctx, c.mlpY.ensure(sizeof(float) * (size_t)(n > 0 ? n : 1)));
    HIP_CHECK(ctx, hdm::mlp_forward(c, c.mlpX.as<float>(), n, c.mlpY.as<float>()));
    if (n > 0) HIP_CHECK(ctx, d2h(y, c.mlpY.ptr, sizeof(float) * (size_t)n, c.stream));
    return HDM_OK;
}

int hdm_mlp_gradient(hdm_ctx *ctx, const hdm_mlp_weights *w, const float *X, const float *target,
                     const uint8_t *censored, int64_t n, float *grad, double *loss) {
    if (!valid(ctx) || !w || n < 1 || !X || !target || !grad) return HDM_E_INVALID;
    Context &c = ctx->c;
    (void)hipSetDevice(c.device);
    int r = mlp_use_weights(ctx, w);
    if (r != HDM_OK) return r;
    const int64_t P = hdm::mlp_param_count(), tiles = (n + 31) / 32;
    HIP_CHECK(ctx, h2d_buf(c.mlpX, X, n * HDM_SCHED_FEATURES, c.stream));
    HIP_CHECK(ctx, h2d_buf(c.mlpT, target, n, c.stream));
    HIP_CHECK(ctx, h2d_buf(c.mlpC, censored, censored ? n : 0, c.stream));
    HIP_CHECK(ctx, c.mlpGrad.ensure(sizeof(float) * (size_t)P));
    HIP_CHECK(ctx, c.mlpLoss.ensure(sizeof(float) * (size_t)tiles));
    HIP_CHECK(ctx, hdm::mlp_gradient(c, c.mlpX.as<float>(), c.mlpT.as<float>(), censored ? c.mlpC.as<uint8_t>() : nullptr,
                                     n, c.mlpGrad.as<float>(), c.mlpLoss.as<float>()));
    HIP_CHECK(ctx, d2h(grad, c.mlpGrad.ptr, sizeof(float) * (size_t)P, c.stream));
    if (loss) {
        std::vector<float> lp((size_t)tiles);
        HIP_CHECK(ctx, d2h(lp.data(), c.mlpLoss.ptr, sizeof(float) * (size_t)tiles, c.stream));
        double s = 0;
        for (float v : lp) s += v;
        *loss = s;
    }
    return HDM_OK;
}

int hdm_mlp_train(hdm_ctx *ctx, hdm_mlp_weights *w, const float *X, const float *target, const uint8_t *censored,
                  int64_t n, const hdm_mlp_train_params *tp, double *loss) {
    if (!valid(ctx) || !w || n < 1 || !X || !target || !tp) return HDM_E_INVALID;
    if (tp->epochs < 0 || tp->batch < 1 || !(tp->lr > 0.0f) || !(tp->eps > 0.0f))
        return fail(ctx, HDM_E_INVALID, "hdm_mlp_train_params: need epochs >= 0, batch >= 1, lr > 0, eps > 0");
    Context &c = ctx->c;
    (void)hipSetDevice(c.device);
    int r = mlp_use_weights(ctx, w);
    if (r != HDM_OK) return r;
    const int64_t P = hdm::mlp_param_count();
    HIP_CHECK(ctx, h2d_buf(c.mlpX, X, n * HDM_SCHED_FEATURES, c.stream));
    HIP_CHECK(ctx, h2d_buf(c.mlpT, target, n, c.stream));
    HIP_CHECK(ctx, h2d_buf(c.mlpC, censored, censored ? n : 0, c.stream));
    HIP_CHECK(ctx, c.mlpGrad.ensure(sizeof(float) * (size_t)P));
    HIP_CHECK(ctx, c.mlpM.ensure(sizeof(float) * (size_t)P));
    HIP_CHECK(ctx, c.mlpV.ensure(sizeof(float) * (size_t)P));
    HIP_CHECK(ctx, c.mlpLoss.ensure(sizeof(float) * (size_t)((n + 31) / 32)));
    HIP_CHECK(ctx, hdm::launch_fill_u32(c.stream, c.mlpM.as<uint32_t>(), 0u, P));
    HIP_CHECK(ctx, hdm::launch_fill_u32(c.stream, c.mlpV.as<uint32_t>(), 0u, P));
    int64_t t = 0;
    for (int32_t ep = 0; ep < tp->epochs; ep++)
        for (int64_t off = 0; off < n; off += tp->batch) {
            const int64_t rows = (n - off < tp->batch) ? n - off : tp->batch;
            HIP_CHECK(ctx, hdm::mlp_gradient(c, c.mlpX.as<float>() + off * HDM_SCHED_FEATURES, c.mlpT.as<float>() + off,
                                             censored ? c.mlpC.as<uint8_t>() + off : nullptr, rows,
                                             c.mlpGrad.as<float>(), c.mlpLoss.as<float>()));
            HIP_CHECK(ctx, hdm::mlp_adam(c, c.mlpGrad.as<float>(), c.mlpM.as<float>(), c.mlpV.as<float>(),
                                         1.0f / (float)rows, tp->lr, tp->beta1, tp->beta2, tp->eps, ++t));
        }
    HIP_CHECK(ctx, d2h(w, c.mlpW.ptr, sizeof(hdm_mlp_weights), c.stream));
    if (loss) {
        HIP_CHECK(ctx, hdm::mlp_gradient(c, c.mlpX.as<float>(), c.mlpT.as<float>(),
                                         censored ? c.mlpC.as<uint8_t>() : nullptr, n, c.mlpGrad.as<float>(),
                                         c.mlpLoss.as<float>()));
        std::vector<float> lp((size_t)((n + 31) / 32));
        HIP_CHECK(ctx, d2h(lp.data(), c.mlpLoss.ptr, sizeof(float) * lp.size(), c.stream));
        double s = 0;
        for (float v : lp) s += v;
        *loss = s / (double)n;
    }
    return HDM_OK;
}

int hdm_sched_stats(hdm_ctx *ctx, int64_t *hostChecks, int64_t *stepsLaunched, int64_t *stepsExecuted) {
    if (!valid(ctx)) return HDM_E_INVALID;
    if (hostChecks) *hostChecks = ctx->c.schedChecks;
    if (stepsLaunched) *stepsLaunched = ctx->c.schedLaunched;
    if (stepsExecuted) *stepsExecuted = ctx->c.schedExecuted;
    return HDM_OK;
}


// ---- execution options -----------------------------------------------------------
int hdm_set_exec(hdm_ctx *ctx, const hdm_exec_params *ep) {
    if (!valid(ctx) || !ep) return HDM_E_INVALID;
    if (ep->variant < -1 || ep->variant >= hdm::STEP_VARIANTS)
        return fail(ctx, HDM_E_INVALID, "hdm_exec_params.variant out of range");
    Context &c = ctx->c;
    if (ep->persistent && !c.coopLaunch)
        return fail(ctx, HDM_E_INVALID, "hdm_exec_params.persistent: the device takes no cooperative launches");
    c.useGraphs = ep->graphs != 0;
    c.usePersistent = ep->persistent != 0;
    if (ep->autotune) {
        c.autotune = true;
        c.tuned = false;
        for (int v = 0; v < 8; v++) {
            c.tuneMs[v] = 0;
            c.tuneCount[v] = 0;
        }
    } else {
        c.autotune = false;
    }
    if (ep->variant >= 0) c.stepVariant = ep->variant;
    c.mergeContacts = ep->listWithPartners == 0;
    if (c.neighBuilt && c.nbrExcludesLinked != hdm::merge_linked_contacts(c)) c.neighBuilt = false;
    return HDM_OK;
}

int hdm_exec_info_get(hdm_ctx *ctx, hdm_exec_info *info) {
    if (!valid(ctx) || !info) return HDM_E_INVALID;
    Context &c = ctx->c;
    std::memset(info, 0, sizeof(*info));
    info->variant = c.stepVariant;
    info->variants = hdm::STEP_VARIANTS;
    info->tuned = c.tuned ? 1 : 0;
    std::snprintf(info->name, sizeof(info->name), "%s", hdm::step_variant_name(c.stepVariant));
    for (int v = 0; v < hdm::STEP_VARIANTS && v < 8; v++) info->tuneMs[v] = c.tuneMs[v];
    info->graphCaptures = c.graphCaptures;
    info->graphReplays = c.graphReplays;
    info->mergedContacts = c.nbrExcludesLinked ? 1 : 0;
    info->persistentLaunches = c.loopLaunches;
    return HDM_OK;
}


int hdm_sched_label_begin(hdm_ctx *ctx, int32_t slot, const hdm_run_params *rp) {
    if (!valid(ctx) || !rp || slot < 0 || slot >= hdm::LABEL_SLOTS) return HDM_E_INVALID;
    Context &c = ctx->c;
    if (c.schedNb == 0 || !c.schedAssigned) return fail(ctx, HDM_E_STATE, "hdm_sched_label_begin: evaluate features first");
    HIP_CHECK(ctx, hdm::sched_label_begin(c, slot, 0.5 * (rp->interactDist - rp->force.attractDistance)));
    return HDM_OK;
}

int hdm_sched_label_track(hdm_ctx *ctx) {
    if (!valid(ctx)) return HDM_E_INVALID;
    HIP_CHECK(ctx, hdm::sched_label_track(ctx->c));
    return HDM_OK;
}

int hdm_sched_label_end(hdm_ctx *ctx, int32_t slot, int32_t capacity, int32_t *firstCross, int32_t *nBlocks) {
    if (!valid(ctx) || slot < 0 || slot >= hdm::LABEL_SLOTS) return HDM_E_INVALID;
    hdm::LabelSlot &s = ctx->c.labels[slot];
    if (!s.active) return fail(ctx, HDM_E_STATE, "hdm_sched_label_end: slot not active");
    if (nBlocks) *nBlocks = s.nb;
    if (firstCross) {
        if (s.nb > capacity) return fail(ctx, HDM_E_CAPACITY, "hdm_sched_label_end: capacity below nBlocks");
        HIP_CHECK(ctx, d2h(firstCross, s.first.ptr, sizeof(int32_t) * (size_t)s.nb, ctx->c.stream));
    }
    s.active = false;
    return HDM_OK;
}

}  // extern "C"
