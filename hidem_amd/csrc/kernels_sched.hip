// Learned rebuild scheduler for gfx950: per-block features of the particle
// state, and a 16-64-64-1 MLP in BF16 on MFMA (v_mfma_f32_32x32x16_bf16,
// FP32 accumulation) with hand-written inference, backward, gradient
// reduction and Adam kernels.
//
// Blocks are cubes of blockCells^3 neighbour-grid cells laid over the
// particles' bounding box at the last rebuild.  The network maps a block's 16
// features to log2 of the number of steps until that block alone would trip
// the exact rebuild test (2 max |x - x0| >= interactDist - attractDistance);
// the minimum over blocks is the scheduler's horizon.  The exact test itself
// keeps running on the device every step (StepGuard), so a short prediction
// costs one extra host check and a long one costs skipped launches, never a
// missed contact.
//
// MFMA operand layout (32x32x16, wave64), lane l, j = 0..7:
//   A (32 x 16):  A[l % 32][8 (l / 32) + j]
//   B (16 x 32):  B[8 (l / 32) + j][l % 32]
//   C (32 x 32):  acc r = 0..15 holds C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
#include "hdm_context.h"
#include "kernel_common.h"

#include <hip/hip_bf16.h>

namespace hdm {

namespace {

constexpr int TPB = 256;
constexpr int NF = HDM_SCHED_FEATURES;   // 16
constexpr int NH = HDM_SCHED_HIDDEN;     // 64
constexpr int NACC = 16;                 // accumulators per block

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Row of the C/D element held in accumulator r by this lane.
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// ---- features ------------------------------------------------------------------
enum {
    A_MAXD = 0, A_SUMD, A_MAXACC, A_SUMACC, A_MAXDISP, A_SUMDISP, A_MAXDZ, A_MAXROT, A_MAXPROJ, A_COUNT,
    A_BELOW, A_SUMZ, A_SUMX, A_BROKEN, A_SIDES, A_SPARE
};
__device__ __constant__ const bool kIsMax[NACC] = {true,  false, true,  false, true,  false, true,  true,
                                                   true,  false, false, false, false, false, false, false};

__global__ void __launch_bounds__(TPB) k_sched_blocks(const double *__restrict__ x, const double *__restrict__ y,
                                                       const double *__restrict__ z, int64_t n, double ox, double oy,
                                                       double oz, double edge, int32_t BX, int32_t BY, int32_t BZ,
                                                       int32_t *__restrict__ blockOf) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    auto cell = [](double v, double o, double s, int32_t N) {
        double f = floor((v - o) / s);
        int32_t i = f < 0.0 ? 0 : (f >= (double)N ? N - 1 : (int32_t)f);
        return i;  // NaN compares false twice -> N - 1
    };
    const int32_t bx = cell(x[p], ox, edge, BX), by = cell(y[p], oy, edge, BY), bz = cell(z[p], oz, edge, BZ);
    blockOf[p] = bx + BX * (by + BY * bz);
}

struct AccArgs {
    const double *x, *y, *z, *x0, *y0, *z0;
    const double *d0, *d1, *d2, *d3, *d4, *d5;
    // Diff before the last fused step (prev and prevLo nullptr: unknown):
    // FP64 form prev = that drec (8 doubles per particle); FP32 form the
    // records, tails and escapes of that state
    const double *prev;
    const MixedRec *prevRec;
    const MixedTail *prevTail;
    const int32_t *blockOf;
    const uint32_t *linkCount;
    const int32_t *slotLink;
    int32_t linkWidth;
    int64_t n;
    double waterLevel;
    float *acc;                        // [nBlocks][NACC]
};

// Per-particle quantities folded into their block's accumulators.  A
// workgroup touches few blocks (particles are stored in spatial order), so it
// first combines its particles in LDS -- a small open-addressed table of block
// ids with LDS atomics -- and then issues one global atomic per accumulator
// and distinct block.  A particle whose block finds the table full goes to
// global memory directly.
constexpr int ACC_TABLE = 32;

__device__ __forceinline__ void acc_add(float *acc, int k, float v) {
    if (kIsMax[k]) atomicMax(reinterpret_cast<unsigned int *>(acc + k), __float_as_uint(v));
    else atomicAdd(acc + k, v);
}

__global__ void __launch_bounds__(TPB) k_sched_accumulate(AccArgs a) {
    __shared__ int32_t keys[ACC_TABLE];
    __shared__ float tab[ACC_TABLE][NACC];
    for (int i = threadIdx.x; i < ACC_TABLE * NACC; i += TPB) {
        if (i < ACC_TABLE) keys[i] = -1;
        tab[i / NACC][i % NACC] = 0.0f;
    }
    __syncthreads();
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p < a.n) {
        float v[NACC];
        const int32_t b = a.blockOf[p];
        const double dx = a.d0[p], dy = a.d1[p], dz = a.d2[p];
        const double D = sqrt(dx * dx + dy * dy + dz * dz);
        double Ac = 0.0;
        if (a.prev) {
            double q[6];
            if (a.prevTail) {
                join_diff6(a.prevRec[p], a.prevTail[p], a.prev, p, q);
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) q[c] = a.prev[p * 8 + c];
            }
            const double ax = dx - q[0], ay = dy - q[1], az = dz - q[2];
            Ac = sqrt(ax * ax + ay * ay + az * az);
        }
        const double ex = a.x[p] - a.x0[p], ey = a.y[p] - a.y0[p], ez = a.z[p] - a.z0[p];
        const double disp = sqrt(ex * ex + ey * ey + ez * ez);
        const double rot = fmax(fabs(a.d3[p]), fmax(fabs(a.d4[p]), fabs(a.d5[p])));
        uint32_t broken = 0;
        const uint32_t cnt = a.linkCount[p];
        const int64_t base = (p >> 6) * (int64_t)64 * a.linkWidth + (p & 63);
        for (uint32_t k = 0; k < cnt; k++) broken += a.slotLink[base + (int64_t)k * 64] < 0 ? 1u : 0u;
        v[A_MAXD] = (float)D;
        v[A_SUMD] = (float)D;
        v[A_MAXACC] = (float)Ac;
        v[A_SUMACC] = (float)Ac;
        v[A_MAXDISP] = (float)disp;
        v[A_SUMDISP] = (float)disp;
        v[A_MAXDZ] = (float)fabs(dz);
        v[A_MAXROT] = (float)rot;
        v[A_MAXPROJ] = (float)(disp + 50.0 * D);
        v[A_COUNT] = 1.0f;
        v[A_BELOW] = a.z[p] <= a.waterLevel ? 1.0f : 0.0f;
        v[A_SUMZ] = (float)a.z[p];
        v[A_SUMX] = (float)a.x[p];
        v[A_BROKEN] = (float)broken;
        v[A_SIDES] = (float)cnt;
        // find or claim the block's LDS slot
        int slot = -1;
        int h = (int)((uint32_t)b * 2654435761u % ACC_TABLE);
        for (int probe = 0; probe < ACC_TABLE; probe++, h = (h + 1) % ACC_TABLE) {
            const int32_t prev = atomicCAS(&keys[h], -1, b);
            if (prev == -1 || prev == b) {
                slot = h;
                break;
            }
        }
        if (slot >= 0) {
            for (int k = 0; k < NACC - 1; k++) acc_add(tab[slot], k, v[k]);
        } else {
            float *acc = a.acc + (int64_t)b * NACC;
            for (int k = 0; k < NACC - 1; k++) acc_add(acc, k, v[k]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ACC_TABLE * (NACC - 1); i += TPB) {
        const int s = i / (NACC - 1), k = i % (NACC - 1);
        const int32_t b = keys[s];
        if (b >= 0) acc_add(a.acc + (int64_t)b * NACC, k, tab[s][k]);
    }
}

struct FeatArgs {
    const float *acc;
    float *brokenPrev;     // [nBlocks]: broken sides at the previous evaluation
    float *feat;           // [nBlocks][NF] raw features
    float *blockDisp;      // [nBlocks] max |x - x0| (metres)
    uint8_t *valid;        // [nBlocks]
    int64_t nBlocks;
    float budget;          // (interactDist - attractDistance) / 2
    float stepsSince, stepsSincePrev;
    float waterLevel, blockEdge, xLo, xSpan;
};

// Raw features of one block (the network standardises them with the
// training-set mean and deviation): speeds and accelerations in units of the
// skin budget per step (log1p-compressed), displacement since the rebuild in
// budget units, recent and total broken-side fractions, steps since rebuild,
// occupancy, submerged fraction, rotation rate, a 50-step projection of the
// displacement and the block's position.
__global__ void __launch_bounds__(TPB) k_sched_features(FeatArgs a) {
    const int64_t b = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (b >= a.nBlocks) return;
    const float *c = a.acc + b * NACC;
    const float cnt = c[A_COUNT];
    float *f = a.feat + b * NF;
    a.blockDisp[b] = c[A_MAXDISP];
    if (!(cnt > 0.0f)) {
        for (int k = 0; k < NF; k++) f[k] = 0.0f;
        a.valid[b] = 0;
        a.brokenPrev[b] = 0.0f;
        return;
    }
    const float ib = 1.0f / a.budget, ic = 1.0f / cnt;
    const float sides = fmaxf(c[A_SIDES], 1.0f);
    const float recent = fmaxf(c[A_BROKEN] - a.brokenPrev[b], 0.0f) / sides * (50.0f / fmaxf(a.stepsSincePrev, 1.0f));
    f[0] = log1pf(1e3f * c[A_MAXD] * ib);
    f[1] = log1pf(1e3f * c[A_SUMD] * ic * ib);
    f[2] = log1pf(1e6f * c[A_MAXACC] * ib);
    f[3] = log1pf(1e6f * c[A_SUMACC] * ic * ib);
    f[4] = c[A_MAXDISP] * ib;
    f[5] = c[A_SUMDISP] * ic * ib;
    f[6] = log1pf(1e3f * c[A_MAXDZ] * ib);
    f[7] = 100.0f * recent;
    f[8] = c[A_BROKEN] / sides;
    f[9] = a.stepsSince * 1e-3f;
    f[10] = cnt * (1.0f / 150.0f);
    f[11] = c[A_BELOW] * ic;
    f[12] = log1pf(1e3f * c[A_MAXROT]);
    f[13] = c[A_MAXPROJ] * ib;
    f[14] = (c[A_SUMZ] * ic - a.waterLevel) / a.blockEdge;
    f[15] = (c[A_SUMX] * ic - a.xLo) / a.xSpan;
    a.valid[b] = 1;
    a.brokenPrev[b] = c[A_BROKEN];
}

// ---- MLP -----------------------------------------------------------------------
// Weight block layout (floats): W1[16][64], b1[64], W2[64][64], b2[64], w3[64],
// b3, featMean[16], featStd[16] -- the hdm_mlp_weights struct.
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NF * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH, OFF_MEAN = OFF_B3 + 1, OFF_STD = OFF_MEAN + NF,
              N_WEIGHTS = OFF_STD + NF, N_PARAMS = OFF_B3 + 1;
static_assert(N_WEIGHTS * sizeof(float) == sizeof(hdm_mlp_weights), "hdm_mlp_weights layout");

struct MlpTile {
    __bf16 X[32 * NF];     // standardised input rows
    __bf16 H1[32 * NH];    // relu(X W1 + b1)
    float H2[32 * NH];     // relu(H1 W2 + b2)
    __bf16 G2[32 * NH];    // dL/d(pre-activation 2)
    __bf16 G1[32 * NH];    // dL/d(pre-activation 1)
    float red[32 * 33];    // row reductions
    float dy[32];
};

// Forward pass of the 32 rows [row0, row0 + 32) of X in one wave; rows >= n
// are zero.  Leaves X, H1, H2 in the tile and returns y in red[0..31].
__device__ void mlp_forward_tile(const float *__restrict__ w, const __bf16 *__restrict__ w1b,
                                 const __bf16 *__restrict__ w2b, const float *__restrict__ X, int64_t row0, int64_t n,
                                 MlpTile &t) {
    const int l = threadIdx.x & 63;
    for (int e = l; e < 32 * NF; e += 64) {
        const int r = e / NF, k = e % NF;
        const int64_t g = row0 + r;
        float v = 0.0f;
        if (g < n) v = (X[g * NF + k] - w[OFF_MEAN + k]) / w[OFF_STD + k];
        t.X[e] = (__bf16)v;
    }
    __syncthreads();
    // layer 1: [32 x 16] x [16 x 64]
    bf16x8 a;
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = t.X[(l & 31) * NF + 8 * (l >> 5) + j];
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; j++) b[j] = w1b[(8 * (l >> 5) + j) * NH + nt * 32 + (l & 31)];
        f32x16 c = {};
        c = mfma32(a, b, c);
        const int col = nt * 32 + (l & 31);
        const float bias = w[OFF_B1 + col];
#pragma unroll
        for (int r = 0; r < 16; r++) t.H1[acc_row(r, l) * NH + col] = (__bf16)fmaxf(c[r] + bias, 0.0f);
    }
    __syncthreads();
    // layer 2: [32 x 64] x [64 x 64]
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
        f32x16 c = {};
#pragma unroll
        for (int kk = 0; kk < NH / 16; kk++) {
            bf16x8 aa, b;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                aa[j] = t.H1[(l & 31) * NH + kk * 16 + 8 * (l >> 5) + j];
                b[j] = w2b[(kk * 16 + 8 * (l >> 5) + j) * NH + nt * 32 + (l & 31)];
            }
            c = mfma32(aa, b, c);
        }
        const int col = nt * 32 + (l & 31);
        const float bias = w[OFF_B2 + col];
#pragma unroll
        for (int r = 0; r < 16; r++) t.H2[acc_row(r, l) * NH + col] = fmaxf(c[r] + bias, 0.0f);
    }
    __syncthreads();
    // output: y[r] = H2[r] . w3 + b3 (FP32), two lanes per row
    {
        const int r = l & 31, half = l >> 5;
        float s = 0.0f;
        for (int k = half * 32; k < half * 32 + 32; k++) s += t.H2[r * NH + k] * w[OFF_W3 + k];
        t.red[r * 33 + half] = s;
    }
    __syncthreads();
    if (l < 32) t.red[l * 33 + 32] = t.red[l * 33] + t.red[l * 33 + 1] + w[OFF_B3];
    __syncthreads();
    if (l < 32) t.dy[l] = t.red[l * 33 + 32];
    __syncthreads();
}

__global__ void __launch_bounds__(64) k_mlp_forward(const float *__restrict__ w, const __bf16 *__restrict__ w1b,
                                                    const __bf16 *__restrict__ w2b, const float *__restrict__ X,
                                                    int64_t n, float *__restrict__ y) {
    __shared__ MlpTile t;
    const int64_t row0 = (int64_t)blockIdx.x * 32;
    mlp_forward_tile(w, w1b, w2b, X, row0, n, t);
    const int l = threadIdx.x;
    if (l < 32 && row0 + l < n) y[row0 + l] = t.dy[l];
}

// Loss per row: squared error in log2 steps; a censored row (the true value
// is only known to exceed the target) is penalised only below its target.
__device__ __forceinline__ float row_loss_grad(float pred, float target, bool censored, float &grad) {
    const float e = pred - target;
    if (censored && e > 0.0f) {
        grad = 0.0f;
        return 0.0f;
    }
    grad = 2.0f * e;
    return e * e;
}

// Forward and backward of the rows [32 blockIdx.x, +32): per-wave gradient
// rows gpart[blockIdx.x][N_PARAMS] (W1, b1, W2, b2, w3, b3 order, summed over
// the tile's rows, not yet divided by the batch size) and the tile's loss sum.
__global__ void __launch_bounds__(64) k_mlp_backward(const float *__restrict__ w, const __bf16 *__restrict__ w1b,
                                                     const __bf16 *__restrict__ w2b, const float *__restrict__ X,
                                                     const float *__restrict__ target,
                                                     const uint8_t *__restrict__ censored, int64_t n,
                                                     float *__restrict__ gpart, float *__restrict__ lossPart) {
    __shared__ MlpTile t;
    const int64_t row0 = (int64_t)blockIdx.x * 32;
    mlp_forward_tile(w, w1b, w2b, X, row0, n, t);
    const int l = threadIdx.x;
    float *g = gpart + (int64_t)blockIdx.x * N_PARAMS;
    // dL/dy
    if (l < 32) {
        float gr = 0.0f, lo = 0.0f;
        if (row0 + l < n) lo = row_loss_grad(t.dy[l], target[row0 + l], censored && censored[row0 + l] != 0, gr);
        t.dy[l] = gr;
        t.red[l * 33] = lo;
    }
    __syncthreads();
    if (l == 0) {
        float s = 0.0f;
        for (int r = 0; r < 32; r++) s += t.red[r * 33];
        lossPart[blockIdx.x] = s;
    }
    // w3, b3, and G2 = dy w3^T masked by relu'(layer 2)
    {
        float s = 0.0f;
        for (int r = 0; r < 32; r++) s += t.H2[r * NH + l] * t.dy[r];
        g[OFF_W3 + l] = s;
        if (l == 0) {
            float sb = 0.0f;
            for (int r = 0; r < 32; r++) sb += t.dy[r];
            g[OFF_B3] = sb;
        }
        float sb2 = 0.0f;
        const float w3 = w[OFF_W3 + l];
        for (int r = 0; r < 32; r++) {
            const float d = t.H2[r * NH + l] > 0.0f ? t.dy[r] * w3 : 0.0f;
            t.G2[r * NH + l] = (__bf16)d;
            sb2 += d;
        }
        g[OFF_B2 + l] = sb2;
    }
    __syncthreads();
    // dW2 = H1^T G2: [64 x 32] x [32 x 64]
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            f32x16 c = {};
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                bf16x8 a, b;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int k = kk * 16 + 8 * (l >> 5) + j;   // sample row
                    a[j] = t.H1[k * NH + mt * 32 + (l & 31)];
                    b[j] = t.G2[k * NH + nt * 32 + (l & 31)];
                }
                c = mfma32(a, b, c);
            }
#pragma unroll
            for (int r = 0; r < 16; r++)
                g[OFF_W2 + (mt * 32 + acc_row(r, l)) * NH + nt * 32 + (l & 31)] = c[r];
        }
    // G1 = (G2 W2^T) masked by relu'(layer 1): [32 x 64] x [64 x 64]
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
        f32x16 c = {};
#pragma unroll
        for (int kk = 0; kk < NH / 16; kk++) {
            bf16x8 a, b;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = kk * 16 + 8 * (l >> 5) + j;        // layer-2 unit
                a[j] = t.G2[(l & 31) * NH + k];
                b[j] = w2b[(nt * 32 + (l & 31)) * NH + k];       // W2^T[k][i] = W2[i][k]
            }
            c = mfma32(a, b, c);
        }
        const int col = nt * 32 + (l & 31);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = acc_row(r, l);
            t.G1[row * NH + col] = (float)t.H1[row * NH + col] > 0.0f ? (__bf16)c[r] : (__bf16)0.0f;
            t.red[row * 33 + (l & 31)] = 0.0f;
        }
        // b1 partial: column sums of the masked FP32 values
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = acc_row(r, l);
            c[r] = (float)t.H1[row * NH + col] > 0.0f ? c[r] : 0.0f;
        }
        float s = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; r++) s += c[r];
        s += __shfl_xor(s, 32, 64);   // the other half-wave holds the other 16 rows of this column
        if (l < 32) g[OFF_B1 + col] = s;
    }
    __syncthreads();
    // dW1 = X^T G1: [16 x 32] x [32 x 64], rows 16..31 of the A tile are zero
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
        f32x16 c = {};
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            bf16x8 a, b;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = kk * 16 + 8 * (l >> 5) + j;   // sample row
                a[j] = (l & 31) < NF ? t.X[k * NF + (l & 31)] : (__bf16)0.0f;
                b[j] = t.G1[k * NH + nt * 32 + (l & 31)];
            }
            c = mfma32(a, b, c);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = acc_row(r, l);
            if (row < NF) g[OFF_W1 + row * NH + nt * 32 + (l & 31)] = c[r];
        }
    }
}

// grad[k] = sum over tiles of gpart[tile][k] in tile order (deterministic).
__global__ void __launch_bounds__(TPB) k_mlp_reduce(const float *__restrict__ gpart, int64_t tiles,
                                                    float *__restrict__ grad) {
    const int k = blockIdx.x * TPB + threadIdx.x;
    if (k >= N_PARAMS) return;
    float s = 0.0f;
    for (int64_t i = 0; i < tiles; i++) s += gpart[i * N_PARAMS + k];
    grad[k] = s;
}

// Adam on the FP32 master weights (gradient scaled by 1/batch), then the BF16
// copies of W1 and W2 the MFMA tiles read.
__global__ void __launch_bounds__(TPB) k_mlp_adam(float *__restrict__ w, const float *__restrict__ grad,
                                                  float *__restrict__ m, float *__restrict__ v, float scale, float lr,
                                                  float beta1, float beta2, float eps, float c1, float c2,
                                                  __bf16 *__restrict__ w1b, __bf16 *__restrict__ w2b) {
    const int k = blockIdx.x * TPB + threadIdx.x;
    if (k >= N_PARAMS) return;
    const float gk = grad[k] * scale;
    const float mk = beta1 * m[k] + (1.0f - beta1) * gk;
    const float vk = beta2 * v[k] + (1.0f - beta2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    w[k] = wk;
    if (k >= OFF_W1 && k < OFF_W1 + NF * NH) w1b[k - OFF_W1] = (__bf16)wk;
    if (k >= OFF_W2 && k < OFF_W2 + NH * NH) w2b[k - OFF_W2] = (__bf16)wk;
}

__global__ void __launch_bounds__(TPB) k_mlp_cast(const float *__restrict__ w, __bf16 *__restrict__ w1b,
                                                  __bf16 *__restrict__ w2b) {
    const int k = blockIdx.x * TPB + threadIdx.x;
    if (k < NF * NH) w1b[k] = (__bf16)w[OFF_W1 + k];
    if (k < NH * NH) w2b[k] = (__bf16)w[OFF_W2 + k];
}

// Training labels: per block, the first step k (since the reference was
// taken) at which some particle of the block has moved `remaining[b]` from its
// reference position; firstCross[b] = -1 until then.  Steps arrive in order,
// so the first CAS from -1 records the first crossing.
__global__ void __launch_bounds__(TPB) k_label_track(const double *__restrict__ x, const double *__restrict__ y,
                                                      const double *__restrict__ z, const double *__restrict__ ref,
                                                      const int32_t *__restrict__ blockOf,
                                                      const float *__restrict__ remaining, int32_t *firstCross,
                                                      int64_t n, int32_t k) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const int32_t b = blockOf[p];
    if (firstCross[b] >= 0) return;
    const double ex = x[p] - ref[3 * p], ey = y[p] - ref[3 * p + 1], ez = z[p] - ref[3 * p + 2];
    if (sqrt(ex * ex + ey * ey + ez * ez) >= (double)remaining[b]) atomicCAS(&firstCross[b], -1, k);
}

__global__ void __launch_bounds__(TPB) k_label_begin(const double *__restrict__ x, const double *__restrict__ y,
                                                      const double *__restrict__ z, const int32_t *__restrict__ blockOf,
                                                      int64_t n, double *__restrict__ ref,
                                                      int32_t *__restrict__ refBlockOf) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    ref[3 * p] = x[p];
    ref[3 * p + 1] = y[p];
    ref[3 * p + 2] = z[p];
    refBlockOf[p] = blockOf[p];
}

__global__ void __launch_bounds__(TPB) k_label_remaining(const float *__restrict__ blockDisp, int64_t nb, float budget,
                                                          float *__restrict__ remaining, int32_t *__restrict__ first) {
    const int64_t b = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (b >= nb) return;
    const float r = budget - blockDisp[b];
    remaining[b] = r > 0.0f ? r : 0.0f;
    first[b] = -1;
}

// Minimum prediction over valid blocks (log2 steps), as an ordered key.
__global__ void __launch_bounds__(TPB) k_min_valid(const float *__restrict__ pred, const uint8_t *__restrict__ valid,
                                                   int64_t nb, uint64_t *out) {
    double m = 1e30;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < nb; i += (int64_t)gridDim.x * TPB)
        if (valid[i]) m = fmin(m, (double)pred[i]);
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) atomicMin((unsigned long long *)out, (unsigned long long)ordered_key(m));
}

inline int blocks_for(int64_t n) { return (int)((n + TPB - 1) / TPB); }

}  // namespace

int64_t mlp_param_count() { return N_PARAMS; }

hipError_t sched_assign_blocks(Context &c, const SchedGeometry &g) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    hipError_t e = c.schedBlockOf.ensure(sizeof(int32_t) * (size_t)(c.n > 0 ? c.n : 1));
    if (e != hipSuccess || c.n == 0) return e;
    hipLaunchKernelGGL(k_sched_blocks, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                       c.y.as<double>(), c.z.as<double>(), c.n, g.ox, g.oy, g.oz, g.edge, g.BX, g.BY, g.BZ,
                       c.schedBlockOf.as<int32_t>());
    return hipGetLastError();
}

hipError_t sched_features(Context &c, const SchedGeometry &g, const SchedFeatureInputs &in) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    const int64_t nb = (int64_t)g.BX * g.BY * g.BZ;
    hipError_t e = c.schedAcc.ensure(sizeof(float) * NACC * (size_t)nb);
    if (e == hipSuccess) e = c.schedFeat.ensure(sizeof(float) * NF * (size_t)nb);
    if (e == hipSuccess) e = c.schedDisp.ensure(sizeof(float) * (size_t)nb);
    if (e == hipSuccess) e = c.schedValid.ensure((size_t)nb);
    if (e == hipSuccess) e = c.schedPred.ensure(sizeof(float) * (size_t)nb);
    if (e != hipSuccess) return e;
    if (c.schedBrokenPrev.bytes < sizeof(float) * (size_t)nb) {
        e = c.schedBrokenPrev.ensure(sizeof(float) * (size_t)nb);
        if (e != hipSuccess) return e;
        e = launch_fill_u32(c.stream, c.schedBrokenPrev.as<uint32_t>(), 0u, nb);
        if (e != hipSuccess) return e;
    }
    e = launch_fill_u32(c.stream, c.schedAcc.as<uint32_t>(), 0u, NACC * nb);
    if (e != hipSuccess) return e;
    AccArgs a;
    a.x = c.x.as<double>(); a.y = c.y.as<double>(); a.z = c.z.as<double>();
    a.x0 = c.x0.as<double>(); a.y0 = c.y0.as<double>(); a.z0 = c.z0.as<double>();
    a.d0 = c.diff[0].as<double>(); a.d1 = c.diff[1].as<double>(); a.d2 = c.diff[2].as<double>();
    a.d3 = c.diff[3].as<double>(); a.d4 = c.diff[4].as<double>(); a.d5 = c.diff[5].as<double>();
    const bool fp32 = c.prec != HDM_PREC_FP64;
    a.prev = c.altHoldsPrev ? c.drecAlt.as<double>() : nullptr;
    a.prevRec = c.altHoldsPrev && fp32 ? c.recAlt.as<MixedRec>() : nullptr;
    a.prevTail = c.altHoldsPrev && fp32 ? c.dtailAlt.as<MixedTail>() : nullptr;
    a.blockOf = c.schedBlockOf.as<int32_t>();
    a.linkCount = c.linkCount.as<uint32_t>();
    a.slotLink = c.slotLink.as<int32_t>();
    a.linkWidth = c.maxLinks;
    a.n = c.n;
    a.waterLevel = in.waterLevel;
    a.acc = c.schedAcc.as<float>();
    if (c.n > 0) hipLaunchKernelGGL(k_sched_accumulate, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, a);
    FeatArgs f;
    f.acc = c.schedAcc.as<float>();
    f.brokenPrev = c.schedBrokenPrev.as<float>();
    f.feat = c.schedFeat.as<float>();
    f.blockDisp = c.schedDisp.as<float>();
    f.valid = c.schedValid.as<uint8_t>();
    f.nBlocks = nb;
    f.budget = (float)in.budget;
    f.stepsSince = (float)in.stepsSince;
    f.stepsSincePrev = (float)in.stepsSincePrev;
    f.waterLevel = (float)in.waterLevel;
    f.blockEdge = (float)g.edge;
    f.xLo = (float)g.ox;
    f.xSpan = (float)(g.edge * g.BX);
    hipLaunchKernelGGL(k_sched_features, dim3(blocks_for(nb)), dim3(TPB), 0, c.stream, f);
    return hipGetLastError();
}

hipError_t mlp_upload(Context &c, const hdm_mlp_weights &w) {
    hipError_t e = c.mlpW.ensure(sizeof(hdm_mlp_weights));
    if (e == hipSuccess) e = c.mlpW1b.ensure(sizeof(__bf16) * NF * NH);
    if (e == hipSuccess) e = c.mlpW2b.ensure(sizeof(__bf16) * NH * NH);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(c.mlpW.ptr, &w, sizeof(hdm_mlp_weights), hipMemcpyHostToDevice, c.stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_cast, dim3(blocks_for(NH * NH)), dim3(TPB), 0, c.stream, c.mlpW.as<float>(),
                       c.mlpW1b.as<__bf16>(), c.mlpW2b.as<__bf16>());
    return hipGetLastError();
}

hipError_t mlp_forward(Context &c, const float *X, int64_t n, float *y) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3((int)((n + 31) / 32)), dim3(64), 0, c.stream, c.mlpW.as<float>(),
                       c.mlpW1b.as<__bf16>(), c.mlpW2b.as<__bf16>(), X, n, y);
    return hipGetLastError();
}

hipError_t mlp_gradient(Context &c, const float *X, const float *target, const uint8_t *censored, int64_t n,
                        float *grad, float *lossPart) {
    const int64_t tiles = (n + 31) / 32;
    hipError_t e = c.mlpGpart.ensure(sizeof(float) * N_PARAMS * (size_t)(tiles > 0 ? tiles : 1));
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(k_mlp_backward, dim3((int)tiles), dim3(64), 0, c.stream, c.mlpW.as<float>(),
                       c.mlpW1b.as<__bf16>(), c.mlpW2b.as<__bf16>(), X, target, censored, n, c.mlpGpart.as<float>(),
                       lossPart);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks_for(N_PARAMS)), dim3(TPB), 0, c.stream, c.mlpGpart.as<float>(),
                       tiles, grad);
    return hipGetLastError();
}

hipError_t mlp_adam(Context &c, const float *grad, float *m, float *v, float scale, float lr, float beta1,
                    float beta2, float eps, int64_t t) {
    const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(N_PARAMS)), dim3(TPB), 0, c.stream, c.mlpW.as<float>(), grad,
                       m, v, scale, lr, beta1, beta2, eps, c1, c2, c.mlpW1b.as<__bf16>(), c.mlpW2b.as<__bf16>());
    return hipGetLastError();
}

hipError_t sched_label_begin(Context &c, int slot, double budget) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    const int64_t n = c.n, nb = c.schedNb;
    LabelSlot &s = c.labels[slot];
    hipError_t e = s.ref.ensure(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    if (e == hipSuccess) e = s.blockOf.ensure(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (e == hipSuccess) e = s.remaining.ensure(sizeof(float) * (size_t)(nb > 0 ? nb : 1));
    if (e == hipSuccess) e = s.first.ensure(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
    if (e != hipSuccess) return e;
    if (n > 0)
        hipLaunchKernelGGL(k_label_begin, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                           c.y.as<double>(), c.z.as<double>(), c.schedBlockOf.as<int32_t>(), n, s.ref.as<double>(),
                           s.blockOf.as<int32_t>());
    if (nb > 0)
        hipLaunchKernelGGL(k_label_remaining, dim3(blocks_for(nb)), dim3(TPB), 0, c.stream, c.schedDisp.as<float>(), nb,
                           (float)budget, s.remaining.as<float>(), s.first.as<int32_t>());
    s.nb = (int32_t)nb;
    s.start = c.ts;
    s.active = true;
    return hipGetLastError();
}

hipError_t sched_label_track(Context &c) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    for (LabelSlot &s : c.labels) {
        if (!s.active || c.n == 0) continue;
        hipLaunchKernelGGL(k_label_track, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                           c.y.as<double>(), c.z.as<double>(), s.ref.as<double>(), s.blockOf.as<int32_t>(),
                           s.remaining.as<float>(), s.first.as<int32_t>(), c.n, (int32_t)(c.ts - s.start));
    }
    return hipGetLastError();
}

hipError_t sched_min_prediction(Context &c, int64_t nb, uint64_t *out) {
    hipError_t e = launch_fill_u64(c.stream, out, ordered_key(1e30), 1);
    if (e != hipSuccess || nb == 0) return e;
    int grid = blocks_for(nb);
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(k_min_valid, dim3(grid), dim3(TPB), 0, c.stream, c.schedPred.as<float>(),
                       c.schedValid.as<uint8_t>(), nb, out);
    return hipGetLastError();
}

}  // namespace hdm
