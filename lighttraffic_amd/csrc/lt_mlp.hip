// lt_mlp.hip — training and inference kernels of the learned partition scorer.
//
// Layers (row-major, input-major weights: y[j] = sum_k x[k] W[k][j] + b[j]):
//   X[M x 16] -> A1 = relu(X W1 + b1) [M x 256] -> A2 = relu(A1 W2 + b2)
//   -> Y = A2 W3 + b3 [M x 2],   M = (decisions in a batch) x P partitions.
// Weights are kept as FP32 masters for Adam and used as their BF16 rounding
// (round to nearest even) in every product; products accumulate in FP32.
// Losses per decision: softmax cross-entropy of output 0 over the partitions
// holding walkers against the partition the plain scheduler picked next, and
// mean squared error of output 1 against the partition's future traffic share.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <random>
#include <vector>

#include "lt_internal.h"
#include "lt_mlp.h"

#define LT_NA_TRY(expr)             \
    do {                            \
        lt_status _s = (expr);      \
        if (_s != LT_OK) return _s; \
    } while (0)

#define MLP_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            lt::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e));            \
            return LT_E_HIP;                                                                           \
        }                                                                                              \
    } while (0)

namespace lt {

// ---- parameter layout ------------------------------------------------------
constexpr size_t kOffW1 = 0;
constexpr size_t kOffB1 = kOffW1 + (size_t)kFeat * kHidden;
constexpr size_t kOffW2 = kOffB1 + kHidden;
constexpr size_t kOffB2 = kOffW2 + (size_t)kHidden * kHidden;
constexpr size_t kOffW3 = kOffB2 + kHidden;
constexpr size_t kOffB3 = kOffW3 + (size_t)kHidden * kOut;
constexpr size_t kParams = kOffB3 + kOut;

__host__ __device__ inline uint16_t f32_to_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float to_f32(float x) { return x; }
__device__ __forceinline__ float to_f32(uint16_t h) { return bf16_to_f32(h); }

// ---- features ----------------------------------------------------------------
// One workgroup per decision.  Feature j of partition q:
//  0 log1p(c)/log1p(W)   1 c/max c   2 log1p(c/mean c)   3 resident
//  4 chosen now          5..11 loaded k decisions ago (k = 0..6)
//  12 a/(a+P), a = HBM walks since q's last one (recency: Belady's victim is
//     the partition needed farthest ahead, and on walk schedules that is
//     mostly the one walked longest ago)
//  13 vertex share * P   14 edge share * P   15 queued walkers / W
constexpr int kRecentFeat = 7;  // recent loads that become features (the state holds kRecent)

__device__ __forceinline__ float age_feature(uint32_t a, uint32_t P) { return (float)a / ((float)a + (float)P); }

__global__ __launch_bounds__(256) void feat_kernel(uint32_t P, const uint32_t *counts, const uint8_t *resident,
                                                   const uint32_t *age, const int32_t *recent, const int32_t *cur,
                                                   const float *w_total, const float *vsh, const float *esh,
                                                   float *X) {
    const uint32_t b = blockIdx.x;
    const uint32_t *c = counts + (size_t)b * P;
    const uint8_t *r = resident + (size_t)b * P;
    const uint32_t *ag = age + (size_t)b * P;
    __shared__ float s_max[256], s_sum[256];
    float mx = 0.f, sm = 0.f;
    for (uint32_t q = threadIdx.x; q < P; q += blockDim.x) {
        const float v = (float)c[q];
        mx = fmaxf(mx, v);
        sm += v;
    }
    s_max[threadIdx.x] = mx;
    s_sum[threadIdx.x] = sm;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s_max[threadIdx.x] = fmaxf(s_max[threadIdx.x], s_max[threadIdx.x + off]);
            s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
        }
        __syncthreads();
    }
    const float cmax = fmaxf(s_max[0], 1.f), csum = s_sum[0];
    const float W = fmaxf(w_total[b], 1.f);
    const float mean = fmaxf(csum / (float)P, 1.f);
    const float lw = log1pf(W);
    for (uint32_t q = threadIdx.x; q < P; q += blockDim.x) {
        const float v = (float)c[q];
        float *x = X + ((size_t)b * P + q) * kFeat;
        x[0] = log1pf(v) / lw;
        x[1] = v / cmax;
        x[2] = log1pf(v / mean);
        x[3] = r[q] ? 1.f : 0.f;
        x[4] = cur[b] == (int32_t)q ? 1.f : 0.f;
#pragma unroll
        for (int k = 0; k < kRecentFeat; k++) x[5 + k] = recent[b * kRecent + k] == (int32_t)q ? 1.f : 0.f;
        x[12] = age_feature(ag[q], P);
        x[13] = vsh[q];
        x[14] = esh[q];
        x[15] = csum / W;
    }
}

// ---- tiled FP32-accumulate GEMM ------------------------------------------------
// C[M x N] = op(A)[M x K] * op(B)[K x N], 64x64 tiles, 256 threads, 4x4 per thread.
// op(A)[m][k] = TA ? A[k*lda+m] : A[m*lda+k];  op(B)[k][n] = TB ? B[n*ldb+k] : B[k*ldb+n].
// EPI 0: store; 1: relu(acc + bias[n]); 2: acc * (mask[m][n] > 0); 3: acc + bias[n].
enum { EPI_STORE = 0, EPI_BIAS_RELU = 1, EPI_RELU_GRAD = 2, EPI_BIAS = 3 };

// With gridDim.z > 1 the K range is split: slice z covers [z*kchunk, (z+1)*kchunk)
// and stores its partial product at C + z*M*ldc (summed by splitk_reduce_kernel).
template <bool TA, bool TB, typename BT, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const float *A, int lda, const BT *B, int ldb,
                                                   float *C, int ldc, const BT *bias, const float *mask, int ldm,
                                                   int kchunk) {
    constexpr int TM = 64, TN = 64, TK = 16;
    __shared__ float As[TK][TM + 4];
    __shared__ float Bs[TK][TN + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kb = blockIdx.z * kchunk;
    const int ke = min(K, kb + kchunk);
    C += (size_t)blockIdx.z * M * ldc;
    float acc[4][4] = {};
    for (int k0 = kb; k0 < ke; k0 += TK) {
        for (int i = threadIdx.x; i < TM * TK; i += 256) {
            int mm, kk;
            if (TA) { mm = i % TM; kk = i / TM; } else { kk = i % TK; mm = i / TK; }
            const int m = m0 + mm, k = k0 + kk;
            float v = 0.f;
            if (m < M && k < ke) v = TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
            As[kk][mm] = v;
        }
        for (int i = threadIdx.x; i < TN * TK; i += 256) {
            int nn, kk;
            if (TB) { kk = i % TK; nn = i / TK; } else { nn = i % TN; kk = i / TN; }
            const int n = n0 + nn, k = k0 + kk;
            float v = 0.f;
            if (n < N && k < ke) v = to_f32(TB ? B[(size_t)n * ldb + k] : B[(size_t)k * ldb + n]);
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk++) {
            float a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) bb[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], bb[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = m0 + ty * 4 + i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + tx * 4 + j;
            if (n >= N) continue;
            float v = acc[i][j];
            if (EPI == EPI_BIAS_RELU) v = fmaxf(v + to_f32(bias[n]), 0.f);
            else if (EPI == EPI_BIAS) v = v + to_f32(bias[n]);
            else if (EPI == EPI_RELU_GRAD) v = mask[(size_t)m * ldm + n] > 0.f ? v : 0.f;
            C[(size_t)m * ldc + n] = v;
        }
    }
}

template <bool TA, bool TB, typename BT, int EPI>
static void gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const BT *B, int ldb, float *C, int ldc,
                 const BT *bias = nullptr, const float *mask = nullptr, int ldm = 0) {
    dim3 grid((N + 63) / 64, (M + 63) / 64, 1);
    gemm_kernel<TA, TB, BT, EPI><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, mask, ldm, K);
}

// out[i] = sum_z part[z*n + i], fixed order (deterministic).
__global__ void splitk_reduce_kernel(size_t n, int splits, const float *part, float *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < splits; z++) s += part[(size_t)z * n + i];
    out[i] = s;
}

// Weight gradient C[M x N] = A^T B with a long K (rows of the batch): K is cut
// into slices of kSplitK, each slice writes a partial product, then a fixed-order sum.
constexpr int kSplitK = 256;
static void gemm_wgrad(hipStream_t s, int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
                       float *partial) {
    const int splits = (K + kSplitK - 1) / kSplitK;
    dim3 grid((N + 63) / 64, (M + 63) / 64, splits);
    gemm_kernel<true, false, float, EPI_STORE><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, partial, N, nullptr,
                                                                   nullptr, 0, kSplitK);
    const size_t n = (size_t)M * N;
    splitk_reduce_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, splits, partial, C);
}

// ---- losses -----------------------------------------------------------------
// One workgroup per decision; writes dY and adds the weighted losses to loss[0..1].
__global__ __launch_bounds__(256) void loss_kernel(uint32_t P, uint32_t B, const float *Y, const uint32_t *counts,
                                                   const int32_t *next_label, const float *keep_label,
                                                   const float *wa, const float *wb, float *dY, float *loss) {
    const uint32_t b = blockIdx.x;
    const float *y = Y + (size_t)b * P * kOut;
    float *dy = dY + (size_t)b * P * kOut;
    const uint32_t *c = counts + (size_t)b * P;
    const int32_t lab = next_label[b];
    const float a_w = lab >= 0 ? wa[b] : 0.f, b_w = wb[b];
    __shared__ float s_red[256];
    // softmax over candidates (walkers queued, or the label itself)
    float mx = -INFINITY;
    for (uint32_t q = threadIdx.x; q < P; q += blockDim.x)
        if (c[q] > 0 || (int32_t)q == lab) mx = fmaxf(mx, y[q * kOut]);
    s_red[threadIdx.x] = mx;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s_red[threadIdx.x] = fmaxf(s_red[threadIdx.x], s_red[threadIdx.x + off]);
        __syncthreads();
    }
    mx = s_red[0];
    __syncthreads();
    float se = 0.f;
    for (uint32_t q = threadIdx.x; q < P; q += blockDim.x)
        if (c[q] > 0 || (int32_t)q == lab) se += expf(y[q * kOut] - mx);
    s_red[threadIdx.x] = se;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
        __syncthreads();
    }
    se = s_red[0];
    __syncthreads();
    float lb = 0.f;
    const float inv_b = 1.f / (float)B;
    for (uint32_t q = threadIdx.x; q < P; q += blockDim.x) {
        const bool cand = c[q] > 0 || (int32_t)q == lab;
        float g0 = 0.f;
        if (cand && a_w > 0.f && se > 0.f) {
            const float p = expf(y[q * kOut] - mx) / se;
            g0 = a_w * (p - ((int32_t)q == lab ? 1.f : 0.f)) * inv_b;
        }
        const float d = y[q * kOut + 1] - keep_label[(size_t)b * P + q];
        lb += b_w * d * d;
        dy[q * kOut] = g0;
        dy[q * kOut + 1] = b_w * 2.f * d * inv_b / (float)P;
    }
    s_red[threadIdx.x] = lb;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (a_w > 0.f && se > 0.f) atomicAdd(&loss[0], a_w * (logf(se) + mx - y[lab * kOut]) * inv_b);
        atomicAdd(&loss[1], s_red[0] * inv_b / (float)P);
    }
}

// db[n] = sum_m dZ[m][n]: slice z of the rows sums into part[z*N + n]
// (reduced afterwards in fixed order).
__global__ void colsum_kernel(int M, int N, const float *dZ, float *part, int rows_per_slice) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int m0 = blockIdx.y * rows_per_slice, m1 = min(M, m0 + rows_per_slice);
    float s = 0.f;
    for (int m = m0; m < m1; m++) s += dZ[(size_t)m * N + n];
    part[(size_t)blockIdx.y * N + n] = s;
}

static void colsum(hipStream_t s, int M, int N, const float *dZ, float *out, float *partial) {
    constexpr int kRows = 64;  // rows per slice: many short slices keep every CU busy
    const int splits = (M + kRows - 1) / kRows;
    colsum_kernel<<<dim3((N + 255) / 256, splits), 256, 0, s>>>(M, N, dZ, partial, kRows);
    splitk_reduce_kernel<<<(N + 255) / 256, 256, 0, s>>>((size_t)N, splits, partial, out);
}

__global__ void adam_kernel(size_t n, float *W, uint16_t *Wb, const float *G, float *Mo, float *Vo, float lr,
                            float b1t, float b2t) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float m = 0.9f * Mo[i] + 0.1f * g;
    const float v = 0.999f * Vo[i] + 0.001f * g * g;
    Mo[i] = m;
    Vo[i] = v;
    const float w = W[i] - lr * (m / (1.f - b1t)) / (sqrtf(v / (1.f - b2t)) + 1e-8f);
    W[i] = w;
    Wb[i] = f32_to_bf16(w);
}

__global__ void round_kernel(size_t n, const float *W, uint16_t *Wb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Wb[i] = f32_to_bf16(W[i]);
}

// ---- fused inference: one decision, 8 partitions per workgroup --------------
// Inference copy of the first two layers, output-major: Wi[j * kFeat + k] =
// W1[k][j] and Wi[kWiW2 + j * kHidden + k] = W2[k][j], so that the thread of
// output j reads its weights with 16-byte loads.
constexpr size_t kWiW2 = (size_t)kHidden * kFeat;
constexpr size_t kWiSize = kWiW2 + (size_t)kHidden * kHidden;

__global__ void transpose_infer_weights_kernel(const uint16_t *Wb, uint16_t *Wi) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kWiW2) {
        const size_t j = i / kFeat, k = i % kFeat;
        Wi[i] = Wb[kOffW1 + k * kHidden + j];
    } else if (i < kWiSize) {
        const size_t t = i - kWiW2, j = t / kHidden, k = t % kHidden;
        Wi[i] = Wb[kOffW2 + k * kHidden + j];
    }
}

// The decision state infer_kernel reads, packed in one buffer (host memory
// mapped for the device, or device memory):
//   i32 recent[8] | i32 cur | f32 w_total | f32 max count | f32 count sum |
//   pad to kStateHead | u32 counts[P] | u32 age[P] | u8 resident[P]
// The host supplies the max and the sum of the counts, so a workgroup reads
// only the header and its own rows: the launch reads O(P) bytes of the state,
// and LDS use does not grow with P.
constexpr size_t kStateHead = 64;
constexpr int kHeadCur = kRecent, kHeadW = kRecent + 1, kHeadMax = kRecent + 2, kHeadSum = kRecent + 3;
static_assert(4 * (kHeadSum + 1) <= (int)kStateHead, "state header");
__host__ __device__ constexpr size_t state_bytes(uint32_t P) { return kStateHead + 9 * (size_t)P; }

// Header of one decision's packed state.
static void pack_state_head(int32_t *head, const uint32_t *counts, uint32_t P, const int32_t *recent, int32_t cur,
                            float w_total) {
    uint32_t mx = 0;
    double sum = 0.0;
    for (uint32_t q = 0; q < P; q++) {
        mx = std::max(mx, counts[q]);
        sum += counts[q];
    }
    memset(head, 0, kStateHead);
    for (int k = 0; k < kRecent; k++) head[k] = recent[k];
    head[kHeadCur] = cur;
    const float f[3] = {w_total, (float)mx, (float)sum};
    memcpy(&head[kHeadW], f, sizeof(f));
}

// Each workgroup builds the features of 8 partitions (rows) and runs the
// three layers on them.  Each output neuron accumulates its inputs in
// increasing order in FP32.
__global__ __launch_bounds__(256) void infer_kernel(uint32_t P, const uint8_t *state, const float *vsh,
                                                    const float *esh, const uint16_t *Wb, const uint16_t *Wi,
                                                    float *Y) {
    constexpr int R = 8;
    __shared__ float s_x[R][kFeat];
    __shared__ float s_h1[R][kHidden];
    __shared__ float s_h2[R][kHidden];
    const int j = threadIdx.x;
    const int32_t *head = (const int32_t *)state;
    const uint32_t *counts = (const uint32_t *)(state + kStateHead);
    const uint32_t *age = counts + P;
    const uint8_t *res = state + kStateHead + 8 * (size_t)P;
    const uint32_t r0 = blockIdx.x * R;
    if (j < R) {
        const uint32_t q = r0 + j;
        float *x = s_x[j];
        if (q < P) {
            const float W = fmaxf(__int_as_float(head[kHeadW]), 1.f);
            const float cmax = fmaxf(__int_as_float(head[kHeadMax]), 1.f), csum = __int_as_float(head[kHeadSum]);
            const float mean = fmaxf(csum / (float)P, 1.f);
            const float v = (float)counts[q];
            x[0] = log1pf(v) / log1pf(W);
            x[1] = v / cmax;
            x[2] = log1pf(v / mean);
            x[3] = res[q] ? 1.f : 0.f;
            x[4] = head[kHeadCur] == (int32_t)q ? 1.f : 0.f;
            for (int k = 0; k < kRecentFeat; k++) x[5 + k] = head[k] == (int32_t)q ? 1.f : 0.f;
            x[12] = age_feature(age[q], P);
            x[13] = vsh[q];
            x[14] = esh[q];
            x[15] = csum / W;
        } else {
            for (int k = 0; k < kFeat; k++) x[k] = 0.f;
        }
    }
    __syncthreads();
    const uint16_t *B1 = Wb + kOffB1, *B2 = Wb + kOffB2, *W3 = Wb + kOffW3, *B3 = Wb + kOffB3;
    float a[R];
#pragma unroll
    for (int r = 0; r < R; r++) a[r] = 0.f;
    {
        const uint4 *w1 = (const uint4 *)(Wi + (size_t)j * kFeat);
        uint4 wv[kFeat / 8];
#pragma unroll
        for (int c = 0; c < kFeat / 8; c++) wv[c] = w1[c];
#pragma unroll
        for (int k = 0; k < kFeat; k++) {
            const uint32_t word = (&wv[k / 8].x)[(k % 8) / 2];
            const float w = __uint_as_float((k & 1) ? (word & 0xFFFF0000u) : (word << 16));
#pragma unroll
            for (int r = 0; r < R; r++) a[r] = fmaf(s_x[r][k], w, a[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) s_h1[r][j] = fmaxf(a[r] + bf16_to_f32(B1[j]), 0.f);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) a[r] = 0.f;
    const uint4 *w2 = (const uint4 *)(Wi + kWiW2 + (size_t)j * kHidden);
    for (int k0 = 0; k0 < kHidden; k0 += 64) {
        uint4 wv[8];
#pragma unroll
        for (int c = 0; c < 8; c++) wv[c] = w2[k0 / 8 + c];
#pragma unroll
        for (int kk = 0; kk < 64; kk++) {
            const uint32_t word = (&wv[kk / 8].x)[(kk % 8) / 2];
            const float w = __uint_as_float((kk & 1) ? (word & 0xFFFF0000u) : (word << 16));
#pragma unroll
            for (int r = 0; r < R; r++) a[r] = fmaf(s_h1[r][k0 + kk], w, a[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) s_h2[r][j] = fmaxf(a[r] + bf16_to_f32(B2[j]), 0.f);
    __syncthreads();
    const int lane = j & 63, wave = j >> 6;
    for (int t = wave; t < R * kOut; t += 4) {
        const int r = t / kOut, o = t % kOut;
        float s = 0.f;
        for (int k = lane; k < kHidden; k += 64) s = fmaf(s_h2[r][k], bf16_to_f32(W3[k * kOut + o]), s);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0 && r0 + r < P) Y[(size_t)(r0 + r) * kOut + o] = s + bf16_to_f32(B3[o]);
    }
}

// ---- host object ---------------------------------------------------------------
template <typename T>
static lt_status dalloc(T **p, size_t n) {
    if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc of %zu bytes failed", sizeof(T) * n);
        return LT_E_NOMEM;
    }
    return LT_OK;
}

struct Scorer {
    uint32_t P = 0;
    hipStream_t stream = nullptr;
    float *d_vsh = nullptr, *d_esh = nullptr;
    std::vector<float> h_vsh, h_esh;
    float *W = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    uint16_t *Wb = nullptr;
    uint16_t *Wi = nullptr;     // inference copy of W1, W2 (output-major)
    uint64_t wver = 0, wi_ver = ~0ull;  // bumped on every weight update; Wi is valid for wi_ver
    uint64_t uid = 0;                    // unique per scorer created in this process
    uint64_t adam_t = 0;
    bool trained = false;
    // training buffers (capacity in decisions)
    uint32_t cap = 0;
    uint32_t *d_counts = nullptr, *d_age = nullptr;
    uint8_t *d_res = nullptr;
    int32_t *d_recent = nullptr, *d_cur = nullptr, *d_next = nullptr;
    float *d_wtot = nullptr, *d_keep = nullptr, *d_wa = nullptr, *d_wb = nullptr;
    float *X = nullptr, *A1 = nullptr, *A2 = nullptr, *Y = nullptr, *dY = nullptr, *dA1 = nullptr, *dA2 = nullptr;
    float *partial = nullptr;  // split-K partial products
    float *d_loss = nullptr;
    // inference buffers
    uint8_t *h_stage = nullptr;  // one decision's packed state, pinned and mapped (see infer_kernel)
    float *h_Y = nullptr;        // its scores, written by the kernel into pinned mapped memory
    uint8_t *d_stage = nullptr;  // device views of h_stage and h_Y
    float *d_Y = nullptr;
    hipStream_t pending = nullptr;  // stream of the launched inference not yet waited for
    Samples data;

    ~Scorer();
    lt_status init(uint32_t P_, const uint32_t *node_off, const uint64_t *edge_off, uint64_t seed);
    lt_status ensure_cap(uint32_t n);
    lt_status forward_backward(uint32_t nb, bool update, float lr);
};

Scorer::~Scorer() {
    if (pending) (void)hipStreamSynchronize(pending);
    if (stream) (void)hipStreamSynchronize(stream);
    void *ptrs[] = {d_vsh, d_esh, W, G, Mo, Vo, Wb, Wi, d_counts, d_age, d_res, d_recent, d_cur, d_next, d_wtot, d_keep, d_wa,
                    d_wb, X, A1, A2, Y, dY, dA1, dA2, partial, d_loss};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (h_stage) (void)hipHostFree(h_stage);
    if (h_Y) (void)hipHostFree(h_Y);
    if (stream) (void)hipStreamDestroy(stream);
}

static std::atomic<uint64_t> g_scorer_uid{0};

lt_status Scorer::init(uint32_t P_, const uint32_t *node_off, const uint64_t *edge_off, uint64_t seed) {
    P = P_;
    uid = ++g_scorer_uid;
    MLP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const double V = std::max<double>(1.0, node_off[P]), E = std::max<double>(1.0, (double)edge_off[P]);
    h_vsh.resize(P);
    h_esh.resize(P);
    for (uint32_t q = 0; q < P; q++) {
        h_vsh[q] = (float)((node_off[q + 1] - node_off[q]) / V * P);
        h_esh[q] = (float)((edge_off[q + 1] - edge_off[q]) / E * P);
    }
    lt_status st;
    if ((st = dalloc(&d_vsh, P)) || (st = dalloc(&d_esh, P))) return st;
    MLP_HIP(hipMemcpy(d_vsh, h_vsh.data(), 4 * P, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(d_esh, h_esh.data(), 4 * P, hipMemcpyHostToDevice));
    if ((st = dalloc(&W, kParams)) || (st = dalloc(&G, kParams)) || (st = dalloc(&Mo, kParams)) ||
        (st = dalloc(&Vo, kParams)) || (st = dalloc(&Wb, kParams)) || (st = dalloc(&Wi, kWiSize)) ||
        (st = dalloc(&d_loss, 2)))
        return st;
    // He-style initialisation from a seeded generator; biases zero.
    std::vector<float> w(kParams, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, int fan_in, size_t n) {
        std::normal_distribution<float> nd(0.f, sqrtf(2.f / (float)fan_in));
        for (size_t i = 0; i < n; i++) w[off + i] = nd(rng);
    };
    fill(kOffW1, kFeat, (size_t)kFeat * kHidden);
    fill(kOffW2, kHidden, (size_t)kHidden * kHidden);
    fill(kOffW3, kHidden, (size_t)kHidden * kOut);
    for (size_t i = 0; i < (size_t)kHidden * kOut; i++) w[kOffW3 + i] *= 0.1f;
    MLP_HIP(hipMemcpy(W, w.data(), 4 * kParams, hipMemcpyHostToDevice));
    MLP_HIP(hipMemset(Mo, 0, 4 * kParams));
    MLP_HIP(hipMemset(Vo, 0, 4 * kParams));
    round_kernel<<<(kParams + 255) / 256, 256, 0, stream>>>(kParams, W, Wb);
    MLP_HIP(hipGetLastError());
    wver++;
    MLP_HIP(hipHostMalloc((void **)&h_stage, state_bytes(P), hipHostMallocMapped | hipHostMallocCoherent));
    MLP_HIP(hipHostMalloc((void **)&h_Y, sizeof(float) * P * kOut, hipHostMallocMapped | hipHostMallocCoherent));
    MLP_HIP(hipHostGetDevicePointer((void **)&d_stage, h_stage, 0));
    MLP_HIP(hipHostGetDevicePointer((void **)&d_Y, h_Y, 0));
    MLP_HIP(hipStreamSynchronize(stream));
    return LT_OK;
}

lt_status Scorer::ensure_cap(uint32_t n) {
    if (n <= cap) return LT_OK;
    void *ptrs[] = {d_counts, d_age, d_res, d_recent, d_cur, d_next, d_wtot, d_keep, d_wa, d_wb, X, A1, A2, Y, dY, dA1, dA2,
                    partial};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    const size_t rows = (size_t)n * P;
    lt_status st;
    if ((st = dalloc(&d_counts, rows)) || (st = dalloc(&d_age, rows)) || (st = dalloc(&d_res, rows)) || (st = dalloc(&d_recent, (size_t)n * kRecent)) ||
        (st = dalloc(&d_cur, n)) || (st = dalloc(&d_next, n)) || (st = dalloc(&d_wtot, n)) ||
        (st = dalloc(&d_keep, rows)) || (st = dalloc(&d_wa, n)) || (st = dalloc(&d_wb, n)) ||
        (st = dalloc(&X, rows * kFeat)) || (st = dalloc(&A1, rows * kHidden)) || (st = dalloc(&A2, rows * kHidden)) ||
        (st = dalloc(&Y, rows * kOut)) || (st = dalloc(&dY, rows * kOut)) || (st = dalloc(&dA1, rows * kHidden)) ||
        (st = dalloc(&dA2, rows * kHidden)) ||
        (st = dalloc(&partial, std::max(((rows + kSplitK - 1) / kSplitK) * (size_t)kHidden * kHidden,
                                         ((rows + 63) / 64) * (size_t)kHidden))))
        return st;
    cap = n;
    return LT_OK;
}

// Forward + backward over the first nb decisions of the device buffers; G
// receives the gradient, d_loss the two losses; with update, one Adam step.
lt_status Scorer::forward_backward(uint32_t nb, bool update, float lr) {
    const int M = (int)(nb * P);
    const int H = kHidden;
    hipStream_t s = stream;
    MLP_HIP(hipMemsetAsync(d_loss, 0, 8, s));
    feat_kernel<<<nb, 256, 0, s>>>(P, d_counts, d_res, d_age, d_recent, d_cur, d_wtot, d_vsh, d_esh, X);
    // forward with the BF16 weights
    gemm<false, false, uint16_t, EPI_BIAS_RELU>(s, M, H, kFeat, X, kFeat, Wb + kOffW1, H, A1, H, Wb + kOffB1);
    gemm<false, false, uint16_t, EPI_BIAS_RELU>(s, M, H, H, A1, H, Wb + kOffW2, H, A2, H, Wb + kOffB2);
    gemm<false, false, uint16_t, EPI_BIAS>(s, M, kOut, H, A2, H, Wb + kOffW3, kOut, Y, kOut, Wb + kOffB3);
    loss_kernel<<<nb, 256, 0, s>>>(P, nb, Y, d_counts, d_next, d_keep, d_wa, d_wb, dY, d_loss);
    // backward
    gemm_wgrad(s, H, kOut, M, A2, H, dY, kOut, G + kOffW3, partial);
    colsum(s, M, kOut, dY, G + kOffB3, partial);
    gemm<false, true, uint16_t, EPI_RELU_GRAD>(s, M, H, kOut, dY, kOut, Wb + kOffW3, kOut, dA2, H, nullptr, A2, H);
    gemm_wgrad(s, H, H, M, A1, H, dA2, H, G + kOffW2, partial);
    colsum(s, M, H, dA2, G + kOffB2, partial);
    gemm<false, true, uint16_t, EPI_RELU_GRAD>(s, M, H, H, dA2, H, Wb + kOffW2, H, dA1, H, nullptr, A1, H);
    gemm_wgrad(s, kFeat, H, M, X, kFeat, dA1, H, G + kOffW1, partial);
    colsum(s, M, H, dA1, G + kOffB1, partial);
    MLP_HIP(hipGetLastError());
    if (update) {
        adam_t++;
        const float b1t = powf(0.9f, (float)adam_t), b2t = powf(0.999f, (float)adam_t);
        adam_kernel<<<(kParams + 255) / 256, 256, 0, s>>>(kParams, W, Wb, G, Mo, Vo, lr, b1t, b2t);
        MLP_HIP(hipGetLastError());
        wver++;
    }
    return LT_OK;
}

Scorer *scorer_new(uint32_t P, const uint32_t *node_off, const uint64_t *edge_off, uint64_t seed, lt_status *st) {
    std::unique_ptr<Scorer> s(new Scorer);
    *st = s->init(P, node_off, edge_off, seed);
    if (*st != LT_OK) return nullptr;
    return s.release();
}

void scorer_free(Scorer *s) { delete s; }
Samples &scorer_samples(Scorer *s) { return s->data; }
uint32_t scorer_partitions(const Scorer *s) { return s->P; }
bool scorer_trained(const Scorer *s) { return s && s->trained; }

// Stage one decision's state in pinned mapped memory in the packed layout of
// infer_kernel, which reads it over the bus, and launch the fused MLP on
// `stream` (nullptr = the scorer's own); scorer_infer_wait returns the scores,
// which come back the same way.  A launch first waits for the one before it,
// whose kernel may still be reading the staged state.
lt_status scorer_infer_launch(Scorer *s, const DecisionState &d, void *stream_v) {
    hipStream_t st = stream_v ? (hipStream_t)stream_v : s->stream;
    if (s->pending) {
        MLP_HIP(hipStreamSynchronize(s->pending));
        s->pending = nullptr;
    }
    const uint32_t P = s->P;
    uint8_t *h = s->h_stage;
    pack_state_head((int32_t *)h, d.counts, P, d.recent, d.cur, d.w_total);
    memcpy(h + kStateHead, d.counts, 4 * (size_t)P);
    memcpy(h + kStateHead + 4 * (size_t)P, d.age, 4 * (size_t)P);
    memcpy(h + kStateHead + 8 * (size_t)P, d.resident, P);
    if (s->wi_ver != s->wver) {
        transpose_infer_weights_kernel<<<(unsigned)((kWiSize + 255) / 256), 256, 0, st>>>(s->Wb, s->Wi);
        MLP_HIP(hipGetLastError());
        s->wi_ver = s->wver;
    }
    infer_kernel<<<(P + 7) / 8, 256, 0, st>>>(P, s->d_stage, s->d_vsh, s->d_esh, s->Wb, s->Wi, s->d_Y);
    MLP_HIP(hipGetLastError());
    s->pending = st;
    return LT_OK;
}

// Scores (out[P*2]) of the last launched inference.  Polls: this wait can be
// on the scheduler's critical path.
lt_status scorer_infer_wait(Scorer *s, float *out) {
    if (!s->pending) {
        set_error("no inference in flight");
        return LT_E_INTERNAL;
    }
    hipError_t e;
    while ((e = hipStreamQuery(s->pending)) == hipErrorNotReady) {
    }
    s->pending = nullptr;
    MLP_HIP(e);
    memcpy(out, s->h_Y, sizeof(float) * s->P * kOut);
    return LT_OK;
}

bool scorer_infer_pending(const Scorer *s) { return s && s->pending; }

lt_status scorer_infer(Scorer *s, const DecisionState &d, void *stream_v, float *out) {
    lt_status st = scorer_infer_launch(s, d, stream_v);
    return st != LT_OK ? st : scorer_infer_wait(s, out);
}

// ---- inference without allocation ----------------------------------------------
// Scratch of lt_scorer_infer_noalloc, one copy per GPU, held until the
// process ends: one decision's packed state (the layout infer_kernel reads),
// its scores, and the output-major copy of the first two layers' weights.
// Sized for kNoAllocMaxP partitions (about 1.4 MB).  In host code the name
// `scratch` is only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr uint32_t kNoAllocMaxP = 65535;  // the largest partition count a .part file holds
constexpr size_t kNaStateBytes = (state_bytes(kNoAllocMaxP) + 255) & ~(size_t)255;
constexpr size_t kNaYBytes = ((sizeof(float) * kNoAllocMaxP * kOut) + 255) & ~(size_t)255;
constexpr size_t kNaWiBytes = ((sizeof(uint16_t) * kWiSize) + 255) & ~(size_t)255;
static __device__ __align__(256) uint8_t scratch[kNaStateBytes + kNaYBytes + kNaWiBytes];

namespace {
constexpr int kNaMaxDevices = 64;
struct NoAllocDevice {
    uint8_t *base = nullptr;          // device address of this GPU's scratch
    uint64_t wi_uid = 0;  // scorer (Scorer::uid) whose weights the scratch's inference copy holds
    uint64_t wi_ver = 0;  // and their version
};
NoAllocDevice g_noalloc[kNaMaxDevices];
std::mutex g_noalloc_mu;
}  // namespace

// This GPU's scratch address, looked up on the first call made on it.
static lt_status noalloc_scratch(int dev, NoAllocDevice **out) {
    if (dev < 0 || dev >= kNaMaxDevices) {
        set_error("lt_scorer_infer_noalloc: device %d beyond the %d supported", dev, kNaMaxDevices);
        return LT_E_ARG;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice &d = g_noalloc[dev];
    if (!d.base) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("lt_scorer_infer_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
            return LT_E_HIP;
        }
        d.base = (uint8_t *)ptr;
    }
    *out = &d;
    return LT_OK;
}

}  // namespace lt

struct lt_scorer : lt::Scorer {};

using namespace lt;

static lt_status upload_samples(Scorer *s, uint64_t off, uint32_t nb, const uint32_t *counts, const uint8_t *resident,
                                const uint32_t *age, const int32_t *recent, const int32_t *cur, const float *w_total,
                                const int32_t *next_label, const float *keep_label, const float *wa, const float *wb,
                                const std::vector<uint64_t> *order) {
    const uint32_t P = s->P;
    std::vector<uint32_t> c((size_t)nb * P), ag((size_t)nb * P);
    std::vector<uint8_t> r((size_t)nb * P);
    std::vector<int32_t> rc((size_t)nb * kRecent), cu(nb), nx(nb);
    std::vector<float> wt(nb), kp((size_t)nb * P), a(nb), b(nb);
    for (uint32_t i = 0; i < nb; i++) {
        const uint64_t j = order ? (*order)[off + i] : off + i;
        memcpy(&c[(size_t)i * P], counts + j * P, 4 * (size_t)P);
        memcpy(&r[(size_t)i * P], resident + j * P, P);
        memcpy(&ag[(size_t)i * P], age + j * P, 4 * (size_t)P);
        memcpy(&rc[(size_t)i * kRecent], recent + j * kRecent, 4 * kRecent);
        cu[i] = cur[j];
        nx[i] = next_label[j];
        wt[i] = w_total[j];
        if (keep_label) memcpy(&kp[(size_t)i * P], keep_label + j * P, 4 * (size_t)P);
        a[i] = wa ? wa[j] : 1.f;
        b[i] = (wb && keep_label) ? wb[j] : 0.f;
    }
    hipStream_t st = s->stream;
    MLP_HIP(hipMemcpyAsync(s->d_counts, c.data(), 4 * c.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_res, r.data(), r.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_age, ag.data(), 4 * ag.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_recent, rc.data(), 4 * rc.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_cur, cu.data(), 4 * cu.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_next, nx.data(), 4 * nx.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_wtot, wt.data(), 4 * wt.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_keep, kp.data(), 4 * kp.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_wa, a.data(), 4 * a.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(s->d_wb, b.data(), 4 * b.size(), hipMemcpyHostToDevice, st));
    MLP_HIP(hipStreamSynchronize(st));  // the host vectors die at return
    return LT_OK;
}

extern "C" {

lt_status lt_scorer_create(const lt_partition *p, uint64_t seed, lt_scorer **out) {
    if (!p || !out) return LT_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("the scorer needs a HIP device");
        return LT_E_NODEVICE;
    }
    std::unique_ptr<lt_scorer> s(new lt_scorer);
    lt_status st = s->init(p->P(), p->node_off.data(), p->edge_off.data(), seed);
    if (st != LT_OK) return st;
    *out = s.release();
    return LT_OK;
}

void lt_scorer_free(lt_scorer *s) { delete s; }

lt_status lt_scorer_loss_grad(lt_scorer *s, uint64_t n, const uint32_t *counts, const uint8_t *resident,
                              const uint32_t *age, const int32_t *recent, const int32_t *cur, const float *w_total,
                              const int32_t *next_label, const float *keep_label, const float *wa, const float *wb,
                              float *loss_out, float *grad_out) {
    if (!s || !n || !counts || !resident || !age || !recent || !cur || !w_total || !next_label) return LT_E_ARG;
    if (n > 512) {
        set_error("at most 512 decisions per call");
        return LT_E_ARG;
    }
    lt_status st = s->ensure_cap((uint32_t)n);
    if (st != LT_OK) return st;
    if ((st = upload_samples(s, 0, (uint32_t)n, counts, resident, age, recent, cur, w_total, next_label, keep_label,
                             wa, wb, nullptr)) != LT_OK)
        return st;
    if ((st = s->forward_backward((uint32_t)n, false, 0.f)) != LT_OK) return st;
    if (loss_out) MLP_HIP(hipMemcpyAsync(loss_out, s->d_loss, 8, hipMemcpyDeviceToHost, s->stream));
    if (grad_out) MLP_HIP(hipMemcpyAsync(grad_out, s->G, 4 * kParams, hipMemcpyDeviceToHost, s->stream));
    MLP_HIP(hipStreamSynchronize(s->stream));
    return LT_OK;
}

lt_status lt_scorer_train(lt_scorer *s, uint64_t n, const uint32_t *counts, const uint8_t *resident,
                          const uint32_t *age, const int32_t *recent, const int32_t *cur, const float *w_total,
                          const int32_t *next_label, const float *keep_label, const float *wa, const float *wb,
                          uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *loss_out) {
    if (!s || !n || !counts || !resident || !age || !recent || !cur || !w_total || !next_label || !batch || !epochs)
        return LT_E_ARG;
    batch = std::min<uint32_t>(batch, (uint32_t)std::min<uint64_t>(n, 512));
    lt_status st = s->ensure_cap(batch);
    if (st != LT_OK) return st;
    std::vector<uint64_t> order(n);
    for (uint64_t i = 0; i < n; i++) order[i] = i;
    std::mt19937_64 rng(seed);
    float last[2] = {0, 0};
    for (uint32_t e = 0; e < epochs; e++) {
        std::shuffle(order.begin(), order.end(), rng);
        double acc[2] = {0, 0};
        uint32_t steps = 0;
        for (uint64_t off = 0; off + batch <= n; off += batch) {
            if ((st = upload_samples(s, off, batch, counts, resident, age, recent, cur, w_total, next_label,
                                     keep_label, wa, wb, &order)) != LT_OK)
                return st;
            if ((st = s->forward_backward(batch, true, lr)) != LT_OK) return st;
            float l[2];
            MLP_HIP(hipMemcpyAsync(l, s->d_loss, 8, hipMemcpyDeviceToHost, s->stream));
            MLP_HIP(hipStreamSynchronize(s->stream));
            acc[0] += l[0];
            acc[1] += l[1];
            steps++;
        }
        if (steps) {
            last[0] = (float)(acc[0] / steps);
            last[1] = (float)(acc[1] / steps);
        }
    }
    for (int i = 0; i < 2; i++)
        if (!std::isfinite(last[i])) {
            set_error("scorer training diverged");
            return LT_E_INTERNAL;
        }
    s->trained = true;
    if (loss_out) {
        loss_out[0] = last[0];
        loss_out[1] = last[1];
    }
    return LT_OK;
}

lt_status lt_scorer_infer(lt_scorer *s, const uint32_t *counts, const uint8_t *resident, const uint32_t *age,
                          const int32_t *recent, int32_t cur, float w_total, float *out) {
    if (!s || !counts || !resident || !age || !recent || !out) return LT_E_ARG;
    DecisionState d;
    d.counts = counts;
    d.resident = resident;
    d.age = age;
    for (int k = 0; k < kRecent; k++) d.recent[k] = recent[k];
    d.cur = cur;
    d.w_total = w_total;
    return scorer_infer(s, d, s->stream, out);
}

// The scores of lt_scorer_infer, computed through the GPU's static scratch
// instead of the scorer's staging buffers: state copied in, infer_kernel,
// scores copied out, all on the scorer's stream; no allocation of any kind.
lt_status lt_scorer_infer_noalloc(lt_scorer *s, const uint32_t *counts, const uint8_t *resident, const uint32_t *age,
                                  const int32_t *recent, int32_t cur, float w_total, float *out) {
    if (!s || !counts || !resident || !age || !recent || !out) return LT_E_ARG;
    const uint32_t P = s->P;
    if (P > kNoAllocMaxP) {
        set_error("lt_scorer_infer_noalloc: %u partitions exceed the scratch's %u", P, kNoAllocMaxP);
        return LT_E_ARG;
    }
    int dev = 0;
    MLP_HIP(hipGetDevice(&dev));
    NoAllocDevice *nd = nullptr;
    LT_NA_TRY(noalloc_scratch(dev, &nd));
    uint8_t *state = nd->base, *Y = nd->base + kNaStateBytes, *Wi = nd->base + kNaStateBytes + kNaYBytes;
    int32_t head[kStateHead / 4];
    pack_state_head(head, counts, P, recent, cur, w_total);
    hipStream_t st = s->stream;
    MLP_HIP(hipMemcpyAsync(state, head, kStateHead, hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(state + kStateHead, counts, 4 * (size_t)P, hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(state + kStateHead + 4 * (size_t)P, age, 4 * (size_t)P, hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(state + kStateHead + 8 * (size_t)P, resident, P, hipMemcpyHostToDevice, st));
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (nd->wi_uid != s->uid || nd->wi_ver != s->wver) {  // the scratch holds other weights
            transpose_infer_weights_kernel<<<(unsigned)((kWiSize + 255) / 256), 256, 0, st>>>(s->Wb,
                                                                                              (uint16_t *)Wi);
            MLP_HIP(hipGetLastError());
            nd->wi_uid = s->uid;
            nd->wi_ver = s->wver;
        }
    }
    infer_kernel<<<(P + 7) / 8, 256, 0, st>>>(P, state, s->d_vsh, s->d_esh, s->Wb, (const uint16_t *)Wi, (float *)Y);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(out, Y, sizeof(float) * P * kOut, hipMemcpyDeviceToHost, st));
    MLP_HIP(hipStreamSynchronize(st));
    return LT_OK;
}

uint64_t lt_scorer_num_params(void) { return kParams; }

lt_status lt_scorer_fit(lt_scorer *s, uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *loss_out) {
    if (!s) return LT_E_ARG;
    Samples &d = s->data;
    if (d.n == 0) {
        set_error("no training samples: add traces first");
        return LT_E_ARG;
    }
    return lt_scorer_train(s, d.n, d.counts.data(), d.resident.data(), d.age.data(), d.recent.data(), d.cur.data(),
                           d.w_total.data(), d.next.data(), d.keep.data(), d.wa.data(), d.wb.data(), epochs, batch, lr,
                           seed, loss_out);
}

lt_status lt_scorer_get_params(const lt_scorer *s, float *out) {
    if (!s || !out) return LT_E_ARG;
    MLP_HIP(hipMemcpy(out, s->W, 4 * kParams, hipMemcpyDeviceToHost));
    return LT_OK;
}

lt_status lt_scorer_set_params(lt_scorer *s, const float *in) {
    if (!s || !in) return LT_E_ARG;
    MLP_HIP(hipMemcpy(s->W, in, 4 * kParams, hipMemcpyHostToDevice));
    round_kernel<<<(kParams + 255) / 256, 256, 0, s->stream>>>(kParams, s->W, s->Wb);
    MLP_HIP(hipGetLastError());
    s->wver++;
    MLP_HIP(hipStreamSynchronize(s->stream));
    s->trained = true;
    return LT_OK;
}

// File: "LTMLP2\0\0" | u32 features | u32 hidden | u32 outputs | u32 pad | u16 bf16 weights[kParams]
lt_status lt_scorer_save(const lt_scorer *s, const char *path) {
    if (!s || !path) return LT_E_ARG;
    std::vector<uint16_t> wb(kParams);
    MLP_HIP(hipMemcpy(wb.data(), s->Wb, 2 * kParams, hipMemcpyDeviceToHost));
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot create %s", path);
        return LT_E_IO;
    }
    const char magic[8] = {'L', 'T', 'M', 'L', 'P', '2', 0, 0};
    const uint32_t hdr[4] = {kFeat, kHidden, kOut, 0};
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(hdr, 4, 4, f) == 4 && fwrite(wb.data(), 2, kParams, f) == kParams;
    ok = fclose(f) == 0 && ok;
    if (!ok) {
        set_error("write to %s failed", path);
        return LT_E_IO;
    }
    return LT_OK;
}

lt_status lt_scorer_load(lt_scorer *s, const char *path) {
    if (!s || !path) return LT_E_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error("cannot open %s", path);
        return LT_E_IO;
    }
    char magic[8];
    uint32_t hdr[4];
    std::vector<uint16_t> wb(kParams);
    bool ok = fread(magic, 1, 8, f) == 8 && fread(hdr, 4, 4, f) == 4 && fread(wb.data(), 2, kParams, f) == kParams;
    fclose(f);
    if (!ok || memcmp(magic, "LTMLP2", 6) != 0 || hdr[0] != kFeat || hdr[1] != kHidden || hdr[2] != kOut) {
        set_error("%s is not a scorer weights file", path);
        return LT_E_FORMAT;
    }
    std::vector<float> w(kParams);
    for (size_t i = 0; i < kParams; i++) {
        const uint32_t u = (uint32_t)wb[i] << 16;
        memcpy(&w[i], &u, 4);
    }
    return lt_scorer_set_params(s, w.data());
}

}  // extern "C"
