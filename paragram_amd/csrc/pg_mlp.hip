// The learned cell radius: a 16 -> 64 -> 64 -> 2 network (ReLU) that
// predicts, from a point's k = 8 nearest-neighbour distances and weight
// statistics, the half-width of its cell (log of the ratio to the 8th
// neighbour distance) and its face count / 16.  The prediction only chooses
// the cell kernel's starting cube (pg_compute_cells_ex); the exact traversal
// decides the cell, so a wrong prediction costs time, never correctness.
//
// Parameters (fp32 master copy, 5378 floats): W1 [16][64], b1 [64],
// W2 [64][64], b2 [64], W3 [64][2], b3 [2], row-major [in][out].  Every
// product uses BF16 operands (weights and layer inputs rounded to nearest
// even) with FP32 accumulation; biases and accumulators are FP32.
//
// Inference runs on the matrix cores: v_mfma_f32_16x16x16_bf16 in the
// transposed form Z^T = W^T X^T, so the accumulator fragment of one layer
// (lane l: rows 4 (l / 16) + j, column l % 16) is already the B operand of
// the next (k = 4 (l / 16) + j, column l % 16) and activations never leave
// registers.  Training (forward + backward, one sample per lane, FP32 VALU
// with the same BF16 roundings) writes per-block gradient sums that a second
// kernel reduces in a fixed order and feeds to Adam.
#include "pg_common.hpp"
#include "pg_kernels.hpp"

#include <float.h>

namespace pg {

constexpr int kIn = 16, kHid = 64, kOut = 2;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kOut, kParams = kOffB3 + kOut;
static_assert(kParams == 5378, "parameter count");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint32_t bf16_bits(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_round(float x) { return __uint_as_float(bf16_bits(x) << 16); }
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 v;
    v[0] = (short)bf16_bits(a);
    v[1] = (short)bf16_bits(b);
    v[2] = (short)bf16_bits(c);
    v[3] = (short)bf16_bits(d);
    return v;
}

// ------------------------------------------------------------------ features
// Per point (original order): the 16 features from its k = 8 nearest
// neighbours (ids and squared distances in original order) and the index.
//   0-6  d_t / d_8 (t = 1..7)        7  |mean of the neighbour offsets| / d_8
//   8    log1p(distance to the cloud's AABB boundary / d_8)
//   9    min(1, that distance / (8 d_8))
//   10   (w_s - mean w_nn) / d_8^2    11  (max w_nn - w_s) / d_8^2
//   12   (w_s - min w_nn) / d_8^2     13  std w_nn / d_8^2
//   14   (d_8 - d_1) / d_8            15  1
// and log(d_8), the scale the prediction is relative to.  In a batch
// (clouds non-null) the AABB is that of the point's own cloud.
__global__ void mlp_features_kernel(int64_t n, const IndexHeader *__restrict__ hdr, const float4 *__restrict__ pts,
                                    const uint32_t *__restrict__ inv, const int32_t *__restrict__ ids,
                                    const float *__restrict__ d2, float *__restrict__ X, float *__restrict__ logd8,
                                    const IndexHeader *__restrict__ clouds, const int64_t *__restrict__ cloud_cells,
                                    int n_clouds) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t si = inv[i];
    const float4 s = pts[si];
    if (clouds != nullptr) {
        int lo = 0, hi = n_clouds - 1;  // largest c with cloud_cells[c] <= si
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cloud_cells[mid] <= (int64_t)si) lo = mid;
            else hi = mid - 1;
        }
        hdr = &clouds[lo];
    }
    float d[8];
    for (int t = 0; t < 8; ++t) d[t] = sqrtf(fminf(d2[i * 8 + t], 1e30f));
    const float d8 = fmaxf(d[7], 1e-30f);
    float cx = 0.f, cy = 0.f, cz = 0.f, wm = 0.f, wmax = -FLT_MAX, wmin = FLT_MAX, w2 = 0.f;
    for (int t = 0; t < 8; ++t) {
        const float4 p = pts[inv[ids[i * 8 + t]]];
        cx += p.x - s.x;
        cy += p.y - s.y;
        cz += p.z - s.z;
        wm += p.w;
        w2 += p.w * p.w;
        wmax = fmaxf(wmax, p.w);
        wmin = fminf(wmin, p.w);
    }
    cx *= 0.125f, cy *= 0.125f, cz *= 0.125f, wm *= 0.125f;
    const float wstd = sqrtf(fmaxf(w2 * 0.125f - wm * wm, 0.0f));
    float b = FLT_MAX;
    const float sc[3] = {s.x, s.y, s.z};
    for (int k = 0; k < 3; ++k) b = fminf(b, fminf(sc[k] - hdr->raw_lo[k], hdr->raw_hi[k] - sc[k]));
    b = fmaxf(b, 0.0f) / d8;
    const float dd = d8 * d8;
    float *x = X + i * kIn;
    for (int t = 0; t < 7; ++t) x[t] = d[t] / d8;
    x[7] = sqrtf(cx * cx + cy * cy + cz * cz) / d8;
    x[8] = log1pf(b);
    x[9] = fminf(1.0f, b * 0.125f);
    x[10] = (s.w - wm) / dd;
    x[11] = (wmax - s.w) / dd;
    x[12] = (s.w - wmin) / dd;
    x[13] = wstd / dd;
    x[14] = (d[7] - d[0]) / d8;
    x[15] = 1.0f;
    logd8[i] = logf(d8);
}

// ------------------------------------------------------------------ inference
// Each wave loads the weight fragments once, then takes 64-point blocks (four
// 16-point tiles) in a grid-stride loop.
__global__ void __launch_bounds__(64) mlp_forward_kernel(int64_t n, const float *__restrict__ X,
                                                         const float *__restrict__ P, float *__restrict__ Y) {
    const int l = threadIdx.x;
    const int col = l & 15, kq = l >> 4;
    // weight fragments (A operands): W^T[out][in] = W[in][out]
    bf16x4 a1[4], a2[4][4], a3[4];
    for (int o = 0; o < 4; ++o) {
        const int out = 16 * o + col;
        a1[o] = pack4(P[kOffW1 + (4 * kq + 0) * kHid + out], P[kOffW1 + (4 * kq + 1) * kHid + out],
                      P[kOffW1 + (4 * kq + 2) * kHid + out], P[kOffW1 + (4 * kq + 3) * kHid + out]);
        for (int k = 0; k < 4; ++k) {
            const int in = 16 * k + 4 * kq;
            a2[o][k] = pack4(P[kOffW2 + (in + 0) * kHid + out], P[kOffW2 + (in + 1) * kHid + out],
                             P[kOffW2 + (in + 2) * kHid + out], P[kOffW2 + (in + 3) * kHid + out]);
        }
    }
    for (int k = 0; k < 4; ++k) {
        const int in = 16 * k + 4 * kq;
        const bool ok = col < kOut;
        a3[k] = pack4(ok ? P[kOffW3 + (in + 0) * kOut + col] : 0.f, ok ? P[kOffW3 + (in + 1) * kOut + col] : 0.f,
                      ok ? P[kOffW3 + (in + 2) * kOut + col] : 0.f, ok ? P[kOffW3 + (in + 3) * kOut + col] : 0.f);
    }
    const int64_t blocks = (n + 63) / 64;
    for (int64_t blk = blockIdx.x; blk < blocks; blk += gridDim.x)
    for (int t = 0; t < 4; ++t) {
        const int64_t pt = blk * 64 + 16 * t + col;
        float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (pt < n) xv = *(const float4 *)(X + pt * kIn + 4 * kq);
        const bf16x4 bx = pack4(xv.x, xv.y, xv.z, xv.w);
        bf16x4 b1v[4];
        for (int o = 0; o < 4; ++o) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[o], bx, acc, 0, 0, 0);
            const int r = 16 * o + 4 * kq;
            b1v[o] = pack4(fmaxf(acc[0] + P[kOffB1 + r + 0], 0.f), fmaxf(acc[1] + P[kOffB1 + r + 1], 0.f),
                           fmaxf(acc[2] + P[kOffB1 + r + 2], 0.f), fmaxf(acc[3] + P[kOffB1 + r + 3], 0.f));
        }
        bf16x4 b2v[4];
        for (int o = 0; o < 4; ++o) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int k = 0; k < 4; ++k) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[o][k], b1v[k], acc, 0, 0, 0);
            const int r = 16 * o + 4 * kq;
            b2v[o] = pack4(fmaxf(acc[0] + P[kOffB2 + r + 0], 0.f), fmaxf(acc[1] + P[kOffB2 + r + 1], 0.f),
                           fmaxf(acc[2] + P[kOffB2 + r + 2], 0.f), fmaxf(acc[3] + P[kOffB2 + r + 3], 0.f));
        }
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 4; ++k) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[k], b2v[k], acc, 0, 0, 0);
        if (kq == 0 && pt < n) {  // rows 0 and 1 of the output tile
            Y[pt * kOut + 0] = acc[0] + P[kOffB3 + 0];
            Y[pt * kOut + 1] = acc[1] + P[kOffB3 + 1];
        }
    }
}

// ------------------------------------------------------------------ training
// One block of kTrainBlock samples, one sample per thread.  Loss: mean over the valid
// samples (mask > 0) of (y0 - t0)^2 + (y1 - t1)^2.  Writes this block's sum of
// per-sample gradients (and loss) to part[block][0 .. kParams].
constexpr int kTrainBlock = 32;  // samples (and threads) per block; LDS holds their activations

__global__ void __launch_bounds__(kTrainBlock)
    mlp_train_kernel(int64_t n, const float *__restrict__ X, const float *__restrict__ T,
                     const float *__restrict__ mask, const float *__restrict__ P, float *__restrict__ part) {
    __shared__ float sx[kTrainBlock][kIn];
    __shared__ float sh1[kTrainBlock][kHid + 1];  // bf16-rounded layer inputs (+1: no bank conflicts)
    __shared__ float sh2[kTrainBlock][kHid + 1];
    __shared__ float sd1[kTrainBlock][kHid + 1];  // dL/dz1
    __shared__ float sd2[kTrainBlock][kHid + 1];  // dL/dz2
    __shared__ float sdy[kTrainBlock][kOut];
    __shared__ float sloss[kTrainBlock];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kTrainBlock + t;
    const bool valid = i < n && mask[i] > 0.0f;
    float x[kIn];
    for (int k = 0; k < kIn; ++k) {
        x[k] = i < n ? bf16_round(X[i * kIn + k]) : 0.0f;
        sx[t][k] = x[k];
    }
    float h1[kHid];
    for (int j = 0; j < kHid; ++j) {
        float z = P[kOffB1 + j];
        for (int k = 0; k < kIn; ++k) z = fmaf(x[k], bf16_round(P[kOffW1 + k * kHid + j]), z);
        h1[j] = z;
    }
    for (int j = 0; j < kHid; ++j) sh1[t][j] = bf16_round(fmaxf(h1[j], 0.f));
    float y0 = P[kOffB3 + 0], y1 = P[kOffB3 + 1];
    float h2[kHid];
    for (int j = 0; j < kHid; ++j) {
        float z = P[kOffB2 + j];
        for (int k = 0; k < kHid; ++k) z = fmaf(sh1[t][k], bf16_round(P[kOffW2 + k * kHid + j]), z);
        h2[j] = z;
        const float a = bf16_round(fmaxf(z, 0.f));
        sh2[t][j] = a;
        y0 = fmaf(a, bf16_round(P[kOffW3 + j * kOut + 0]), y0);
        y1 = fmaf(a, bf16_round(P[kOffW3 + j * kOut + 1]), y1);
    }
    const float e0 = valid ? y0 - T[i * kOut + 0] : 0.f, e1 = valid ? y1 - T[i * kOut + 1] : 0.f;
    sloss[t] = e0 * e0 + e1 * e1;
    sdy[t][0] = 2.f * e0;
    sdy[t][1] = 2.f * e1;
    // dz2 = (W3 dy) * relu'(z2), dz1 = (W2 dz2) * relu'(z1); straight through the roundings
    for (int j = 0; j < kHid; ++j) {
        const float g = h2[j] > 0.f ? bf16_round(P[kOffW3 + j * kOut + 0]) * sdy[t][0] +
                                          bf16_round(P[kOffW3 + j * kOut + 1]) * sdy[t][1]
                                    : 0.f;
        sd2[t][j] = g;
    }
    for (int k = 0; k < kHid; ++k) {
        float g = 0.f;
        if (h1[k] > 0.f)
            for (int j = 0; j < kHid; ++j) g = fmaf(bf16_round(P[kOffW2 + k * kHid + j]), sd2[t][j], g);
        sd1[t][k] = g;
    }
    __syncthreads();
    // block sums of the outer products, thread t owns parameter columns
    float *out = part + (int64_t)blockIdx.x * (kParams + 1);
    for (int p = t; p < kParams; p += kTrainBlock) {
        float g = 0.f;
        if (p < kOffB1) {
            const int k = p / kHid, j = p % kHid;
            for (int s = 0; s < kTrainBlock; ++s) g = fmaf(sx[s][k], sd1[s][j], g);
        } else if (p < kOffW2) {
            const int j = p - kOffB1;
            for (int s = 0; s < kTrainBlock; ++s) g += sd1[s][j];
        } else if (p < kOffB2) {
            const int k = (p - kOffW2) / kHid, j = (p - kOffW2) % kHid;
            for (int s = 0; s < kTrainBlock; ++s) g = fmaf(sh1[s][k], sd2[s][j], g);
        } else if (p < kOffW3) {
            const int j = p - kOffB2;
            for (int s = 0; s < kTrainBlock; ++s) g += sd2[s][j];
        } else if (p < kOffB3) {
            const int k = (p - kOffW3) / kOut, j = (p - kOffW3) % kOut;
            for (int s = 0; s < kTrainBlock; ++s) g = fmaf(sh2[s][k], sdy[s][j], g);
        } else {
            const int j = p - kOffB3;
            for (int s = 0; s < kTrainBlock; ++s) g += sdy[s][j];
        }
        out[p] = g;
    }
    if (t == 0) {
        float L = 0.f;
        for (int s = 0; s < kTrainBlock; ++s) L += sloss[s];
        out[kParams] = L;
    }
}

// Reduces the block sums in block order, divides by the valid count, and
// (lr > 0) applies one Adam step; grad receives the mean gradient and the
// mean loss (grad[kParams]).
__global__ void mlp_adam_kernel(int blocks, const float *__restrict__ part, float count, float *__restrict__ P,
                                float *__restrict__ grad, float *__restrict__ m, float *__restrict__ v, float lr,
                                float b1, float b2, float eps, int step) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kParams) return;
    float g = 0.f;
    for (int b = 0; b < blocks; ++b) g += part[(int64_t)b * (kParams + 1) + p];
    g /= fmaxf(count, 1.0f);
    grad[p] = g;
    if (p == kParams || lr <= 0.f) return;
    const float mm = b1 * m[p] + (1.f - b1) * g, vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)step)), vh = vv / (1.f - powf(b2, (float)step));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

hipError_t mlp_features(int64_t n, const void *index_ws, const int32_t *ids, const float *d2, float *X,
                        float *logd8, hipStream_t stream, int n_clouds) {
    if (n <= 0) return hipSuccess;
    const IndexLayout L = make_index_layout(n);
    const char *ws = (const char *)index_ws;
    const IndexHeader *clouds = nullptr;
    const int64_t *cloud_cells = nullptr;
    if (n_clouds > 0) {
        const BatchLayout B = make_batch_layout(n, n_clouds);
        clouds = (const IndexHeader *)(ws + B.clouds);
        cloud_cells = (const int64_t *)(ws + B.offsets);
    }
    hipLaunchKernelGGL(mlp_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n,
                       (const IndexHeader *)(ws + L.header), (const float4 *)(ws + L.points),
                       (const uint32_t *)(ws + L.inv), ids, d2, X, logd8, clouds, cloud_cells, n_clouds);
    return hipGetLastError();
}

hipError_t mlp_forward(int64_t n, const float *X, const float *params, float *Y, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + 63) / 64;
    const int64_t grid = blocks < 8192 ? blocks : 8192;  // 32 waves per compute unit, each over many blocks
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)grid), dim3(64), 0, stream, n, X, params, Y);
    return hipGetLastError();
}

int mlp_train_blocks(int64_t n) { return (int)((n + kTrainBlock - 1) / kTrainBlock); }

hipError_t mlp_train_step(int64_t n, const float *X, const float *T, const float *mask, float count, float *params,
                          float *grad, float *m, float *v, float *part, float lr, int step, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int blocks = mlp_train_blocks(n);
    hipLaunchKernelGGL(mlp_train_kernel, dim3(blocks), dim3(kTrainBlock), 0, stream, n, X, T, mask, params, part);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 1 + 255) / 256), dim3(256), 0, stream, blocks, part, count,
                       params, grad, m, v, lr, 0.9f, 0.999f, 1e-8f, step);
    return hipGetLastError();
}

}  // namespace pg
