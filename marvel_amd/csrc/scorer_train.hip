// Training of the candidate scorer on the GPU (no counterpart in the
// reference): the MLP 16 -> 64 -> 64 -> 1 of scorer.hpp, bf16 operands and
// fp32 accumulation on v_mfma_f32_16x16x32_bf16, weighted binary
// cross-entropy on the logits, Adam on fp32 master weights.
//
// One wave handles tiles of 32 samples:
//   forward   H1 = relu(X W1 + b1), H2 = relu(H1 W2 + b2), z = H2 w3 + b3
//             (two 16-row halves per layer, X / H1 / H2 rounded to bf16);
//   loss      L = sum_s w_s (softplus(z_s) - y_s z_s) / sum_s w_s
//   backward  dz_s = w_s (sigmoid(z_s) - y_s) / sum w;
//             dH2 = bf16(dz w3^T * [H2 > 0]);  dH1 = bf16((dH2 W2^T) * [H1 > 0])
//             dW3 = H2^T dz, db3 = sum dz, db2 = colsum dH2, db1 = colsum dH1,
//             dW2 = H1^T dH2 and dW1 = X^T dH1 as MFMA products whose
//             reduction dimension (k = 32) is the tile's samples.
// Activations and error signals are staged in LDS in both row-major and
// transposed order, so every MFMA operand is one 16-byte LDS read per lane.
// Each wave keeps its gradient sums in registers across its tiles and
// writes them once; a second kernel adds the waves' partial sums in a fixed
// order (the result does not depend on timing), and a third applies Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "mi_common.h"
#include "scorer.hpp"

namespace mi {

// parameter / gradient vector layout (fp32): W1 [16][64] input-major, b1,
// W2 [64][64] input-major, b2, w3, b3; the loss sum after it
constexpr int kP_W1 = 0, kP_B1 = 1024, kP_W2 = 1088, kP_B2 = 5184, kP_W3 = 5248, kP_B3 = 5312, kP_N = 5313;
constexpr int kP_LOSS = 5313, kP_STRIDE = 5376;
constexpr int kTrainRows = 32;
constexpr int kTrainWaves = 2;   // waves per workgroup

// bf16 operands of one training step, regenerated from the fp32 masters
struct TrainOps {
    bf16x8 w1b[4 * 64];       // W1 as B operand (k = feature, column = hidden 1 unit), k >= 16 zero
    bf16x8 w2b[2 * 4 * 64];   // W2 as B operand (k = hidden 1 unit, column = hidden 2 unit)
    bf16x8 w3b[2 * 64];       // w3 in column 0
    bf16x8 w2t[2 * 4 * 64];   // W2^T as B operand (k = hidden 2 unit, column = hidden 1 unit)
};

__device__ __forceinline__ float bf(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__global__ void k_train_pack(const float* __restrict__ P, TrainOps* __restrict__ ops)
{
    uint16_t* w1b = reinterpret_cast<uint16_t*>(ops->w1b);
    uint16_t* w2b = reinterpret_cast<uint16_t*>(ops->w2b);
    uint16_t* w3b = reinterpret_cast<uint16_t*>(ops->w3b);
    uint16_t* w2t = reinterpret_cast<uint16_t*>(ops->w2t);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * 4 * 64 * 8; q += gridDim.x * blockDim.x) {
        const int e = q & 7, l = (q >> 3) & 63, cb = (q >> 9) & 3, t = q >> 11;
        const int kk = 32 * t + 8 * (l >> 4) + e;   // k index of this operand element
        const int col = 16 * cb + (l & 15);
        w2b[q] = bf16_bits(P[kP_W2 + kk * kHid + col]);
        w2t[q] = bf16_bits(P[kP_W2 + col * kHid + kk]);
        if (t == 0) {
            const int k1 = 8 * (l >> 4) + e;
            w1b[q] = k1 < kFeat ? bf16_bits(P[kP_W1 + k1 * kHid + col]) : (uint16_t)0;
        }
        if (cb == 0) w3b[(t * 64 + l) * 8 + e] = (l & 15) == 0 ? bf16_bits(P[kP_W3 + kk]) : (uint16_t)0;
    }
}

struct TrainLds {
    uint16_t x[kTrainRows][kFeat];     // X row-major (bf16)
    uint16_t xt[kFeat][kTrainRows];    // X^T
    uint16_t h1[kTrainRows][kHid];
    uint16_t h1t[kHid][kTrainRows];
    uint16_t h2[kTrainRows][kHid];
    uint16_t dh2[kTrainRows][kHid];
    uint16_t dh2t[kHid][kTrainRows];
    uint16_t dh1t[kHid][kTrainRows];
    float y[kTrainRows], w[kTrainRows], dz[kTrainRows];
};

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

// Gradient sums of the rows perm[base .. base + nb) (perm == null: the rows
// base .. base + nb themselves) into part[wave][kP_STRIDE].
__global__ __launch_bounds__(64 * kTrainWaves) void k_train_grad(const TrainOps* __restrict__ ops,
                                                                 const float* __restrict__ P,
                                                                 const float* __restrict__ feat,
                                                                 const float* __restrict__ label,
                                                                 const float* __restrict__ weight,
                                                                 const int* __restrict__ perm, long long base,
                                                                 long long nb, float inv_wsum,
                                                                 float* __restrict__ part)
{
    __shared__ __attribute__((aligned(16))) TrainLds lds_all[kTrainWaves];
    const int l = __lane_id();
    const int wv = threadIdx.x >> 6;
    TrainLds& S = lds_all[wv];
    const long long gw = (long long)blockIdx.x * kTrainWaves + wv;
    const long long nw = (long long)gridDim.x * kTrainWaves;
    const f32x4 zero = {};
    f32x4 gW2[4][4], gW1[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gW1[a] = zero;
#pragma unroll
        for (int b = 0; b < 4; b++) gW2[a][b] = zero;
    }
    float gw3 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
    float gb1[4] = {0.f, 0.f, 0.f, 0.f};
    const float b3 = P[kP_B3];
    const float w3j = P[kP_W3 + l];
    const long long ntile = (nb + kTrainRows - 1) / kTrainRows;
    for (long long tile = gw; tile < ntile; tile += nw) {
        // ---- load 32 samples: lane l holds sample l >> 1, features 8 (l & 1) .. + 7
        {
            const int s = l >> 1, f0 = 8 * (l & 1);
            const long long r = tile * kTrainRows + s;
            const bool in = r < nb;
            const long long row = in ? (perm ? (long long)perm[base + r] : base + r) : 0;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const uint16_t v = in ? bf16_bits(feat[row * kFeat + f0 + e]) : (uint16_t)0;
                S.x[s][f0 + e] = v;
                S.xt[f0 + e][s] = v;
            }
            if ((l & 1) == 0) {
                S.y[s] = in ? label[row] : 0.f;
                S.w[s] = in ? weight[row] : 0.f;
            }
        }
        mlp_wave_sync();
        // ---- forward, layer 1
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const bf16x8 xa = l < 32 ? ld8(&S.x[16 * h + (l & 15)][8 * (l >> 4)]) : bf16x8{};
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                const f32x4 acc = mfma16(xa, ops->w1b[cb * 64 + l], zero);
                const int col = 16 * cb + (l & 15);
                const float bias = P[kP_B1 + col];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = 16 * h + 4 * (l >> 4) + v;
                    const float a = acc[v] + bias;
                    const uint16_t q = bf16_bits(a > 0.f ? a : 0.f);
                    S.h1[row][col] = q;
                    S.h1t[col][row] = q;
                }
            }
        }
        mlp_wave_sync();
        // ---- forward, layer 2
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                f32x4 acc = zero;
#pragma unroll
                for (int t = 0; t < 2; t++)
                    acc = mfma16(ld8(&S.h1[16 * h + (l & 15)][32 * t + 8 * (l >> 4)]), ops->w2b[(t * 4 + cb) * 64 + l],
                                 acc);
                const int col = 16 * cb + (l & 15);
                const float bias = P[kP_B2 + col];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float a = acc[v] + bias;
                    S.h2[16 * h + 4 * (l >> 4) + v][col] = bf16_bits(a > 0.f ? a : 0.f);
                }
            }
        }
        mlp_wave_sync();
        // ---- forward, layer 3, loss and dz
#pragma unroll
        for (int h = 0; h < 2; h++) {
            f32x4 o = zero;
#pragma unroll
            for (int t = 0; t < 2; t++) o = mfma16(ld8(&S.h2[16 * h + (l & 15)][32 * t + 8 * (l >> 4)]), ops->w3b[t * 64 + l], o);
            if ((l & 15) == 0) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int s = 16 * h + 4 * (l >> 4) + v;
                    const float z = o[v] + b3;
                    const float ws = S.w[s], ys = S.y[s];
                    const float p = 1.f / (1.f + expf(-z));
                    S.dz[s] = ws * (p - ys) * inv_wsum;
                    loss += ws * (fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))) - ys * z);
                }
            }
        }
        mlp_wave_sync();
        // ---- dH2, dW3, db2, db3: lane l owns hidden-2 unit l
        for (int s = 0; s < kTrainRows; s++) {
            const float d = S.dz[s];
            const float h2v = bf(S.h2[s][l]);
            gw3 += d * h2v;
            const uint16_t q = bf16_bits(h2v > 0.f ? d * w3j : 0.f);
            S.dh2[s][l] = q;
            S.dh2t[l][s] = q;
            gb2 += bf(q);
            if (l == 0) gb3 += d;
        }
        mlp_wave_sync();
        // ---- dW2 = H1^T dH2 (k = samples)
#pragma unroll
        for (int ib = 0; ib < 4; ib++) {
            const bf16x8 a = ld8(&S.h1t[16 * ib + (l & 15)][8 * (l >> 4)]);
#pragma unroll
            for (int jb = 0; jb < 4; jb++) gW2[ib][jb] = mfma16(a, ld8(&S.dh2t[16 * jb + (l & 15)][8 * (l >> 4)]), gW2[ib][jb]);
        }
        // ---- dH1 = (dH2 W2^T) * [H1 > 0], db1
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                f32x4 acc = zero;
#pragma unroll
                for (int t = 0; t < 2; t++)
                    acc = mfma16(ld8(&S.dh2[16 * h + (l & 15)][32 * t + 8 * (l >> 4)]), ops->w2t[(t * 4 + cb) * 64 + l],
                                 acc);
                const int col = 16 * cb + (l & 15);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = 16 * h + 4 * (l >> 4) + v;
                    const uint16_t q = bf16_bits(S.h1[row][col] != 0 ? acc[v] : 0.f);
                    S.dh1t[col][row] = q;
                    gb1[cb] += bf(q);
                }
            }
        }
        mlp_wave_sync();
        // ---- dW1 = X^T dH1 (k = samples); lanes 16.. of the A operand carry feature rows 0..15 again
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
            gW1[cb] = mfma16(ld8(&S.xt[l & 15][8 * (l >> 4)]), ld8(&S.dh1t[16 * cb + (l & 15)][8 * (l >> 4)]), gW1[cb]);
        mlp_wave_sync();
    }
    // ---- this wave's partial sums
    float* out = part + gw * kP_STRIDE;
#pragma unroll
    for (int ib = 0; ib < 4; ib++)
#pragma unroll
        for (int jb = 0; jb < 4; jb++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                out[kP_W2 + (16 * ib + 4 * (l >> 4) + v) * kHid + 16 * jb + (l & 15)] = gW2[ib][jb][v];
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[kP_W1 + (4 * (l >> 4) + v) * kHid + 16 * cb + (l & 15)] = gW1[cb][v];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        float t = gb1[cb];
        t += __shfl_xor(t, 16, 64);
        t += __shfl_xor(t, 32, 64);
        if (l < 16) out[kP_B1 + 16 * cb + l] = t;
    }
    out[kP_B2 + l] = gb2;
    out[kP_W3 + l] = gw3;
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (l == 0) {
        out[kP_B3] = gb3;
        out[kP_LOSS] = loss * inv_wsum;
    }
}

// grad[p] = sum over waves of part[w][p], in wave order
__global__ void k_train_reduce(const float* __restrict__ part, int nw, float* __restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kP_LOSS) return;
    float s = 0.f;
    for (int w = 0; w < nw; w++) s += part[(long long)w * kP_STRIDE + p];
    grad[p] = s;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) with bias correction factors c1 = 1 - beta1^t, c2 = 1 - beta2^t
__global__ void k_train_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
                             float* __restrict__ v, float lr, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kP_N) return;
    const float gp = g[p];
    const float mp = 0.9f * m[p] + 0.1f * gp;
    const float vp = 0.999f * v[p] + 0.001f * gp * gp;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + 1e-8f);
}

namespace {

struct TrainBufs {
    float *feat = nullptr, *label = nullptr, *weight = nullptr, *P = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    float* part = nullptr;
    int* perm = nullptr;
    TrainOps* ops = nullptr;
    ~TrainBufs()
    {
        for (void* p : {(void*)feat, (void*)label, (void*)weight, (void*)P, (void*)g, (void*)m, (void*)v, (void*)part,
                        (void*)perm, (void*)ops})
            if (p) (void)hipFree(p);
    }
};

constexpr int kGradGrid = 256;   // workgroups of the gradient kernel (512 waves)

int upload(TrainBufs& b, const float* feat, const float* label, const float* weight, long long n, const float* params)
{
    if (hipMalloc(&b.feat, sizeof(float) * n * kFeat) != hipSuccess ||
        hipMalloc(&b.label, sizeof(float) * n) != hipSuccess || hipMalloc(&b.weight, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&b.P, sizeof(float) * kP_STRIDE) != hipSuccess ||
        hipMalloc(&b.g, sizeof(float) * kP_STRIDE) != hipSuccess ||
        hipMalloc(&b.m, sizeof(float) * kP_STRIDE) != hipSuccess ||
        hipMalloc(&b.v, sizeof(float) * kP_STRIDE) != hipSuccess ||
        hipMalloc(&b.part, sizeof(float) * kP_STRIDE * kGradGrid * kTrainWaves) != hipSuccess ||
        hipMalloc(&b.ops, sizeof(TrainOps)) != hipSuccess)
        return MI_ERR_DEVICE;
    if (hipMemcpy(b.feat, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.label, label, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.weight, weight, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.P, params, sizeof(float) * kP_N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b.m, 0, sizeof(float) * kP_STRIDE) != hipSuccess ||
        hipMemset(b.v, 0, sizeof(float) * kP_STRIDE) != hipSuccess)
        return MI_ERR_DEVICE;
    return 0;
}

// one gradient evaluation over rows perm[base .. base + nb)
int grad_step(TrainBufs& b, const int* perm, long long base, long long nb, float inv_wsum)
{
    hipLaunchKernelGGL(k_train_pack, dim3(8), dim3(256), 0, 0, b.P, b.ops);
    const long long ntile = (nb + kTrainRows - 1) / kTrainRows;
    const int grid = (int)std::max<long long>(1, std::min<long long>(kGradGrid, (ntile + kTrainWaves - 1) / kTrainWaves));
    hipLaunchKernelGGL(k_train_grad, dim3(grid), dim3(64 * kTrainWaves), 0, 0, b.ops, b.P, b.feat, b.label, b.weight,
                       perm, base, nb, inv_wsum, b.part);
    hipLaunchKernelGGL(k_train_reduce, dim3((kP_LOSS + 256) / 256), dim3(256), 0, 0, b.part, grid * kTrainWaves,
                       b.g);
    return hipGetLastError() == hipSuccess ? 0 : MI_ERR_DEVICE;
}

}  // namespace

}  // namespace mi

using namespace mi;

extern "C" {

int mi_scorer_grad(const float* feat, const float* label, const float* weight, long long n, const float* params,
                   float* grad, float* loss)
{
    if (!feat || !label || !weight || !params || !grad || n <= 0) return MI_ERR_ARG;
    double ws = 0;
    for (long long i = 0; i < n; i++) ws += weight[i];
    if (!(ws > 0)) return MI_ERR_ARG;
    TrainBufs b;
    int rc = upload(b, feat, label, weight, n, params);
    if (rc == 0) rc = grad_step(b, nullptr, 0, n, (float)(1.0 / ws));
    std::vector<float> g(kP_STRIDE);
    if (rc == 0 && (hipMemcpy(g.data(), b.g, sizeof(float) * (kP_LOSS + 1), hipMemcpyDeviceToHost) != hipSuccess))
        rc = MI_ERR_DEVICE;
    if (rc != 0) {
        mi_set_error("mi_scorer_grad: device error");
        return rc;
    }
    memcpy(grad, g.data(), sizeof(float) * kP_N);
    if (loss) *loss = g[kP_LOSS];
    return 0;
}

int mi_scorer_train(const float* feat, const float* label, const float* weight, long long n, int epochs, int batch,
                    float lr, unsigned long long seed, float* params, float* epoch_loss)
{
    if (!feat || !label || !weight || !params || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0)) return MI_ERR_ARG;
    TrainBufs b;
    int rc = upload(b, feat, label, weight, n, params);
    if (rc == 0 && hipMalloc(&b.perm, sizeof(int) * n) != hipSuccess) rc = MI_ERR_DEVICE;
    std::vector<int> perm(n);
    for (long long i = 0; i < n; i++) perm[i] = (int)i;
    std::mt19937_64 rng(seed);
    long long t = 0;
    for (int ep = 0; ep < epochs && rc == 0; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpy(b.perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess) {
            rc = MI_ERR_DEVICE;
            break;
        }
        double lsum = 0, wsum_all = 0;
        for (long long s0 = 0; s0 < n && rc == 0; s0 += batch) {
            const long long nb = std::min<long long>(batch, n - s0);
            double ws = 0;
            for (long long i = s0; i < s0 + nb; i++) ws += weight[perm[i]];
            if (!(ws > 0)) continue;
            rc = grad_step(b, b.perm, s0, nb, (float)(1.0 / ws));
            if (rc) break;
            t += 1;
            const float c1 = 1.f - powf(0.9f, (float)t), c2 = 1.f - powf(0.999f, (float)t);
            hipLaunchKernelGGL(k_train_adam, dim3((kP_N + 255) / 256), dim3(256), 0, 0, b.P, b.g, b.m, b.v, lr, c1, c2);
            if (epoch_loss) {
                float l = 0.f;
                if (hipMemcpy(&l, b.g + kP_LOSS, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
                    rc = MI_ERR_DEVICE;
                    break;
                }
                lsum += (double)l * ws;
                wsum_all += ws;
            }
        }
        if (epoch_loss) epoch_loss[ep] = wsum_all > 0 ? (float)(lsum / wsum_all) : 0.f;
    }
    if (rc == 0 && (hipDeviceSynchronize() != hipSuccess ||
                    hipMemcpy(params, b.P, sizeof(float) * kP_N, hipMemcpyDeviceToHost) != hipSuccess))
        rc = MI_ERR_DEVICE;
    if (rc != 0) mi_set_error("mi_scorer_train: device error");
    return rc;
}

}  // extern "C"
