// Candidate scorer: weights on the device, capture store, batched MFMA
// inference (host side and C-ABI).  No counterpart in the reference.
// A candidate is a read pair past the report gate pre-check; its 16
// features come from k_job_features (filter_kernels.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/marvl_filter.h"
#include "mi_common.h"
#include "scorer.hpp"

namespace mi {

namespace {

struct Scorer {
    std::mutex mu;
    int mode = 0;            // 0 exact, 1 capture, 2 gate
    bool loaded = false;
    void* dev = nullptr;     // packed weights + biases, one allocation
    MlpW w{};
    float thr = 0.f;
    std::vector<float> cap_feat;   // captured candidates (kFeat floats each)
    std::vector<int> cap_nrec;     // records of the candidate's exact run
    std::vector<unsigned long long> cap_work;   // band cells + snake steps of that run
};

Scorer& S()
{
    static Scorer s;
    return s;
}

// W1 [16][64], W2 [64][64], W3 [64] (input-major) into B operands of
// v_mfma_f32_16x16x32_bf16 (lane l: column l & 15, k = 8*(l >> 4) + e)
void pack(const float* w1, const float* w2, const float* w3, std::vector<uint16_t>& p1, std::vector<uint16_t>& p2,
          std::vector<uint16_t>& p3)
{
    p1.assign(4 * 64 * 8, 0);
    p2.assign(2 * 4 * 64 * 8, 0);
    p3.assign(2 * 64 * 8, 0);
    for (int cb = 0; cb < 4; cb++)
        for (int l = 0; l < 64; l++)
            for (int e = 0; e < 8; e++) {
                const int k = 8 * (l >> 4) + e, j = 16 * cb + (l & 15);
                p1[(cb * 64 + l) * 8 + e] = k < kFeat ? bf16_bits(w1[k * kHid + j]) : 0;
            }
    for (int t = 0; t < 2; t++)
        for (int cb = 0; cb < 4; cb++)
            for (int l = 0; l < 64; l++)
                for (int e = 0; e < 8; e++) {
                    const int k = 32 * t + 8 * (l >> 4) + e, j = 16 * cb + (l & 15);
                    p2[((t * 4 + cb) * 64 + l) * 8 + e] = bf16_bits(w2[k * kHid + j]);
                }
    for (int t = 0; t < 2; t++)
        for (int l = 0; l < 64; l++)
            for (int e = 0; e < 8; e++) {
                const int k = 32 * t + 8 * (l >> 4) + e;
                p3[(t * 64 + l) * 8 + e] = (l & 15) == 0 ? bf16_bits(w3[k]) : 0;
            }
}

}  // namespace

// Batched inference: one wave per tile of 16 candidates (grid-stride over tiles).
__global__ __launch_bounds__(256) void k_mlp_batch(const float* __restrict__ feat, long long n, MlpW w,
                                                   float* __restrict__ out)
{
    __shared__ __attribute__((aligned(16))) uint16_t lds[4][2 * kMlpLdsHalves];
    const int l = __lane_id();
    const int wv = threadIdx.x >> 6;
    const long long ntile = (n + kMlpRows - 1) / kMlpRows;
    for (long long tile = (long long)blockIdx.x * 4 + wv; tile < ntile; tile += (long long)gridDim.x * 4) {
        const long long row = tile * kMlpRows + (l & 15);
        const bool in = row < n && l < 32;   // lanes 32..63 carry k >= 16: zero
        float f[8];
#pragma unroll
        for (int e = 0; e < 8; e++) f[e] = in ? feat[row * kFeat + 8 * (l >> 4) + e] : 0.f;
        const bf16x8 x = pack8(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]);
        const f32x4 o = mlp_tile(x, w, lds[wv]);
        if ((l & 15) == 0) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const long long r = tile * kMlpRows + 4 * (l >> 4) + v;
                if (r < n) out[r] = o[v] + w.b3;
            }
        }
    }
}

int scorer_mode() { return S().mode; }

bool scorer_weights(MlpW* w, float* thr)
{
    Scorer& s = S();
    if (!s.loaded) return false;
    *w = s.w;
    *thr = s.thr;
    return true;
}

void scorer_logits(const float* dfeat, long long n, const MlpW& w, float* dout, hipStream_t st)
{
    if (n <= 0) return;
    const long long ntile = (n + kMlpRows - 1) / kMlpRows;
    const int grid = (int)std::min<long long>((ntile + 3) / 4, 8192);
    hipLaunchKernelGGL(k_mlp_batch, dim3(grid), dim3(256), 0, st, dfeat, n, w, dout);
}

void scorer_capture_append(const float* feat, const int* nrec, const unsigned long long* work, long long n)
{
    Scorer& s = S();
    std::lock_guard<std::mutex> g(s.mu);
    s.cap_feat.insert(s.cap_feat.end(), feat, feat + n * kFeat);
    s.cap_nrec.insert(s.cap_nrec.end(), nrec, nrec + n);
    s.cap_work.insert(s.cap_work.end(), work, work + n);
}

}  // namespace mi

using namespace mi;

extern "C" {

int mi_scorer_load(const float* w1, const float* b1, const float* w2, const float* b2, const float* w3, float b3,
                   float threshold)
{
    if (!w1 || !b1 || !w2 || !b2 || !w3) {
        mi_set_error("mi_scorer_load: null weight array");
        return MI_ERR_ARG;
    }
    Scorer& s = S();
    std::vector<uint16_t> p1, p2, p3;
    pack(w1, w2, w3, p1, p2, p3);
    const size_t n1 = p1.size() * 2, n2 = p2.size() * 2, n3 = p3.size() * 2, nb = 2 * kHid * sizeof(float);
    if (s.dev == nullptr && hipMalloc(&s.dev, n1 + n2 + n3 + nb) != hipSuccess) {
        mi_set_error("mi_scorer_load: device allocation failed");
        return MI_ERR_DEVICE;
    }
    char* d = (char*)s.dev;
    std::vector<float> bias(2 * kHid);
    memcpy(bias.data(), b1, kHid * sizeof(float));
    memcpy(bias.data() + kHid, b2, kHid * sizeof(float));
    if (hipMemcpy(d, p1.data(), n1, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n1, p2.data(), n2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n1 + n2, p3.data(), n3, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n1 + n2 + n3, bias.data(), nb, hipMemcpyHostToDevice) != hipSuccess) {
        mi_set_error("mi_scorer_load: copy to the device failed");
        return MI_ERR_DEVICE;
    }
    s.w.w1 = (const bf16x8*)d;
    s.w.w2 = (const bf16x8*)(d + n1);
    s.w.w3 = (const bf16x8*)(d + n1 + n2);
    s.w.b1 = (const float*)(d + n1 + n2 + n3);
    s.w.b2 = s.w.b1 + kHid;
    s.w.b3 = b3;
    s.thr = threshold;
    s.loaded = true;
    return 0;
}

int mi_scorer_mode(int mode)
{
    if (mode < 0 || mode > 2) {
        mi_set_error("mi_scorer_mode: mode must be 0 (exact), 1 (capture) or 2 (gate)");
        return MI_ERR_ARG;
    }
    if (mode == 2 && !S().loaded) {
        mi_set_error("mi_scorer_mode: gating needs weights (mi_scorer_load)");
        return MI_ERR_ARG;
    }
    S().mode = mode;
    return 0;
}

long long mi_scorer_captured(void)
{
    std::lock_guard<std::mutex> g(S().mu);
    return (long long)S().cap_nrec.size();
}

long long mi_scorer_take(float* feat, int* nrec, unsigned long long* work, long long cap)
{
    Scorer& s = S();
    std::lock_guard<std::mutex> g(s.mu);
    const long long n = std::min<long long>(cap, (long long)s.cap_nrec.size());
    if (n > 0) {
        memcpy(feat, s.cap_feat.data(), sizeof(float) * n * kFeat);
        memcpy(nrec, s.cap_nrec.data(), sizeof(int) * n);
        if (work) memcpy(work, s.cap_work.data(), sizeof(unsigned long long) * n);
    }
    s.cap_feat.clear();
    s.cap_nrec.clear();
    s.cap_work.clear();
    return n;
}

int mi_mlp_score(const float* feat, long long n, float* logits)
{
    Scorer& s = S();
    if (!s.loaded) {
        mi_set_error("mi_mlp_score: no weights loaded");
        return MI_ERR_ARG;
    }
    if (n <= 0) return 0;
    float *df = nullptr, *dout = nullptr;
    int rc = 0;
    if (hipMalloc(&df, sizeof(float) * n * kFeat) != hipSuccess || hipMalloc(&dout, sizeof(float) * n) != hipSuccess) {
        mi_set_error("mi_mlp_score: device allocation failed");
        rc = MI_ERR_DEVICE;
    } else {
        const long long ntile = (n + kMlpRows - 1) / kMlpRows;
        const int grid = (int)std::min<long long>((ntile + 3) / 4, 4096);
        if (hipMemcpy(df, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice) != hipSuccess) rc = MI_ERR_DEVICE;
        if (rc == 0) {
            hipLaunchKernelGGL(k_mlp_batch, dim3(grid), dim3(256), 0, 0, df, n, s.w, dout);
            if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
                hipMemcpy(logits, dout, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
                rc = MI_ERR_DEVICE;
        }
        if (rc) mi_set_error("mi_mlp_score: device error");
    }
    if (df) (void)hipFree(df);
    if (dout) (void)hipFree(dout);
    return rc;
}

}  // extern "C"
