// Device self-tests of the wave64 building blocks the aligner relies on
// (cross-lane scans, neighbour gathers, ballots/mbcnt, the 2-bit snake
// matcher, the persistent task loop).  Each test writes plain per-lane /
// per-task outputs that the host checks against a CPU computation.
#include <hip/hip_runtime.h>

#include "la_wave.hpp"
#include "mi_common.h"

namespace mi {

// test 0: persistent loop scaffold -- outs[t] = 3*t + 1 for every task
__global__ __launch_bounds__(64) void st_persistent(int n, int* counter, int* outs, const int* abort_flag)
{
    for (;;) {
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(counter, __lane_id() == 0 ? 1 : 0));
        const int ab = __builtin_amdgcn_readfirstlane(__atomic_load_n(abort_flag, __ATOMIC_RELAXED));
        if (t >= n || ab) break;
        outs[t] = 3 * t + 1;   // every lane stores the same value
    }
}

// test 1: per wave, input v[64]; outputs per lane:
//   o[0] = exclusive suffix max (lanes above), o[1] = exclusive prefix min (lanes below),
//   o[2] = value of lane-1 (wrap), o[3] = value of lane+1 (wrap),
//   o[4] = readlane(v, lowest lane with v odd), o[5] = mbcnt rank among odd lanes,
//   o[6] = __any(v > thr), o[7] = popcount of ballot(v odd)
__global__ __launch_bounds__(64) void st_lanes(const int* in, int thr, int* out)
{
    const int lane = __lane_id();
    const int v = in[blockIdx.x * 64 + lane];
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int w = __shfl_down(incl, o, 64);
        if (lane + o < 64) incl = max(incl, w);
    }
    int excl = __shfl_down(incl, 1, 64);
    if (lane == 63) excl = -1;
    int inc2 = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int w = __shfl_up(inc2, o, 64);
        if (lane - o >= 0) inc2 = min(inc2, w);
    }
    int exc2 = __shfl_up(inc2, 1, 64);
    if (lane == 0) exc2 = INT32_MAX;
    const int up = shfl_i(v, (lane + 63) & 63);
    const int dn = shfl_i(v, (lane + 1) & 63);
    const unsigned long long odd = __ballot(v & 1);
    int rl = -7;
    if (odd) rl = rdl(v, lowest_lane(odd));
    const int rank = prefix_rank(odd);
    const int any = __any(v > thr);
    int* o = out + (blockIdx.x * 64 + lane) * 8;
    o[0] = excl; o[1] = exc2; o[2] = up; o[3] = dn; o[4] = rl; o[5] = rank; o[6] = any; o[7] = __popcll(odd);
}

// test 2: snake matcher; per lane one query (x, y, lim, dir) on packed A/B
__global__ __launch_bounds__(64) void st_match(const uint64_t* A, const uint64_t* B, const int* q, int n, int* out)
{
    const int i = blockIdx.x * 64 + __lane_id();
    if (i >= n) return;
    const int x = q[4 * i], y = q[4 * i + 1], lim = q[4 * i + 2], dir = q[4 * i + 3];
    bool mism;
    const int s = dir > 0 ? match_fwd(A, B, x, y, lim, mism) : match_rev(A, B, x, y, lim, mism);
    out[2 * i] = s;
    out[2 * i + 1] = mism ? 1 : 0;
}

// test 3: DPP lane moves; per lane: [lane_below(v), lane_above(v), the
// scans, the edge moves, the DPP suffix max, pick3 with sel = L % 3 - 1]
__global__ __launch_bounds__(64) void st_dpp(const int* in, int* out)
{
    const int lane = __lane_id();
    const int v = in[blockIdx.x * 64 + lane];
    int* o = out + (blockIdx.x * 64 + lane) * 8;
    o[0] = lane_below(v);
    o[1] = lane_above(v);
    // the aligner's scans: suffix max over values >= -1, prefix min over values >= 0
    o[2] = wave_suffix_max<true>(v);
    o[3] = wave_prefix_min(v < 0 ? 2147483647 : v);
    o[4] = lane_above_or(v, -7);
    o[5] = lane_below_or(v, -9);
    o[6] = wave_suffix_max<false>(v);
    o[7] = pick3(v, lane % 3 - 1);
}

}  // namespace mi

using namespace mi;

extern "C" int mi_selftest(int which, const void* in, long long nin, const void* in2, long long nin2,
                           const void* q, int nq, int arg, void* out, long long nout)
{
    void *din = nullptr, *din2 = nullptr, *dq = nullptr, *dout = nullptr;
    int* dctr = nullptr;
    int rc = MI_ERR_DEVICE;
    const int* abort_flag = mi_abort_flag_device();
    if (abort_flag == nullptr) return MI_ERR_DEVICE;
    if (nin > 0 && (hipMalloc(&din, nin) != hipSuccess || hipMemcpy(din, in, nin, hipMemcpyHostToDevice) != hipSuccess)) goto done;
    if (nin2 > 0 && (hipMalloc(&din2, nin2) != hipSuccess || hipMemcpy(din2, in2, nin2, hipMemcpyHostToDevice) != hipSuccess)) goto done;
    if (nq > 0 && (hipMalloc(&dq, 16LL * nq) != hipSuccess || hipMemcpy(dq, q, 16LL * nq, hipMemcpyHostToDevice) != hipSuccess)) goto done;
    if (hipMalloc(&dout, nout) != hipSuccess || hipMemset(dout, 0, nout) != hipSuccess) goto done;
    if (hipMalloc(&dctr, sizeof(int)) != hipSuccess || hipMemset(dctr, 0, sizeof(int)) != hipSuccess) goto done;
    switch (which) {
    case 0:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(st_persistent), dim3(arg), dim3(64), 0, 0, (int)(nout / 4), dctr, (int*)dout, abort_flag);
        break;
    case 1:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(st_lanes), dim3((int)(nin / 256)), dim3(64), 0, 0, (const int*)din, arg, (int*)dout);
        break;
    case 2:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(st_match), dim3((nq + 63) / 64), dim3(64), 0, 0, (const uint64_t*)din + 1, (const uint64_t*)din2 + 1, (const int*)dq, nq,
                                         (int*)dout);
        break;
    case 3:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(st_dpp), dim3((int)(nin / 256)), dim3(64), 0, 0, (const int*)din, (int*)dout);
        break;
    default:
        rc = MI_ERR_ARG;
        goto done;
    }
    if (hipGetLastError() != hipSuccess) goto done;
    if (mi_wait_watchdog(0, "mi_selftest") != 0) { rc = MI_ERR_TIMEOUT; goto done; }
    if (hipMemcpy(out, dout, nout, hipMemcpyDeviceToHost) != hipSuccess) goto done;
    rc = 0;
done:
    (void)hipFree(din);
    (void)hipFree(din2);
    (void)hipFree(dq);
    (void)hipFree(dout);
    (void)hipFree(dctr);
    return rc;
}
