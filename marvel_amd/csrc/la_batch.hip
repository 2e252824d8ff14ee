// Batched Local_Alignment on gfx950: a persistent grid of single-wave
// workgroups pulls alignment tasks from a device counter; every task is one
// daligner Local_Alignment (dalign/align.c:1922-2060) computed by one wave64
// (la_wave.hpp).  Tasks that outgrow the 64-diagonal register window or the
// per-wave pebble pool are re-run by a wider instantiation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "la_wave.hpp"
#include "mi_common.h"

namespace mi {

struct LaTask {
    long long apk, bpk;       // word offset of word 0 of the A / B read in the packed buffer
    long long tr;             // u16 offset of this task's trace area (4*maxtp entries)
    int alen, blen, comp, low, hgh, anti, minp, maxp, maxtp;
};

// Host view of a task's output record (first 18 of its 64 ints).
struct LaOut {
    int abpos, bbpos, aepos, bepos, diffs;
    int b_abpos, b_bbpos, b_aepos, b_bepos;
    int atlen, btlen, afirst, bfirst;   // traces start at origin + first (origins: maxtp, 3*maxtp)
    int status;
    unsigned cells_lo;
    int cells_hi;
    unsigned snake_lo;
    int snake_hi;
    int pad[46];
};
static_assert(sizeof(LaOut) == 64 * sizeof(int), "output record is one int per lane");

// Output record of one task as 64 ints, one per lane, so the whole wave
// stores it without a lane-dependent branch.
constexpr int kOutInts = 64;

template <int NCH>
__global__ __launch_bounds__(64, NCH == 1 ? 8 : 1) void la_batch_kernel(const LaTask* __restrict__ tasks, const int* __restrict__ order,
                                                      int ntasks, const uint64_t* __restrict__ pk, DevSpec sp,
                                                      uint16_t* __restrict__ trbuf, int* __restrict__ outs,
                                                      Peb* __restrict__ pools, int pool_cap, Peb* __restrict__ ptsbuf,
                                                      int pts_cap, int* __restrict__ counter,
                                                      const int* abort_flag)
{
    const int wave = blockIdx.x;
    const int lane = __lane_id();
    Peb* pool = pools + (size_t)wave * pool_cap;
    Peb* pts = ptsbuf + (size_t)wave * 2 * pts_cap;
    for (;;) {
        // Every lane executes the atomic (only lane 0 adds 1), so no
        // lane-dependent branch precedes the broadcast and the loop exit is
        // wave-uniform; the first active lane is lane 0.
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(counter, lane == 0 ? 1 : 0));
        const int ab = __builtin_amdgcn_readfirstlane(__atomic_load_n(abort_flag, __ATOMIC_RELAXED));
        if (t >= ntasks || ab) break;
        const int ti = order ? order[t] : t;
        const LaTask tk = tasks[ti];
        uint16_t* atr = trbuf + tk.tr + tk.maxtp;
        uint16_t* btr = trbuf + tk.tr + 3 * (long long)tk.maxtp;
        WaveCount wc;
        LaResult r = local_alignment<NCH>(sp, pk + tk.apk, tk.alen, pk + tk.bpk, tk.blen, tk.comp, tk.low, tk.hgh,
                                          tk.anti, tk.minp, tk.maxp, pool, pool_cap, pts, pts_cap, atr, btr, tk.maxtp,
                                          abort_flag, wc);
        const int vals[18] = {r.abpos, r.bbpos, r.aepos, r.bepos, r.diffs, r.b_abpos, r.b_bbpos, r.b_aepos,
                              r.b_bepos, r.atlen, r.btlen, r.afirst, r.bfirst, r.status,
                              (int)(wc.cells & 0xffffffff), (int)(wc.cells >> 32),
                              (int)(wc.snake & 0xffffffff), (int)(wc.snake >> 32)};
        int v = 0;
#pragma unroll
        for (int q = 0; q < 18; q++) v = (lane == q) ? vals[q] : v;
        outs[(size_t)ti * kOutInts + lane] = v;
    }
}

static int launch_la(int nch, int grid, const LaTask* tasks, const int* order, int n, const uint64_t* pk,
                     const DevSpec& sp, uint16_t* trbuf, int* outs, Peb* pools, int pool_cap, Peb* pts,
                     int pts_cap, int* counter, const int* abort_flag, hipStream_t st)
{
    if (hipMemsetAsync(counter, 0, sizeof(int), st) != hipSuccess) return -1;
    switch (nch) {
    case 1:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(la_batch_kernel<1>), dim3(grid), dim3(64), 0, st, tasks, order, n, pk, sp, trbuf, outs, pools, pool_cap, pts, pts_cap, counter, abort_flag);
        break;
    case 4:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(la_batch_kernel<4>), dim3(grid), dim3(64), 0, st, tasks, order, n, pk, sp, trbuf, outs, pools, pool_cap, pts, pts_cap, counter, abort_flag);
        break;
    case 16:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(la_batch_kernel<16>), dim3(grid), dim3(64), 0, st, tasks, order, n, pk, sp, trbuf, outs, pools, pool_cap, pts, pts_cap, counter, abort_flag);
        break;
    default:
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mi

using namespace mi;

// ---- C-ABI -----------------------------------------------------------------

extern "C" {

// Batched Local_Alignment over host arrays (see include/marvl_align.h).
int mi_la_batch_host(const MiAlignParams* prm, const uint64_t* packed, long long npacked, const MiLaTask* tasks,
                     int ntasks, MiLaResult* results, uint16_t* traces, long long ntrace)
{
    if (ntasks <= 0) return 0;
    DevSpecHost* hs = mi_spec_device(prm);
    if (hs == nullptr) return MI_ERR_DEVICE;
    std::vector<LaTask> tk(ntasks);
    long long tro = 0;
    for (int i = 0; i < ntasks; i++) {
        const MiLaTask& s = tasks[i];
        const int mx = std::max(s.alen, s.blen);
        const int maxtp = 2 * (mx / prm->trace_space + 2);
        if (s.aword < 1 || s.bword < 1 || s.aword + (s.alen + 31) / 32 + 2 > npacked ||
            s.bword + (s.blen + 31) / 32 + 2 > npacked || s.alen < 0 || s.blen < 0 || s.hgh < s.low)
            return MI_ERR_ARG;
        tk[i].apk = s.aword;
        tk[i].bpk = s.bword;
        tk[i].alen = s.alen;
        tk[i].blen = s.blen;
        tk[i].comp = s.comp;
        tk[i].low = s.low;
        tk[i].hgh = s.hgh;
        tk[i].anti = s.anti;
        const bool selfie = (s.aword == s.bword && s.selfie);
        tk[i].minp = (s.lbord < 0) ? ((selfie && s.low >= 0) ? 1 : -INT32_MAX) : s.low - s.lbord;
        tk[i].maxp = (s.hbord < 0) ? ((selfie && s.hgh <= 0) ? -1 : INT32_MAX) : s.hgh + s.hbord;
        tk[i].maxtp = maxtp;
        tk[i].tr = tro;
        tro += 4LL * maxtp;
    }
    int grid = mi_num_cus() * 8;
    if (grid > ntasks) grid = ntasks;
    int pool_cap = 1 << 17;
    int pts_cap = 8192;

    LaTask* dtk = nullptr;
    int* dorder = nullptr;
    uint64_t* dpk = nullptr;
    uint16_t* dtr = nullptr;
    LaOut* dout = nullptr;
    Peb* dpool = nullptr;
    Peb* dpts = nullptr;
    int* dctr = nullptr;
    int rc = MI_ERR_DEVICE;
    std::vector<LaOut> ho(ntasks);
    hipStream_t st = 0;
    if (hipMalloc(&dtk, sizeof(LaTask) * ntasks) != hipSuccess) goto done;
    if (hipMalloc(&dorder, sizeof(int) * ntasks) != hipSuccess) goto done;
    if (hipMalloc(&dpk, sizeof(uint64_t) * npacked) != hipSuccess) goto done;
    if (hipMalloc(&dtr, sizeof(uint16_t) * std::max(tro, 1LL)) != hipSuccess) goto done;
    if (hipMalloc(&dout, sizeof(LaOut) * ntasks) != hipSuccess) goto done;
    if (hipMalloc(&dctr, sizeof(int)) != hipSuccess) goto done;
    if (hipMemcpy(dtk, tk.data(), sizeof(LaTask) * ntasks, hipMemcpyHostToDevice) != hipSuccess) goto done;
    if (hipMemcpy(dpk, packed, sizeof(uint64_t) * npacked, hipMemcpyHostToDevice) != hipSuccess) goto done;
    if (hipMemset(dtr, 0, sizeof(uint16_t) * std::max(tro, 1LL)) != hipSuccess) goto done;
    {
        // pass 1: 64-wide window for every task; then wider windows / pools for the ones that need it
        std::vector<int> todo(ntasks);
        for (int i = 0; i < ntasks; i++) todo[i] = i;
        const int nchs[3] = {1, 4, 16};
        int level = 0;
        while (!todo.empty()) {
            const int nch = nchs[std::min(level, 2)];
            // level 0 (64-lane window) runs 8 waves per SIMD with smaller pools;
            // tasks that fill them re-run at the next level
            const int pc = (level == 0) ? pool_cap >> 2 : ((level >= 3) ? pool_cap << 3 : pool_cap);
            int g = std::min<int>(level == 0 ? grid * 4 : grid, (int)todo.size());
            if (level >= 3) g = std::min(g, 256);
            if (dpool) { (void)hipFree(dpool); dpool = nullptr; }
            if (dpts) { (void)hipFree(dpts); dpts = nullptr; }
            if (hipMalloc(&dpool, sizeof(Peb) * (size_t)pc * g) != hipSuccess) goto done;
            if (hipMalloc(&dpts, sizeof(Peb) * (size_t)2 * pts_cap * g) != hipSuccess) goto done;
            if (hipMemcpy(dorder, todo.data(), sizeof(int) * todo.size(), hipMemcpyHostToDevice) != hipSuccess) goto done;
            if (launch_la(nch, g, dtk, dorder, (int)todo.size(), dpk, hs->spec, dtr, (int*)dout, dpool, pc, dpts, pts_cap,
                          dctr, mi_abort_flag_device(), st))
                goto done;
            if (mi_wait_watchdog(st, "mi_la_batch_host") != 0) { rc = MI_ERR_TIMEOUT; goto done; }
            if (hipMemcpy(ho.data(), dout, sizeof(LaOut) * ntasks, hipMemcpyDeviceToHost) != hipSuccess) goto done;
            std::vector<int> again;
            for (int i : todo)
                if (ho[i].status == LA_WIDE || ho[i].status == LA_POOL) again.push_back(i);
            todo.swap(again);
            level += 1;
            if (level > 3 && !todo.empty()) break;
        }
    }
    {
        std::vector<uint16_t> htr(std::max(tro, 1LL));
        if (hipMemcpy(htr.data(), dtr, sizeof(uint16_t) * std::max(tro, 1LL), hipMemcpyDeviceToHost) != hipSuccess) goto done;
        long long o = 0;
        for (int i = 0; i < ntasks; i++) {
            const LaOut& h = ho[i];
            MiLaResult& r = results[i];
            r.status = h.status;
            r.abpos = h.abpos; r.bbpos = h.bbpos; r.aepos = h.aepos; r.bepos = h.bepos; r.diffs = h.diffs;
            r.b_abpos = h.b_abpos; r.b_bbpos = h.b_bbpos; r.b_aepos = h.b_aepos; r.b_bepos = h.b_bepos;
            r.atlen = h.status == 0 ? h.atlen : 0;
            r.btlen = h.status == 0 ? h.btlen : 0;
            r.cells = ((long long)h.cells_hi << 32) | h.cells_lo;
            r.snake = ((long long)h.snake_hi << 32) | h.snake_lo;
            r.atrace = o;
            r.btrace = o + r.atlen;
            if (o + r.atlen + r.btlen > ntrace) { rc = MI_ERR_ARG; goto done; }
            const uint16_t* base = htr.data() + tk[i].tr;
            if (r.atlen) memcpy(traces + o, base + tk[i].maxtp + h.afirst, sizeof(uint16_t) * r.atlen);
            if (r.btlen) memcpy(traces + o + r.atlen, base + 3LL * tk[i].maxtp + h.bfirst, sizeof(uint16_t) * r.btlen);
            o += r.atlen + r.btlen;
        }
    }
    rc = 0;
done:
    if (rc == MI_ERR_DEVICE) mi_set_error("mi_la_batch_host: HIP error");
    if (rc == MI_ERR_TIMEOUT) mi_set_error("mi_la_batch_host: kernel exceeded the watchdog deadline and was stopped");
    (void)hipFree(dtk);
    (void)hipFree(dorder);
    (void)hipFree(dpk);
    (void)hipFree(dtr);
    (void)hipFree(dout);
    (void)hipFree(dpool);
    (void)hipFree(dpts);
    (void)hipFree(dctr);
    return rc;
}

}  // extern "C"
