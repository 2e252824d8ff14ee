// Batched Local_Alignment on gfx950: a persistent grid of single-wave
// workgroups pulls alignment tasks from a device counter; every task is one
// daligner Local_Alignment (dalign/align.c:1922-2060) computed by one wave64
// (la_wave.hpp).
//
// Device-resident batch (mi_la_batch_device):
//   * inputs stay in HBM: 2-bit packed reads and MiLaTask records;
//   * each finished task reserves its A and B traces in a compact trace arena
//     with one 64-bit atomic and copies them there, so the output is exactly
//     the traces produced (no worst-case per-task slots, no clearing);
//   * tasks that outgrow the 64-diagonal register window or the per-wave
//     pebble pool append themselves to a device retry list and are re-run by
//     a wider instantiation (128 / 256 / 1,024 diagonals, larger pools); only
//     the retry count comes back to the host between levels;
//   * before a launch every task it covers is marked LA_UNRUN, so a launch
//     stopped early (watchdog flag) can never leave a stale record behind;
//   * the per-wave scratch (pebble pools, chain points, trace staging) is one
//     workspace carved per level.  The library keeps a grown workspace per
//     device, so repeated calls allocate nothing.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "la_wave.hpp"
#include "mi_common.h"

namespace mi {

constexpr int LA_SPACE = 6;   // the call's trace arena is full (task not written)
constexpr int LA_UNRUN = 7;   // the launch stopped before reaching the task

constexpr int kResInts = (int)(sizeof(MiLaResult) / sizeof(int));
static_assert(sizeof(MiLaResult) == 80 && kResInts <= 64, "result record is written one int per lane");

struct BatchScratch {
    Peb* pools;
    int pool_cap;
    Peb* pts;
    int pts_cap;
    uint16_t* latr;   // per wave: 4 * maxtp u16 (A and B traces, each with maxtp before and after its origin)
    int latr_cap;
};

template <int NCH>
__global__ __launch_bounds__(64, NCH == 1 ? 8 : 1) void la_batch_kernel(
    const MiLaTask* __restrict__ tasks, const int* __restrict__ order, int norder, const uint64_t* __restrict__ pk,
    DevSpec sp, MiLaResult* __restrict__ res, uint16_t* __restrict__ arena, long long arena_cap,
    unsigned long long* __restrict__ arena_top, int* __restrict__ retry, int* __restrict__ nretry, BatchScratch ws,
    int* __restrict__ counter, unsigned long long* __restrict__ work, const int* abort_flag)
{
    const int wave = blockIdx.x;
    const int lane = __lane_id();
    Peb* pool = ws.pools + (size_t)wave * ws.pool_cap;
    Peb* pts = ws.pts + (size_t)wave * 2 * ws.pts_cap;
    uint16_t* latr = ws.latr + (size_t)wave * ws.latr_cap;
    const int TS = sp.tspace;
    long long tot_cells = 0, tot_snake = 0;
    for (;;) {
        // Every lane executes the atomic (only lane 0 adds 1), so no
        // lane-dependent branch precedes the broadcast and the loop exit is
        // wave-uniform; the first active lane is lane 0.
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(counter, lane == 0 ? 1 : 0));
        const int ab = __builtin_amdgcn_readfirstlane(__atomic_load_n(abort_flag, __ATOMIC_RELAXED));
        if (t >= norder || ab) break;
        const int ti = order ? order[t] : t;
        const MiLaTask tk = tasks[ti];
        const int mx = tk.alen > tk.blen ? tk.alen : tk.blen;
        const int maxtp = 2 * (mx / TS + 2);
        // Local_Alignment's borders (dalign/align.c:1966-1985): lbord/hbord < 0
        // leave the wave unbounded, except that a read aligned to itself may not
        // reach the main diagonal
        const bool selfie = (tk.aword == tk.bword && tk.selfie);
        const int minp = (tk.lbord < 0) ? ((selfie && tk.low >= 0) ? 1 : -INT32_MAX) : tk.low - tk.lbord;
        const int maxp = (tk.hbord < 0) ? ((selfie && tk.hgh <= 0) ? -1 : INT32_MAX) : tk.hgh + tk.hbord;
        WaveCount wc;
        LaResult r;
        if (4 * maxtp > ws.latr_cap) {
            r.status = LA_POOL;
        } else {
            uint16_t* atr = latr + maxtp;
            uint16_t* btr = latr + 3 * maxtp;
            r = local_alignment<NCH>(sp, pk + tk.aword, tk.alen, pk + tk.bword, tk.blen, tk.comp, tk.low, tk.hgh,
                                     tk.anti, minp, maxp, pool, ws.pool_cap, pts, ws.pts_cap, atr, btr, maxtp,
                                     abort_flag, wc);
        }
        tot_cells += wc.cells;
        tot_snake += wc.snake;
        long long off = 0;
        int status = r.status;
        if (status == LA_OK) {
            const int need = r.atlen + r.btlen;
            unsigned long long o = atomicAdd(arena_top, lane == 0 ? (unsigned long long)need : 0ull);
            off = (long long)__shfl(o, 0, 64);
            if (off + need > arena_cap) {
                status = LA_SPACE;
            } else {
                const uint16_t* ta = latr + maxtp + r.afirst;
                const uint16_t* tb = latr + 3 * maxtp + r.bfirst;
                for (int i = lane; i < r.atlen; i += 64) arena[off + i] = ta[i];
                for (int i = lane; i < r.btlen; i += 64) arena[off + r.atlen + i] = tb[i];
            }
        } else if (status == LA_WIDE || status == LA_POOL) {
            const int q = __shfl(atomicAdd(nretry, lane == 0 ? 1 : 0), 0, 64);
            if (lane == 0) retry[q] = ti;
        }
        const bool ok = status == LA_OK;
        const int vals[kResInts] = {status,
                                    ok ? r.abpos : 0, ok ? r.bbpos : 0, ok ? r.aepos : 0, ok ? r.bepos : 0,
                                    ok ? r.diffs : 0,
                                    ok ? r.b_abpos : 0, ok ? r.b_bbpos : 0, ok ? r.b_aepos : 0, ok ? r.b_bepos : 0,
                                    ok ? r.atlen : 0, ok ? r.btlen : 0,
                                    (int)(off & 0xffffffff), (int)(off >> 32),
                                    (int)((off + (ok ? r.atlen : 0)) & 0xffffffff),
                                    (int)((off + (ok ? r.atlen : 0)) >> 32),
                                    (int)(wc.cells & 0xffffffff), (int)(wc.cells >> 32),
                                    (int)(wc.snake & 0xffffffff), (int)(wc.snake >> 32)};
        int v = 0;
#pragma unroll
        for (int q = 0; q < kResInts; q++) v = (lane == q) ? vals[q] : v;
        if (lane < kResInts) reinterpret_cast<int*>(res + ti)[lane] = v;
    }
    if (lane == 0) {
        atomicAdd(&work[0], (unsigned long long)tot_cells);
        atomicAdd(&work[1], (unsigned long long)tot_snake);
    }
}

// status of every task a launch covers := LA_UNRUN
static __global__ void k_la_mark_unrun(const int* __restrict__ order, int n, MiLaResult* __restrict__ res)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) res[order ? order[i] : (int)i].status = LA_UNRUN;
}

// out[0] = max read length, out[1] = tasks whose reads leave the packed
// buffer or whose band is empty, out[2] = tasks still LA_UNRUN (results)
static __global__ void k_la_check(const MiLaTask* __restrict__ tasks, int n, long long npacked, int* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MiLaTask s = tasks[i];
    const bool bad = s.aword < 1 || s.bword < 1 || s.alen < 0 || s.blen < 0 || s.hgh < s.low ||
                     s.aword + (s.alen + 31) / 32 + 2 > npacked || s.bword + (s.blen + 31) / 32 + 2 > npacked;
    if (bad) atomicAdd(&out[1], 1);
    else atomicMax(&out[0], s.alen > s.blen ? s.alen : s.blen);
}

static __global__ void k_la_count_unrun(const MiLaResult* __restrict__ res, int n, int* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && res[i].status == LA_UNRUN) atomicAdd(&out[2], 1);
}

// ---- workspace ----------------------------------------------------------------
// One batch call needs: a few counters, two retry lists of ntasks ints, and per
// level the per-wave scratch of that level's grid.  Level l runs window NCH[l]
// with pool PCAP[l] pebbles on at most GRID[l] waves.

struct Level {
    int nch;
    int pool_cap;
    int waves;   // upper bound; the scratch may cap it
};

static const int kCounterBytes = 256;

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t wave_bytes(int pool_cap, int pts_cap, int latr_cap)
{
    return (size_t)pool_cap * sizeof(Peb) + (size_t)2 * pts_cap * sizeof(Peb) + (size_t)latr_cap * sizeof(uint16_t);
}

// carve `waves` waves of scratch at p (256-byte aligned pieces)
static BatchScratch carve(char* p, int waves, int pool_cap, int pts_cap, int latr_cap)
{
    BatchScratch s;
    s.pool_cap = pool_cap;
    s.pts_cap = pts_cap;
    s.latr_cap = latr_cap;
    s.pools = (Peb*)p;
    p += align_up((size_t)waves * pool_cap * sizeof(Peb));
    s.pts = (Peb*)p;
    p += align_up((size_t)waves * 2 * pts_cap * sizeof(Peb));
    s.latr = (uint16_t*)p;
    return s;
}

static size_t carve_bytes(int waves, int pool_cap, int pts_cap, int latr_cap)
{
    return align_up((size_t)waves * pool_cap * sizeof(Peb)) + align_up((size_t)waves * 2 * pts_cap * sizeof(Peb)) +
           align_up((size_t)waves * latr_cap * sizeof(uint16_t));
}

static int launch_level(int nch, int grid, const MiLaTask* tasks, const int* order, int n, const uint64_t* pk,
                        const DevSpec& sp, MiLaResult* res, uint16_t* arena, long long arena_cap,
                        unsigned long long* arena_top, int* retry, int* nretry, const BatchScratch& ws, int* counter,
                        unsigned long long* work, const int* abort_flag, hipStream_t st)
{
    if (hipMemsetAsync(counter, 0, sizeof(int), st) != hipSuccess) return -1;
    if (hipMemsetAsync(nretry, 0, sizeof(int), st) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_la_mark_unrun, dim3((n + 255) / 256), dim3(256), 0, st, order, n, res);
    switch (nch) {
    case 1:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(la_batch_kernel<1>), dim3(grid), dim3(64), 0, st, tasks, order, n, pk, sp,
                           res, arena, arena_cap, arena_top, retry, nretry, ws, counter, work, abort_flag);
        break;
    case 2:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(la_batch_kernel<2>), dim3(grid), dim3(64), 0, st, tasks, order, n, pk, sp,
                           res, arena, arena_cap, arena_top, retry, nretry, ws, counter, work, abort_flag);
        break;
    case 4:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(la_batch_kernel<4>), dim3(grid), dim3(64), 0, st, tasks, order, n, pk, sp,
                           res, arena, arena_cap, arena_top, retry, nretry, ws, counter, work, abort_flag);
        break;
    case 16:
        hipLaunchKernelGGL(HIP_KERNEL_NAME(la_batch_kernel<16>), dim3(grid), dim3(64), 0, st, tasks, order, n, pk,
                           sp, res, arena, arena_cap, arena_top, retry, nretry, ws, counter, work, abort_flag);
        break;
    default:
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Where a batch call's scratch lives: a buffer of `bytes` bytes.
struct ScratchArea {
    char* base;
    size_t bytes;
};

struct BatchPlan {
    int pts_cap, latr_cap;
    Level lv[4];
};

static BatchPlan make_plan(int maxlen, int tspace, int cus)
{
    BatchPlan p;
    // a chain holds one pebble per trace point of its read plus the root
    p.pts_cap = maxlen / tspace + 64;
    p.latr_cap = 4 * (2 * (maxlen / tspace + 2));
    const int w0 = mi_tuning().batch_waves_per_cu;
    const int w1 = mi_tuning().batch_wide_waves_per_cu;
    // windows of 64, 128, 256 and 1,024 diagonals; a task re-runs at the
    // next level when its band outgrows the window or its pool fills
    p.lv[0] = Level{1, 1 << 15, cus * (w0 ? w0 : 32)};   // default 8 waves per SIMD
    p.lv[1] = Level{2, 1 << 17, cus * (w1 ? w1 : 16)};
    p.lv[2] = Level{4, 1 << 17, cus * (w1 ? w1 : 16)};
    p.lv[3] = Level{16, 1 << 20, 256};
    return p;
}

// Validates the tasks on the device: the longest read (sizes the per-wave
// scratch) or MI_ERR_ARG when a task's reads leave the packed buffer or its
// band is empty.  d_tmp: 3 device ints.
static int check_tasks(const MiLaTask* d_tasks, int ntasks, long long npacked, int* d_tmp, int* maxlen)
{
    hipStream_t st = 0;
    if (hipMemsetAsync(d_tmp, 0, 3 * sizeof(int), st) != hipSuccess) return MI_ERR_DEVICE;
    hipLaunchKernelGGL(k_la_check, dim3((ntasks + 255) / 256), dim3(256), 0, st, d_tasks, ntasks, npacked, d_tmp);
    int chk[3] = {0, 0, 0};
    if (hipMemcpyAsync(chk, d_tmp, sizeof(chk), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return MI_ERR_DEVICE;
    if (chk[1] != 0) {
        mi_set_error("batch: a task's reads lie outside the packed buffer or its band is empty");
        return MI_ERR_ARG;
    }
    *maxlen = chk[0];
    return 0;
}

// The level loop.  `area` holds the counters, the retry lists and the
// per-level wave scratch; levels whose full grid does not fit run fewer waves.
static int run_batch(const DevSpec& sp, const uint64_t* pk, const MiLaTask* tasks, int ntasks, int maxlen,
                     const int* order, MiLaResult* res, uint16_t* arena, long long arena_cap,
                     unsigned long long* d_arena_top, const ScratchArea& area, const int* abort_flag, bool watchdog,
                     long long* work, const char* who)
{
    char* p = area.base;
    int* ctr = (int*)p;   // [0] queue, [1] retry count, [2] unrun count
    unsigned long long* dwork = (unsigned long long*)(p + 64);
    p += kCounterBytes;
    int* lists[2];
    lists[0] = (int*)p;
    p += align_up((size_t)ntasks * sizeof(int));
    lists[1] = (int*)p;
    p += align_up((size_t)ntasks * sizeof(int));
    const size_t fixed = (size_t)(p - area.base);
    if (fixed > area.bytes) {
        mi_set_error("batch: task count exceeds the scratch");
        return MI_ERR_ARG;
    }
    hipStream_t st = 0;
    if (hipMemsetAsync(area.base, 0, kCounterBytes, st) != hipSuccess) return MI_ERR_DEVICE;
    int chk[3] = {0, 0, 0};
    // kernel time of each level (HIP events on the launch stream) when work is asked for
    static hipEvent_t ev[2] = {nullptr, nullptr};
    if (work) {
        for (int i = 0; i < 8; i++) work[i] = 0;
        for (int i = 0; i < 2; i++)
            if (ev[i] == nullptr && hipEventCreate(&ev[i]) != hipSuccess) return MI_ERR_DEVICE;
    }
    const BatchPlan plan = make_plan(maxlen, sp.tspace, mi_num_cus());
    const int* cur = order;
    int n = ntasks;
    int which = 0;
    for (int level = 0; level < 4 && n > 0; level++) {
        const Level& L = plan.lv[level];
        const size_t per = wave_bytes(L.pool_cap, plan.pts_cap, plan.latr_cap) + 3 * 256;
        long long g = std::min<long long>(L.waves, n);
        const long long fit = (area.bytes - fixed) / per;
        if (g > fit) g = fit;
        while (g > 0 && carve_bytes((int)g, L.pool_cap, plan.pts_cap, plan.latr_cap) > area.bytes - fixed) g--;
        if (g <= 0) {
            mi_set_error("batch: scratch too small for one wave of this level");
            return MI_ERR_ARG;
        }
        const BatchScratch ws = carve(area.base + fixed, (int)g, L.pool_cap, plan.pts_cap, plan.latr_cap);
        int* retry = lists[which];
        if (work && hipEventRecord(ev[0], st) != hipSuccess) return MI_ERR_DEVICE;
        if (launch_level(L.nch, (int)g, tasks, cur, n, pk, sp, res, arena, arena_cap, d_arena_top, retry, ctr + 1, ws,
                         ctr, dwork, abort_flag, st))
            return MI_ERR_DEVICE;
        if (work && hipEventRecord(ev[1], st) != hipSuccess) return MI_ERR_DEVICE;
        if (watchdog) {
            const int w = mi_wait_watchdog(st, who);
            if (w > 0) {
                mi_set_error("batch: kernel exceeded the watchdog deadline and was stopped");
                return MI_ERR_TIMEOUT;
            }
            if (w < 0) return MI_ERR_DEVICE;
        }
        int nr = 0;
        if (hipMemcpyAsync(&nr, ctr + 1, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return MI_ERR_DEVICE;
        if (work) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) return MI_ERR_DEVICE;
            work[2 + level] = (long long)(ms * 1000.0);
            work[6] += nr;
        }
        cur = retry;
        n = nr;
        which ^= 1;
    }
    // tasks still wider than the widest level keep LA_WIDE / LA_POOL; a task
    // left LA_UNRUN means a launch was stopped from outside
    hipLaunchKernelGGL(k_la_count_unrun, dim3((ntasks + 255) / 256), dim3(256), 0, st, res, ntasks, ctr + 2);
    unsigned long long wk[2] = {0, 0};
    if (hipMemcpyAsync(chk, ctr + 2, sizeof(chk), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(wk, dwork, sizeof(wk), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return MI_ERR_DEVICE;
    if (work) {
        work[0] = (long long)wk[0];
        work[1] = (long long)wk[1];
    }
    if (chk[2] != 0) {
        mi_set_error("batch: a launch stopped before reaching every task (abort flag raised)");
        return MI_ERR_TIMEOUT;
    }
    return 0;
}

}  // namespace mi
