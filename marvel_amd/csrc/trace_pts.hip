// Compute_Trace_PTS (dalign/align.c:5286-5397) on gfx950: the exact
// alignment between consecutive trace points of a daligner path, by the
// O(NP) wave of iter_np (dalign/align.c:4556-4916), in the LOWERMOST,
// GREEDIEST or UPPERMOST tie-breaking mode.
//
// One wave per path, one lane per trace-point segment: the segments of a
// path are independent alignments of at most trace_spacing A bases, so
// lane i runs the whole sequential O(NP) wave of segment i (64 segments at a
// time) on its own rows in HBM.  The rows keep the reference's layout (rows
// of trace_spacing + nmax + 3 ints, column 0 at offset nmax + 1, forward
// rows d = -2..dmax followed by the back-pointer rows), sized from the
// path-wide maxima of the trace points like the reference.  Each lane leaves
// its segment's indel positions in a staging area; a wave prefix sum places
// them in path order, and the difference counts are summed across lanes.
// Bases come from the 2-bit packed reads, compared 32 at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "la_wave.hpp"
#include "mi_common.h"

namespace mi {

constexpr int TR_OK = 0;
constexpr int TR_BIG = 1;    // rows do not fit this level's per-lane area
constexpr int TR_SELF = 2;   // self comparison on diagonal 0 (the reference exits)

struct TrTask {
    long long aword, bword;     // word 0 of the A / B read in the packed buffer
    long long pts;              // offset of the u16 trace points
    long long out;              // offset of this path's int trace (capacity: its A + B span + 16)
    int abpos, bbpos, aepos, bepos, tlen;
    int selfie;                 // A and B are the same sequence object
};

struct TrOut {
    int status, diffs, ntrace, pad;
};

// rows d = -2..dmax of one table, `s` ints apart; row d, column k at base[(d+2)*s + k]
struct Rows {
    int* base;
    long long s;
    __device__ __forceinline__ int* operator[](int d) const { return base + (d + 2) * s; }
};

// the O(NP) wave of one segment; returns its difference count or -1
__device__ int np_segment_lane(const uint64_t* A, int sa, int M, const uint64_t* B, int sb, int N, int mode, int selfie,
                               Rows PVF, Rows PHF, int* stage, int* nstage)
{
    const int del = M - N;
    int D, low, hgh, posl = -INT32_MAX, posh = INT32_MAX;
    if (del >= 0) {
        low = 0;
        hgh = del;
    } else {
        low = del;
        hgh = 0;
    }
    if (selfie) {   // B and A are the same read: stay off its main diagonal
        if (sb == sa) return -1;
        if (sb < sa) posl = (sb - sa) + 1;
        else posh = (sb - sa) - 1;
    }
    int *F2, *F1 = PVF[-2], *F0 = PVF[-1], *HF;
    for (int k = low - 1; k <= hgh + 1; k++) F1[k] = F0[k] = -2;
    F0[0] = -1;
    low += 1;
    hgh -= 1;
    for (D = 0;; D++) {
        F2 = F1;
        F1 = F0;
        F0 = PVF[D];
        HF = PHF[D];
        if ((D & 1) == 0) {
            if (low > posl) low -= 1;
            if (hgh < posh) hgh += 1;
        }
        F0[hgh + 1] = F0[low - 1] = -2;
        // furthest reaching point on diagonal k (A index = B index + k) from
        // its three predecessors, then the snake
        auto step = [&](int k, int i, int am, int ap, int mdir, int pdir) -> int {
            const int ac = F1[k] + 1;
            int j;
            if (ac < am) {
                if (ap < am) { HF[k] = mdir; j = am; }
                else { HF[k] = pdir; j = ap; }
            } else if (ap < ac) { HF[k] = 0; j = ac; }
            else { HF[k] = pdir; j = ap; }
            const int lim = (N < i) ? N : i;
            if (j < lim) {
                bool mism;
                j += match_fwd(A, B, sa + j + k, sb + j, lim - j, mism);
            }
            F0[k] = j;
            return j;
        };
        int j = -2, i = M - hgh;
        for (int k = hgh; k > del; k--) {
            j = step(k, i, F2[k - 1], j + 1, -1, 4);
            i += 1;
        }
        j = -2;
        i = M - low;
        for (int k = low; k < del; k++) {
            j = step(k, i, j, F2[k + 1] + 1, 2, 1);
            i -= 1;
        }
        step(del, M - del, j, F0[del + 1] + 1, 2, 4);
        if (F0[del] >= N) break;
    }

    // trace back, turning the back-pointers into forward links
    int k, h, m, e, c;
    const int ap0 = -sa - 1, bp0 = sb + 1;
    PHF[0][0] = 3;
    c = N;
    k = del;
    e = PHF[D][k];
    PHF[D][k] = 3;
    auto back_match = [&](int kk, int cc, int mm) -> int {   // c after skipping matches A[sa+kk+c] == B[sb+c], c >= mm
        if (cc < mm) return cc;
        bool mism;
        return cc - match_rev(A, B, sa + kk + cc + 1, sb + cc + 1, cc - mm + 1, mism);
    };
    while (e != 3) {
        h = k + e;
        if (e > 1) h -= 3;
        else if (e == 0) D -= 1;
        else D -= 2;
        if (mode == 1 && h < k) {   // UPPERMOST
            m = (k < 0) ? -k : 0;
            if (PVF[D][h] <= c) c = PVF[D][h] - 1;
            c = back_match(k, c, m);
            if (e == -1) {
                if (c <= PVF[D + 2][k + 1]) { e = 4; h = k + 1; D = D + 2; }
                else if (c == PVF[D + 1][k]) { e = 0; h = k; D = D + 1; }
                else PVF[D][h] = c + 1;
            } else {
                m = (k == del) ? D : D - 2;
                if (c <= PVF[m][k + 1]) { e = (k == del) ? 4 : 1; h = k + 1; D = m; }
                else if (c == PVF[D - 1][k]) { e = 0; h = k; D = D - 1; }
                else PVF[D][h] = c + 1;
            }
        } else if (mode == -1 && h > k) {   // LOWERMOST
            m = (k < 0) ? -k : 0;
            if (PVF[D][h] < c) c = PVF[D][h];
            c = back_match(k, c, m);
            if (e == 1) {
                if (c < PVF[D + 2][k - 1]) { e = 2; h = k - 1; D = D + 2; }
                else if (c == PVF[D + 1][k]) { e = 0; h = k; D = D + 1; }
                else PVF[D][h] = c--;
            } else {
                m = (k == del) ? D : D - 2;
                if (c < PVF[m][k - 1]) { e = (k == del) ? 2 : -1; h = k - 1; D = m; }
                else if (c == PVF[D - 1][k]) { e = 0; h = k; D = D - 1; }
                else PVF[D][h] = c--;
            }
        }
        m = PHF[D][h];
        PHF[D][h] = e;
        e = m;
        k = h;
    }
    int n = 0;
    k = D = 0;
    e = PHF[D][k];
    while (e != 3) {
        h = k - e;
        c = PVF[D][k];
        if (e > 1) h += 3;
        else if (e == 0) D += 1;
        else D += 2;
        if (h > k) stage[n++] = bp0 + c;
        else if (h < k) stage[n++] = ap0 - (c + k);
        k = h;
        e = PHF[D][h];
    }
    *nstage = n;
    return D + abs(del);
}

// inclusive prefix sum over the wave
__device__ __forceinline__ int wave_incl_sum(int v)
{
    const int lane = __lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        const int w = __shfl_up(v, o, 64);
        if (lane >= o) v += w;
    }
    return v;
}

__global__ __launch_bounds__(64) void k_trace_pts(const TrTask* __restrict__ tasks, const int* __restrict__ order,
                                                  int ntasks, const uint64_t* __restrict__ pk,
                                                  const uint16_t* __restrict__ points, int TS, int mode,
                                                  int* __restrict__ trace, TrOut* __restrict__ outs,
                                                  int* __restrict__ scratch, long long lane_cap,
                                                  int* __restrict__ counter, const int* abort_flag)
{
    const int lane = __lane_id();
    int* mine = scratch + ((size_t)blockIdx.x * 64 + lane) * (size_t)lane_cap;
    for (;;) {
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(counter, lane == 0 ? 1 : 0));
        const int ab = __builtin_amdgcn_readfirstlane(__atomic_load_n(abort_flag, __ATOMIC_RELAXED));
        if (t >= ntasks || ab) break;
        const int ti = order ? order[t] : t;
        const TrTask tk = tasks[ti];
        const uint16_t* P = points + tk.pts;
        const uint64_t* A = pk + tk.aword;
        const uint64_t* B = pk + tk.bword;
        const int nseg = tk.tlen >= 2 ? tk.tlen / 2 : 1;
        // path-wide maxima of the recorded differences and B advances (dalign/align.c:5321-5331)
        int dmax = 0, nmax = 0;
        for (int s = lane; 2 * s + 1 < tk.tlen; s += 64) {
            dmax = max(dmax, (int)P[2 * s]);
            nmax = max(nmax, (int)P[2 * s + 1]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            dmax = max(dmax, __shfl_xor(dmax, o, 64));
            nmax = max(nmax, __shfl_xor(nmax, o, 64));
        }
        if (tk.tlen <= 1) nmax = tk.bepos - tk.bbpos;
        const long long s_row = TS + nmax + 3;
        const long long rows = 2LL * (dmax + 3) + 2;
        const long long need = rows * s_row + (TS + nmax + 16);
        int status = TR_OK, diffs = 0;
        int placed = 0;   // trace words of earlier batches
        if (need > lane_cap) status = TR_BIG;
        else {
            // forward rows after one spare row, then the back-pointer rows;
            // column 0 at offset nmax + 1 (dalign/align.c:5330-5345)
            const Rows PVF{mine + s_row + (nmax + 1), s_row};
            const Rows PHF{mine + s_row + (nmax + 1) + (dmax + 3) * s_row, s_row};
            int* stage = mine + rows * s_row;
            const int a0 = (tk.abpos / TS) * TS;
            int bsum = 0;   // B offset of this batch of 64 segments
            for (int s0 = 0; s0 < nseg; s0 += 64) {
                const int s = s0 + lane;
                const bool act = s < nseg;
                // B advance of segment s (the last one ends at bepos) and its prefix in the batch
                const int badv = (act && s < nseg - 1) ? (int)P[2 * s + 1] : 0;
                const int incl = wave_incl_sum(badv);
                int n = 0, d = 0;
                if (act) {
                    const int sa = (s == 0) ? tk.abpos : a0 + s * TS;
                    const int ea = (s == nseg - 1) ? tk.aepos : a0 + (s + 1) * TS;
                    const int sb = tk.bbpos + bsum + (incl - badv);
                    const int eb = (s == nseg - 1) ? tk.bepos : sb + badv;
                    d = np_segment_lane(A, sa, ea - sa, B, sb, eb - sb, mode, tk.selfie, PVF, PHF, stage, &n);
                }
                if (__any(d < 0)) {
                    status = TR_SELF;
                    break;
                }
                // place the batch's indel positions in segment order
                const int ninc = wave_incl_sum(n);
                int* dst = trace + tk.out + placed + (ninc - n);
                for (int q = 0; q < n; q++) dst[q] = stage[q];
                placed += __shfl(ninc, 63, 64);
                int dsum = d;
                for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o, 64);
                diffs += dsum;
                bsum += __shfl(incl, 63, 64);
            }
        }
        if (lane == 0) outs[ti] = TrOut{status, diffs, placed, 0};
        wave_sync();
    }
}

constexpr int TR_UNRUN = 7;   // the launch stopped before reaching the task

__global__ void k_tr_mark_unrun(const int* __restrict__ order, int n, TrOut* __restrict__ outs)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) outs[order[i]].status = TR_UNRUN;
}

// device buffers of mi_trace_pts_host, kept between calls (one set per device)
struct TrBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool grow(size_t need)
    {
        if (need <= bytes) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        const size_t b = std::max(need, (size_t)1 << 16);
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = b;
        return true;
    }
};
struct TrWorkspace {
    TrBuf tasks, order, packed, points, trace, outs, ctr, scratch;
};
static std::mutex g_tr_lock;
static TrWorkspace& tr_workspace()
{
    static TrWorkspace ws[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    return ws[dev];
}

}  // namespace mi

using namespace mi;

extern "C" {

// Batched Compute_Trace_PTS over host arrays (see include/marvl_align.h).
// The per-lane row area of each task is known on the host from its trace
// points (the path-wide maxima the rows are sized by), so tasks are split
// into a level that fits 2^15 ints per lane and a level sized by the largest
// remaining task, with grids capped so the scratch stays within a budget.
// Device buffers are kept between calls (grown, never shrunk).
int mi_trace_pts_host(int trace_spacing, int mode, const uint64_t* packed, long long npacked, const uint16_t* points,
                      long long npoints, const MiTraceTask* tasks, int ntasks, MiTraceResult* results, int* trace,
                      long long ntrace)
{
    if (ntasks <= 0) return 0;
    if (trace_spacing <= 0 || mode < -1 || mode > 1) return MI_ERR_ARG;
    std::vector<TrTask> tk(ntasks);
    std::vector<long long> need(ntasks);
    long long tro = 0;
    for (int i = 0; i < ntasks; i++) {
        const MiTraceTask& s = tasks[i];
        if (s.aword < 1 || s.bword < 1 || s.aword + (s.alen + 31) / 32 + 2 > npacked ||
            s.bword + (s.blen + 31) / 32 + 2 > npacked || s.points < 0 || s.points + s.tlen > npoints ||
            s.abpos < 0 || s.aepos > s.alen || s.abpos > s.aepos || s.bbpos < 0 || s.bepos > s.blen ||
            s.bbpos > s.bepos || s.tlen < 0 || (s.tlen & 1))
            return MI_ERR_ARG;
        // segment boundaries must follow the trace points (dalign/align.c:5360-5375)
        const int nseg = s.tlen >= 2 ? s.tlen / 2 : 1;
        if (nseg > 1 && (s.abpos / trace_spacing + nseg - 1) * trace_spacing >= s.aepos) return MI_ERR_ARG;
        long long bsum = s.bbpos;
        int dmax = 0, nmax = 0;
        for (int q = 0; q + 1 < s.tlen; q += 2) {
            dmax = std::max(dmax, (int)points[s.points + q]);
            nmax = std::max(nmax, (int)points[s.points + q + 1]);
            if (q + 2 < s.tlen - 1) bsum += points[s.points + q + 1];
        }
        if (bsum > s.bepos) return MI_ERR_ARG;
        if (s.tlen <= 1) nmax = s.bepos - s.bbpos;
        // the kernel's row area: (2 (dmax + 3) + 2) rows of (TS + nmax + 3) ints plus the staging
        const long long s_row = trace_spacing + nmax + 3;
        need[i] = (2LL * (dmax + 3) + 2) * s_row + (trace_spacing + nmax + 16);
        tk[i].aword = s.aword;
        tk[i].bword = s.bword;
        tk[i].pts = s.points;
        tk[i].abpos = s.abpos;
        tk[i].bbpos = s.bbpos;
        tk[i].aepos = s.aepos;
        tk[i].bepos = s.bepos;
        tk[i].tlen = s.tlen;
        tk[i].selfie = (s.aword == s.bword && s.selfie) ? 1 : 0;
        tk[i].out = tro;
        tro += (long long)(s.aepos - s.abpos) + (s.bepos - s.bbpos) + 16;
    }
    // two levels: tasks whose rows fit 2^15 ints per lane, then the rest
    const long long kSmall = 1LL << 15;
    const long long kBudget = 4LL << 30;   // scratch bytes per level
    std::vector<int> lv[2];
    long long cap[2] = {0, 0};
    for (int i = 0; i < ntasks; i++) {
        const int l = need[i] <= kSmall ? 0 : 1;
        lv[l].push_back(i);
        cap[l] = std::max(cap[l], need[i]);
    }
    std::lock_guard<std::mutex> g(g_tr_lock);
    TrWorkspace& W = tr_workspace();
    long long scr_ints = 0;
    int grids[2] = {0, 0};
    for (int l = 0; l < 2; l++) {
        if (lv[l].empty()) continue;
        const long long per_wave = cap[l] * 64;
        long long gmax = std::max(1LL, kBudget / (per_wave * (long long)sizeof(int)));
        grids[l] = (int)std::min<long long>(std::min<long long>(mi_num_cus() * 4, (long long)lv[l].size()), gmax);
        scr_ints = std::max(scr_ints, per_wave * grids[l]);
    }
    if (!W.tasks.grow(sizeof(TrTask) * ntasks) || !W.order.grow(sizeof(int) * ntasks) ||
        !W.packed.grow(sizeof(uint64_t) * npacked) || !W.points.grow(sizeof(uint16_t) * std::max(npoints, 1LL)) ||
        !W.trace.grow(sizeof(int) * std::max(tro, 1LL)) || !W.outs.grow(sizeof(TrOut) * ntasks) ||
        !W.ctr.grow(256) || !W.scratch.grow(sizeof(int) * (size_t)scr_ints)) {
        mi_set_error("mi_trace_pts_host: device memory");
        return MI_ERR_DEVICE;
    }
    TrTask* dtk = (TrTask*)W.tasks.p;
    int* dorder = (int*)W.order.p;
    TrOut* dout = (TrOut*)W.outs.p;
    int* dtr = (int*)W.trace.p;
    int* dctr = (int*)W.ctr.p;
    hipStream_t st = 0;
    const int* flag = mi_abort_flag_device();
    if (flag == nullptr) return MI_ERR_DEVICE;
    if (hipMemcpy(dtk, tk.data(), sizeof(TrTask) * ntasks, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(W.packed.p, packed, sizeof(uint64_t) * npacked, hipMemcpyHostToDevice) != hipSuccess ||
        (npoints > 0 &&
         hipMemcpy(W.points.p, points, sizeof(uint16_t) * npoints, hipMemcpyHostToDevice) != hipSuccess)) {
        mi_set_error("mi_trace_pts_host: HIP error");
        return MI_ERR_DEVICE;
    }
    std::vector<TrOut> ho(ntasks);
    for (int l = 0; l < 2; l++) {
        const int n = (int)lv[l].size();
        if (n == 0) continue;
        if (hipMemcpy(dorder, lv[l].data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(dctr, 0, sizeof(int)) != hipSuccess) {
            mi_set_error("mi_trace_pts_host: HIP error");
            return MI_ERR_DEVICE;
        }
        hipLaunchKernelGGL(k_tr_mark_unrun, dim3((n + 255) / 256), dim3(256), 0, st, dorder, n, dout);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace_pts), dim3(grids[l]), dim3(64), 0, st, dtk, dorder, n,
                           (const uint64_t*)W.packed.p, (const uint16_t*)W.points.p, trace_spacing, mode, dtr, dout,
                           (int*)W.scratch.p, cap[l], dctr, flag);
        if (hipGetLastError() != hipSuccess) {
            mi_set_error("mi_trace_pts_host: launch failed");
            return MI_ERR_DEVICE;
        }
        const int wd = mi_wait_watchdog(st, "mi_trace_pts_host");
        if (wd > 0) {
            mi_set_error("mi_trace_pts_host: kernel exceeded the watchdog deadline and was stopped");
            return MI_ERR_TIMEOUT;
        }
        if (wd < 0) {
            mi_set_error("mi_trace_pts_host: HIP error");
            return MI_ERR_DEVICE;
        }
    }
    std::vector<int> htr(std::max(tro, 1LL));
    if (hipMemcpy(ho.data(), dout, sizeof(TrOut) * ntasks, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(htr.data(), dtr, sizeof(int) * std::max(tro, 1LL), hipMemcpyDeviceToHost) != hipSuccess) {
        mi_set_error("mi_trace_pts_host: HIP error");
        return MI_ERR_DEVICE;
    }
    long long o = 0;
    for (int i = 0; i < ntasks; i++) {
        const TrOut& h = ho[i];
        if (h.status == TR_UNRUN || h.status == TR_BIG) {
            mi_set_error(h.status == TR_UNRUN ? "mi_trace_pts_host: a launch stopped before reaching every task"
                                              : "mi_trace_pts_host: a task's rows exceed its level's area");
            return MI_ERR_TIMEOUT;
        }
        MiTraceResult& r = results[i];
        r.status = h.status;
        r.diffs = h.diffs;
        r.ntrace = h.status == TR_OK ? h.ntrace : 0;
        r.trace = o;
        // a path's trace holds at most one entry per base of its A and B spans
        const long long slice = (long long)(tk[i].aepos - tk[i].abpos) + (tk[i].bepos - tk[i].bbpos) + 16;
        if (r.ntrace < 0 || r.ntrace > slice) {
            mi_set_error("mi_trace_pts_host: trace longer than its slice");
            return MI_ERR_DEVICE;
        }
        if (o + r.ntrace > ntrace) return MI_ERR_ARG;
        memcpy(trace + o, htr.data() + tk[i].out, sizeof(int) * r.ntrace);
        o += r.ntrace;
    }
    return 0;
}

}  // extern "C"
