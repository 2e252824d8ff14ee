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

}  // namespace mi

using namespace mi;

extern "C" {

// Batched Compute_Trace_PTS over host arrays (see include/marvl_align.h).
int mi_trace_pts_host(int trace_spacing, int mode, const uint64_t* packed, long long npacked, const uint16_t* points,
                      long long npoints, const MiTraceTask* tasks, int ntasks, MiTraceResult* results, int* trace,
                      long long ntrace)
{
    if (ntasks <= 0) return 0;
    if (trace_spacing <= 0 || mode < -1 || mode > 1) return MI_ERR_ARG;
    std::vector<TrTask> tk(ntasks);
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
        for (int q = 1; q < s.tlen - 2; q += 2) bsum += points[s.points + q];
        if (bsum > s.bepos) return MI_ERR_ARG;
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
    const int cus = mi_num_cus();
    struct Lv {
        long long lane_cap;
        int waves;
    };
    const Lv levels[2] = {{1LL << 15, cus * 4}, {1LL << 22, 16}};
    TrTask* dtk = nullptr;
    int* dorder = nullptr;
    uint64_t* dpk = nullptr;
    uint16_t* dpts = nullptr;
    int* dtr = nullptr;
    TrOut* dout = nullptr;
    int* dscr = nullptr;
    int* dctr = nullptr;
    int rc = MI_ERR_DEVICE;
    std::vector<TrOut> ho(ntasks);
    std::vector<int> todo(ntasks);
    for (int i = 0; i < ntasks; i++) todo[i] = i;
    hipStream_t st = 0;
    if (hipMalloc(&dtk, sizeof(TrTask) * ntasks) != hipSuccess) goto done;
    if (hipMalloc(&dorder, sizeof(int) * ntasks) != hipSuccess) goto done;
    if (hipMalloc(&dpk, sizeof(uint64_t) * npacked) != hipSuccess) goto done;
    if (hipMalloc(&dpts, sizeof(uint16_t) * std::max(npoints, 1LL)) != hipSuccess) goto done;
    if (hipMalloc(&dtr, sizeof(int) * std::max(tro, 1LL)) != hipSuccess) goto done;
    if (hipMalloc(&dout, sizeof(TrOut) * ntasks) != hipSuccess) goto done;
    if (hipMalloc(&dctr, sizeof(int)) != hipSuccess) goto done;
    if (hipMemcpy(dtk, tk.data(), sizeof(TrTask) * ntasks, hipMemcpyHostToDevice) != hipSuccess) goto done;
    if (hipMemcpy(dpk, packed, sizeof(uint64_t) * npacked, hipMemcpyHostToDevice) != hipSuccess) goto done;
    if (npoints > 0 && hipMemcpy(dpts, points, sizeof(uint16_t) * npoints, hipMemcpyHostToDevice) != hipSuccess) goto done;
    for (int level = 0; level < 2 && !todo.empty(); level++) {
        const Lv& L = levels[level];
        const int grid = (int)std::min<long long>(L.waves, (long long)todo.size());
        if (dscr) (void)hipFree(dscr);
        dscr = nullptr;
        if (hipMalloc(&dscr, sizeof(int) * (size_t)L.lane_cap * 64 * grid) != hipSuccess) goto done;
        if (hipMemcpy(dorder, todo.data(), sizeof(int) * todo.size(), hipMemcpyHostToDevice) != hipSuccess) goto done;
        if (hipMemset(dctr, 0, sizeof(int)) != hipSuccess) goto done;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace_pts), dim3(grid), dim3(64), 0, st, dtk, dorder, (int)todo.size(),
                           dpk, dpts, trace_spacing, mode, dtr, dout, dscr, L.lane_cap, dctr, mi_abort_flag_device());
        if (hipGetLastError() != hipSuccess) goto done;
        {
            const int wd = mi_wait_watchdog(st, "mi_trace_pts_host");
            if (wd > 0) { rc = MI_ERR_TIMEOUT; goto done; }
            if (wd < 0) goto done;
        }
        if (hipMemcpy(ho.data(), dout, sizeof(TrOut) * ntasks, hipMemcpyDeviceToHost) != hipSuccess) goto done;
        std::vector<int> again;
        for (int i : todo)
            if (ho[i].status == TR_BIG) again.push_back(i);
        todo.swap(again);
    }
    {
        std::vector<int> htr(std::max(tro, 1LL));
        if (hipMemcpy(htr.data(), dtr, sizeof(int) * std::max(tro, 1LL), hipMemcpyDeviceToHost) != hipSuccess) goto done;
        long long o = 0;
        for (int i = 0; i < ntasks; i++) {
            MiTraceResult& r = results[i];
            r.status = ho[i].status;
            r.diffs = ho[i].diffs;
            r.ntrace = ho[i].status == TR_OK ? ho[i].ntrace : 0;
            r.trace = o;
            if (o + r.ntrace > ntrace) { rc = MI_ERR_ARG; goto done; }
            memcpy(trace + o, htr.data() + tk[i].out, sizeof(int) * r.ntrace);
            o += r.ntrace;
        }
    }
    rc = 0;
done:
    if (rc == MI_ERR_DEVICE) mi_set_error("mi_trace_pts_host: HIP error");
    if (rc == MI_ERR_TIMEOUT) mi_set_error("mi_trace_pts_host: kernel exceeded the watchdog deadline and was stopped");
    (void)hipFree(dtk);
    (void)hipFree(dorder);
    (void)hipFree(dpk);
    (void)hipFree(dpts);
    (void)hipFree(dtr);
    (void)hipFree(dout);
    (void)hipFree(dscr);
    (void)hipFree(dctr);
    return rc;
}

}  // extern "C"
