// report_thread (dalign/filter.c:1617-1989) on gfx950: one wave per read pair.
//
// The reference walks a pair's seeds (sorted by apos) in 50 kb panels that
// overlap by 10 kb.  Per panel it scores diagonal bands (band = diag >> w),
// then visits the seeds in order and aligns a seed when its band has enough
// score and the seed lies beyond the end of the last alignment that covered
// its band (lasta).  Each alignment then raises lasta over its diagonal span.
//
// Here one wave owns the pair:
//   * band scores: 64 seeds at a time; lanes in the same band find each other
//     with ballots, take the previous seed's apos from the lane below (or
//     the band's lastp), add min(K, apos - prev) with atomics, and the last
//     lane of the band stores lastp -- the order the sequential loop implies;
//   * gating: the 64 seeds of a chunk test the condition together, the first
//     passing lane is aligned by the whole wave (local_alignment), lasta is
//     raised in parallel over the span, and testing resumes after that seed;
//   * redundant alignments of the pair are fused by lane 0
//     (dalign/filter.c:1249-1566) and the records are appended to the output.
// The band arrays (score | lastp | lasta) of a wave use the reference's
// per-thread layout, so neighbour reads at the outer bands see the same
// cells; any access outside [mindiag, maxdiag] is counted as an edge event.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "filter_kernels.hpp"
#include "la_wave.hpp"
#include "records.h"

namespace mi {

// statuses of a pair
constexpr int RP_OK = 0;
constexpr int RP_RETRY = 1;    // needs a wider band window or bigger pools
constexpr int RP_UNDEF = 3;    // an alignment reached a reference-undefined state
constexpr int RP_ARENA = 4;    // per-wave match arena too small
constexpr int RP_OUT = 5;      // output arena full (records not written)

struct MPath {   // a kept alignment of the pair (dalign/align.h:76-83 fields)
    int tlen, diffs, abpos, bbpos, aepos, bepos;
    long long toff;   // offset of its u16 trace in the wave's trace pool
};


struct ReportArgs {
    const Seed* hits;
    const long long* pstart;   // seed range of pair p: [pstart[p], pstart[p+1])
    const int* jobs;           // pair index of every job
    const int* order;          // jobs of this launch
    int norder;
    const uint64_t* pkA;
    const long long* w0A;
    const int* lenA;
    const uint64_t* pkB;
    const long long* w0B;
    const int* lenB;
    int afirst, bfirst, comp, self, symmetric, hgap_min;
    int K, binshift, hitmin, minhit, minover, tspace, small_trace;
    int mindiag, maxdiag, w;
    DevSpec sp;
    // per-wave scratch
    int* bands;         // (3w + 2) ints per wave, zero on entry, left zero on exit
    Peb* pools;
    int pool_cap;
    Peb* pts;
    int pts_cap;
    uint16_t* latr;     // alignment trace scratch: 4 * maxtp_all per wave
    int maxtp_all;
    MPath* mpaths;      // 2 * mcap per wave (A list then B list)
    int mcap;
    uint16_t* mtrace;   // mtcap per wave
    long long mtcap;
    // the band cell shared by lastp[maxdiag] and lasta[mindiag-1]: its value
    // when the job starts (null: 0), at the end, and whether the job used it
    const int* edge_in;
    int* edge_out;
    int* touch;
    // outputs
    OutRec* orecs;
    long long ocap;
    uint8_t* otr;
    long long otcap;
    unsigned long long* ocount;   // [0] records, [1] trace bytes
    int* status;                  // per job
    int* final_gen;               // per job: launch of its last successful run
    int gen;                      // this launch
    int* counter;
    const int* abort_flag;
    unsigned long long* stats;    // [0] seeds aligned, [1] cells, [2] snake, [3] edge events, [4] wave steps
    unsigned long long* job_work; // optional, per job: band cells + snake steps of its last successful run
    unsigned* redund;             // optional, per job: kRedundBranches counters of handle_redundancies
};

// ---- redundancy handling (dalign/filter.c:1249-1566), lane 0 ------------------

__device__ int entwine(const MPath* jp, const MPath* kp, const uint16_t* T, int TS, int* where)
{
    const uint16_t* jt = T + jp->toff;
    const uint16_t* kt = T + kp->toff;
    int mn = 10000, den = 0, y2 = jp->bbpos, b2 = kp->bbpos, ac, i, j, k;
    const int jseg = jp->abpos / TS, kseg = kp->abpos / TS;
    if (jseg < kseg) {
        ac = kseg * TS;
        j = 1 + 2 * (kseg - jseg);
        k = 1;
        for (i = 1; i < j; i += 2) y2 += jt[i];
    } else {
        ac = jseg * TS;
        k = 1 + 2 * (jseg - kseg);
        j = 1;
        for (i = 1; i < k; i += 2) b2 += kt[i];
    }
    const int ae = jp->aepos < kp->aepos ? jp->aepos : kp->aepos;
    for (;;) {
        ac += TS;
        if (ac >= ae) break;
        y2 += jt[j];
        b2 += kt[k];
        j += 2;
        k += 2;
        i = y2 > b2 ? y2 - b2 : b2 - y2;
        if (i <= mn) {
            mn = i;
            if (i == 0) *where = ac;
        }
        den += 1;
    }
    return den == 0 ? -1 : mn;
}

// p1 up to A coordinate `at`, then p2; appends the new trace at *top
__device__ bool fusion(MPath* p1, int at, const MPath* p2, uint16_t* T, long long* top, long long cap, int TS)
{
    const int k1 = 2 * ((at / TS) - (p1->abpos / TS));
    const int k2 = 2 * ((at / TS) - (p2->abpos / TS));
    const int len = k1 + (p2->tlen - k2);
    if (*top + len > cap) return false;
    uint16_t* d = T + *top;
    const uint16_t* s1 = T + p1->toff;
    const uint16_t* s2 = T + p2->toff;
    int n = 0, diff = 0;
    for (int k = 0; k < k1; k += 2) {
        d[n++] = s1[k];
        d[n++] = s1[k + 1];
        diff += s1[k];
    }
    for (int k = k2; k < p2->tlen; k += 2) {
        d[n++] = s2[k];
        d[n++] = s2[k + 1];
        diff += s2[k];
    }
    p1->aepos = p2->aepos;
    p1->bepos = p2->bepos;
    p1->diffs = diff;
    p1->toff = *top;
    p1->tlen = n;
    *top += len;
    return true;
}

// Branch counters of handle_redundancies (pass statistics):
//   0 j starts first, k reaches further: fused, A and B views, COMP      4 j contains k: k replaced
//   1 the same, not COMP                                                  5 same start, j longer: k replaced
//   2 the same, A view only                                               6 k starts first, j reaches further: fused, A and B, COMP
//   3 (unused)                                                            7 the same, not COMP; 8 the same, A view only
constexpr int kRedundBranches = 9;

__device__ int handle_redundancies(MPath* am, int n, MPath* bm, uint16_t* T, long long* top, long long cap, int TS,
                                   int comp, bool* ok, unsigned (&br)[kRedundBranches])
{
    int no = 0, j, k, aw = 0, bw = 0;
    for (j = 1; j < n; j++) {
        MPath* jp = am + j;
        for (k = no; k >= 0; k--) {
            MPath* kp = am + k;
            if (jp->abpos < kp->abpos) {
                if (!(kp->abpos <= jp->aepos && kp->bbpos <= jp->bepos)) continue;
                if (entwine(jp, kp, T, TS, &aw) != 0) continue;
                if (kp->aepos > jp->aepos) {
                    if (bm) {
                        if (comp) {
                            if (entwine(bm + k, bm + j, T, TS, &bw) != 0) continue;
                            *ok &= fusion(jp, aw, kp, T, top, cap, TS);
                            am[k] = *jp;
                            *ok &= fusion(bm + k, bw, bm + j, T, top, cap, TS);
                            br[0]++;
                        } else {
                            if (entwine(bm + j, bm + k, T, TS, &bw) != 0) continue;
                            *ok &= fusion(jp, aw, kp, T, top, cap, TS);
                            am[k] = *jp;
                            *ok &= fusion(bm + j, bw, bm + k, T, top, cap, TS);
                            bm[k] = bm[j];
                            br[1]++;
                        }
                    } else {
                        *ok &= fusion(jp, aw, kp, T, top, cap, TS);
                        am[k] = *jp;
                        br[2]++;
                    }
                } else {
                    am[k] = *jp;
                    if (bm) bm[k] = bm[j];
                    br[4]++;
                }
                break;
            } else {
                if (!(jp->abpos <= kp->aepos && jp->bbpos <= kp->bepos)) continue;
                if (entwine(kp, jp, T, TS, &aw) != 0) continue;
                if (kp->abpos == jp->abpos) {
                    if (kp->aepos < jp->aepos) {
                        am[k] = *jp;
                        if (bm) bm[k] = bm[j];
                        br[5]++;
                    }
                } else if (jp->aepos > kp->aepos) {
                    if (bm) {
                        if (comp) {
                            if (entwine(bm + j, bm + k, T, TS, &bw) != 0) continue;
                            *ok &= fusion(kp, aw, jp, T, top, cap, TS);
                            *ok &= fusion(bm + j, bw, bm + k, T, top, cap, TS);
                            bm[k] = bm[j];
                            br[6]++;
                        } else {
                            if (entwine(bm + k, bm + j, T, TS, &bw) != 0) continue;
                            *ok &= fusion(kp, aw, jp, T, top, cap, TS);
                            *ok &= fusion(bm + k, bw, bm + j, T, top, cap, TS);
                            br[7]++;
                        }
                    } else {
                        *ok &= fusion(kp, aw, jp, T, top, cap, TS);
                        br[8]++;
                    }
                }
                break;
            }
        }
        if (k < 0) {
            no += 1;
            am[no] = *jp;
            if (bm) bm[no] = bm[j];
        }
    }
    return no + 1;
}

// ---- the pair kernel --------------------------------------------------------------

// The 64-lane window (NCH = 1) runs at 8 waves per SIMD (64 VGPRs, a few
// spills): the aligner is bound by VALU issue and dependent latencies, and
// measured on the 5 Mbp / 30x workload 3 -> 4 -> 5 -> 6 -> 8 waves per SIMD
// took the kernel from 380 to 305, 277, 258 and 248 ms per step.
// AUX: the instantiation that also fills job_work / redund (capture mode and
// the branch statistics); the default one keeps those pointers out of its
// register budget.
// The kernel's arguments read where they are used: a scalar load through the
// kernarg segment pointer, made opaque to the compiler so that no argument is
// held in (or spilled from) SGPRs across the aligner's loop.
#ifndef MI_CPU_WAVE_EMU
typedef const ReportArgs __attribute__((address_space(4))) KernArgs;
__device__ __forceinline__ KernArgs* kargs()
{
    KernArgs* p = (KernArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return p;
}
#else
// the CPU wave emulator (tests/wave_emu) hands the launch's arguments over
typedef const ReportArgs KernArgs;
inline KernArgs* kargs() { return (KernArgs*)emu::kernarg(); }
#endif
#define KA(f) (kargs()->f)

template <int NCH, bool AUX>
__global__ __launch_bounds__(64, NCH == 1 ? 8 : 1) void k_report(ReportArgs)
{
    const int lane = __lane_id();
    const int wave = blockIdx.x;
    const int W3 = 3 * KA(w) + 2;
    int* F = KA(bands) + (size_t)wave * W3;
    int* score = F + 1 - KA(mindiag);             // score[band], band in [mindiag-1, maxdiag+1]
    int* lastp = score + KA(w);
    int* lasta = lastp + KA(w);
    int* const edge = F + 2 * KA(w);              // lastp[maxdiag] == lasta[mindiag-1]
    Peb* pool = KA(pools) + (size_t)wave * KA(pool_cap);
    Peb* pts = KA(pts) + (size_t)wave * 2 * KA(pts_cap);
    uint16_t* latr = KA(latr) + (size_t)wave * 4 * KA(maxtp_all);
    MPath* am = KA(mpaths) + (size_t)wave * 2 * KA(mcap);
    MPath* bm = am + KA(mcap);
    uint16_t* T = KA(mtrace) + (size_t)wave * KA(mtcap);
    const int TS = KA(tspace);

    for (;;) {
        // all lanes take part in the atomic (lane 0 adds 1), so the loop exit is uniform
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(KA(counter), lane == 0 ? 1 : 0));
        const int ab = __builtin_amdgcn_readfirstlane(__atomic_load_n(KA(abort_flag), __ATOMIC_RELAXED));
        if (t >= KA(norder) || ab) break;
        const int job = KA(order)[t];
        const int p = KA(jobs)[job];
        const long long ps = KA(pstart)[p], pe = KA(pstart)[p + 1];
        const int ar = KA(hits)[ps].aread, br = KA(hits)[ps].bread;
        const int alen = KA(lenA)[ar], blen = KA(lenB)[br];
        const uint64_t* Aseq = KA(pkA) + KA(w0A)[ar];
        const uint64_t* Bseq = KA(pkB) + KA(w0B)[br];
        const bool doA = alen >= KA(hgap_min);
        const bool doB = KA(symmetric) && blen >= KA(hgap_min) && (ar != br || !KA(self) || !KA(comp));
        const int maxtp = 2 * ((alen > blen ? alen : blen) / TS + 2);
        uint16_t* atr = latr + maxtp;
        uint16_t* btr = latr + 3 * maxtp;
        int novla = 0, novlb = 0;
        long long ttop = 0;
        int pstatus = RP_OK;
        int tmin = INT32_MAX, tmax = -INT32_MAX;   // lasta cells raised by this pair
        bool touched = false;
        long long aligned = 0, edges = 0;
        WaveCount wc;
        {
            const int e0 = KA(edge_in) ? KA(edge_in)[job] : 0;
            if (lane == 0) *edge = e0;
            wave_sync();
        }

        long long nidx = ps;
        int amark2 = 0;
        while (nidx < pe && pstatus == RP_OK) {
            const int amark = amark2 + 50000;
            amark2 = amark - 10000;
            const long long lidx = nidx;
            // panel (dalign/filter.c:1740-1755): up to and including the first
            // seed beyond amark; the next panel starts after the last seed at
            // or before amark2
            long long pend = pe, h2 = lidx;
            for (long long c = lidx; c < pe; c += 64) {
                const long long f = c + lane;
                const int ap = (f < pe) ? KA(hits)[f].apos : INT32_MAX;
                const unsigned long long beyond = __ballot(f < pe && ap > amark);
                const unsigned long long within2 = __ballot(f < pe && ap <= amark2);
                if (within2) h2 = c + highest_lane(within2) + 1;
                if (beyond) {
                    pend = c + lowest_lane(beyond) + 1;
                    break;
                }
            }
            nidx = pend;
            if (nidx - lidx >= KA(minhit)) {
                // band scores (dalign/filter.c:1757-1766): each seed adds
                // min(K, apos - apos of the band's previous seed)
                for (long long c = lidx; c < nidx; c += 64) {
                    const long long f = c + lane;
                    const bool v = f < nidx;
                    int ap = 0, band = KA(mindiag);
                    if (v) {
                        ap = KA(hits)[f].apos;
                        band = KA(hits)[f].diag >> KA(binshift);
                    }
                    const unsigned bi = (unsigned)(band - KA(mindiag));
                    unsigned long long peers = __ballot(v);
#pragma unroll
                    for (int b = 0; b < 16; b++) {
                        const unsigned long long m = __ballot((bi >> b) & 1);
                        peers &= ((bi >> b) & 1) ? m : ~m;
                    }
                    const unsigned long long below = peers & ((1ull << lane) - 1);
                    const int prevl = below ? highest_lane(below) : lane;
                    const int pap = __shfl(ap, prevl, 64);
                    int prev = 0;
                    if (v) prev = below ? pap : lastp[band];
                    if (__any(v && band == KA(maxdiag))) touched = true;
                    edges += __popcll(__ballot(v && (band == KA(mindiag) || band == KA(maxdiag))));
                    wave_sync();
                    if (v) {
                        const int d = ap - prev;
                        atomicAdd(&score[band], d >= KA(K) ? KA(K) : d);
                        if (lane == highest_lane(peers)) lastp[band] = ap;
                    }
                    wave_sync();
                }
                // gating in seed order (dalign/filter.c:1768-1893): a seed is
                // aligned when it lies beyond its band's lasta and the band
                // pair score reaches hitmin; 64 seeds are tested at once and
                // testing resumes after each aligned seed
                long long f0 = lidx;
                while (f0 < nidx && pstatus == RP_OK) {
                    const long long f = f0 + lane;
                    const bool v = f < nidx;
                    bool go = false;
                    if (v) {
                        const int ap = KA(hits)[f].apos, band = KA(hits)[f].diag >> KA(binshift);
                        go = ap > lasta[band] && (score[band] + score[band + 1] >= KA(hitmin) ||
                                                  score[band] + score[band - 1] >= KA(hitmin));
                    }
                    const unsigned long long gm = __ballot(go);
                    if (!gm) {
                        f0 += 64;
                        continue;
                    }
                    const long long fh = f0 + lowest_lane(gm);
                    f0 = fh + 1;
                    const Seed sd = KA(hits)[fh];
                    const int ap = sd.apos, dg = sd.diag, bp = ap - dg, band0 = dg >> KA(binshift);
                    // a read against itself (-I, self forward pass: the same bases): the
                    // wave may not reach the main diagonal (Local_Alignment's selfie
                    // bounds, dalign/align.c:1966-1985)
                    const bool selfie = KA(self) && !KA(comp) && ar == br;
                    DevSpec spv;   // by value: the aligner reads the spec from SGPRs
                    spv.table = KA(sp.table);
                    spv.score = KA(sp.score);
                    spv.tspace = KA(sp.tspace);
                    spv.ave_path = KA(sp.ave_path);
                    LaResult r = local_alignment<NCH>(spv, Aseq, alen, Bseq, blen, KA(comp), dg, dg, ap + bp,
                                                      (selfie && dg >= 0) ? 1 : -INT32_MAX,
                                                      (selfie && dg <= 0) ? -1 : INT32_MAX, pool, KA(pool_cap), pts,
                                                      KA(pts_cap), atr, btr, maxtp, KA(abort_flag), wc);
                    if (r.status != LA_OK) {
                        pstatus = (r.status == LA_WIDE || r.status == LA_POOL || r.status == LA_ABORT) ? RP_RETRY
                                                                                                        : RP_UNDEF;
                        break;
                    }
                    aligned++;
                    const uint16_t* ta = atr + r.afirst;
                    const uint16_t* tb = btr + r.bfirst;
                    // Diagonal_Span (dalign/filter.c:1568-1599) widened by one
                    // band on each side and to the seed's band
                    int lo = r.abpos - r.bbpos, hi = lo;
                    {
                        int dd = r.aepos - r.bepos;
                        if (dd < lo) lo = dd;
                        else if (dd > hi) hi = dd;
                        dd = (r.abpos / TS) * TS - r.bbpos;
                        for (int i = 1; i < r.atlen - 2; i += 2) {   // same walk on every lane
                            dd += TS - ta[i];
                            if (dd < lo) lo = dd;
                            else if (dd > hi) hi = dd;
                        }
                        lo = (lo >> KA(binshift)) - 1;
                        hi = (hi >> KA(binshift)) + 1;
                        if (band0 < lo) lo = band0;
                        else if (band0 > hi) hi = band0;
                    }
                    if (lo < KA(mindiag) || hi > KA(maxdiag)) edges++;
                    if (lo < KA(mindiag)) touched = true;
                    // a consistent trace keeps the span within one band of the
                    // array on each side; the clamp only guards the memory
                    lo = max(lo, KA(mindiag) - 1);
                    hi = min(hi, KA(maxdiag) + 1);
                    tmin = min(tmin, lo);
                    tmax = max(tmax, hi);
                    wave_sync();
                    for (int i = lane; i <= hi - lo; i += 64)
                        if (r.aepos > lasta[lo + i]) lasta[lo + i] = r.aepos;
                    wave_sync();
                    if ((r.aepos - r.abpos) + (r.bepos - r.bbpos) >= KA(minover)) {
                        const long long need = (doA ? r.atlen : 0) + (doB ? r.btlen : 0);
                        if (novla >= KA(mcap) || novlb >= KA(mcap) || ttop + need > KA(mtcap)) {
                            pstatus = RP_ARENA;
                            break;
                        }
                        if (doA) {
                            for (int i = lane; i < r.atlen; i += 64) T[ttop + i] = ta[i];
                            if (lane == 0)
                                am[novla] = MPath{r.atlen, r.diffs, r.abpos, r.bbpos, r.aepos, r.bepos, ttop};
                            ttop += r.atlen;
                            novla++;
                        }
                        if (doB) {
                            for (int i = lane; i < r.btlen; i += 64) T[ttop + i] = tb[i];
                            if (lane == 0)
                                bm[novlb] = MPath{r.btlen, r.diffs, r.b_abpos, r.b_bbpos, r.b_aepos, r.b_bepos, ttop};
                            ttop += r.btlen;
                            novlb++;
                        }
                        wave_sync();
                    }
                }
                // panel reset of the scored bands (dalign/filter.c:1895-1899)
                wave_sync();
                for (long long c = lidx; c < nidx; c += 64) {
                    const long long f = c + lane;
                    if (f < nidx) {
                        const int band = KA(hits)[f].diag >> KA(binshift);
                        score[band] = 0;
                        lastp[band] = 0;
                    }
                }
                wave_sync();
            }
            nidx = h2;
        }
        // per-pair lasta reset (dalign/filter.c:1905-1920): every raised cell
        // inside [mindiag, maxdiag] belongs to a run of non-zero cells through
        // a seed band of this pair, so the reference's walk clears exactly the
        // raised in-range cells; cells outside the range keep their value
        if (tmin <= tmax) {   // some alignment raised cells (tmin / tmax start empty)
            const int lo = max(tmin, KA(mindiag)), hi = min(tmax, KA(maxdiag));
            for (int i = lane; i <= hi - lo; i += 64) lasta[lo + i] = 0;
            wave_sync();
        }
        if (pstatus == RP_OK) {
            // fuse redundant alignments (lane 0) and append the records
            bool ok = true;
            int na = novla, nb = novlb;
            if (lane == 0) {
                unsigned br[kRedundBranches] = {};
                if (na > 1) {
                    if (nb > 1) na = nb = handle_redundancies(am, na, bm, T, &ttop, KA(mtcap), TS, KA(comp), &ok, br);
                    else na = handle_redundancies(am, na, nullptr, T, &ttop, KA(mtcap), TS, KA(comp), &ok, br);
                } else if (nb > 1)
                    nb = handle_redundancies(bm, nb, nullptr, T, &ttop, KA(mtcap), TS, KA(comp), &ok, br);
                if constexpr (AUX) {
                    if (KA(redund))   // the job's last successful run overwrites earlier ones
                        for (int q = 0; q < kRedundBranches; q++) KA(redund)[(size_t)job * kRedundBranches + q] = br[q];
                }
            }
            ok = __builtin_amdgcn_readfirstlane(ok ? 1 : 0) != 0;
            na = __builtin_amdgcn_readfirstlane(na);
            nb = __builtin_amdgcn_readfirstlane(nb);
            if (!ok) pstatus = RP_ARENA;
            else if (na + nb > 0) {
                wave_sync();
                long long tl = 0;
                for (int i = 0; i < na; i++) tl += am[i].tlen;
                for (int i = 0; i < nb; i++) tl += bm[i].tlen;
                const int tbytes = KA(small_trace) ? 1 : 2;
                unsigned long long r0 = 0, t0 = 0;
                r0 = atomicAdd(&KA(ocount)[0], lane == 0 ? (unsigned long long)(na + nb) : 0ull);
                t0 = atomicAdd(&KA(ocount)[1], lane == 0 ? (unsigned long long)(tl * tbytes) : 0ull);
                const unsigned long long rbase = __shfl(r0, 0, 64);
                const unsigned long long tbase = __shfl(t0, 0, 64);
                if (rbase + na + nb > (unsigned long long)KA(ocap) || tbase + tl * tbytes > (unsigned long long)KA(otcap))
                    pstatus = RP_OUT;
                else {
                    unsigned long long to = tbase;
                    for (int i = 0; i < na + nb; i++) {
                        const bool isA = i < na;
                        const MPath m = isA ? am[i] : bm[i - na];
                        if (lane == 0) {
                            OutRec o;
                            o.tlen = m.tlen;
                            o.diffs = m.diffs;
                            o.abpos = m.abpos;
                            o.bbpos = m.bbpos;
                            o.aepos = m.aepos;
                            o.bepos = m.bepos;
                            o.flags = (uint32_t)KA(comp);
                            o.aread = isA ? ar + KA(afirst) : br + KA(bfirst);
                            o.bread = isA ? br + KA(bfirst) : ar + KA(afirst);
                            o.local = isA ? i : i - na;
                            o.job = job;
                            o.toff = (long long)to;
                            o.gen = KA(gen);
                            o.pad = 0;
                            KA(orecs)[rbase + i] = o;
                        }
                        // Compress_TraceTo8 (dalign/align.c:3352-3360) when the spacing is <= 125
                        for (int q = lane; q < m.tlen; q += 64) {
                            const uint16_t x = T[m.toff + q];
                            if (tbytes == 1) KA(otr)[to + q] = (uint8_t)x;
                            else {
                                KA(otr)[to + 2 * q] = (uint8_t)(x & 0xff);
                                KA(otr)[to + 2 * q + 1] = (uint8_t)(x >> 8);
                            }
                        }
                        to += (unsigned long long)m.tlen * tbytes;
                    }
                }
            }
        }
        wave_sync();
        const int eout = *edge;
        if (lane == 0) {
            KA(status)[job] = pstatus;
            if (pstatus == RP_OK) KA(final_gen)[job] = KA(gen);
            KA(edge_out)[job] = eout;
            KA(touch)[job] = touched ? 1 : 0;
            if (pstatus == RP_OK) {
                atomicAdd(&KA(stats)[0], (unsigned long long)aligned);
                atomicAdd(&KA(stats)[1], (unsigned long long)wc.cells);
                atomicAdd(&KA(stats)[2], (unsigned long long)wc.snake);
                atomicAdd(&KA(stats)[3], (unsigned long long)edges);
                atomicAdd(&KA(stats)[4], (unsigned long long)wc.steps);
                if constexpr (AUX) {
                    if (KA(job_work)) KA(job_work)[job] = (unsigned long long)(wc.cells + wc.snake);
                }
            }
        }
        // leave the band arrays zero for the next job: the edge cell and the
        // pad always, everything when the pair stopped early
        if (pstatus != RP_OK) {
            wave_sync();
            for (int i = lane; i < W3; i += 64) F[i] = 0;
        } else if (lane == 0) {
            *edge = 0;
            F[W3 - 1] = 0;
        }
        wave_sync();
    }
}
#undef KA

}  // namespace mi
