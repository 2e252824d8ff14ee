// DBdust on the GPU: the SDUST low-complexity scan of db/DBdust.c:236-470
// (Morgulis et al., JCB 13(5) 2006) over every read of a block.  The scan is
// sequential along a read.  With the reference's unbiased scores (the
// default) each read is cut into segments of 384 steps, one lane each, and
// a lane reaches its segment's true starting state by a warm-up (the
// segmented scan below); the biased scores (-b) and windows below 3 keep one
// lane per read.  Each lane keeps its triplet counts and a ring of triplet
// codes in LDS ([item][lane], a bank per lane); the candidate lists sit in
// LDS (sequential scan, windows up to 64) or HBM.
//
// Per read (triplet codes t[0 .. len-3] of bases j, j+1, j+2):
//   * the window counts cover (wb, j], the scanned run (lb, j]; wsqr / lsqr
//     are their sums of count*(count-1)/2 (or of the biased weights);
//   * a triplet seen thresh2 times in the run cuts the run back past its
//     previous occurrence;
//   * when the window is dense enough, the run is walked back from lb and
//     every prefix scoring above the candidates it covers becomes (or
//     updates) a candidate interval [c, j];
//   * a candidate whose start falls out of the window is retired onto the
//     interval stack, merged with the top interval when they overlap or touch;
//   * at the end the remaining candidates are retired, and intervals of
//     at least `mlen` (the -m option less one) become the points (b, e+1).
// Score arithmetic is the reference's: doubles, no fused multiply-add.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/marvl_filter.h"
#include "mi_common.h"

#pragma clang fp contract(off)

namespace mi {

constexpr int kDustMaxWindow = 256;

struct DustArgs {
    const uint8_t* bases;        // bytes 0..3, read r at bases + boff[r]
    const long long* boff;
    const int* rlen;
    int nreads;
    int window, thresh2i, mlen, biased;
    double thresh, thresh2r;
    const double* skew;          // [64] biased triplet weights
    // per-lane candidate nodes, [node][lane slot] with nslots = grid lanes
    int* cbeg;
    int* cend;
    double* cscore;
    short* cnext;
    short* cprev;
    long long nslots;
    int* stack;                  // per read: the interval stack at soff[r] (rlen[r] + 4 ints)
    const long long* soff;
    int* npts;                   // out: points kept per read (written to stack + soff[r])
    int* status;                 // out: nonzero when a read ran out of candidate nodes
};

// Candidate nodes live in LDS when the window is at most kDustLdsWindow
// (the default 64): every step of the scan touches them, and HBM latency
// would serialise each lane's walk.  Larger windows keep them in HBM.
constexpr int kDustLdsWindow = 64;

template <bool LDS_NODES>
__global__ __launch_bounds__(64) void k_sdust(DustArgs a)
{
    __shared__ int wcnt[64 * 64], lcnt[64 * 64];   // [triplet][lane]
    constexpr int kN = LDS_NODES ? (kDustLdsWindow + 1) * 64 : 1;
    __shared__ int sbeg[kN], send[kN];
    __shared__ double ssc[kN];
    __shared__ short snx[kN], spv[kN];
    // triplet codes of the positions the scan can still reach: a ring of
    // kRing per lane ([slot][lane]), refilled kRing/2 positions ahead.  The
    // scan reads only positions in (j - window, j], so a slot is rewritten
    // only after its position left the window.
    constexpr int kRing = LDS_NODES ? 128 : 2 * kDustMaxWindow;
    __shared__ uint8_t ring[kRing * 64];
    const int lane = threadIdx.x;
    const long long slot = (long long)blockIdx.x * 64 + lane;
    const long long S = a.nslots;
    for (int r = (int)slot; r < a.nreads; r += (int)S) {
        for (int e = 0; e < 64; e++) wcnt[e * 64 + lane] = lcnt[e * 64 + lane] = 0;
        const uint8_t* b = a.bases + a.boff[r];
        const int L = a.rlen[r] - 2;
        auto T = [&](int i) { return (int)ring[(i & (kRing - 1)) * 64 + lane]; };
        auto fill = [&](int j0) {   // codes of positions [j0, j0 + kRing/2) that exist
            const int j1 = min(j0 + kRing / 2, L);
#pragma unroll 8
            for (int i = j0; i < j1; i++)
                ring[(i & (kRing - 1)) * 64 + lane] = (uint8_t)((b[i] << 4) | (b[i + 1] << 2) | b[i + 2]);
        };
        // candidate nodes: 0 is the list head (beg = -2), 1..window the free list
        auto BEG = [&](int n) -> int& { return LDS_NODES ? sbeg[n * 64 + lane] : a.cbeg[n * S + slot]; };
        auto END = [&](int n) -> int& { return LDS_NODES ? send[n * 64 + lane] : a.cend[n * S + slot]; };
        auto SC = [&](int n) -> double& { return LDS_NODES ? ssc[n * 64 + lane] : a.cscore[n * S + slot]; };
        auto NX = [&](int n) -> short& { return LDS_NODES ? snx[n * 64 + lane] : a.cnext[n * S + slot]; };
        auto PV = [&](int n) -> short& { return LDS_NODES ? spv[n * 64 + lane] : a.cprev[n * S + slot]; };
        NX(0) = 0;
        PV(0) = 0;
        BEG(0) = -2;
        for (int n = 1; n < a.window; n++) NX(n) = (short)(n + 1);
        NX(a.window) = -1;
        int freel = 1;
        int* stk = a.stack + a.soff[r];
        int top = 0;   // stk[0] = -2 sentinel, intervals as pairs from stk[1]
        stk[0] = -2;
        bool bad = false;
        auto retire = [&](int J) {
            const int e = END(J) + 2;
            if (stk[top] + 1 >= BEG(J)) {
                if (stk[top] < e) stk[top] = e;
            } else {
                stk[++top] = BEG(J);
                stk[++top] = e;
            }
            const int p = PV(J);
            PV(0) = (short)p;
            NX(p) = 0;
            NX(J) = (short)freel;
            freel = J;
        };
        int wb = -1, lb = -1;
        if (!a.biased) {
            int wsqr = 0, lsqr = 0;
            for (int j = 0; j < L && !bad; j++) {
                if ((j & (kRing / 2 - 1)) == 0) fill(j);
                int c = T(j);
                if (j > a.window - 3) {
                    const int d = T(++wb);
                    wsqr -= --wcnt[d * 64 + lane];
                }
                wsqr += wcnt[c * 64 + lane]++;
                if (lb < wb) {
                    const int d = T(++lb);
                    lsqr -= --lcnt[d * 64 + lane];
                }
                const int trun = lcnt[c * 64 + lane]++;
                lsqr += trun;
                if (trun >= a.thresh2i)
                    while (lb < j) {
                        const int d = T(++lb);
                        lsqr -= --lcnt[d * 64 + lane];
                        if (d == c) break;
                    }
                int J = PV(0);
                if (J != 0 && BEG(J) <= wb) retire(J);
                if (wsqr <= lsqr * a.thresh) continue;
                J = NX(0);
                int Lp = 0;
                double mscore = 0.;
                for (c = lb; c > wb; c--) {
                    const int d = T(c);
                    lsqr += lcnt[d * 64 + lane]++;
                    if (lsqr >= a.thresh * (j - c)) {
                        for (int g = 0; BEG(J) >= c; Lp = J, J = NX(J)) {
                            if (++g > a.window + 1) {   // a list longer than its nodes: stop the read
                                bad = true;
                                break;
                            }
                            if (SC(J) > mscore) mscore = SC(J);
                        }
                        if (bad) break;
                        if (lsqr >= mscore * (j - c)) {
                            mscore = (1. * lsqr) / (j - c);
                            if (BEG(Lp) == c) {
                                END(Lp) = j;
                                SC(Lp) = mscore;
                            } else {
                                const int A = freel;
                                if (A < 0) {
                                    bad = true;
                                    break;
                                }
                                freel = NX(A);
                                BEG(A) = c;
                                END(A) = j;
                                SC(A) = mscore;
                                PV(A) = (short)Lp;
                                NX(Lp) = (short)A;
                                Lp = A;
                                PV(J) = (short)A;
                                NX(A) = (short)J;
                            }
                        }
                    }
                }
                for (c++; c <= lb; c++) {
                    const int d = T(c);
                    lsqr -= --lcnt[d * 64 + lane];
                }
            }
        } else {
            double wsqr = 0., lsqr = 0.;
            for (int j = 0; j < L && !bad; j++) {
                if ((j & (kRing / 2 - 1)) == 0) fill(j);
                int c = T(j);
                if (j > a.window - 3) {
                    const int d = T(++wb);
                    wsqr -= (--wcnt[d * 64 + lane]) * a.skew[d];
                }
                wsqr += (wcnt[c * 64 + lane]++) * a.skew[c];
                if (lb < wb) {
                    const int d = T(++lb);
                    lsqr -= (--lcnt[d * 64 + lane]) * a.skew[d];
                }
                const double trun = (lcnt[c * 64 + lane]++) * a.skew[c];
                lsqr += trun;
                if (trun >= a.thresh2r)
                    while (lb < j) {
                        const int d = T(++lb);
                        lsqr -= (--lcnt[d * 64 + lane]) * a.skew[d];
                        if (d == c) break;
                    }
                int J = PV(0);
                if (J != 0 && BEG(J) <= wb) retire(J);
                if (wsqr <= lsqr * a.thresh) continue;
                J = NX(0);
                int Lp = 0;
                double mscore = 0.;
                for (c = lb; c > wb; c--) {
                    const int d = T(c);
                    lsqr += (lcnt[d * 64 + lane]++) * a.skew[d];
                    if (lsqr >= a.thresh * (j - c)) {
                        for (int g = 0; BEG(J) >= c; Lp = J, J = NX(J)) {
                            if (++g > a.window + 1) {   // a list longer than its nodes: stop the read
                                bad = true;
                                break;
                            }
                            if (SC(J) > mscore) mscore = SC(J);
                        }
                        if (bad) break;
                        if (lsqr >= mscore * (j - c)) {
                            mscore = lsqr / (j - c);
                            if (BEG(Lp) == c) {
                                END(Lp) = j;
                                SC(Lp) = mscore;
                            } else {
                                const int A = freel;
                                if (A < 0) {
                                    bad = true;
                                    break;
                                }
                                freel = NX(A);
                                BEG(A) = c;
                                END(A) = j;
                                SC(A) = mscore;
                                PV(A) = (short)Lp;
                                NX(Lp) = (short)A;
                                Lp = A;
                                PV(J) = (short)A;
                                NX(A) = (short)J;
                            }
                        }
                    }
                }
                for (c++; c <= lb; c++) {
                    const int d = T(c);
                    lsqr -= (--lcnt[d * 64 + lane]) * a.skew[d];
                }
            }
        }
        if (bad) {
            a.status[r] = 1;
            a.npts[r] = 0;
            continue;
        }
        for (int J = PV(0), g = 0; J != 0 && g <= a.window; J = PV(0), g++) retire(J);
        if (PV(0) != 0) {
            a.status[r] = 1;
            a.npts[r] = 0;
            continue;
        }
        // keep intervals of at least mlen as points (b, e + 1), in place
        int n = 0;
        for (int k = 1; k < top; k += 2)
            if (stk[k + 1] - stk[k] >= a.mlen) {
                const int x = stk[k], y = stk[k + 1] + 1;
                stk[n++] = x;
                stk[n++] = y;
            }
        a.npts[r] = n;
        a.status[r] = 0;
    }
}

// ---- segmented scan (unbiased scores) ------------------------------------------
// Before step j the unbiased scan's state is (lb, the candidate list): wb is
// max(-1, j - 1 - window + 2), both count arrays and their sums follow from
// wb, lb and the read, and every live candidate begins inside the window.
// So a read is cut into segments of `seglen` steps, one lane each.  A
// segment's lane starts `warm` steps early from the smallest state (lb = wb,
// no candidate; lb only ever moves forward and the map from the previous lb
// to the next is monotone, so the warm-up's lb never passes the true one and
// usually meets it within the warm-up: on the 5 Mbp reads every boundary
// after 128 steps), records the state it reached at its first step,
// then runs its own steps, folding the intervals it retires as the interval
// stack does and appending them, and records its final state.  A second
// kernel compares each segment's first state with its predecessor's final
// state.  Where they differ, one lane per read re-runs from the true state
// and carries on until a final state matches the next segment's first.
// Retirement order is increasing begin (one retirement per step, of the
// tail, whose begin is then wb), so folding each segment's intervals and
// then the segments' results in order gives the sequential stack.  The
// biased scores are sums of doubles whose rounding depends on the whole
// read, so they keep the sequential kernel above.

struct DustEnt {
    int beg, end;
    double sc;
};
struct DustHdr {
    int lb, n;
};
struct DustIv {   // a folded interval: segment, pass (0 main, 1 repair), begin, end + 2
    int seg, pass, b, e;
};

struct DustSegArgs {
    const uint8_t* bases;
    const long long* boff;
    const int* rlen;
    const int* seg_read;   // [nseg] read of the segment (segments of a read are consecutive)
    const int* seg_s0;     // [nseg] its first step
    int nseg, seglen, warm, window, thresh2i;
    double thresh;
    // candidate nodes, [node][slot], slot < nstride
    int* cbeg;
    int* cend;
    double* csc;
    short* cnx;
    short* cpv;
    long long nstride;
    // boundary states, record 2s = the state segment s began its own steps
    // in, 2s + 1 = the state after its last step: hdr = (lb, n) and n
    // candidates (head first, decreasing begin) at ent[rec * (window + 1)]
    DustHdr* hdr;
    DustEnt* ent;
    uint8_t* flag;       // compare: segment s began in another state than s - 1 ended in
    uint8_t* repaired;   // repair: segments re-run
    const int* rep_seg;  // repair: per listed read, its first flagged segment
    int nrep;
    int* mismatch;       // compare: flagged segments
    DustIv* out;         // the folded intervals
    int* out_cnt;
    int out_cap;
    int* bad;            // per segment: a candidate list outgrew its nodes
};

__device__ __forceinline__ bool dust_rec_equal(const DustSegArgs& a, long long x, long long y)
{
    const DustHdr hx = a.hdr[x], hy = a.hdr[y];
    if (hx.lb != hy.lb || hx.n != hy.n) return false;
    const int E = a.window + 1;
    for (int i = 0; i < hx.n; i++) {   // scores are finite and positive: == is bit equality
        const DustEnt ex = a.ent[x * E + i], ey = a.ent[y * E + i];
        if (ex.beg != ey.beg || ex.end != ey.end || ex.sc != ey.sc) return false;
    }
    return true;
}

// RING: triplet-code ring per lane, at least 2 x window slots.  REPAIR:
// one lane per listed read, from its first flagged segment.
template <int RING, bool REPAIR>
__global__ __launch_bounds__(64) void k_sdust_seg(DustSegArgs a)
{
    __shared__ uint8_t wcnt[64 * 64], lcnt[64 * 64];   // [triplet][lane]
    __shared__ uint8_t ring[RING * 64];                // [slot][lane]
    const int lane = threadIdx.x;
    const long long slot = (long long)blockIdx.x * 64 + lane;
    const long long S = a.nstride;
    const int W = a.window, E = W + 1;
    auto BEG = [&](int n) -> int& { return a.cbeg[n * S + slot]; };
    auto END = [&](int n) -> int& { return a.cend[n * S + slot]; };
    auto SC = [&](int n) -> double& { return a.csc[n * S + slot]; };
    auto NX = [&](int n) -> short& { return a.cnx[n * S + slot]; };
    auto PV = [&](int n) -> short& { return a.cpv[n * S + slot]; };
    auto WC = [&](int e) -> uint8_t& { return wcnt[e * 64 + lane]; };
    auto LC = [&](int e) -> uint8_t& { return lcnt[e * 64 + lane]; };
    auto T = [&](int i) { return (int)ring[(i & (RING - 1)) * 64 + lane]; };
    if (slot >= (REPAIR ? a.nrep : a.nseg)) return;
    int s = REPAIR ? a.rep_seg[slot] : (int)slot;
    const int r = a.seg_read[s];
    const uint8_t* b = a.bases + a.boff[r];
    const int L = a.rlen[r] - 2;
    int j = 0, wb = -1, lb = -1, wsqr = 0, lsqr = 0;
    int freel = 1, tail = 0, tailbeg = 0;
    int topb = 0, tope = 0;   // the segment's top interval, not yet appended
    bool has = false;
    // codes of positions [lo, min(hi, L)) into the ring
    auto fill = [&](int lo, int hi) {
        hi = min(hi, L);
        if (lo >= hi) return;
        int c = (b[lo] << 2) | b[lo + 1];
        for (int i = lo; i < hi; i++) {
            c = ((c << 2) & 63) | b[i + 2];
            ring[(i & (RING - 1)) * 64 + lane] = (uint8_t)c;
        }
    };
    // the state before step p: lb and the candidates of record `rec`, or
    // (rec < 0) lb = wb and no candidate
    auto setup = [&](int p, long long rec) {
        j = p;
        wb = (p - 1 > W - 3) ? p - W + 1 : -1;
        lb = rec < 0 ? wb : a.hdr[rec].lb;
        const int n = rec < 0 ? 0 : a.hdr[rec].n;
        fill(max(0, wb + 1), (p + RING / 2 - 1) & ~(RING / 2 - 1));
        for (int e = 0; e < 64; e++) WC(e) = LC(e) = 0;
        wsqr = lsqr = 0;
        for (int i = wb + 1; i < p; i++) wsqr += WC(T(i))++;
        for (int i = lb + 1; i < p; i++) lsqr += LC(T(i))++;
        BEG(0) = -2;
        NX(0) = (short)(n ? 1 : 0);
        PV(0) = (short)n;
        for (int i = 0; i < n; i++) {
            const DustEnt e = a.ent[rec * E + i];
            BEG(i + 1) = e.beg;
            END(i + 1) = e.end;
            SC(i + 1) = e.sc;
            NX(i + 1) = (short)(i + 1 < n ? i + 2 : 0);
            PV(i + 1) = (short)i;
        }
        for (int k = n + 1; k < W; k++) NX(k) = (short)(k + 1);
        if (n < W) NX(W) = -1;
        freel = n < W ? n + 1 : -1;
        tail = n;
        tailbeg = n ? a.ent[rec * E + n - 1].beg : 0;
    };
    auto save = [&](long long rec) {
        int n = 0;
        for (int J = NX(0); J != 0 && n < E; J = NX(J), n++) a.ent[rec * E + n] = DustEnt{BEG(J), END(J), SC(J)};
        a.hdr[rec] = DustHdr{lb, n};
    };
    auto emit = [&]() {
        const int k = atomicAdd(a.out_cnt, 1);
        if (k < a.out_cap) a.out[k] = DustIv{s, REPAIR ? 1 : 0, topb, tope};
    };
    // retire the tail candidate (its begin has left the window), folding it
    // into the top interval when `rec`
    auto retire = [&](bool rec) {
        const int J = tail;
        if (rec) {
            const int bb = tailbeg, e = END(J) + 2;
            if (has && tope + 1 >= bb) {
                if (tope < e) tope = e;
            } else {
                if (has) emit();
                topb = bb;
                tope = e;
                has = true;
            }
        }
        const int p = PV(J);
        PV(0) = (short)p;
        NX(p) = 0;
        NX(J) = (short)freel;
        freel = J;
        tail = p;
        tailbeg = p ? BEG(p) : 0;
    };
    // one step of db/DBdust.c:386-466; false when a list outgrew its nodes
    auto step = [&](bool rec) -> bool {
        if ((j & (RING / 2 - 1)) == 0) fill(j, j + RING / 2);
        int c = T(j);
        if (j > W - 3) {
            const int d = T(++wb);
            wsqr -= --WC(d);
        }
        wsqr += WC(c)++;
        if (lb < wb) {
            const int d = T(++lb);
            lsqr -= --LC(d);
        }
        const int trun = LC(c)++;
        lsqr += trun;
        // the loops over the run read the next position's code one iteration
        // ahead (ring and counts are separate LDS arrays, so the read overlaps
        // the count's round trip; the ring index wraps, so it stays in bounds)
        if (trun >= a.thresh2i)
            for (int dn = T(lb + 1); lb < j;) {
                const int d = dn;
                dn = T(++lb + 1);
                lsqr -= --LC(d);
                if (d == c) break;
            }
        if (tail != 0 && tailbeg <= wb) retire(rec);
        if (wsqr <= lsqr * a.thresh) {
            j++;
            return true;
        }
        int J = NX(0), Lp = 0;
        int jbeg = J ? BEG(J) : -2, lpbeg = -2;
        double mscore = 0.;
        int dn = T(lb);
        for (c = lb; c > wb; c--) {
            // the run's sum only grows toward the window's (at c = wb + 1 the
            // run is the window), and thresh * (j - c) only grows as c falls:
            // once that exceeds wsqr no position from here on can pass the
            // test below, and the positions not added need no undo (most
            // walks end here at c = lb)
            if (a.thresh * (j - c) > wsqr) break;
            const int d = dn;
            dn = T(c - 1);
            lsqr += LC(d)++;
            if (lsqr >= a.thresh * (j - c)) {
                for (int g = 0; jbeg >= c;) {
                    if (++g > W + 1) return false;
                    const double sc = SC(J);
                    if (sc > mscore) mscore = sc;
                    Lp = J;
                    lpbeg = jbeg;
                    J = NX(J);
                    jbeg = J ? BEG(J) : -2;
                }
                if (lsqr >= mscore * (j - c)) {
                    mscore = (1. * lsqr) / (j - c);
                    if (lpbeg == c) {
                        END(Lp) = j;
                        SC(Lp) = mscore;
                    } else {
                        const int A = freel;
                        if (A < 0) return false;
                        freel = NX(A);
                        BEG(A) = c;
                        END(A) = j;
                        SC(A) = mscore;
                        PV(A) = (short)Lp;
                        NX(Lp) = (short)A;
                        PV(J) = (short)A;
                        NX(A) = (short)J;
                        if (J == 0) {
                            tail = A;
                            tailbeg = c;
                        }
                        Lp = A;
                        lpbeg = c;
                    }
                }
            }
        }
        for (dn = T(++c); c <= lb; c++) {
            const int d = dn;
            dn = T(c + 1);
            lsqr -= --LC(d);
        }
        j++;
        return true;
    };
    auto drain = [&]() -> bool {
        for (int g = 0; tail != 0 && g <= W; g++) retire(true);
        return tail == 0;
    };
    if (!REPAIR) {
        const int s0 = a.seg_s0[s], s1 = min(s0 + a.seglen, L);
        setup(s0 == 0 ? 0 : max(0, s0 - a.warm), -1);
        bool ok = true;
        while (ok && j < s0) ok = step(false);
        if (s0 > 0) save(2LL * s);
        while (ok && j < s1) ok = step(true);
        if (ok) {
            if (s1 == L) ok = drain();
            else save(2LL * s + 1);
        }
        if (has) emit();
        if (!ok) a.bad[s] = 1;
        return;
    }
    // re-run from the true final state of the segment before
    setup(a.seg_s0[s], 2LL * (s - 1) + 1);
    for (;;) {
        const int s1 = min(a.seg_s0[s] + a.seglen, L);
        a.repaired[s] = 1;
        has = false;
        bool ok = true;
        while (ok && j < s1) ok = step(true);
        if (ok && s1 == L) ok = drain();
        if (has) emit();
        if (!ok) {
            a.bad[s] = 1;
            return;
        }
        if (s1 == L) return;
        save(2LL * s + 1);
        if (!dust_rec_equal(a, 2LL * s + 1, 2LL * (s + 1))) {
            s++;   // carry on into the next segment with the state in hand
            continue;
        }
        int t = s + 2;   // the next segment that began in another state than its predecessor ended in
        while (t < a.nseg && a.seg_read[t] == r && !a.flag[t]) t++;
        if (t >= a.nseg || a.seg_read[t] != r) return;
        s = t;
        setup(a.seg_s0[t], 2LL * (t - 1) + 1);
    }
}

__global__ void k_dust_compare(DustSegArgs a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nseg) return;
    const bool f = a.seg_s0[s] > 0 && !dust_rec_equal(a, 2LL * s, 2LL * (s - 1) + 1);
    a.flag[s] = f ? 1 : 0;
    if (f) atomicAdd(a.mismatch, 1);
}

}  // namespace mi

using namespace mi;

namespace {

int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

// The segmented scan of the unbiased scores over the reads of `db`:
// anno/points as mi_dust.  Bases already on the device at dbases (read i at
// boff[i]).
int dust_segmented(const HITS_DB* db, const void* dbases, const std::vector<long long>& boff,
                   const std::vector<int>& rlen, int window, double threshold, int mlen, long long* anno, int* points)
{
    const int n = db->nreads;
    const int seglen = std::max(64, env_int("MI_DUST_SEGMENT", std::max(384, 4 * window)));   // warm-up <= half
    const int warm = std::max(0, env_int("MI_DUST_WARMUP", std::max(64, 2 * window)));
    std::vector<int> seg_read, seg_s0;
    for (int i = 0; i < n; i++)
        for (int s0 = 0; s0 < rlen[i] - 2; s0 += seglen) {
            seg_read.push_back(i);
            seg_s0.push_back(s0);
        }
    const int nseg = (int)seg_read.size();
    if (nseg == 0) {
        for (int i = 0; i <= n; i++) anno[i] = 0;
        return 0;
    }
    const long long E = window + 1;
    const long long nstride = (nseg + 63LL) / 64 * 64;
    long long cap = 4LL * nseg + 4096;
    std::vector<void*> bufs;
    int rc = 0;
    const char* err = "mi_dust: device error";
    auto M = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (rc == 0 && hipMalloc(&p, std::max<size_t>(bytes, 8)) == hipSuccess) {
            bufs.push_back(p);
            return p;
        }
        if (rc == 0) err = "mi_dust: device allocation failed";
        rc = MI_ERR_DEVICE;
        return nullptr;
    };
    DustSegArgs a{};
    a.bases = (const uint8_t*)dbases;
    a.boff = (const long long*)M(8 * (size_t)n);
    a.rlen = (const int*)M(4 * (size_t)n);
    a.seg_read = (const int*)M(4 * (size_t)nseg);
    a.seg_s0 = (const int*)M(4 * (size_t)nseg);
    a.nseg = nseg;
    a.seglen = seglen;
    a.warm = warm;
    a.window = window;
    a.thresh = threshold;
    a.thresh2i = (int)ceil(2. * threshold);
    a.cbeg = (int*)M(4 * E * nstride);
    a.cend = (int*)M(4 * E * nstride);
    a.csc = (double*)M(8 * E * nstride);
    a.cnx = (short*)M(2 * E * nstride);
    a.cpv = (short*)M(2 * E * nstride);
    a.nstride = nstride;
    a.hdr = (DustHdr*)M(sizeof(DustHdr) * 2 * (size_t)nseg);
    a.ent = (DustEnt*)M(sizeof(DustEnt) * 2 * E * nseg);
    a.flag = (uint8_t*)M(nseg);
    a.repaired = (uint8_t*)M(nseg);
    int* counters = (int*)M(3 * sizeof(int));   // mismatches, intervals, unused
    a.bad = (int*)M(4 * (size_t)nseg);
    if (rc == 0) {
        a.mismatch = counters;
        a.out_cnt = counters + 1;
        const bool ok =
            hipMemcpy((void*)a.boff, boff.data(), 8 * (size_t)n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy((void*)a.rlen, rlen.data(), 4 * (size_t)n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy((void*)a.seg_read, seg_read.data(), 4 * (size_t)nseg, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy((void*)a.seg_s0, seg_s0.data(), 4 * (size_t)nseg, hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) rc = MI_ERR_DEVICE;
    }
    const bool large = window > 64;
    std::vector<DustIv> out;
    std::vector<uint8_t> repaired;
    int mism = 0, nrep = 0;
    for (int attempt = 0; rc == 0 && attempt < 2; attempt++) {
        a.out = (DustIv*)M(sizeof(DustIv) * (size_t)cap);
        a.out_cap = (int)std::min<long long>(cap, INT32_MAX);
        if (rc) break;
        if (hipMemset(counters, 0, 3 * sizeof(int)) != hipSuccess ||
            hipMemset(a.bad, 0, 4 * (size_t)nseg) != hipSuccess ||
            hipMemset(a.repaired, 0, nseg) != hipSuccess) {
            rc = MI_ERR_DEVICE;
            break;
        }
        const dim3 grid((unsigned)(nstride / 64));
        if (large) hipLaunchKernelGGL((k_sdust_seg<512, false>), grid, dim3(64), 0, 0, a);
        else hipLaunchKernelGGL((k_sdust_seg<128, false>), grid, dim3(64), 0, 0, a);
        hipLaunchKernelGGL(k_dust_compare, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, 0, a);
        int cnt[2];
        if (hipGetLastError() != hipSuccess || hipMemcpy(cnt, counters, 2 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = MI_ERR_DEVICE;
            break;
        }
        mism = cnt[0];
        nrep = 0;
        void* drep = nullptr;
        if (mism > 0) {
            // one repair lane per read with a flagged segment, from its first
            std::vector<uint8_t> flag(nseg);
            if (hipMemcpy(flag.data(), a.flag, nseg, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = MI_ERR_DEVICE;
                break;
            }
            std::vector<int> rep;
            for (int s = 0; s < nseg; s++)
                if (flag[s] && (rep.empty() || seg_read[rep.back()] != seg_read[s])) rep.push_back(s);
            nrep = (int)rep.size();
            drep = M(4 * rep.size());
            if (rc || hipMemcpy(drep, rep.data(), 4 * rep.size(), hipMemcpyHostToDevice) != hipSuccess) {
                rc = MI_ERR_DEVICE;
                break;
            }
            a.rep_seg = (const int*)drep;
            a.nrep = nrep;
            const dim3 rgrid((unsigned)((nrep + 63) / 64));
            if (large) hipLaunchKernelGGL((k_sdust_seg<512, true>), rgrid, dim3(64), 0, 0, a);
            else hipLaunchKernelGGL((k_sdust_seg<128, true>), rgrid, dim3(64), 0, 0, a);
        }
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(cnt, counters, 2 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = MI_ERR_DEVICE;
            break;
        }
        if (cnt[1] > a.out_cap) {   // more intervals than the buffer: again with room for all
            cap = (long long)cnt[1] + 4096;
            continue;
        }
        std::vector<int> bad(nseg);
        out.resize(cnt[1]);
        repaired.resize(nseg);
        if (hipMemcpy(bad.data(), a.bad, 4 * (size_t)nseg, hipMemcpyDeviceToHost) != hipSuccess ||
            (cnt[1] && hipMemcpy(out.data(), a.out, sizeof(DustIv) * (size_t)cnt[1], hipMemcpyDeviceToHost) != hipSuccess) ||
            hipMemcpy(repaired.data(), a.repaired, nseg, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = MI_ERR_DEVICE;
            break;
        }
        for (int s = 0; s < nseg; s++)
            if (bad[s]) {
                rc = MI_ERR_ARG;
                err = "mi_dust: a read needs more candidate intervals than the window holds";
                break;
            }
        break;
    }
    for (void* p : bufs) (void)hipFree(p);
    if (rc) {
        mi_set_error(err);
        return rc;
    }
    if (getenv("MI_DUST_STATS"))
        fprintf(stderr, "mi_dust: %d segments of %d steps (warm-up %d), %d began in another state, %d reads re-run\n",
                nseg, seglen, warm, mism, nrep);
    // the intervals of each segment in append order (one lane appends a
    // segment's intervals in order), from the pass that ran it last
    std::vector<int> first(nseg + 1, 0);
    for (const DustIv& e : out)
        if (e.pass == repaired[e.seg]) first[e.seg + 1]++;
    for (int s = 0; s < nseg; s++) first[s + 1] += first[s];
    struct Iv {
        int b, e;
    };
    std::vector<Iv> iv(first[nseg]);
    std::vector<int> at(first.begin(), first.end() - 1);
    for (const DustIv& e : out)
        if (e.pass == repaired[e.seg]) iv[at[e.seg]++] = Iv{e.b, e.e};
    // fold the segments' intervals read by read, as the interval stack does,
    // and keep those of at least mlen as points (b, e + 1)
    long long o = 0;
    anno[0] = 0;
    std::vector<Iv> stk;
    for (int i = 0, s = 0; i < n; i++) {
        stk.clear();
        for (; s < nseg && seg_read[s] == i; s++)
            for (int k = first[s]; k < first[s + 1]; k++) {
                const Iv v = iv[k];
                if (!stk.empty() && stk.back().e + 1 >= v.b) {
                    if (stk.back().e < v.e) stk.back().e = v.e;
                } else {
                    stk.push_back(v);
                }
            }
        for (const Iv& v : stk)
            if (v.e - v.b >= mlen) {
                if (points) {
                    points[o] = v.b;
                    points[o + 1] = v.e + 1;
                }
                o += 2;
            }
        anno[i + 1] = 4 * o;
    }
    return 0;
}

}  // namespace

extern "C" int mi_dust(const HITS_DB* db, int window, double threshold, int mlen, int biased, long long* anno,
                       int* points)
{
    if (!db || !anno || (db->nreads > 0 && (!db->bases || !db->reads))) {
        mi_set_error("mi_dust: null block, bases or output");
        return MI_ERR_ARG;
    }
    if (window <= 0 || window > kDustMaxWindow || !(threshold > 0.)) {
        mi_set_error("mi_dust: window must be in [1, 256] and the threshold positive");
        return MI_ERR_ARG;
    }
    const int n = db->nreads;
    anno[0] = 0;
    if (n == 0) return 0;
    std::vector<long long> boff(n), soff(n + 1);
    std::vector<int> rlen(n);
    long long lo = INT64_MAX, hi = 0, tot = 0;
    for (int i = 0; i < n; i++) {
        rlen[i] = db->reads[i].rlen;
        if (rlen[i] < 0) {
            mi_set_error("mi_dust: negative read length");
            return MI_ERR_ARG;
        }
        lo = std::min<long long>(lo, db->reads[i].boff);
        hi = std::max<long long>(hi, db->reads[i].boff + rlen[i] + 2);
        soff[i] = tot;
        tot += rlen[i] + 4;
    }
    soff[n] = tot;
    for (int i = 0; i < n; i++) boff[i] = db->reads[i].boff - lo;
    if (!biased && window >= 3) {
        void* dbases = nullptr;
        if (hipMalloc(&dbases, std::max<long long>(hi - lo, 8)) != hipSuccess) {
            mi_set_error("mi_dust: device allocation failed");
            return MI_ERR_DEVICE;
        }
        int rc = MI_ERR_DEVICE;
        if (hipMemcpy(dbases, (const char*)db->bases + lo, hi - lo, hipMemcpyHostToDevice) == hipSuccess)
            rc = dust_segmented(db, dbases, boff, rlen, window, threshold, mlen, anno, points);
        else
            mi_set_error("mi_dust: device error");
        (void)hipFree(dbases);
        return rc;
    }
    // the biased weights, with the reference's float product (db/DBdust.c:250-257)
    double skew[64];
    for (int p = 0, x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++)
            for (int z = 0; z < 4; z++) {
                const float f = db->freq[x] * db->freq[y] * db->freq[z];
                skew[p++] = .015625 / f;
            }
    const int cus = std::max(1, mi_num_cus());
    const int grid = (int)std::min<long long>((n + 63) / 64, (long long)cus * 8);
    const long long nslots = (long long)grid * 64;
    const bool lds_nodes = window <= kDustLdsWindow;
    const long long nodes = lds_nodes ? 1 : (long long)(window + 1) * nslots;
    void *dbases = nullptr, *dboff = nullptr, *drlen = nullptr, *dskew = nullptr, *dbeg = nullptr, *dend = nullptr,
         *dsc = nullptr, *dnx = nullptr, *dpv = nullptr, *dstk = nullptr, *dsoff = nullptr, *dnp = nullptr,
         *dst = nullptr;
    int rc = 0;
    auto M = [&](void** p, size_t b) { return rc == 0 && hipMalloc(p, std::max<size_t>(b, 8)) == hipSuccess; };
    if (!M(&dbases, hi - lo) || !M(&dboff, 8 * n) || !M(&drlen, 4 * n) || !M(&dskew, sizeof skew) ||
        !M(&dbeg, 4 * nodes) || !M(&dend, 4 * nodes) || !M(&dsc, 8 * nodes) || !M(&dnx, 2 * nodes) ||
        !M(&dpv, 2 * nodes) || !M(&dstk, 4 * tot) || !M(&dsoff, 8 * (n + 1)) || !M(&dnp, 4 * n) || !M(&dst, 4 * n)) {
        mi_set_error("mi_dust: device allocation failed");
        rc = MI_ERR_DEVICE;
    }
    std::vector<int> np(n), st(n), stk;
    if (rc == 0) {
        DustArgs a{};
        a.bases = (const uint8_t*)dbases;
        a.boff = (const long long*)dboff;
        a.rlen = (const int*)drlen;
        a.nreads = n;
        a.window = window;
        a.thresh = threshold;
        a.thresh2r = 2. * threshold;
        a.thresh2i = (int)ceil(a.thresh2r);
        a.mlen = mlen;
        a.biased = biased;
        a.skew = (const double*)dskew;
        a.cbeg = (int*)dbeg;
        a.cend = (int*)dend;
        a.cscore = (double*)dsc;
        a.cnext = (short*)dnx;
        a.cprev = (short*)dpv;
        a.nslots = nslots;
        a.stack = (int*)dstk;
        a.soff = (const long long*)dsoff;
        a.npts = (int*)dnp;
        a.status = (int*)dst;
        const bool ok =
            hipMemcpy(dbases, (const char*)db->bases + lo, hi - lo, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dboff, boff.data(), 8 * n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(drlen, rlen.data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dskew, skew, sizeof skew, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dsoff, soff.data(), 8 * (n + 1), hipMemcpyHostToDevice) == hipSuccess;
        if (ok) {
            if (lds_nodes) hipLaunchKernelGGL(k_sdust<true>, dim3(grid), dim3(64), 0, 0, a);
            else hipLaunchKernelGGL(k_sdust<false>, dim3(grid), dim3(64), 0, 0, a);
        }
        stk.resize(tot);
        if (!ok || hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(np.data(), dnp, 4 * n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(st.data(), dst, 4 * n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(stk.data(), dstk, 4 * tot, hipMemcpyDeviceToHost) != hipSuccess) {
            mi_set_error("mi_dust: device error");
            rc = MI_ERR_DEVICE;
        }
    }
    for (void* p : {dbases, dboff, drlen, dskew, dbeg, dend, dsc, dnx, dpv, dstk, dsoff, dnp, dst})
        if (p) (void)hipFree(p);
    if (rc) return rc;
    long long o = 0;
    for (int i = 0; i < n; i++) {
        if (st[i]) {
            mi_set_error("mi_dust: a read needs more candidate intervals than the window holds");
            return MI_ERR_ARG;
        }
        if (points && np[i] > 0) memcpy(points + o, stk.data() + soff[i], 4 * (size_t)np[i]);
        o += np[i];
        anno[i + 1] = 4 * o;
    }
    return 0;
}
