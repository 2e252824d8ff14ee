// Wave64 furthest-reaching-point aligner for gfx950 (CDNA4).
//
// One 64-lane wavefront computes one daligner Local_Alignment
// (dalign/align.c:1922-2060): the forward pass (dalign/align.c:369-1110)
// then the reverse pass (dalign/align.c:1114-1916).  The band of diagonals of
// the current wave lives in registers: lane L of chunk j holds diagonal
// k = base + 64*j + L, and lanes outside the band hold the pass's sentinel
// (-1 forward, INT32_MAX reverse), exactly the values the reference stores at
// V[low-1] / V[hgh+1].  Neighbour diagonals k-1 / k+1 of the previous wave are
// fetched with cross-lane gathers (ds_bpermute), the record / trim / clip
// decisions of the sequential reference loop become wave-wide prefix scans
// and 64-bit ballots, and match extension ("snake") compares 16 bases per
// step from 2-bit packed reads.
//
// Pebbles (trace-point cells, dalign/align.c:359-365) are appended to a
// per-wave pool in HBM; lanes that cross a trace-point boundary in the same
// wave allocate with a ballot + mbcnt prefix.  A diagonal carries only the
// index of its current A and B pebble: the reference's `cells[ha].mark` test
// reads the pebble's mark back from the pool in the (rare) waves where a
// lane reaches a trace point, and the trace is rebuilt by one walk per chain
// that also counts its length.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Optional per-wave trace hook for the CPU wave emulator used by the tests;
// compiles to nothing in the library.
#ifndef MI_WAVE_TRACE
#define MI_WAVE_TRACE(...)
#endif

namespace mi {

constexpr int kPathLen = 60;
constexpr uint64_t kPathTop = 1ull << 60;
constexpr uint64_t kPathInt = kPathTop - 1;
constexpr int kTrimLen = 15;
constexpr int kTrimMask = 0x7fff;
constexpr int kTrimMLag = 200;
constexpr int kWaveLag = 30;
// windows of several 64-lane chunks: a band of at most kOneChunk diagonals
// that spans two chunks is moved into the middle of the first, so the
// other chunks' steps are skipped (at least 12 diagonals of slack each side)
constexpr int kOneChunk = 40;

// status codes of one alignment
constexpr int LA_OK = 0;
constexpr int LA_WIDE = 1;    // band wider than this instantiation's register window
constexpr int LA_POOL = 2;    // pebble pool or point scratch exhausted
constexpr int LA_UNDEF = 3;   // input drives the wave into a state the reference leaves undefined
constexpr int LA_CHAIN = 4;   // internal consistency check of a pebble chain failed
constexpr int LA_ABORT = 5;   // host watchdog asked the kernel to stop

struct DevSpec {
    const int16_t* table;   // [32768] suffix-positivity table, then the [32768] score table (one base register)
    const int16_t* score;   // == table + 32768
    int tspace;
    int ave_path;
};

struct Peb {
    int up, diag, dif, mark;
};

// ---- sequence access -------------------------------------------------------
// 2-bit packed read: base i at bits 2*(i&31) of word (i>>5); one guard word
// before word 0 and two after the last word, so windows starting in
// [-32, len+32] stay inside the allocation.
__device__ __forceinline__ uint64_t win32(const uint64_t* pk, int p)
{
    int q = p >> 5;
    int r = p & 31;
    uint64_t lo = pk[q];
    if (r == 0) return lo;
    uint64_t hi = pk[q + 1];
    return (lo >> (2 * r)) | (hi << (64 - 2 * r));
}

// 16-base window starting at base p, read as 32-bit words of the same
// buffer (base i at bits 2*(i&15) of word i>>4): one funnel shift of the two
// words holding it.  Windows starting in [-32, len+32] stay inside the
// guard words.
__device__ __forceinline__ uint32_t win16(const uint32_t* pk, int p)
{
    const int q = p >> 4;
    return __builtin_amdgcn_alignbit(pk[q + 1], pk[q], (uint32_t)(2 * (p & 15)));
}

// forward: count of equal bases A[x+i] == B[y+i] for i < lim
__device__ __forceinline__ int match_fwd(const uint64_t* A, const uint64_t* B, int x, int y, int lim, bool& mism)
{
    const uint32_t* a = reinterpret_cast<const uint32_t*>(A);
    const uint32_t* b = reinterpret_cast<const uint32_t*>(B);
    int s = 0;
    mism = false;
    while (s < lim) {
        const uint32_t d = win16(a, x + s) ^ win16(b, y + s);
        if (d) {
            s += __builtin_ctz(d) >> 1;
            if (s < lim) mism = true;
            else s = lim;
            return s;
        }
        s += 16;
    }
    return lim;
}

// reverse: count of equal bases A[x-1-i] == B[y-1-i] for i < lim
__device__ __forceinline__ int match_rev(const uint64_t* A, const uint64_t* B, int x, int y, int lim, bool& mism)
{
    const uint32_t* a = reinterpret_cast<const uint32_t*>(A);
    const uint32_t* b = reinterpret_cast<const uint32_t*>(B);
    int s = 0;
    mism = false;
    while (s < lim) {
        const uint32_t d = win16(a, x - s - 16) ^ win16(b, y - s - 16);
        if (d) {
            s += __builtin_clz(d) >> 1;
            if (s < lim) mism = true;
            else s = lim;
            return s;
        }
        s += 16;
    }
    return lim;
}

// ---- cross-lane helpers ----------------------------------------------------
__device__ __forceinline__ int shfl_i(int v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    int lo = __shfl((int)(uint32_t)v, src, 64);
    int hi = __shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ int rdl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
// value of lane (L-1) & 63 / (L+1) & 63: DPP wave_ror:1 / wave_rol:1, full-rate
// VALU moves (no LDS traffic).  All 64 lanes must be active.
// (every lane has a source, so no "old" value is needed: one v_mov_b32_dpp)
__device__ __forceinline__ int lane_below(int v) { return __builtin_amdgcn_mov_dpp(v, 0x13C, 0xF, 0xF, false); }
__device__ __forceinline__ int lane_above(int v) { return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, false); }
__device__ __forceinline__ uint64_t lane_below64(uint64_t v)
{
    return ((uint64_t)(uint32_t)lane_below((int)(uint32_t)(v >> 32)) << 32) | (uint32_t)lane_below((int)(uint32_t)v);
}
__device__ __forceinline__ uint64_t lane_above64(uint64_t v)
{
    return ((uint64_t)(uint32_t)lane_above((int)(uint32_t)(v >> 32)) << 32) | (uint32_t)lane_above((int)(uint32_t)v);
}
// Wave-wide scans from DPP row moves (no LDS): row_shl:n / row_shr:n read
// lane L+n / L-n of the same 16-lane row, row_bcast:15 / :31 hand the last
// lane of a row (of two rows) to the next row(s).  Lanes without a source
// read 0 (bound_ctrl), the identity of max over non-negative values, so each
// level is a single v_max_i32 with a DPP operand.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp0(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, true);
}
// inclusive max over lanes >= L of x >= 0
__device__ __forceinline__ int suffix_max_nn(int x)
{
    x = max(x, dpp0<0x101>(x));   // row_shl:1
    x = max(x, dpp0<0x102>(x));   // row_shl:2
    x = max(x, dpp0<0x104>(x));   // row_shl:4
    x = max(x, dpp0<0x108>(x));   // row_shl:8
    // lane 16r now holds row r's maximum; lanes of row r add the rows above
    // it, selected with arithmetic masks (-1 / 0) rather than lane masks
    const int s1 = __builtin_amdgcn_readlane(x, 16), s2 = __builtin_amdgcn_readlane(x, 32);
    const int s3 = __builtin_amdgcn_readlane(x, 48);
    int row = __lane_id() >> 4;
    asm volatile("" : "+v"(row));   // keeps the row masks from being hoisted into (spilled) SGPR lane masks
    const int m1 = (row - 1) >> 31, m2 = (row - 2) >> 31, m3 = (row - 3) >> 31;
    return max(max(x, s1 & m1), max(s2 & m2, s3 & m3));
}
// inclusive max over lanes <= L of x >= 0
__device__ __forceinline__ int prefix_max_nn(int x)
{
    x = max(x, dpp0<0x111>(x));        // row_shr:1
    x = max(x, dpp0<0x112>(x));        // row_shr:2
    x = max(x, dpp0<0x114>(x));        // row_shr:4
    x = max(x, dpp0<0x118>(x));        // row_shr:8
    x = max(x, dpp0<0x142, 0xA>(x));   // row_bcast:15 into rows 1, 3
    x = max(x, dpp0<0x143, 0xC>(x));   // row_bcast:31 into rows 2, 3
    return x;
}
// value of lane L+1 / L-1; the lane without one (63 / 0) gets `edge`
// (DPP wave_shl:1 / wave_shr:1)
__device__ __forceinline__ int lane_above_or(int v, int edge) { return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ int lane_below_or(int v, int edge) { return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xF, 0xF, false); }
// inclusive max over lanes >= L (x >= -1).  GATHER: the lanes reversed by one
// ds_bpermute each way around the DPP prefix scan, for passes that run many
// waves per SIMD (the LDS pipe is idle there, the VALU is not); otherwise
// the DPP suffix scan, which keeps the two LDS round trips off the path of a
// wave that runs alone (the wider windows).
template <bool GATHER>
__device__ __forceinline__ int wave_suffix_max(int x)
{
    if constexpr (GATHER) {
        const int r = 63 - __lane_id();
        return __shfl(prefix_max_nn(__shfl(x + 1, r, 64)), r, 64) - 1;
    } else {
        return suffix_max_nn(x + 1) - 1;
    }
}
// inclusive min over lanes <= L (any x): the max of INT32_MAX - x >= 0
__device__ __forceinline__ int wave_prefix_min(int x) { return 2147483647 - prefix_max_nn(2147483647 - x); }

__device__ __forceinline__ uint64_t rdl64(uint64_t v, int lane)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}
// Order this wave's earlier global stores before its later loads (all
// lanes of one wave; no workgroup barrier, so several waves per workgroup may
// run independent alignments).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ int lowest_lane(unsigned long long m) { return __ffsll((long long)m) - 1; }
__device__ __forceinline__ int highest_lane(unsigned long long m) { return 63 - __clzll(m); }
__device__ __forceinline__ int prefix_rank(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Per-diagonal state (one register per field per lane).
struct Diag {
    int V, M, NA, NB;
    uint64_t T;
    int HA, HB;   // current A / B pebble (index in the pool)
};

__device__ __forceinline__ Diag diag_zero(int sent)
{
    Diag d;
    d.V = sent; d.M = 0; d.NA = 0; d.NB = 0; d.T = 0;
    d.HA = d.HB = 0;
    return d;
}

__device__ __forceinline__ Diag diag_shfl(const Diag& D, int sl)
{
    Diag g;
    g.V = shfl_i(D.V, sl); g.M = shfl_i(D.M, sl); g.NA = shfl_i(D.NA, sl); g.NB = shfl_i(D.NB, sl);
    g.T = shfl_u64(D.T, sl);
    g.HA = shfl_i(D.HA, sl); g.HB = shfl_i(D.HB, sl);
    return g;
}

// fields of lane L + sel (sel in {-1, 0, +1}, wrapping like (L + sel) & 63)
__device__ __forceinline__ int pick3(int v, int sel)
{
    const int b = lane_below(v), a = lane_above(v);
    return sel < 0 ? b : (sel > 0 ? a : v);
}
// The predecessor's fields the pass carries on (M, T, HA, HB) from lane
// (L + sel) & 63; V, NA and NB stay the lane's own.  GATHER: one ds_bpermute
// per field, on the LDS pipe rather than as DPP moves and selects on the
// VALU; otherwise the DPP moves (shorter latency for a wave that runs alone).
template <bool GATHER>
__device__ __forceinline__ Diag diag_pick(const Diag& D, int sel)
{
    Diag g;
    g.V = D.V; g.NA = D.NA; g.NB = D.NB;
    if constexpr (GATHER) {
        const int src = (__lane_id() + sel) & 63;
        g.M = shfl_i(D.M, src);
        g.T = shfl_u64(D.T, src);
        g.HA = shfl_i(D.HA, src); g.HB = shfl_i(D.HB, src);
    } else {
        g.M = pick3(D.M, sel);
        g.T = ((uint64_t)(uint32_t)pick3((int)(uint32_t)(D.T >> 32), sel) << 32) |
              (uint32_t)pick3((int)(uint32_t)D.T, sel);
        g.HA = pick3(D.HA, sel); g.HB = pick3(D.HB, sel);
    }
    return g;
}

__device__ __forceinline__ Diag diag_rdl(const Diag& D, int ln)
{
    Diag g;
    g.V = rdl(D.V, ln); g.M = rdl(D.M, ln); g.NA = rdl(D.NA, ln); g.NB = rdl(D.NB, ln);
    g.T = rdl64(D.T, ln);
    g.HA = rdl(D.HA, ln); g.HB = rdl(D.HB, ln);
    return g;
}

// Work counters of one wave: band cells visited, snake compare steps, wave steps.
struct WaveCount {
    long long cells = 0, snake = 0, steps = 0;
};

struct PassOut {
    int trimx, trimy, trimd;
    int tha, thb;   // tail pebbles of the A and B chains
    int used;       // pebbles allocated by the pass
    int status;
};

// Uniform read of diagonal kk's field set (kk inside the window).
template <int NCH>
__device__ __forceinline__ Diag fetch_diag(const Diag (&bd)[NCH], int base, int kk)
{
    int jj = (kk - base) >> 6, ln = (kk - base) & 63;
    Diag r = diag_zero(0);
#pragma unroll
    for (int j = 0; j < NCH; j++)
        if (j == jj) r = diag_rdl(bd[j], ln);
    return r;
}

// One pass of the wave (forward DIR=+1, reverse DIR=-1) from the band
// [low, hgh] on anti-diagonal mida.  Pebbles go to pool[0, pool_cap).
template <int NCH, int DIR>
__device__ PassOut wave_pass(const DevSpec& sp, const uint64_t* A, int alen, const uint64_t* B, int blen,
                             int comp, int low, int hgh, int mida, int minp, int maxp, Peb* pool,
                             int pool_cap, const int* abort_flag, WaveCount& wc)
{
    constexpr int W = 64 * NCH;
    constexpr int SENT = (DIR > 0) ? -1 : INT32_MAX;
    const int lane = __lane_id();
    const int TS = sp.tspace;
    const int boff = comp ? blen % TS : 0;
    PassOut out;
    out.status = LA_OK;

    Diag bd[NCH];
    if (hgh - low + 1 > W) {
        out.status = LA_WIDE;
        return out;
    }
    int base = low - (W - (hgh - low + 1)) / 2;
    int navail = 0;

    // state of the pass (names follow dalign/align.c:386-401)
    int besta = mida, besty = (mida - hgh) >> 1;
    int trima = mida, trimy = besty, trimd = 0;
    int trimha = 0, trimhb = 1;   // pebbles 0/1 = roots of the first visited diagonal
    int morea = mida, morey = besty, mored = 0, moreha = 0, morehb = 1;
    int morem = -1, lasta = mida, more = 1;
    int dif = 0;
    int my_snake = 0;   // this lane's snake compare steps in the pass

    int cN[NCH], yN[NCH], mN[NCH];
    uint64_t bN[NCH];
    bool stA[NCH], stB[NCH];
    Diag P[NCH];   // per diagonal: the predecessor's fields, updated by this wave's pebbles
    // Chunk j holds diagonals [base + 64 j, base + 64 j + 63].  A chunk wholly
    // outside the band [low, hgh] holds only sentinels (V = SENT, no record,
    // no clip, no pebble), so the per-chunk steps skip it: a wave-uniform
    // test, always true for the 64-diagonal window.
    auto live = [&](int j) { return NCH == 1 || (base + 64 * j + 63 >= low && base + 64 * j <= hgh); };

    // ---- wave 0 (dalign/align.c:455-587, :1193-1322) --------------------
    {
        const int nband = hgh - low + 1;
        if (2 * nband > pool_cap) {
            out.status = LA_POOL;
            return out;
        }
        navail = 2 * nband;
        bool undef = false;
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            const int k = base + 64 * j + lane;
            const bool in = (k >= low && k <= hgh);
            bd[j] = diag_zero(SENT);
            P[j] = diag_zero(0);
            cN[j] = SENT;
            yN[j] = 0;
            mN[j] = kPathLen;
            bN[j] = kPathInt;
            stA[j] = stB[j] = false;
            if (!in) continue;
            int y = (mida - k) >> 1;
            const int idx = (DIR > 0) ? (hgh - k) : (k - low);   // visiting order
            int na, nb, hamark, hbmark;
            if (DIR > 0) {
                na = ((y + k) / TS) * TS;
                nb = ((y + (TS - boff)) / TS - 1) * TS + boff;
                hamark = na;
                hbmark = nb;
                na += TS;
                nb += TS;
            } else {
                na = ((y + k + TS - 1) / TS - 1) * TS;
                nb = ((y + (TS - boff) - 1) / TS - 1) * TS + boff;
                hamark = y + k;
                hbmark = y;
            }
            const int ha = 2 * idx, hb = ha + 1;
            pool[ha] = Peb{-1, k, 0, hamark};
            pool[hb] = Peb{-1, k, 0, hbmark};
            P[j].HA = ha;
            P[j].HB = hb;
            bd[j].NA = na;
            bd[j].NB = nb;
            const int x = y + k;
            const bool outside = (y < 0 || y > blen || x < 0 || x > alen);
            if (outside) undef = true;
            bool mism;
            int s;
            if (DIR > 0) {
                s = match_fwd(A, B, x, y, outside ? 0 : min(alen - x, blen - y), mism);
                y += s;
                if (!mism) { if (y == blen) stB[j] = true; else stA[j] = true; }
            } else {
                s = match_rev(A, B, x, y, outside ? 0 : min(x, y), mism);
                y -= s;
                if (!mism) { if (y == 0) stB[j] = true; else stA[j] = true; }
            }
            my_snake += s + 1;
            cN[j] = 2 * y + k;
            yN[j] = y;
        }
        if (__any(undef)) {
            out.status = LA_UNDEF;
            return out;
        }
    }

    for (;;) {
        // ---- trace-point pebbles (dalign/align.c:521-563, :826-878, :1258-1297, :1562-1612)
        // Wave 0 pushes unconditionally; later waves only when the inherited
        // pebble does not already sit on this trace point.
        for (int rounds = 0;; rounds++) {
            if (rounds > (alen + blen) / TS + 8) {   // more trace points than the reads hold
                out.status = LA_UNDEF;
                return out;
            }
            bool needA[NCH], needB[NCH], any = false;
#pragma unroll
            for (int j = 0; j < NCH; j++) {
                needA[j] = needB[j] = false;
                if (!live(j)) continue;
                const int k = base + 64 * j + lane;
                const bool in = (k >= low && k <= hgh);
                const int y = yN[j];
                needA[j] = in && ((DIR > 0) ? (y + k >= bd[j].NA) : (y + k <= bd[j].NA));
                needB[j] = in && ((DIR > 0) ? (y >= bd[j].NB) : (y <= bd[j].NB));
                any |= needA[j] | needB[j];
            }
            if (!__any(any)) break;
            wave_sync();   // the pebbles pushed so far are visible to the mark reads below
#pragma unroll
            for (int j = 0; j < NCH; j++) {
                if (!live(j)) continue;
                const int k = base + 64 * j + lane;
                Diag& D = bd[j];
                Diag& G = P[j];
                // marks of the current pebbles (a pebble pushed in an earlier
                // round of this wave is read back by the lane that wrote it)
                int gam = 0, gbm = 0;
                if (needA[j] && dif != 0) gam = pool[G.HA].mark;
                if (needB[j] && dif != 0) gbm = pool[G.HB].mark;
                const bool pushA = needA[j] && (dif == 0 || ((DIR > 0) ? (gam < D.NA) : (gam > D.NA)));
                const bool pushB = needB[j] && (dif == 0 || ((DIR > 0) ? (gbm < D.NB) : (gbm > D.NB)));
                const unsigned long long ba = __ballot(pushA), bb = __ballot(pushB);
                const int na_ = __popcll(ba), nb_ = __popcll(bb);
                if (navail + na_ + nb_ > pool_cap) {
                    out.status = LA_POOL;
                    return out;
                }
                if (pushA) {
                    const int idx = navail + prefix_rank(ba);
                    pool[idx] = Peb{G.HA, k, dif, D.NA};
                    G.HA = idx;
                }
                navail += na_;
                if (pushB) {
                    const int idx = navail + prefix_rank(bb);
                    pool[idx] = Peb{G.HB, k, dif, D.NB};
                    G.HB = idx;
                }
                navail += nb_;
                if (needA[j]) D.NA += DIR * TS;
                if (needB[j]) D.NB += DIR * TS;
            }
        }

        // ---- records: strict improvements of besta in visiting order ------
        // forward visits k descending (exclusive suffix max over lanes),
        // reverse visits k ascending (exclusive prefix min).
        {
            int carry = besta;
            int best_k = 0, last_c = 0, tr_k = 0;
            bool have_best = false, have_last = false, have_tr = false;
#pragma unroll
            for (int jj = 0; jj < NCH; jj++) {
                const int j = (DIR > 0) ? (NCH - 1 - jj) : jj;
                if (!live(j)) continue;
                const int c = cN[j];
                const int incl = (DIR > 0) ? wave_suffix_max<NCH == 1>(c) : wave_prefix_min(c);
                const int excl = (DIR > 0) ? lane_above_or(incl, SENT) : lane_below_or(incl, SENT);
                const int prior = (DIR > 0) ? max(excl, carry) : min(excl, carry);
                const bool rec = (DIR > 0) ? (c > prior) : (c < prior);
                // masks straight from the compares; the trim tables are read by
                // the lanes of bm only
                const unsigned long long brec = __ballot(rec);
                unsigned long long bm = brec, bt = brec;   // wave 0 records set besta, lasta and the trim point
                if (dif != 0) {
                    bm = brec & __ballot(mN[j] >= sp.ave_path);
                    bt = 0;
                    if (bm) {
                        int t = 0;
                        if (rec && mN[j] >= sp.ave_path) {
                            const uint64_t b = bN[j];
                            if (sp.table[b & kTrimMask] >= 0)
                                t = (sp.table[(b >> kTrimLen) & kTrimMask] + sp.table[32768 + (b & kTrimMask)] >= 0);
                        }
                        bt = __ballot(t != 0);
                    }
                }
                if (brec) {
                    const int ln = (DIR > 0) ? lowest_lane(brec) : highest_lane(brec);
                    best_k = base + 64 * j + ln;
                    have_best = true;
                }
                if (bm) {
                    const int ln = (DIR > 0) ? lowest_lane(bm) : highest_lane(bm);
                    last_c = rdl(c, ln);
                    have_last = true;
                }
                if (bt) {
                    const int ln = (DIR > 0) ? lowest_lane(bt) : highest_lane(bt);
                    tr_k = base + 64 * j + ln;
                    have_tr = true;
                }
                const int cm = (DIR > 0) ? rdl(incl, 0) : rdl(incl, 63);
                carry = (DIR > 0) ? max(carry, cm) : min(carry, cm);
            }
            if (have_best) {
                const int jj = (best_k - base) >> 6, ln = (best_k - base) & 63;
#pragma unroll
                for (int j = 0; j < NCH; j++)
                    if (j == jj) { besta = rdl(cN[j], ln); besty = rdl(yN[j], ln); }
            }
            if (have_last) lasta = last_c;
            if (have_tr) {
                const int jj = (tr_k - base) >> 6, ln = (tr_k - base) & 63;
#pragma unroll
                for (int j = 0; j < NCH; j++)
                    if (j == jj) {
                        trima = rdl(cN[j], ln);
                        trimy = rdl(yN[j], ln);
                        trimha = rdl(P[j].HA, ln);
                        trimhb = rdl(P[j].HB, ln);
                    }
                trimd = dif;
            }
        }

        // ---- commit the wave ------------------------------------------------
        // Only V needs the band test: a lane outside the band holds SENT, and
        // an in-band V is an antidiagonal >= 0 (forward) / < SENT (reverse),
        // so the predecessor choice never takes an out-of-band lane's other
        // fields.
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            if (!live(j)) continue;   // its V is SENT since the last wave's sentinel step
            const int k = base + 64 * j + lane;
            Diag& D = bd[j];
            D.V = (k >= low && k <= hgh) ? cN[j] : SENT;
            D.M = mN[j]; D.T = bN[j];
            D.HA = P[j].HA; D.HB = P[j].HB;
        }
        wc.cells += hgh - low + 1;
        wc.steps += 1;

        // ---- clips (dalign/align.c:589-625, :912-944, :1327-1361, :1650-1686)
        {
            int aclip_k = (DIR > 0) ? INT32_MAX : -INT32_MAX;
            int bclip_k = (DIR > 0) ? -INT32_MAX : INT32_MAX;
#pragma unroll
            for (int j = 0; j < NCH; j++) {
                if (!live(j)) continue;
                const unsigned long long ba = __ballot(stA[j]), bb = __ballot(stB[j]);
                if (ba) {
                    more = 0;   // forward keeps the last visited = lowest k; reverse the highest
                    const int k = base + 64 * j + ((DIR > 0) ? lowest_lane(ba) : highest_lane(ba));
                    if (DIR > 0) aclip_k = min(aclip_k, k);
                    else aclip_k = max(aclip_k, k);
                }
                if (bb) {
                    more = 0;
                    const int k = base + 64 * j + ((DIR > 0) ? highest_lane(bb) : lowest_lane(bb));
                    if (DIR > 0) bclip_k = max(bclip_k, k);
                    else bclip_k = min(bclip_k, k);
                }
            }
            if (more == 0) {
                const int x = besta - besty, y = besty;
                const bool atB = (DIR > 0) ? (y == blen) : (y == 0);
                const bool atA = (DIR > 0) ? (x == alen) : (x == 0);
                if (!atB && !atA) more = 1;
                int cl[2] = {aclip_k, bclip_k};
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int kk = cl[q];
                    bool hit;
                    if (DIR > 0) hit = (q == 0) ? (hgh >= kk) : (low <= kk);
                    else hit = (q == 0) ? (low <= kk) : (hgh >= kk);
                    if (!hit) continue;
                    if ((DIR > 0) == (q == 0)) hgh = kk - 1;
                    else low = kk + 1;
                    const Diag g = fetch_diag<NCH>(bd, base, kk);
                    if (morem <= g.M) {
                        morem = g.M; morea = g.V; morey = (morea - kk) / 2; mored = dif;
                        moreha = g.HA; morehb = g.HB;
                    }
                }
            }
        }

        // ---- WAVE_LAG trim of the band ends (dalign/align.c:946-955, :1688-1697)
        if (dif > 0) {
            const int n = (DIR > 0) ? besta - kWaveLag : besta + kWaveLag;
            int nlo = INT32_MAX, nhi = -INT32_MAX;
#pragma unroll
            for (int j = 0; j < NCH; j++) {
                if (!live(j)) continue;
                const int k = base + 64 * j + lane;
                const bool keep = (k >= low && k <= hgh) && ((DIR > 0) ? (bd[j].V >= n) : (bd[j].V <= n));
                const unsigned long long bk = __ballot(keep);
                if (bk) {
                    nlo = min(nlo, base + 64 * j + lowest_lane(bk));
                    nhi = max(nhi, base + 64 * j + highest_lane(bk));
                }
            }
            if (nhi == -INT32_MAX) hgh = low - 1;
            else { low = nlo; hgh = nhi; }
        }
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            if (!live(j)) {
                bd[j].V = SENT;
                continue;
            }
            const int k = base + 64 * j + lane;
            if (!(k >= low && k <= hgh)) bd[j].V = SENT;
        }

        MI_WAVE_TRACE(DIR, dif, low, hgh, base, besta, lasta, more, navail);
        // ---- next wave? --------------------------------------------------
        if (DIR > 0) { if (!(more && lasta >= besta - kTrimMLag)) break; }
        else { if (!(more && lasta <= besta + kTrimMLag)) break; }
        if (hgh < low || dif > 2 * (alen + blen) + 1024) {
            out.status = LA_UNDEF;   // empty band with the loop still running
            return out;
        }
        if ((dif & 255) == 255 &&
            __builtin_amdgcn_readfirstlane(__atomic_load_n(abort_flag, __ATOMIC_RELAXED))) {
            out.status = LA_ABORT;
            return out;
        }

        // extend the band by one diagonal on each side (dalign/align.c:645-743)
        low -= 1;
        hgh += 1;
        const bool elo = (low >= minp);
        if (!elo) low += 1;
        const bool ehi = (hgh <= maxp);
        if (!ehi) hgh -= 1;
        const int wd = hgh - low + 1;
        bool rebase = (low < base || hgh > base + W - 1);
        if (rebase && wd > W) {
            out.status = LA_WIDE;
            return out;
        }
        if constexpr (NCH > 1)
            if (wd <= kOneChunk && ((low - base) >> 6) != ((hgh - base) >> 6)) rebase = true;
        if (rebase) {
            const int nbase = (NCH > 1 && wd <= kOneChunk) ? low - (64 - wd) / 2 : low - (W - wd) / 2;
            const int delta = nbase - base;
            Diag nb[NCH];
#pragma unroll
            for (int j = 0; j < NCH; j++) {
                const int s = 64 * j + lane + delta;   // index in the old window
                const int sl = s & 63;
                Diag nd = diag_zero(SENT);
#pragma unroll
                for (int q = 0; q < NCH; q++) {
                    const Diag g = diag_shfl(bd[q], sl);
                    if (s >= 0 && s < W && (s >> 6) == q) nd = g;
                }
                nb[j] = nd;
            }
#pragma unroll
            for (int j = 0; j < NCH; j++) bd[j] = nb[j];
            base = nbase;
        }
        dif += 1;

        // previous-wave values of k-1 and k+1 (lanes outside the band hold SENT)
        int Vm[NCH], Vp[NCH];
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            if (!live(j)) continue;
            int up = lane_below(bd[j].V);
            int dn = lane_above(bd[j].V);
            const int upc = (j > 0) ? rdl(bd[j - 1].V, 63) : SENT;       // wave-uniform reads
            const int dnc = (j < NCH - 1) ? rdl(bd[j + 1].V, 0) : SENT;
            if (lane == 0) up = upc;
            if (lane == 63) dn = dnc;
            Vm[j] = up;
            Vp[j] = dn;
        }
        // new end diagonals take NA/NB of their inner neighbour
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            if (!live(j)) continue;   // the chunks holding low and hgh are live
            const int k = base + 64 * j + lane;
            if constexpr (NCH == 1) {
                // the new end lanes read their inner neighbour (one lane up /
                // down, always inside the window), every other lane itself
                const int src = lane + ((elo && k == low) ? 1 : 0) - ((ehi && k == hgh) ? 1 : 0);
                bd[j].NA = shfl_i(bd[j].NA, src);
                bd[j].NB = shfl_i(bd[j].NB, src);
                continue;
            }
            int naU = lane_below(bd[j].NA), nbU = lane_below(bd[j].NB);
            int naD = lane_above(bd[j].NA), nbD = lane_above(bd[j].NB);
            if (j > 0) {
                const int a = rdl(bd[j - 1].NA, 63), b = rdl(bd[j - 1].NB, 63);
                if (lane == 0) { naU = a; nbU = b; }
            }
            if (j < NCH - 1) {
                const int a = rdl(bd[j + 1].NA, 0), b = rdl(bd[j + 1].NB, 0);
                if (lane == 63) { naD = a; nbD = b; }
            }
            if (elo && k == low) { bd[j].NA = naD; bd[j].NB = nbD; }
            if (ehi && k == hgh) { bd[j].NA = naU; bd[j].NB = nbU; }
        }

        // predecessor choice (dalign/align.c:760-791, :1498-1529) and gather
        bool undef = false;
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            if (!live(j)) {
                cN[j] = SENT;
                stA[j] = stB[j] = false;
                continue;
            }
            const int k = base + 64 * j + lane;
            const bool in = (k >= low && k <= hgh);
            const int am = Vm[j], ac = bd[j].V, ap = Vp[j];
            int sel;
            if (DIR > 0) {
                if (ac < am && am >= ap) sel = -1;
                else if (ac < am || ac < ap) sel = +1;
                else sel = 0;
            } else {
                if (ac > ap && ap <= am) sel = +1;
                else if (ac > ap || ac > am) sel = -1;
                else sel = 0;
            }
            Diag g = diag_pick<NCH == 1>(bd[j], sel);
            if (j > 0) {
                const Diag e = diag_rdl(bd[j - 1], 63);
                if (lane == 0 && sel == -1) g = e;
            }
            if (j < NCH - 1) {
                const Diag e = diag_rdl(bd[j + 1], 0);
                if (lane == 63 && sel == +1) g = e;
            }
            P[j] = g;
            // lanes outside the band carry sentinels: no arithmetic on them
            int c = in ? ((sel == 0) ? ac + 2 * DIR : ((sel < 0) ? am : ap) + DIR) : k;

            // mismatch column, then the snake with s match columns at once
            int m = g.M;
            uint64_t b = g.T;
            if (b & kPathTop) m -= 1;
            b <<= 1;
            int y = (c - k) >> 1;
            stA[j] = stB[j] = false;
            if (in) {
                const int x = y + k;
                const bool outside = (y < 0 || y > blen || x < 0 || x > alen);
                if (outside) undef = true;
                bool mism;
                int s;
                if (DIR > 0) {
                    s = match_fwd(A, B, x, y, outside ? 0 : min(alen - x, blen - y), mism);
                    y += s;
                    if (!mism) { if (y == blen) stB[j] = true; else stA[j] = true; }
                } else {
                    s = match_rev(A, B, x, y, outside ? 0 : min(x, y), mism);
                    y -= s;
                    if (!mism) { if (y == 0) stB[j] = true; else stA[j] = true; }
                }
                my_snake += s + 1;
                // s unit steps of dalign/align.c:799-823 in closed form
                const int sc = s < 61 ? s : 61;
                const uint64_t msk = (sc == 0) ? 0 : (((1ull << sc) - 1) << (61 - sc));
                m += sc - __popcll(b & msk);
                b = (s >= 64) ? ~0ull : ((b << s) | ((1ull << s) - 1));
                c = 2 * y + k;
            } else
                c = SENT;
            cN[j] = c;
            yN[j] = y;
            mN[j] = m;
            bN[j] = b;
        }
        if (__any(undef)) {
            out.status = LA_UNDEF;
            return out;
        }
    }

    if (morem >= 0) {
        out.trimx = morea - morey;
        out.trimy = morey;
        out.trimd = mored;
        out.tha = moreha;
        out.thb = morehb;
    } else {
        out.trimx = trima - trimy;
        out.trimy = trimy;
        out.trimd = trimd;
        out.tha = trimha;
        out.thb = trimhb;
    }
    out.used = navail;
    long long t = my_snake;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
    wc.snake += t;
    return out;
}

// ---- trace assembly ----------------------------------------------------------
// Lane 0 rebuilds the A-view chain and lane 1 the B-view chain of one pass
// (the same code runs on both lanes with per-lane coordinates).  The chain
// from the tail pebble back to its root is copied tail first to the end of
// pts[0..pts_cap), so it reads root to tail from pts + pts_cap - (depth + 1),
// then walked like dalign/align.c:970-1107 (forward) or :1709-1901 (reverse).
//   tr     : trace origin (forward writes tr[0..], reverse prepends tr[-1..]);
//            it holds maxtp values on either side
//   fwdlen : (reverse only) length of the forward part already at tr[0..]
// Returns the new trace length / first index through *len / *first, or
// false for a chain that leaves the pool, outgrows pts_cap or needs more
// than maxtp trace values.
template <int DIR>
__device__ bool assemble_chain(const Peb* pool, int pool_used, Peb* pts_all, int pts_cap, int tail, bool isB, int mida,
                               int TS, int boff, int trimx, int trimy, int trimd, uint16_t* tr, int maxtp, int fwdlen,
                               int* len, int* first, int* root_diag)
{
    // copy the chain, tail first, until its root (up == -1); a chain that
    // leaves the pool or does not end within pts_cap links is reported,
    // never followed further
    int h = tail, nl = 0;
    while (h != -1) {
        if (h < 0 || h >= pool_used || nl >= pts_cap) return false;
        const Peb p = pool[h];
        pts_all[pts_cap - 1 - nl] = p;
        h = p.up;
        nl++;
    }
    if (nl == 0) return false;
    const int depth = nl - 1;
    if (2 * depth + 2 > maxtp) return false;
    const Peb* pts = pts_all + (pts_cap - nl);
    const int tprim = isB ? trimy : trimx;   // end-point coordinate along the trace axis
    const int tsec = isB ? trimx : trimy;    // and across it
    // secondary coordinate of a pebble: A-view -> y, B-view -> x
    auto sec = [&](const Peb& p) { return isB ? p.mark + p.diag : p.mark - p.diag; };
    *root_diag = pts[0].diag;
    if (DIR > 0) {
        const int k0 = pts[0].diag;
        int b = isB ? (mida + k0) / 2 : (mida - k0) / 2;
        int e = 0, n = 0, kk = k0;
        for (int i = 1; i <= depth; i++) {
            kk = pts[i].diag;
            const int a = sec(pts[i]);
            const int dd = pts[i].dif;
            tr[n++] = (uint16_t)(dd - e);
            tr[n++] = (uint16_t)(a - b);
            b = a;
            e = dd;
        }
        const int prim = isB ? b - kk : b + kk;
        if (prim != tprim) {
            tr[n++] = (uint16_t)(trimd - e);
            tr[n++] = (uint16_t)(tsec - b);
        } else if (b != tsec && n > 0) {
            tr[n - 1] = (uint16_t)(tr[n - 1] + (tsec - b));
            tr[n - 2] = (uint16_t)(tr[n - 2] + (trimd - e));
        }
        *len = n;
        *first = 0;
    } else {
        int kk = pts[0].diag;
        int b = sec(pts[0]);
        int e = 0, n = 0, i = 0;
        bool live = true;   // reference: h >= 0
        const int prim0 = isB ? b - kk : b + kk;
        if (prim0 % TS != (isB ? boff : 0)) {
            int a, dd;
            i = 1;
            if (i > depth) {
                a = tsec;
                dd = trimd;
                live = false;
            } else {
                kk = pts[i].diag;
                a = sec(pts[i]);
                dd = pts[i].dif;
            }
            if (fwdlen == 0) {
                tr[--n] = (uint16_t)(b - a);
                tr[--n] = isB ? (uint16_t)(b - a) : (uint16_t)(dd - e);   // dalign/align.c:1831-1832
            } else {
                tr[1] = (uint16_t)(tr[1] + (b - a));
                tr[0] = (uint16_t)(tr[0] + (dd - e));
            }
            b = a;
            e = dd;
        }
        if (live) {
            for (int q = i + 1; q <= depth; q++) {
                kk = pts[q].diag;
                const int a = sec(pts[q]);
                tr[--n] = (uint16_t)(b - a);
                const int dd = pts[q].dif;
                tr[--n] = (uint16_t)(dd - e);
                b = a;
                e = dd;
            }
            const int prim = isB ? b - kk : b + kk;
            if (prim != tprim) {
                tr[--n] = (uint16_t)(b - tsec);
                tr[--n] = (uint16_t)(trimd - e);
            } else if (b != tsec && (n < 0 || fwdlen > 0)) {
                tr[n + 1] = (uint16_t)(tr[n + 1] + (b - tsec));
                tr[n] = (uint16_t)(tr[n] + (trimd - e));
            }
        }
        *len = -n;
        *first = n;
    }
    return true;
}

struct LaResult {
    int abpos, bbpos, aepos, bepos, diffs;     // A-view path
    int b_abpos, b_bbpos, b_aepos, b_bepos;    // B-view path (diffs equal)
    int atlen, btlen;
    int afirst, bfirst;                        // trace start relative to the trace origins
    int status;
};

// Full Local_Alignment (dalign/align.c:1922-2060) by one wave.
//   pool / pool_cap : pebble pool of this wave
//   pts / pts_cap   : point scratch of this wave (2 chains)
//   atr / btr       : trace origins (each needs maxtp entries before and after)
template <int NCH>
__device__ LaResult local_alignment(const DevSpec& sp, const uint64_t* A, int alen, const uint64_t* B, int blen,
                                    int comp, int low, int hgh, int anti, int minp, int maxp, Peb* pool,
                                    int pool_cap, Peb* pts, int pts_cap, uint16_t* atr, uint16_t* btr, int maxtp,
                                    const int* abort_flag, WaveCount& wc)
{
    LaResult r;
    const int lane = __lane_id();
    const int TS = sp.tspace;
    const int boff = comp ? blen % TS : 0;
    r.status = LA_OK;

    PassOut f = wave_pass<NCH, +1>(sp, A, alen, B, blen, comp, low, hgh, anti, minp, maxp, pool, pool_cap, abort_flag,
                                   wc);
    if (f.status != LA_OK) { r.status = f.status; return r; }
    wave_sync();
    int len = 0, first = 0, rootk = 0;
    bool ok = true;
    if (lane < 2) {
        const bool isB = (lane == 1);
        ok = assemble_chain<+1>(pool, f.used, pts + (isB ? pts_cap : 0), pts_cap, isB ? f.thb : f.tha, isB, anti, TS,
                                boff, f.trimx, f.trimy, f.trimd, isB ? btr : atr, maxtp, 0, &len, &first, &rootk);
    }
    if (__any(!ok)) { r.status = LA_CHAIN; return r; }
    const int fa = rdl(len, 0), fb = rdl(len, 1);
    const int mind = rdl(rootk, 1);   // diagonal of the B chain root (dalign/align.c:1059-1061)
    r.aepos = f.trimx;
    r.bepos = f.trimy;
    r.diffs = f.trimd;
    if (comp) { r.b_abpos = blen - f.trimy; r.b_bbpos = alen - f.trimx; }
    else { r.b_aepos = f.trimy; r.b_bepos = f.trimx; }
    wave_sync();   // the pool is reused by the reverse pass

    PassOut v = wave_pass<NCH, -1>(sp, A, alen, B, blen, comp, mind, mind, anti, minp, maxp, pool, pool_cap,
                                   abort_flag, wc);
    if (v.status != LA_OK) { r.status = v.status; return r; }
    wave_sync();
    len = 0;
    first = 0;
    if (lane < 2) {
        const bool isB = (lane == 1);
        ok = assemble_chain<-1>(pool, v.used, pts + (isB ? pts_cap : 0), pts_cap, isB ? v.thb : v.tha, isB, anti, TS,
                                boff, v.trimx, v.trimy, v.trimd, isB ? btr : atr, maxtp, isB ? fb : fa, &len, &first,
                                &rootk);
    }
    if (__any(!ok)) { r.status = LA_CHAIN; return r; }
    const int ra = rdl(len, 0), rb = rdl(len, 1);
    r.afirst = rdl(first, 0);
    r.bfirst = rdl(first, 1);
    r.atlen = fa + ra;
    r.btlen = fb + rb;
    r.abpos = v.trimx;
    r.bbpos = v.trimy;
    r.diffs += v.trimd;
    if (comp) { r.b_aepos = blen - v.trimy; r.b_bepos = alen - v.trimx; }
    else { r.b_abpos = v.trimy; r.b_bbpos = v.trimx; }
    wave_sync();

    if (comp) {   // B-view trace runs the other way (dalign/align.c:2004-2023)
        uint16_t* t = btr + r.bfirst;
        const int npair = r.btlen / 2;
        for (int p = lane; p < npair / 2; p += 64) {
            const int q = npair - 1 - p;
            const uint16_t a0 = t[2 * p], a1 = t[2 * p + 1];
            const uint16_t b0 = t[2 * q], b1 = t[2 * q + 1];
            t[2 * p] = b0; t[2 * p + 1] = b1;
            t[2 * q] = a0; t[2 * q + 1] = a1;
        }
        wave_sync();
    }
    return r;
}

}  // namespace mi
