// DBdust on the GPU: the SDUST low-complexity scan of db/DBdust.c:236-470
// (Morgulis et al., JCB 13(5) 2006) over every read of a block, one lane per
// read.  The scan is sequential along a read, so the parallelism is across
// reads; each lane keeps its triplet counts and (for windows up to 64) its
// candidate list in LDS ([item][lane], a bank per lane), its interval stack
// in global memory.
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

}  // namespace mi

using namespace mi;

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
