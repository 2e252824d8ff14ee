// k-mer match filter kernels for gfx950 (dalign/filter.c restated for HBM):
// packed k-mer keys, their radix sort, the merge-join of equal codes into
// seed pairs, the seed sort and the split of the seeds into read pairs.
//
// Key of an index entry (dalign/filter.c:69-74 KmerPos, restated):
//   key = code << (rbits + pbits) | read << pbits | rpos
// with rbits / pbits the bit widths of the block's read count and longest
// read, so sorting keys orders by (code, read, rpos), the order Sort_Kmers'
// stable byte sort leaves (dalign/filter.c:698-903).  2K + rbits + pbits <= 64.
// Seed record (dalign/filter.c:76-82 SeedPair): diag, apos, aread, bread.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

struct KeyFmt {
    int rbits, pbits;
    __device__ __host__ __forceinline__ int cshift() const { return rbits + pbits; }
    __device__ __forceinline__ uint64_t code(uint64_t k) const { return k >> cshift(); }
    __device__ __forceinline__ int read(uint64_t k) const { return (int)((k >> pbits) & ((1ull << rbits) - 1)); }
    __device__ __forceinline__ int pos(uint64_t k) const { return (int)(k & ((1ull << pbits) - 1)); }
};

struct Seed {
    int diag, apos, aread, bread;
};

struct KeyU64 {
    __device__ __forceinline__ uint64_t operator()(uint64_t k) const { return k; }
};

// (bread, aread, apos) key of a seed, the order of the pair sort
// (dalign/filter.c:2030-2054: stable LSD on apos, aread, bread bytes)
struct KeySeed {
    int sa, sb;   // aread shift (= apos bits), bread shift (= apos + aread bits)
    __device__ __forceinline__ uint64_t operator()(const Seed& s) const
    {
        return ((uint64_t)(uint32_t)s.bread << sb) | ((uint64_t)(uint32_t)s.aread << sa) | (uint32_t)s.apos;
    }
};

// 2-bit packing of a block: read r's bases (values 0..3, one byte each) at
// bases[boff[r]] go to words w0[r] .. ; base i at bits 2*(i&31) of word i>>5
__global__ void k_pack(const uint8_t* __restrict__ bases, const long long* __restrict__ boff,
                       const long long* __restrict__ w0, const int* __restrict__ rlen, int nreads,
                       uint64_t* __restrict__ pk)
{
    const int r = blockIdx.x;
    if (r >= nreads) return;
    const int len = rlen[r];
    const uint8_t* s = bases + boff[r];
    uint64_t* d = pk + w0[r];
    const int nw = (len + 31) >> 5;
    for (int wi = threadIdx.x; wi < nw; wi += blockDim.x) {
        uint64_t v = 0;
        const int j0 = wi * 32;
        const int m = min(32, len - j0);
        for (int b = 0; b < m; b++) v |= (uint64_t)(s[j0 + b] & 3) << (2 * b);
        d[wi] = v;
    }
}

// reverse complement of every read (dalign/daligner.c:456-573): base j of
// the output read = 3 - base (len-1-j) of the input read
__global__ void k_revcomp(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                          const long long* __restrict__ w0, const int* __restrict__ rlen, int nreads)
{
    const int r = blockIdx.x;
    if (r >= nreads) return;
    const int len = rlen[r];
    const uint64_t* s = in + w0[r];
    uint64_t* d = out + w0[r];
    const int nw = (len + 31) >> 5;
    for (int wi = threadIdx.x; wi < nw; wi += blockDim.x) {
        uint64_t v = 0;
        const int j0 = wi * 32;
        const int m = min(32, len - j0);
        for (int b = 0; b < m; b++) {
            const int q = len - 1 - (j0 + b);
            const uint64_t base = (s[q >> 5] >> (2 * (q & 31))) & 3;
            v |= (3 - base) << (2 * b);
        }
        d[wi] = v;
    }
}

// Order of the 2-bit groups of a word reversed (group 0 becomes group 31):
// pairs swapped within nibbles, nibbles within bytes, then the bytes.
__device__ __forceinline__ uint64_t reverse_bases(uint64_t x)
{
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return __builtin_bswap64(x);
}

// last run s in [lo, hi) with kofs[s] <= i (kofs[lo] <= i holds)
__device__ __forceinline__ int run_of(const long long* __restrict__ kofs, int lo, int hi, long long i)
{
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (kofs[m] <= i) lo = m;
        else hi = m;
    }
    return lo;
}

// one thread per k-mer: index -> (read, rpos) -> key.  The k-mers come in
// runs of consecutive positions: kofs[s] = k-mers before run s; run s is
// read seg_read[s] from last-base position seg_pos0[s] on (without a mask,
// seg_read == nullptr: run r is all of read r from K-1 on).  rpos is the
// position of the k-mer's last base (dalign/filter.c:400-490).  Thread 0
// finds the runs of the block's first and last k-mer; the other threads
// search only between them (one run, mostly).
__global__ __launch_bounds__(256) void k_kmers(const uint64_t* __restrict__ pk, const long long* __restrict__ w0,
                                               const long long* __restrict__ kofs, int nrun,
                                               const int* __restrict__ seg_read, const int* __restrict__ seg_pos0,
                                               long long nk, int K, KeyFmt fmt, uint64_t* __restrict__ keys)
{
    __shared__ int span[2];
    const long long i0 = (long long)blockIdx.x * blockDim.x;
    if (threadIdx.x == 0) {
        const long long i1 = min(i0 + (long long)blockDim.x, nk) - 1;
        span[0] = run_of(kofs, 0, nrun, i0);
        span[1] = run_of(kofs, span[0], nrun, i1);
    }
    __syncthreads();
    const long long i = i0 + threadIdx.x;
    if (i >= nk) return;
    const int lo = run_of(kofs, span[0], span[1] + 1, i);
    const int r = seg_read ? seg_read[lo] : lo;
    const int rpos = (int)(i - kofs[lo]) + (seg_pos0 ? seg_pos0[lo] : K - 1);
    const uint64_t* w = pk + w0[r];
    // bases rpos-K+1 .. rpos, first base most significant
    const int b0 = rpos - K + 1;
    const int wi = b0 >> 5, sh = 2 * (b0 & 31);
    // the 2K <= 62 bits of a k-mer (K <= 31) span at most two words
    uint64_t win = w[wi] >> sh;
    if (sh) win |= w[wi + 1] << (64 - sh);
    // base j of the window at bits 2j: reversed, base 0 is the top group
    const uint64_t code = reverse_bases(win) >> (64 - 2 * K);
    keys[i] = (code << fmt.cshift()) | ((uint64_t)r << fmt.pbits) | (uint64_t)rpos;
}

// -b k-mers (biased_tuple_thread, dalign/filter.c:492-633), one lane per
// run of positions [p, q) of a read: a window grows forward until its weight
// (the sum of its bases' ceil(-10000 log4 freq)) reaches 10000 K, sheds its
// oldest bases while it stays there, and yields its code left-aligned to 2K
// bits, at the window's last position, when its weight exceeds 10000 (K-2).
// Unmasked runs are whole reads, and a window that runs into the read's end
// ends it; masked runs stop at q and may still yield a k-mer at position q.
// Pass 1 (EMIT false) counts each run's k-mers, pass 2 writes their keys
// from out[run] on.
struct BiasParams {
    int lb[4];
    int lognorm, logthresh, masked;
};

template <bool EMIT>
__global__ void k_biased(const uint64_t* __restrict__ pk, const long long* __restrict__ w0, const int* __restrict__ rread,
                         const int* __restrict__ rp, const int* __restrict__ rq, int nrun, int K, BiasParams bp,
                         KeyFmt fmt, const long long* __restrict__ out, int* __restrict__ counts,
                         uint64_t* __restrict__ keys)
{
    const uint64_t kmask = (1ull << (2 * K)) - 1;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nrun; s += gridDim.x * blockDim.x) {
        const int r = rread[s], q = rq[s];
        int p = rp[s];
        const uint64_t* w = pk + w0[r];
        auto base = [&](int i) { return (int)((w[i >> 5] >> (2 * (i & 31))) & 3); };
        long long o = EMIT ? out[s] : 0;
        int n = 0, a = 0, k = 1;
        uint64_t c = 0;
        bool end = false;
        while (p < q) {
            int x = base(p);
            a += bp.lb[x];
            c = (c << 2) | (uint64_t)x;
            while (a < bp.lognorm && k < K) {
                if (++p >= q) {
                    end = !bp.masked;
                    break;
                }
                k += 1;
                x = base(p);
                a += bp.lb[x];
                c = (c << 2) | (uint64_t)x;
            }
            if (end) break;
            while (k > 1) {   // (positive weights never shed the last base)
                const int u = a - bp.lb[base(p - k + 1)];
                if (u < bp.lognorm) break;
                a = u;
                k -= 1;
            }
            if (a > bp.logthresh) {
                if (EMIT)
                    keys[o + n] = (((c << (2 * K - 2 * k)) & kmask) << fmt.cshift()) | ((uint64_t)r << fmt.pbits) |
                                  (uint64_t)p;
                n++;
            }
            p += 1;
            a -= bp.lb[base(p - k)];
        }
        if (!EMIT) counts[s] = n;
    }
}

// group starts of a sorted key list (by code)
__global__ void k_group_flags(const uint64_t* __restrict__ keys, long long n, KeyFmt fmt,
                              unsigned long long* __restrict__ flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || fmt.code(keys[i]) != fmt.code(keys[i - 1])) ? 1 : 0;
}

// inclusive count of set flags minus one: the group of every entry; and
// the start of every group
__global__ void k_group_index(const unsigned long long* __restrict__ flag, const unsigned long long* __restrict__ excl,
                              long long n, long long* __restrict__ gidx, long long* __restrict__ gstart)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long g = (long long)(excl[i] + flag[i]) - 1;
    if (gidx) gidx[i] = g;
    if (flag[i]) gstart[g] = i;
}

__global__ void k_set_ll(long long* __restrict__ p, long long v) { *p = v; }

// -t suppression (dalign/filter.c:645-696): entries of groups of size >= t are dropped
__global__ void k_keep_small(const long long* __restrict__ gidx, const long long* __restrict__ gstart, long long n,
                             long long t, unsigned long long* __restrict__ keep)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long g = gidx[i];
    keep[i] = (gstart[g + 1] - gstart[g] < t) ? 1 : 0;
}

__global__ void k_compact_u64(const uint64_t* __restrict__ in, const unsigned long long* __restrict__ keep,
                              const unsigned long long* __restrict__ pos, long long n, uint64_t* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (keep[i]) out[pos[i]] = in[i];
}

__device__ __forceinline__ long long lower_code(const uint64_t* keys, long long n, uint64_t code, const KeyFmt& f)
{
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long m = (lo + hi) >> 1;
        if (f.code(keys[m]) < code) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// per A group: the B range with the same code (find_tuple, dalign/filter.c:911-930),
// for B another index than A (with B = A the range is the group itself and
// needs no kernel).  The groups of a 256-thread block have ascending codes, so their starts lie in
// the B range [first code's start, start of the code after the last one),
// which k_brange_bounds finds for every block (one thread per block, all
// blocks' searches in flight at once); each group then bisects only inside
// its block's range and gallops to its end from its start (most codes have a
// handful of entries, so the end is in the start's cache line).
__global__ void k_brange_bounds(const uint64_t* __restrict__ ak, KeyFmt fa, const long long* __restrict__ gstart,
                                long long ng, const uint64_t* __restrict__ bk, KeyFmt fb, long long nb, int nblock,
                                long long* __restrict__ bound)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblock) return;
    const long long g0 = (long long)t * 256, g1 = min(g0 + 256, ng) - 1;
    const long long lo = lower_code(bk, nb, fa.code(ak[gstart[g0]]), fb);
    bound[2 * t] = lo;
    bound[2 * t + 1] = lo + lower_code(bk + lo, nb - lo, fa.code(ak[gstart[g1]]) + 1, fb);
}

__global__ __launch_bounds__(256) void k_group_brange(const uint64_t* __restrict__ ak, KeyFmt fa,
                                                      const long long* __restrict__ gstart, long long ng,
                                                      const uint64_t* __restrict__ bk, KeyFmt fb, long long nb,
                                                      const long long* __restrict__ bound,
                                                      long long* __restrict__ bs, long long* __restrict__ be)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const uint64_t code = fa.code(ak[gstart[g]]);
    const long long b0 = bound[2 * blockIdx.x];
    const long long s = b0 + lower_code(bk + b0, bound[2 * blockIdx.x + 1] - b0, code, fb);
    // first entry past the code: gallop to a bracket [lo, hi), then bisect
    long long lo = s, step = 1, hi = s;
    while (hi < nb && fb.code(bk[hi]) == code) {
        lo = hi + 1;
        hi = (step < nb - hi) ? hi + step : nb;
        step <<= 1;
    }
    while (lo < hi) {
        const long long m = (lo + hi) >> 1;
        if (fb.code(bk[m]) == code) lo = m + 1;
        else hi = m;
    }
    bs[g] = s;
    be[g] = lo;
}

// per A entry: number of B entries it pairs with (count_thread,
// dalign/filter.c:948-1056): in self mode only B reads below the A read,
// otherwise the whole B group
__global__ void k_counts(const uint64_t* __restrict__ ak, KeyFmt fa, long long na, const long long* __restrict__ gidx,
                         const long long* __restrict__ bs, const long long* __restrict__ be,
                         const uint64_t* __restrict__ bk, KeyFmt fb, int self, unsigned long long* __restrict__ cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    const long long g = gidx[i];
    long long s = bs[g], e = be[g];
    // self: B entries of reads below A's (1); with -I also A's own read, at
    // positions below A's (2) or, against the complement, wholly (3)
    // (dalign/filter.c:984-1001)
    if (self) {
        const uint64_t ka = ak[i];
        const int ar = fa.read(ka), ap = fa.pos(ka);
        long long lo = s, hi = e;
        while (lo < hi) {
            const long long m = (lo + hi) >> 1;
            const int br = fb.read(bk[m]);
            const bool below = self == 3 ? br <= ar : (br < ar || (self == 2 && br == ar && fb.pos(bk[m]) < ap));
            if (below) lo = m + 1;
            else hi = m;
        }
        e = lo;
    }
    cnt[i] = (unsigned long long)(e - s);
}

// group totals (from the exclusive scan of counts) and their histogram
// (hitgram).  Most groups have tiny totals, so the low bins are counted in
// LDS per workgroup and flushed once; only totals >= kLdsBins go straight
// to the global histogram.
constexpr int kLdsBins = 1024;

__global__ __launch_bounds__(256) void k_group_totals(const long long* __restrict__ gstart, long long ng,
                                                      const unsigned long long* __restrict__ cscan,
                                                      unsigned long long* __restrict__ gtot,
                                                      unsigned long long* __restrict__ hist, int maxgram)
{
    __shared__ unsigned h[kLdsBins];
    for (int i = threadIdx.x; i < kLdsBins; i += 256) h[i] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * 256 * 16;
    for (int k = 0; k < 16; k++) {
        const long long g = base + k * 256 + threadIdx.x;
        if (g < ng) {
            const unsigned long long t = cscan[gstart[g + 1]] - cscan[gstart[g]];
            gtot[g] = t;
            if (t < (unsigned long long)kLdsBins && t < (unsigned long long)maxgram) atomicAdd(&h[t], 1u);
            else if (t < (unsigned long long)maxgram) atomicAdd(&hist[t], 1ull);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsBins && i < maxgram; i += 256)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// groups whose total reaches the cap are dropped (merge_thread, dalign/filter.c:1148-1154)
__global__ void k_apply_limit(const long long* __restrict__ gidx, const unsigned long long* __restrict__ gtot,
                              long long na, unsigned long long limit, unsigned long long* __restrict__ cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    if (gtot[gidx[i]] >= limit) cnt[i] = 0;
}

// seed emission (merge_thread, dalign/filter.c:1061-1231): seeds in merge
// order, for every A entry its B partners in index order.  A wave takes 64
// consecutive A entries; each lane loads one entry (its seed range in the
// exclusive scan `off`, na + 1 entries, its A position and read, the start
// of its B partners) into the wave's LDS slots, then the wave writes the
// entries' seeds in order, consecutive lanes to consecutive slots: lane q
// finds the entry of seed q by bisecting the 64 ranges in LDS.
__global__ __launch_bounds__(256) void k_emit(const uint64_t* __restrict__ ak, KeyFmt fa, long long na,
                                              const long long* __restrict__ gidx, const long long* __restrict__ bs,
                                              const uint64_t* __restrict__ bk, KeyFmt fb,
                                              const unsigned long long* __restrict__ off, Seed* __restrict__ out)
{
    __shared__ unsigned long long so[4][65];   // seed range starts of the wave's entries (and the end)
    __shared__ long long sb[4][64];            // first B partner
    __shared__ int sap[4][64], sar[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long e0 = ((long long)blockIdx.x * 4 + w) * 64;
    if (e0 >= na) return;   // wave-uniform
    const long long e = e0 + lane;
    const int n = (int)min(64ll, na - e0);
    if (lane < n) {
        const uint64_t a = ak[e];
        const unsigned long long o = off[e], o1 = off[e + 1];
        so[w][lane] = o;
        sb[w][lane] = (o1 > o) ? bs[gidx[e]] : 0;
        sap[w][lane] = fa.pos(a);
        sar[w][lane] = fa.read(a);
        if (lane == n - 1) so[w][n] = o1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const unsigned long long first = so[w][0], total = so[w][n] - first;
    for (unsigned long long q = lane; q < total; q += 64) {
        const unsigned long long g = first + q;
        int lo = 0, hi = n;   // last entry with so <= g
        while (hi - lo > 1) {
            const int m = (lo + hi) >> 1;
            if (so[w][m] <= g) lo = m;
            else hi = m;
        }
        const uint64_t b = bk[sb[w][lo] + (long long)(g - so[w][lo])];
        const int ap = sap[w][lo];
        out[g] = Seed{ap - fb.pos(b), ap, sar[w][lo], fb.read(b)};
    }
}

// read-pair boundaries of the sorted seeds
__global__ void k_pair_flags(const Seed* __restrict__ h, long long n, unsigned long long* __restrict__ flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || h[i].aread != h[i - 1].aread || h[i].bread != h[i - 1].bread) ? 1 : 0;
}

// report_thread chunk starts (dalign/filter.c:2282-2294): p = (n*i) >> nshift
// moved past the run of seeds with the B read of seed p-1
__global__ void k_chunk_bounds(const Seed* __restrict__ h, long long n, int nthreads, int nshift,
                               long long* __restrict__ cb)
{
    const int i = threadIdx.x;
    if (i > nthreads) return;
    if (i == 0) { cb[0] = 0; return; }
    if (i == nthreads) { cb[i] = n; return; }
    long long p = (long long)(((unsigned long long)n * (unsigned long long)i) >> nshift);
    if (p > 0 && p < n) {
        const int d = h[p - 1].bread;
        long long lo = p, hi = n;   // first index >= p whose B read differs (seeds are sorted by B read)
        while (lo < hi) {
            const long long m = (lo + hi) >> 1;
            if (h[m].bread == d) lo = m + 1;
            else hi = m;
        }
        p = lo;
    }
    cb[i] = p;
}

// pairs report_thread aligns (dalign/filter.c:1699-1731): at least minhit
// seeds, starting before the chunk's end - minhit, and not both reads
// shorter than the -H cutoff
__global__ void k_select(const long long* __restrict__ pstart, long long np, const Seed* __restrict__ h,
                         const int* __restrict__ lenA, const int* __restrict__ lenB, const long long* __restrict__ cb,
                         int nthreads, int minhit, int hgap, unsigned long long* __restrict__ sel)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const long long s = pstart[p], e = pstart[p + 1];
    int c = 0;   // chunk holding s: last c with cb[c] <= s
    while (c + 1 < nthreads && cb[c + 1] <= s) c++;
    bool ok = (e - s >= minhit) && (s < cb[c + 1] - minhit);
    if (ok) {
        const Seed x = h[s];
        if (lenA[x.aread] < hgap && lenB[x.bread] < hgap) ok = false;
    }
    sel[p] = ok ? 1 : 0;
}

__global__ void k_compact_idx(const unsigned long long* __restrict__ sel, const unsigned long long* __restrict__ pos,
                              long long n, int* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (sel[i]) out[pos[i]] = (int)i;
}

}  // namespace mi

namespace mi {

// Exact pre-check of a read pair against report_thread's gate
// (dalign/filter.c:1740-1778): a pair can only align a seed if, in some
// panel with at least minhit seeds, a seed's band score plus a neighbour
// band's score reaches hitmin.  Scores are computed as the report kernel
// computes them, with the band arrays' edge cells (score[maxdiag+1] is
// lastp[mindiag], score[mindiag-1] is the zero pad).  lasta only makes the
// gate harder and is 0 when a pair starts, and a non-zero carried
// lastp[maxdiag] only lowers scores, so a pair that fails here aligns
// nothing.  Pairs with more than kPreMax seeds are always kept.
// Per job: keep (1: run the report kernel), touch (1: some scored panel has
// a seed in band maxdiag, i.e. the pair reads and clears the shared cell).
constexpr int kPreMax = 32;

__global__ void k_prefilter(const int* __restrict__ jobs, int njobs, const long long* __restrict__ pstart,
                            const Seed* __restrict__ h, int K, int binshift, int hitmin, int minhit, int mindiag,
                            int maxdiag, unsigned long long* __restrict__ keep, int* __restrict__ touch)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const int p = jobs[j];
    const long long s = pstart[p], e = pstart[p + 1];
    if (e - s > kPreMax) {
        keep[j] = 1;
        touch[j] = 0;
        return;
    }
    int band[kPreMax], contrib[kPreMax];
    bool pass = false, tch = false;
    long long nidx = s;
    int amark2 = 0;
    while (nidx < e && !pass) {
        const int amark = amark2 + 50000;
        amark2 = amark - 10000;
        const long long lidx = nidx;
        long long h2 = lidx;
        int apos;
        do {
            apos = h[nidx].apos;
            nidx++;
            if (apos <= amark2) h2 = nidx;
        } while (nidx < e && apos <= amark);
        const int n = (int)(nidx - lidx);
        if (n >= minhit) {
            int lastp_min = 0;
            for (int f = 0; f < n; f++) {
                const Seed x = h[lidx + f];
                band[f] = x.diag >> binshift;
                int prev = 0;
                for (int g = f - 1; g >= 0; g--)
                    if (band[g] == band[f]) {
                        prev = h[lidx + g].apos;
                        break;
                    }
                const int d = x.apos - prev;
                contrib[f] = d >= K ? K : d;
                if (band[f] == mindiag) lastp_min = x.apos;
                if (band[f] == maxdiag) tch = true;
            }
            auto S = [&](int b) -> int {
                if (b == maxdiag + 1) return lastp_min;
                if (b == mindiag - 1) return 0;
                int t = 0;
                for (int g = 0; g < n; g++)
                    if (band[g] == b) t += contrib[g];
                return t;
            };
            for (int f = 0; f < n && !pass; f++) {
                const int sb = S(band[f]);
                if (sb + S(band[f] + 1) >= hitmin || sb + S(band[f] - 1) >= hitmin) pass = true;
            }
        }
        nidx = h2;
    }
    keep[j] = pass ? 1 : 0;
    touch[j] = tch ? 1 : 0;
}

// ---- candidate scorer features (no counterpart in the reference) ------------
// A candidate is a read pair that passed the gate pre-check (a job of the
// report kernel).  Its seeds are sorted by apos; one thread streams them (at
// most kFeatSeeds of them, evenly spaced, for pairs with more) into
// a table of up to kBinSlots diagonal bins (band = diag >> binshift, the
// report_thread bands), each with its seed count, first and last apos, the
// sum and sum of squares of the apos spacing of consecutive seeds, and its
// seeds whose k-mer is a short-period repeat (the 14 bases ending at apos in
// A repeat with a period of 1..6).  A bin not in a full table replaces the
// one with the fewest seeds.  The 16 features describe the strongest bin
// (most seeds), the runner-up and the pair:
//   0 log2(1 + seeds) / 16            8 log2(1 + alen) / 16
//   1 best bin / seeds examined       9 log2(1 + blen) / 16
//   2 log2(1 + best bin seeds) / 16  10 log2(1 + overlap the best bin's diagonal implies) / 16
//   3 log2(1 + runner-up seeds) / 16 11 best bin apos span / that overlap (capped at 2)
//   4 best bin apos span / alen      12 COMP
//   5 log2(1 + best bin span) / 16   13 short-period k-mers in the best bin / its seeds
//   6 log2(1 + mean spacing) / 16    14 bins with >= minhit seeds / 8
//   7 log2(1 + spacing sd) / 16      15 short-period k-mers / seeds
constexpr int kBinSlots = 8;
constexpr int kJobFeat = 16;
constexpr int kFeatSeeds = 256;   // seeds examined per pair at most (evenly spaced)

__device__ __forceinline__ bool short_period_kmer(const uint64_t* A, int apos, int K)
{
    if (apos < K - 1) return false;
    // 16 bases ending at apos (base i at bits 2i), 2-bit packed A
    const uint32_t* a = reinterpret_cast<const uint32_t*>(A);
    const int p0 = apos - 15;
    const int q = p0 >> 4;   // p0 >= -15: word q >= -1 is the guard word before the read
    const uint32_t w = __builtin_amdgcn_alignbit(a[q + 1], a[q], (uint32_t)(2 * (p0 & 15)));
    const int kk = K < 16 ? K : 16;
    const uint32_t km = (kk == 16) ? 0xffffffffu : (w >> (2 * (16 - kk)));   // the k-mer's bases, first at bit 0
    for (int per = 1; per <= 6 && per < kk; per++) {
        const int nb = kk - per;
        const uint32_t mask = (nb >= 16) ? 0xffffffffu : ((1u << (2 * nb)) - 1);
        if (((km >> (2 * per)) & mask) == (km & mask)) return true;
    }
    return false;
}

__global__ __launch_bounds__(128) void k_job_features(const int* __restrict__ jobs, int njobs,
                                                      const unsigned long long* __restrict__ keep,
                                                      const long long* __restrict__ pstart, const Seed* __restrict__ h,
                                                      const uint64_t* __restrict__ pkA, const long long* __restrict__ w0A,
                                                      const int* __restrict__ lenA, const int* __restrict__ lenB,
                                                      int K, int binshift, int minhit, int comp,
                                                      float* __restrict__ feat)
{
    __shared__ int tb[128][kBinSlots][5];   // bin, seeds, first apos, last apos, short-period seeds
    __shared__ float tf[128][kBinSlots][2];  // spacing sum, spacing square sum
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    float* out = feat + (long long)j * kJobFeat;
    if (!keep[j]) {
        for (int i = 0; i < kJobFeat; i++) out[i] = 0.f;
        return;
    }
    int (*T)[5] = tb[threadIdx.x];
    float (*F)[2] = tf[threadIdx.x];
    int used = 0;
    const int p = jobs[j];
    const long long s = pstart[p], e = pstart[p + 1];
    const int ar = h[s].aread, br = h[s].bread;
    const int alen = lenA[ar], blen = lenB[br];
    const uint64_t* A = pkA + w0A[ar];
    int lc_all = 0;
    // pairs with many seeds (repeats) are sampled evenly: the pass stays short
    const long long step = (e - s > kFeatSeeds) ? (e - s + kFeatSeeds - 1) / kFeatSeeds : 1;
    int nsamp = 0;
    for (long long i = s; i < e; i += step) {
        nsamp++;
        const Seed x = h[i];
        const int bin = x.diag >> binshift;
        const int lc = short_period_kmer(A, x.apos, K) ? 1 : 0;
        lc_all += lc;
        int slot = -1, minc = 0x7fffffff, mins = 0;
        for (int q = 0; q < used; q++) {
            if (T[q][0] == bin) slot = q;
            if (T[q][1] < minc) { minc = T[q][1]; mins = q; }
        }
        if (slot >= 0) {
            const float d = (float)(x.apos - T[slot][3]);
            T[slot][1] += 1;
            T[slot][3] = x.apos;
            T[slot][4] += lc;
            F[slot][0] += d;
            F[slot][1] += d * d;
        } else {
            const int q = used < kBinSlots ? used++ : mins;
            T[q][0] = bin; T[q][1] = 1; T[q][2] = x.apos; T[q][3] = x.apos; T[q][4] = lc;
            F[q][0] = 0.f; F[q][1] = 0.f;
        }
    }
    int b1 = 0, b2 = -1, strong = 0;
    for (int q = 0; q < used; q++) {
        if (T[q][1] >= minhit) strong++;
        if (q == 0) continue;
        if (T[q][1] > T[b1][1]) { b2 = b1; b1 = q; }
        else if (b2 < 0 || T[q][1] > T[b2][1]) b2 = q;
    }
    const float n = (float)nsamp;
    const float c = (float)T[b1][1];
    const float span = (float)(T[b1][3] - T[b1][2]);
    const float mean = c > 1.f ? F[b1][0] / (c - 1.f) : 0.f;
    const float var = c > 2.f ? fmaxf(F[b1][1] / (c - 1.f) - mean * mean, 0.f) : 0.f;
    const int dg = T[b1][0] * (1 << binshift) + (1 << binshift) / 2;
    const int ov = max(min(alen, blen + dg) - max(0, dg), 0);
    auto lg = [](float v) { return __log2f(1.f + v) * (1.f / 16.f); };
    out[0] = lg((float)(e - s));
    out[1] = c / n;
    out[2] = lg(c);
    out[3] = lg(b2 >= 0 ? (float)T[b2][1] : 0.f);
    out[4] = span / (float)max(alen, 1);
    out[5] = lg(span);
    out[6] = lg(mean);
    out[7] = lg(sqrtf(var));
    out[8] = lg((float)alen);
    out[9] = lg((float)blen);
    out[10] = lg((float)ov);
    out[11] = fminf(span / (float)max(ov, 1), 2.f);
    out[12] = (float)comp;
    out[13] = (float)T[b1][4] / c;
    out[14] = (float)min(strong, 8) * 0.125f;
    out[15] = (float)lc_all / n;
}

// gating: jobs whose logit is below the threshold are not aligned;
// counts[0] += jobs scored, counts[1] += jobs dropped
__global__ void k_gate_jobs(unsigned long long* __restrict__ keep, const float* __restrict__ logit, int njobs,
                            float thr, unsigned long long* __restrict__ counts)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs || !keep[j]) return;
    atomicAdd(&counts[0], 1ull);
    if (logit[j] < thr) {
        keep[j] = 0;
        atomicAdd(&counts[1], 1ull);
    }
}

}  // namespace mi
