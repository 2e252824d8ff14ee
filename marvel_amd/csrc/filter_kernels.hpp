// k-mer match filter kernels for gfx950 (dalign/filter.c restated for HBM):
// packed k-mer keys, their radix sort, the merge-join of equal codes into
// seed pairs, the seed sort and the split of the seeds into read pairs.
//
// Key of an index entry (dalign/filter.c:69-74 KmerPos, restated):
//   key = code << 36 | read << 18 | rpos       (code = 2K bits, K <= 14)
// so sorting keys orders by (code, read, rpos), the order Sort_Kmers'
// stable byte sort leaves (dalign/filter.c:698-903).
// Seed record (dalign/filter.c:76-82 SeedPair): diag, apos, aread, bread.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

constexpr int kReadBits = 18;
constexpr int kPosBits = 18;
constexpr int kCodeShift = kReadBits + kPosBits;   // 36

struct Seed {
    int diag, apos, aread, bread;
};

struct KeyU64 {
    __device__ __forceinline__ uint64_t operator()(uint64_t k) const { return k; }
};

// (bread, aread, apos) key of a seed, the order of the pair sort
// (dalign/filter.c:2030-2054: LSD on apos, aread, bread bytes).
struct KeySeed {
    __device__ __forceinline__ uint64_t operator()(const Seed& s) const
    {
        return ((uint64_t)(uint32_t)s.bread << 36) | ((uint64_t)(uint32_t)s.aread << 18) | (uint32_t)s.apos;
    }
};

// one thread per k-mer: idx -> (read, rpos) -> key
__global__ void k_kmers(const uint64_t* __restrict__ pk, const long long* __restrict__ word0,
                        const long long* __restrict__ kofs, int nreads, long long nk, int K,
                        uint64_t* __restrict__ keys)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    int lo = 0, hi = nreads;   // last read r with kofs[r] <= i
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (kofs[m] <= i) lo = m;
        else hi = m;
    }
    const int r = lo;
    const int rpos = (int)(i - kofs[r]) + K - 1;   // position of the k-mer's last base
    const uint64_t* w = pk + word0[r];
    uint64_t code = 0;
    for (int j = rpos - K + 1; j <= rpos; j++) code = (code << 2) | ((w[j >> 5] >> (2 * (j & 31))) & 3);
    keys[i] = (code << kCodeShift) | ((uint64_t)r << kPosBits) | (uint64_t)rpos;
}

// reverse complement of every read (dalign/daligner.c:456-573): base j of
// the output read = 3 - base (len-1-j) of the input read
__global__ void k_revcomp(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                          const long long* __restrict__ word0, const int* __restrict__ rlen, int nreads)
{
    const int r = blockIdx.x;
    if (r >= nreads) return;
    const int len = rlen[r];
    const uint64_t* s = in + word0[r];
    uint64_t* d = out + word0[r];
    const int nw = (len + 31) >> 5;
    for (int wi = threadIdx.x; wi < nw; wi += blockDim.x) {
        uint64_t v = 0;
        for (int b = 0; b < 32; b++) {
            const int j = wi * 32 + b;
            if (j >= len) break;
            const int q = len - 1 - j;
            const uint64_t base = (s[q >> 5] >> (2 * (q & 31))) & 3;
            v |= (3 - base) << (2 * b);
        }
        d[wi] = v;
    }
}

// group starts of a sorted key list (by code)
__global__ void k_group_flags(const uint64_t* __restrict__ keys, long long n, unsigned long long* __restrict__ flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || (keys[i] >> kCodeShift) != (keys[i - 1] >> kCodeShift)) ? 1 : 0;
}

// group index of every entry: inclusive count of group starts, minus one
__global__ void k_group_index(const unsigned long long* __restrict__ flag, const unsigned long long* __restrict__ excl,
                              long long n, long long* __restrict__ gidx)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gidx[i] = (long long)(excl[i] + flag[i]) - 1;
}

__global__ void k_group_starts(const unsigned long long* __restrict__ flag, const long long* __restrict__ gidx,
                               long long n, long long* __restrict__ gstart)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) gstart[gidx[i]] = i;
}

__device__ __forceinline__ long long lower_code(const uint64_t* keys, long long n, uint64_t code)
{
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long m = (lo + hi) >> 1;
        if ((keys[m] >> kCodeShift) < code) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// per A group: the matching B range
__global__ void k_group_brange(const uint64_t* __restrict__ ak, const long long* __restrict__ gstart, long long ng,
                               const uint64_t* __restrict__ bk, long long nb, long long* __restrict__ bs,
                               long long* __restrict__ be)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const uint64_t code = ak[gstart[g]] >> kCodeShift;
    bs[g] = lower_code(bk, nb, code);
    be[g] = lower_code(bk, nb, code + 1);
}

// per A entry: number of B entries it pairs with (dalign/filter.c:948-1056):
// self mode pairs only with B reads below the A read, otherwise the whole group
__global__ void k_counts(const uint64_t* __restrict__ ak, long long na, const long long* __restrict__ gidx,
                         const long long* __restrict__ bs, const long long* __restrict__ be,
                         const uint64_t* __restrict__ bk, int self, unsigned long long* __restrict__ cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    const long long g = gidx[i];
    long long s = bs[g], e = be[g];
    if (self) {
        const uint64_t ar = (ak[i] >> kPosBits) & ((1ull << kReadBits) - 1);
        long long lo = s, hi = e;
        while (lo < hi) {
            const long long m = (lo + hi) >> 1;
            if (((bk[m] >> kPosBits) & ((1ull << kReadBits) - 1)) < ar) lo = m + 1;
            else hi = m;
        }
        e = lo;
    }
    cnt[i] = (unsigned long long)(e - s);
}

// group totals (from the exclusive scan of counts) and their histogram
__global__ void k_group_totals(const long long* __restrict__ gstart, long long ng, long long na,
                               const unsigned long long* __restrict__ cscan, const unsigned long long* __restrict__ cnt,
                               unsigned long long* __restrict__ gtot, unsigned long long* __restrict__ hist, int maxgram)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const long long s = gstart[g], e = (g + 1 < ng) ? gstart[g + 1] : na;
    const unsigned long long t = cscan[e - 1] + cnt[e - 1] - cscan[s];
    gtot[g] = t;
    if (t < (unsigned long long)maxgram) atomicAdd(&hist[t], 1ull);
}

// counts of entries whose group reaches the cap are dropped (dalign/filter.c:1148-1154)
__global__ void k_apply_limit(const long long* __restrict__ gidx, const unsigned long long* __restrict__ gtot,
                              long long na, unsigned long long limit, unsigned long long* __restrict__ cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    if (gtot[gidx[i]] >= limit) cnt[i] = 0;
}

// seeds in merge order (dalign/filter.c:1210-1222)
__global__ void k_emit(const uint64_t* __restrict__ ak, long long na, const long long* __restrict__ gidx,
                       const long long* __restrict__ bs, const uint64_t* __restrict__ bk,
                       const unsigned long long* __restrict__ cnt, const unsigned long long* __restrict__ off,
                       Seed* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    const unsigned long long c = cnt[i];
    if (c == 0) return;
    const uint64_t a = ak[i];
    const int ar = (int)((a >> kPosBits) & ((1ull << kReadBits) - 1));
    const int ap = (int)(a & ((1ull << kPosBits) - 1));
    const long long s = bs[gidx[i]];
    Seed* o = out + off[i];
    for (unsigned long long j = 0; j < c; j++) {
        const uint64_t b = bk[s + j];
        const int br = (int)((b >> kPosBits) & ((1ull << kReadBits) - 1));
        const int bp = (int)(b & ((1ull << kPosBits) - 1));
        o[j] = Seed{ap - bp, ap, ar, br};
    }
}

// read-pair boundaries of the sorted seeds
__global__ void k_pair_flags(const Seed* __restrict__ h, long long n, unsigned long long* __restrict__ flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || h[i].aread != h[i - 1].aread || h[i].bread != h[i - 1].bread) ? 1 : 0;
}

__global__ void k_pair_starts(const unsigned long long* __restrict__ flag, const long long* __restrict__ pidx,
                              long long n, long long* __restrict__ pstart)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) pstart[pidx[i]] = i;
}

}  // namespace mi
