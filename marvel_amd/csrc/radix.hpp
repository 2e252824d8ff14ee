// Stable LSD radix sort and exclusive scan for gfx950 (used by the k-mer
// index and the seed sort of the daligner filter).
//
// One pass sorts by an 8-bit digit of a 64-bit key:
//   1. rs_hist    : one 256-thread workgroup per tile of kTile records counts
//                   the digits of its tile in LDS (LDS atomics) and writes
//                   counts[digit * ntiles + tile];
//   2. rs_digit_totals + rs_digit_offsets : each (digit, tile) gets its
//                   output base: the totals of the smaller digits plus the
//                   digit's counts in the earlier tiles;
//   3. rs_scatter : one workgroup per tile ranks the tile stably (ballot
//                   peer groups + per-wave digit counts), stages it in LDS
//                   in digit order and writes each digit's run contiguously.
// The key of a record is produced by a functor, so 8-byte keys and 16-byte
// seed records share the code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

constexpr int kTile = 4096;   // records per tile (64 lanes x 64 rounds)

template <class Rec, class KeyOf>
__global__ __launch_bounds__(256) void rs_hist(const Rec* __restrict__ in, long long n, int shift, KeyOf key,
                                               unsigned* __restrict__ counts, int ntiles)
{
    __shared__ unsigned h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kTile;
    for (int i = tid; i < kTile; i += 256) {
        const long long g = base + i;
        if (g < n) atomicAdd(&h[(key(in[g]) >> shift) & 255], 1u);
    }
    __syncthreads();
    counts[(size_t)tid * ntiles + blockIdx.x] = h[tid];   // digit-major: scanning gives each (digit, tile) its base
}

constexpr int kTileWaves = 4;                        // waves ranking one tile
constexpr int kTileRounds = kTile / 64 / kTileWaves;   // 64-record rounds per wave

// One 256-thread workgroup per tile.  Wave w ranks records
// [w*1024, (w+1)*1024) of the tile in order (8 ballots find a lane's peers
// with the same digit, mbcnt gives its rank among them, a per-wave LDS count
// per digit carries ranks across rounds); the per-wave counts then give every
// record its position in the tile sorted by digit, the tile is staged in LDS
// in that order and written out with consecutive threads storing consecutive
// records of each digit's run.  Ranks follow input order, so the sort is
// stable.
template <class Rec, class KeyOf>
__global__ __launch_bounds__(256) void rs_scatter(const Rec* __restrict__ in, Rec* __restrict__ out, long long n,
                                                  int shift, KeyOf key, const unsigned long long* __restrict__ offs,
                                                  int ntiles)
{
    __shared__ unsigned wcnt[kTileWaves][256];   // per-wave digit counts, then per-wave digit offsets
    __shared__ unsigned tstart[256];             // first tile position of each digit
    __shared__ unsigned long long gbase[256];    // output position of the tile's first record of each digit
    __shared__ Rec buf[kTile];                   // the tile in digit order
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long long base = (long long)blockIdx.x * kTile;
#pragma unroll
    for (int q = 0; q < kTileWaves; q++) wcnt[q][tid] = 0;
    gbase[tid] = offs[(size_t)tid * ntiles + blockIdx.x];
    __syncthreads();

    Rec v[kTileRounds];
    unsigned rk[kTileRounds];
#pragma unroll
    for (int r = 0; r < kTileRounds; r++) {
        const long long g = base + (long long)(w * kTileRounds + r) * 64 + lane;
        if (g < n) v[r] = in[g];
    }
#pragma unroll
    for (int r = 0; r < kTileRounds; r++) {
        const long long g = base + (long long)(w * kTileRounds + r) * 64 + lane;
        const bool valid = g < n;
        const unsigned dg = valid ? (unsigned)(key(v[r]) >> shift) & 255u : 0u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const unsigned long long m = __ballot((dg >> b) & 1);
            peers &= ((dg >> b) & 1) ? m : ~m;
        }
        const unsigned rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0));
        const unsigned prev = valid ? wcnt[w][dg] : 0u;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        __builtin_amdgcn_wave_barrier();   // every lane has read the count before the group's last lane advances it
        if (valid && lane == 63 - __clzll(peers)) wcnt[w][dg] = prev + (unsigned)__popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        __builtin_amdgcn_wave_barrier();
        rk[r] = prev + rank;
    }
    __syncthreads();
    // thread tid owns digit tid: per-wave exclusive offsets and the digit's tile total
    unsigned tot = 0;
#pragma unroll
    for (int q = 0; q < kTileWaves; q++) {
        const unsigned c = wcnt[q][tid];
        wcnt[q][tid] = tot;
        tot += c;
    }
    // exclusive scan of the 256 digit totals
    tstart[tid] = tot;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const unsigned a = tid >= o ? tstart[tid - o] : 0u;
        __syncthreads();
        tstart[tid] += a;
        __syncthreads();
    }
    const unsigned incl = tstart[tid];
    __syncthreads();
    tstart[tid] = incl - tot;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kTileRounds; r++) {
        const long long g = base + (long long)(w * kTileRounds + r) * 64 + lane;
        if (g < n) {
            const unsigned dg = (unsigned)(key(v[r]) >> shift) & 255u;
            buf[tstart[dg] + wcnt[w][dg] + rk[r]] = v[r];
        }
    }
    __syncthreads();
    const int cnt = (int)(n - base < kTile ? n - base : kTile);
    for (int i = tid; i < cnt; i += 256) {
        const Rec x = buf[i];
        const unsigned dg = (unsigned)(key(x) >> shift) & 255u;
        out[gbase[dg] + (unsigned long long)(i - (int)tstart[dg])] = x;
    }
}

// ---- output bases of a pass: offs[d][t] = Σ_{d' < d} total(d') + Σ_{t' < t} counts[d][t'] ----

// total count of each digit (one workgroup per digit)
__global__ __launch_bounds__(256) void rs_digit_totals(const unsigned* __restrict__ counts, int ntiles,
                                                       unsigned long long* __restrict__ tot)
{
    __shared__ unsigned long long red[256];
    const int tid = threadIdx.x;
    const unsigned* row = counts + (size_t)blockIdx.x * ntiles;
    unsigned long long s = 0;
    for (int t = tid; t < ntiles; t += 256) s += row[t];
    red[tid] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) tot[blockIdx.x] = red[0];
}

// one workgroup per digit: its base from the totals of the smaller digits,
// then an exclusive scan of its tile counts in chunks of 4,096
__global__ __launch_bounds__(256) void rs_digit_offsets(const unsigned* __restrict__ counts, int ntiles,
                                                        const unsigned long long* __restrict__ tot,
                                                        unsigned long long* __restrict__ offs)
{
    __shared__ unsigned long long red[256];
    __shared__ unsigned sc[256];
    const int tid = threadIdx.x, d = blockIdx.x;
    red[tid] = tid < d ? tot[tid] : 0ull;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    unsigned long long carry = red[0];
    const unsigned* row = counts + (size_t)d * ntiles;
    unsigned long long* orow = offs + (size_t)d * ntiles;
    for (int c0 = 0; c0 < ntiles; c0 += 256 * 16) {
        unsigned v[16], loc = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int t = c0 + tid * 16 + i;
            v[i] = t < ntiles ? row[t] : 0u;
            loc += v[i];
        }
        sc[tid] = loc;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const unsigned a = tid >= o ? sc[tid - o] : 0u;
            __syncthreads();
            sc[tid] += a;
            __syncthreads();
        }
        unsigned long long run = carry + (sc[tid] - loc);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int t = c0 + tid * 16 + i;
            if (t < ntiles) orow[t] = run;
            run += v[i];
        }
        carry += sc[255];
        __syncthreads();
    }
}

// ---- exclusive scan (device, multi-level) -----------------------------------

template <class T>
__global__ __launch_bounds__(256) void scan_block(const T* __restrict__ in, T* __restrict__ out, long long n,
                                                  T* __restrict__ sums)
{
    // 256 threads x 4 items = 1024 items per block
    __shared__ T s[256];
    const int tid = threadIdx.x;
    const long long b0 = (long long)blockIdx.x * 1024;
    T v[4];
    T loc = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const long long g = b0 + tid * 4 + i;
        v[i] = (g < n) ? in[g] : (T)0;
        loc += v[i];
    }
    s[tid] = loc;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        T add = (tid >= o) ? s[tid - o] : (T)0;
        __syncthreads();
        s[tid] += add;
        __syncthreads();
    }
    T run = s[tid] - loc;   // exclusive prefix of this thread
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const long long g = b0 + tid * 4 + i;
        if (g < n) out[g] = run;
        run += v[i];
    }
    if (tid == 255 && sums) sums[blockIdx.x] = s[255];
}

template <class T>
__global__ __launch_bounds__(256) void scan_add(T* __restrict__ out, long long n, const T* __restrict__ sums)
{
    const long long g = (long long)blockIdx.x * 1024 + threadIdx.x;
    const T a = sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const long long q = g + 256 * i;
        if (q < n) out[q] += a;
    }
}

}  // namespace mi
