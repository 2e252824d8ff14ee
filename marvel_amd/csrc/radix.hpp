// Stable LSD radix sort and exclusive scan for gfx950 (used by the k-mer
// index and the seed sort of the daligner filter).
//
// One pass sorts by an 8-bit digit of a 64-bit key:
//   1. rs_hist    : one 256-thread workgroup per tile of kTile records counts
//                   the digits of its tile in LDS (LDS atomics) and writes
//                   counts[digit * ntiles + tile];
//   2. scan       : exclusive prefix over counts (digit-major), giving each
//                   (digit, tile) its output base;
//   3. rs_scatter : one wave per tile walks the tile in order, 64 records a
//                   round; lanes with equal digits find each other with 8
//                   ballots, take ranks from an mbcnt over the peer mask and
//                   a per-digit LDS cursor, and store to base + rank.  Walking
//                   the tile in order keeps the sort stable.
// The key of a record is produced by a functor, so 8-byte keys and 16-byte
// seed records share the code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

constexpr int kTile = 4096;   // records per tile (64 lanes x 64 rounds)

template <class Rec, class KeyOf>
__global__ __launch_bounds__(256) void rs_hist(const Rec* __restrict__ in, long long n, int shift, KeyOf key,
                                               unsigned long long* __restrict__ counts, int ntiles)
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

template <class Rec, class KeyOf>
__global__ __launch_bounds__(64) void rs_scatter(const Rec* __restrict__ in, Rec* __restrict__ out, long long n,
                                                 int shift, KeyOf key, const unsigned long long* __restrict__ offs,
                                                 int ntiles)
{
    __shared__ unsigned long long cur[256];
    const int lane = __lane_id();
    for (int d = lane; d < 256; d += 64) cur[d] = offs[(size_t)d * ntiles + blockIdx.x];
    __syncthreads();
    const long long base = (long long)blockIdx.x * kTile;
    for (int r = 0; r < kTile / 64; r++) {
        const long long g = base + r * 64 + lane;
        const bool valid = g < n;
        Rec v;
        unsigned dg = 0;
        if (valid) {
            v = in[g];
            dg = (unsigned)(key(v) >> shift) & 255u;
        }
        // peers: valid lanes with the same digit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const unsigned long long m = __ballot((dg >> b) & 1);
            peers &= ((dg >> b) & 1) ? m : ~m;
        }
        const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0));
        unsigned long long pos = 0;
        if (valid) pos = cur[dg] + rank;
        __syncthreads();   // all lanes read cur before the group leader advances it
        if (valid) {
            out[pos] = v;
            const int last = 63 - __clzll(peers);
            if (lane == last) cur[dg] += __popcll(peers);
        }
        __syncthreads();
        if (base + (r + 1) * 64 >= n) break;
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
