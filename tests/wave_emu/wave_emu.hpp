// CPU emulation of the wave64 intrinsics used by marvel_amd/csrc/la_wave.hpp.
// TEST INFRASTRUCTURE ONLY: runs the device code unchanged on 64 host
// threads (one per lane) so its control flow and indexing can be checked and
// debugged on the CPU.  Every cross-lane operation is a two-phase barrier
// exchange, so all 64 lanes must reach the same sequence of them (true for
// the aligner: its cross-lane operations sit in wave-uniform control flow).
#pragma once
#include <sched.h>

#include <atomic>
#include <climits>
#include <cstdint>
#include <cstring>

#define __device__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __host__

namespace emu {
// Generation barrier that yields while waiting: the lanes outnumber the host
// cores, and a yield hands the core to a lane that has not arrived yet.
struct YieldBarrier {
    explicit YieldBarrier(int parties) : n(parties) {}
    int n;
    std::atomic<int> count{0};
    std::atomic<int> gen{0};
    void arrive_and_wait()
    {
        const int g0 = gen.load(std::memory_order_acquire);
        if (count.fetch_add(1, std::memory_order_acq_rel) == n - 1) {
            count.store(0, std::memory_order_relaxed);
            gen.fetch_add(1, std::memory_order_acq_rel);
        } else
            while (gen.load(std::memory_order_acquire) == g0) sched_yield();
    }
};
struct Ctx {
    YieldBarrier bar{64};
    long long slot[64];
};
extern thread_local Ctx* g;   // the calling lane's wave
extern thread_local int lane;
template <class T>
inline T xchg(T v, int src)
{
    long long x = 0;
    memcpy(&x, &v, sizeof(T));
    g->slot[lane] = x;
    g->bar.arrive_and_wait();
    long long r = g->slot[src & 63];
    g->bar.arrive_and_wait();
    T out;
    memcpy(&out, &r, sizeof(T));
    return out;
}
}  // namespace emu

inline int __lane_id() { return emu::lane; }
template <class T>
inline T __shfl(T v, int src, int = 64) { return emu::xchg(v, src); }
template <class T>
inline T __shfl_up(T v, unsigned d, int = 64)
{
    int s = emu::lane - (int)d;
    T r = emu::xchg(v, s < 0 ? emu::lane : s);
    return r;
}
template <class T>
inline T __shfl_down(T v, unsigned d, int = 64)
{
    int s = emu::lane + (int)d;
    return emu::xchg(v, s > 63 ? emu::lane : s);
}
template <class T>
inline T __shfl_xor(T v, int m, int = 64) { return emu::xchg(v, emu::lane ^ m); }
inline unsigned long long __ballot(int p)
{
    emu::g->slot[emu::lane] = p ? 1 : 0;
    emu::g->bar.arrive_and_wait();
    unsigned long long m = 0;
    for (int i = 0; i < 64; i++)
        if (emu::g->slot[i]) m |= 1ull << i;
    emu::g->bar.arrive_and_wait();
    return m;
}
inline int __any(int p) { return __ballot(p) != 0; }
inline int __builtin_amdgcn_readlane(int v, int l) { return emu::xchg(v, l); }
// DPP moves: 0x13C wave_ror:1, 0x134 wave_rol:1, 0x138 wave_shr:1, 0x130 wave_shl:1,
// 0x101-0x10F row_shl:n, 0x111-0x11F row_shr:n, 0x142 row_bcast:15, 0x143 row_bcast:31
// (bank_mask 0xF only; lanes without a source read 0 with bound_ctrl, else keep `old`)
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int, bool bound_ctrl)
{
    const int l = emu::lane, row = l >> 4, rl = l & 15;
    int from = -1;
    switch (ctrl) {
    case 0x13C: from = (l + 63) & 63; break;
    case 0x134: from = (l + 1) & 63; break;
    case 0x138: from = l == 0 ? -1 : l - 1; break;
    case 0x130: from = l == 63 ? -1 : l + 1; break;
    case 0x142: from = row == 0 ? -1 : 16 * row - 1; break;
    case 0x143: from = row < 2 ? -1 : 31; break;
    default:
        if (ctrl >= 0x101 && ctrl <= 0x10F) from = rl + (ctrl - 0x100) < 16 ? l + (ctrl - 0x100) : -1;
        else if (ctrl >= 0x111 && ctrl <= 0x11F) from = rl - (ctrl - 0x110) >= 0 ? l - (ctrl - 0x110) : -1;
        else __builtin_trap();
    }
    const int v = emu::xchg(src, from < 0 ? l : from);   // every lane takes part in the exchange
    if (!((row_mask >> row) & 1)) return old;
    if (from < 0) return bound_ctrl ? 0 : old;
    return v;
}
inline int __builtin_amdgcn_mov_dpp(int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl)
{
    return __builtin_amdgcn_update_dpp(0, src, ctrl, row_mask, bank_mask, bound_ctrl);
}
inline int __builtin_amdgcn_readfirstlane(int v) { return emu::xchg(v, 0); }
// v_alignbit_b32: low 32 bits of {hi, lo} >> (s & 31)
inline unsigned __builtin_amdgcn_alignbit(unsigned hi, unsigned lo, unsigned s)
{
    return (unsigned)(((((unsigned long long)hi) << 32) | lo) >> (s & 31));
}
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned m, unsigned acc)
{
    int l = emu::lane < 32 ? emu::lane : 32;
    unsigned long long below = (l >= 32) ? 0xffffffffull : ((1ull << l) - 1);
    return acc + __builtin_popcountll(m & below);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned m, unsigned acc)
{
    int l = emu::lane < 32 ? 0 : emu::lane - 32;
    unsigned long long below = (1ull << l) - 1;
    return acc + __builtin_popcountll(m & below);
}
#define __builtin_amdgcn_fence(a, b) __atomic_thread_fence(__ATOMIC_SEQ_CST)
inline void __builtin_amdgcn_wave_barrier() { emu::g->bar.arrive_and_wait(); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long)x) : 64; }
using std::max;
using std::min;
