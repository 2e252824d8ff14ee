// Synthetic Local_Alignment workload generated in HBM: read pairs of the
// batched-aligner benchmark (template length log-uniform in [lmin, lmax],
// pair error uniform in [emin, emax] split half into each read with
// insertion : deletion : substitution = 9 : 4 : 2, seed point offset along B
// by up to +-offmax, COMP flag on half of the pairs).
//
// Every random decision is a counter-based hash of (seed, task, stream,
// position), all in integer arithmetic, so a CPU restatement reproduces the
// reads and tasks bit for bit and the task set is the same whatever the grid.
//   key(task)        = mix64(seed ^ mix64(task))
//   h(stream, p)     = mix64(key ^ stream << 48 ^ p)
//   stream 0         : the pair's parameters (p = 0 length, 1 error, 2 seed
//                      position, 3 offset, 4 COMP)
//   stream 1, pos p  : template base (bits 0-1) and read A's decisions
//   stream 2, pos p  : read B's decisions
//   decisions of x   : u_ds = bits 2-25, u_ins = bits 26-49, insertion base =
//                      bits 50-51, substitution offset = 1 + (bits 52-63) % 3;
//                      insertion before the base when u_ins < t_ins, base
//                      dropped when u_ds < t_del, substituted when
//                      t_del <= u_ds < t_del + t_sub (24-bit thresholds).
//   length           : L = T[q] + ((T[q+1] - T[q]) * f >> 32), q = r0 >> 52,
//                      f = bits 20-51 of r0, T[j] = floor(lmin (lmax/lmin)^(j/4096))
//   seed position    : template position p = L/4 + r2 % max(1, L/2); apos / bpos
//                      = bases emitted before p in A / B, bpos += r3 % (2 offmax + 1)
//                      - offmax, clamped to [0, blen].
//
// Layout: reads 2-bit packed as MiLaTask describes, A then B of every task,
// one zero word before each read and two after (shared with the next read).
// One 256-thread workgroup builds one pair: each thread owns a contiguous
// run of template positions, a block scan gives every thread its first
// output base in A and B, and whole 32-bit words are stored plainly while
// the words shared with a neighbouring thread are OR-ed in atomically.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "mi_common.h"

namespace mi {

constexpr int kGenThreads = 256;
constexpr int kLenTab = 4096;

__device__ __host__ __forceinline__ uint64_t gen_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct GenPair {
    uint64_t key;
    int L, seedp, off, comp;
    uint32_t t_ins, t_del, t_sub;
};

struct GenArgs {
    uint64_t seed;
    const int* lentab;   // kLenTab + 1 entries
    int emin_ppm, emax_ppm, offmax;
};

__device__ __forceinline__ GenPair gen_pair(const GenArgs& g, int task)
{
    GenPair q;
    q.key = gen_mix64(g.seed ^ gen_mix64((uint64_t)task));
    const uint64_t r0 = gen_mix64(q.key ^ 0), r1 = gen_mix64(q.key ^ 1), r2 = gen_mix64(q.key ^ 2);
    const uint64_t r3 = gen_mix64(q.key ^ 3), r4 = gen_mix64(q.key ^ 4);
    const int j = (int)(r0 >> 52);
    const uint64_t f = (r0 >> 20) & 0xffffffffull;
    const int t0 = g.lentab[j], t1 = g.lentab[j + 1];
    q.L = t0 + (int)(((uint64_t)(t1 - t0) * f) >> 32);
    const uint64_t e = (uint64_t)g.emin_ppm + r1 % (uint64_t)(g.emax_ppm - g.emin_ppm + 1);
    // half of the pair's error per read, 9 : 4 : 2, in units of 2^-24
    q.t_ins = (uint32_t)(e * 9 * 16777216ull / 30000000ull);
    q.t_del = (uint32_t)(e * 4 * 16777216ull / 30000000ull);
    q.t_sub = (uint32_t)(e * 2 * 16777216ull / 30000000ull);
    const int half = q.L / 2 > 1 ? q.L / 2 : 1;
    q.seedp = q.L / 4 + (int)(r2 % (uint64_t)half);
    q.off = (int)(r3 % (uint64_t)(2 * g.offmax + 1)) - g.offmax;
    q.comp = (int)(r4 & 1);
    return q;
}

// bases read A (x) or B (y) emits for one template position
__device__ __forceinline__ int gen_emits(uint64_t x, const GenPair& q)
{
    const uint32_t uds = (uint32_t)((x >> 2) & 0xffffff), uins = (uint32_t)((x >> 26) & 0xffffff);
    return (uins < q.t_ins ? 1 : 0) + (uds < q.t_del ? 0 : 1);
}

// inclusive block sum of a 64-bit value over the 256 threads (4 waves)
__device__ __forceinline__ uint64_t block_incl_scan(uint64_t v, uint64_t* lds)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) lds[wv] = v;
    __syncthreads();
    uint64_t add = 0;
    for (int w = 0; w < wv; w++) add += lds[w];
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(kGenThreads) void k_gen_count(GenArgs g, int ntasks, int* __restrict__ lens)
{
    __shared__ uint64_t lds[4];
    const int task = blockIdx.x;
    if (task >= ntasks) return;
    const GenPair q = gen_pair(g, task);
    const int per = (q.L + kGenThreads - 1) / kGenThreads;
    const int c0 = min(q.L, (int)threadIdx.x * per), c1 = min(q.L, c0 + per);
    uint32_t na = 0, nb = 0;
    for (int p = c0; p < c1; p++) {
        na += gen_emits(gen_mix64(q.key ^ (1ull << 48) ^ (uint64_t)p), q);
        nb += gen_emits(gen_mix64(q.key ^ (2ull << 48) ^ (uint64_t)p), q);
    }
    const uint64_t tot = block_incl_scan(((uint64_t)nb << 32) | na, lds);
    if (threadIdx.x == kGenThreads - 1) {
        lens[2 * task] = (int)(uint32_t)tot;
        lens[2 * task + 1] = (int)(tot >> 32);
    }
}

// 2-bit base writer for one thread's run of output bases [start, ...)
struct BaseOut {
    uint32_t* w;   // read's words as 32-bit halves (base i at bits 2*(i&15) of w[i>>4])
    int start, pos;
    uint32_t acc;
    __device__ void put(uint32_t b)
    {
        acc |= b << (2 * (pos & 15));
        pos += 1;
        if ((pos & 15) == 0) {
            const int wi = (pos - 1) >> 4;
            if (wi * 16 >= start) w[wi] = acc;   // the whole word is this thread's
            else atomicOr(&w[wi], acc);
            acc = 0;
        }
    }
    __device__ void flush()
    {
        if (pos & 15) atomicOr(&w[pos >> 4], acc);
    }
};

__global__ __launch_bounds__(kGenThreads) void k_gen_emit(GenArgs g, int ntasks, const int* __restrict__ lens,
                                                          const long long* __restrict__ word0,
                                                          uint64_t* __restrict__ packed, MiLaTask* __restrict__ tasks)
{
    __shared__ uint64_t lds[4];
    const int task = blockIdx.x;
    if (task >= ntasks) return;
    const GenPair q = gen_pair(g, task);
    const int per = (q.L + kGenThreads - 1) / kGenThreads;
    const int c0 = min(q.L, (int)threadIdx.x * per), c1 = min(q.L, c0 + per);
    uint32_t na = 0, nb = 0;
    for (int p = c0; p < c1; p++) {
        na += gen_emits(gen_mix64(q.key ^ (1ull << 48) ^ (uint64_t)p), q);
        nb += gen_emits(gen_mix64(q.key ^ (2ull << 48) ^ (uint64_t)p), q);
    }
    const uint64_t incl = block_incl_scan(((uint64_t)nb << 32) | na, lds);
    const uint64_t excl = incl - (((uint64_t)nb << 32) | na);
    const long long wa = word0[2 * task], wb = word0[2 * task + 1];
    BaseOut A{reinterpret_cast<uint32_t*>(packed + wa), (int)(uint32_t)excl, (int)(uint32_t)excl, 0u};
    BaseOut B{reinterpret_cast<uint32_t*>(packed + wb), (int)(excl >> 32), (int)(excl >> 32), 0u};
    for (int p = c0; p < c1; p++) {
        if (p == q.seedp) {
            const int alen = lens[2 * task], blen = lens[2 * task + 1];
            const int apos = A.pos;
            int bpos = B.pos + q.off;
            bpos = bpos < 0 ? 0 : (bpos > blen ? blen : bpos);
            MiLaTask t;
            t.aword = wa;
            t.bword = wb;
            t.alen = alen;
            t.blen = blen;
            t.comp = q.comp;
            t.low = t.hgh = apos - bpos;
            t.anti = apos + bpos;
            t.lbord = t.hbord = -1;
            t.selfie = 0;
            tasks[task] = t;
        }
        const uint64_t x = gen_mix64(q.key ^ (1ull << 48) ^ (uint64_t)p);
        const uint64_t y = gen_mix64(q.key ^ (2ull << 48) ^ (uint64_t)p);
        const uint32_t tb = (uint32_t)(x & 3);
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const uint64_t z = r ? y : x;
            BaseOut& o = r ? B : A;
            const uint32_t uds = (uint32_t)((z >> 2) & 0xffffff), uins = (uint32_t)((z >> 26) & 0xffffff);
            if (uins < q.t_ins) o.put((uint32_t)((z >> 50) & 3));
            if (uds >= q.t_del) {
                const bool sub = uds < q.t_del + q.t_sub;
                const uint32_t so = 1 + (uint32_t)(((z >> 52) & 0xfff) % 3);
                o.put(sub ? ((tb + so) & 3) : tb);
            }
        }
    }
    A.flush();
    B.flush();
}

static void length_table(int lmin, int lmax, std::vector<int>& t)
{
    t.resize(kLenTab + 1);
    for (int j = 0; j <= kLenTab; j++) t[j] = (int)floor((double)lmin * pow((double)lmax / lmin, j / (double)kLenTab));
}

}  // namespace mi

using namespace mi;

extern "C" {

int mi_gen_la_tasks(const MiGenTaskParams* gp, int ntasks, uint64_t* d_packed, long long packed_cap,
                    long long* npacked, MiLaTask* d_tasks)
{
    if (gp == nullptr || npacked == nullptr || ntasks <= 0 || gp->lmin < 16 || gp->lmax <= gp->lmin ||
        gp->lmax > (1 << 24) || gp->emin_ppm < 0 || gp->emax_ppm < gp->emin_ppm || gp->emax_ppm > 1000000 ||
        gp->offmax < 0 || gp->offmax > (1 << 24))
        return MI_ERR_ARG;
    std::vector<int> tab;
    length_table(gp->lmin, gp->lmax, tab);
    int* dtab = nullptr;
    int* dlens = nullptr;
    long long* dw0 = nullptr;
    int rc = MI_ERR_DEVICE;
    std::vector<int> lens(2 * (size_t)ntasks);
    std::vector<long long> w0(2 * (size_t)ntasks);
    long long total = 1;
    GenArgs g;
    if (hipMalloc(&dtab, sizeof(int) * tab.size()) != hipSuccess) goto done;
    if (hipMalloc(&dlens, sizeof(int) * 2 * (size_t)ntasks) != hipSuccess) goto done;
    if (hipMemcpy(dtab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice) != hipSuccess) goto done;
    g.seed = gp->seed;
    g.lentab = dtab;
    g.emin_ppm = gp->emin_ppm;
    g.emax_ppm = gp->emax_ppm;
    g.offmax = gp->offmax;
    hipLaunchKernelGGL(k_gen_count, dim3(ntasks), dim3(kGenThreads), 0, 0, g, ntasks, dlens);
    if (hipGetLastError() != hipSuccess) goto done;
    if (hipMemcpy(lens.data(), dlens, sizeof(int) * lens.size(), hipMemcpyDeviceToHost) != hipSuccess) goto done;
    for (size_t r = 0; r < lens.size(); r++) {
        w0[r] = total;
        total += (lens[r] + 31) / 32 + 2;
    }
    total += 1;
    *npacked = total;
    if (d_packed == nullptr) {
        rc = 0;
        goto done;
    }
    if (total > packed_cap || d_tasks == nullptr) {
        rc = MI_ERR_ARG;
        goto done;
    }
    if (hipMalloc(&dw0, sizeof(long long) * w0.size()) != hipSuccess) goto done;
    if (hipMemcpy(dw0, w0.data(), sizeof(long long) * w0.size(), hipMemcpyHostToDevice) != hipSuccess) goto done;
    if (hipMemset(d_packed, 0, sizeof(uint64_t) * total) != hipSuccess) goto done;
    hipLaunchKernelGGL(k_gen_emit, dim3(ntasks), dim3(kGenThreads), 0, 0, g, ntasks, dlens, dw0, d_packed, d_tasks);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) goto done;
    rc = 0;
done:
    if (rc == MI_ERR_DEVICE) mi_set_error("mi_gen_la_tasks: HIP error");
    if (rc == MI_ERR_ARG) mi_set_error("mi_gen_la_tasks: packed buffer too small");
    (void)hipFree(dtab);
    (void)hipFree(dlens);
    (void)hipFree(dw0);
    return rc;
}

}  // extern "C"
