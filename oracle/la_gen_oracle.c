/* CPU restatement of the batched-aligner workload generator
 * (marvel_amd/csrc/la_gen.hip), one pair at a time, position by position.
 *
 * Test infrastructure only: the tests regenerate sampled pairs here and
 * compare them byte for byte with the reads and tasks the GPU generator
 * wrote, then align them with the oracle's Local_Alignment.
 *
 * The generator (no reference counterpart; it produces inputs):
 *   key          = mix64(seed ^ mix64(task)), h(s, p) = mix64(key ^ s << 48 ^ p)
 *   parameters   : h(0, 0..4) -> length, pair error, seed position, offset, COMP
 *   position p   : h(1, p) -> template base (bits 0-1) and read A's
 *                  decisions, h(2, p) -> read B's decisions
 * Pinned by: every position's output is emitted in order here (insertion,
 * then the kept or substituted base), so an off-by-one in the GPU's per-thread
 * offsets, a dropped atomic word or a wrong seed position shows as a byte
 * difference; the distribution checks (lengths log-uniform, error rates) are
 * in tests/test_la_gen.py. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static uint64_t gmix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Template length of the pair: piecewise-linear inverse of the log-uniform
 * law over 4096 quantiles floor(lmin (lmax/lmin)^(j/4096)). */
static int gen_length(uint64_t r0, int lmin, int lmax)
{
    const int j = (int)(r0 >> 52);
    const double ratio = (double)lmax / lmin;
    const int t0 = (int)floor((double)lmin * pow(ratio, j / 4096.0));
    const int t1 = (int)floor((double)lmin * pow(ratio, (j + 1) / 4096.0));
    const uint64_t f = (r0 >> 20) & 0xffffffffull;
    return t0 + (int)(((uint64_t)(t1 - t0) * f) >> 32);
}

/* Writes the pair's reads (bases 0..3) and its task parameters
 * par = {alen, blen, comp, low, hgh, anti, lbord, hbord}; stats (optional) =
 * {template length, pair error ppm, seed offset, template seed position,
 * insertions, deletions, substitutions of A, the same of B}.  Returns 0, or
 * -1 when a read would exceed its capacity. */
int or_gen_la_pair(uint64_t seed, int task, int lmin, int lmax, int emin_ppm, int emax_ppm, int offmax, uint8_t *a,
                   int acap, uint8_t *b, int bcap, int32_t *par, int32_t *stats)
{
    int32_t ev[6] = {0, 0, 0, 0, 0, 0};
    const uint64_t key = gmix(seed ^ gmix((uint64_t)task));
    const uint64_t r0 = gmix(key ^ 0), r1 = gmix(key ^ 1), r2 = gmix(key ^ 2), r3 = gmix(key ^ 3), r4 = gmix(key ^ 4);
    const int L = gen_length(r0, lmin, lmax);
    const uint64_t e = (uint64_t)emin_ppm + r1 % (uint64_t)(emax_ppm - emin_ppm + 1);
    const uint32_t t_ins = (uint32_t)(e * 9 * 16777216ull / 30000000ull);
    const uint32_t t_del = (uint32_t)(e * 4 * 16777216ull / 30000000ull);
    const uint32_t t_sub = (uint32_t)(e * 2 * 16777216ull / 30000000ull);
    const int half = L / 2 > 1 ? L / 2 : 1;
    const int seedp = L / 4 + (int)(r2 % (uint64_t)half);
    const int off = (int)(r3 % (uint64_t)(2 * offmax + 1)) - offmax;
    const int comp = (int)(r4 & 1);
    int na = 0, nb = 0, apos = 0, bpos = 0;
    for (int p = 0; p < L; p++) {
        const uint64_t x = gmix(key ^ (1ull << 48) ^ (uint64_t)p);
        const uint64_t y = gmix(key ^ (2ull << 48) ^ (uint64_t)p);
        const int tb = (int)(x & 3);
        if (p == seedp) {
            apos = na;
            bpos = nb;
        }
        for (int r = 0; r < 2; r++) {
            const uint64_t z = r ? y : x;
            uint8_t *o = r ? b : a;
            int *n = r ? &nb : &na;
            const int cap = r ? bcap : acap;
            const uint32_t uds = (uint32_t)((z >> 2) & 0xffffff), uins = (uint32_t)((z >> 26) & 0xffffff);
            if (uins < t_ins) {
                if (*n >= cap) return -1;
                o[(*n)++] = (uint8_t)((z >> 50) & 3);
                ev[3 * r] += 1;
            }
            if (uds >= t_del) {
                int v = tb;
                if (uds < t_del + t_sub) {
                    v = (tb + 1 + (int)(((z >> 52) & 0xfff) % 3)) & 3;
                    ev[3 * r + 2] += 1;
                }
                if (*n >= cap) return -1;
                o[(*n)++] = (uint8_t)v;
            } else
                ev[3 * r + 1] += 1;
        }
    }
    bpos += off;
    if (bpos < 0) bpos = 0;
    if (bpos > nb) bpos = nb;
    par[0] = na;
    par[1] = nb;
    par[2] = comp;
    par[3] = par[4] = apos - bpos;
    par[5] = apos + bpos;
    par[6] = par[7] = -1;
    if (stats) {
        stats[0] = L;
        stats[1] = (int32_t)e;
        stats[2] = off;
        stats[3] = seedp;
        memcpy(stats + 4, ev, sizeof ev);
    }
    return 0;
}
