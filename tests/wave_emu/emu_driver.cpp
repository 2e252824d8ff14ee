// Runs marvel_amd/csrc/la_wave.hpp's local_alignment on the CPU wave
// emulator over a la_vectors task file and writes a la_vectors result file
// (plus a status word per task).  TEST INFRASTRUCTURE ONLY.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "hip_cpu.hpp"
#include "../../marvel_amd/csrc/la_wave.hpp"

static int g_abort = 0;
static const double kBias[10] = {.690, .690, .690, .690, .780, .850, .900, .933, .966, 1.000};

int main(int argc, char** argv)
{
    if (argc < 3) return 1;
    FILE* in = fopen(argv[1], "rb");
    FILE* out = fopen(argv[2], "wb");
    int only = argc > 3 ? atoi(argv[3]) : -1;
    int nch = argc > 4 ? atoi(argv[4]) : 1;
    int32_t n, ts;
    double corr;
    float freq[4];
    int64_t nb;
    if (fread(&n, 4, 1, in) != 1 || fread(&corr, 8, 1, in) != 1 || fread(&ts, 4, 1, in) != 1 ||
        fread(freq, 16, 1, in) != 1 || fread(&nb, 8, 1, in) != 1)
        return 1;
    std::vector<char> seq(nb);
    if (fread(seq.data(), 1, nb, in) != (size_t)nb) return 1;
    // alignment constants (same expressions as New_Align_Spec)
    double match = (float)(freq[0] + freq[3]);
    if (match > .5) match = 1. - match;
    int bias = (int)((match + .025) * 20. - 1.);
    if (match < .2) bias = 3;
    int ave_path = (int)(60 * (1. - kBias[bias] * (1. - corr)));
    int mscore = (int)(1000 * kBias[bias] * (1. - corr)), dscore = 1000 - mscore;
    // the device reads the score table right after the suffix table (one base pointer)
    std::vector<int16_t> tables(65536);
    int16_t* table = tables.data();
    int16_t* score = tables.data() + 32768;
    for (int x = 0; x < 32768; x++) {
        int run = 0, top = 0;
        for (int i = 14; i >= 0; i--) {
            top = std::max(top, run);
            run += ((x >> i) & 1) ? mscore : -dscore;
        }
        table[x] = (int16_t)(run - top);
        score[x] = (int16_t)run;
    }
    mi::DevSpec sp{table, score, ts, ave_path};
    emu::Ctx ctx;
    emu::g = &ctx;
    for (int i = 0; i < n; i++) {
        int64_t ao, bo;
        int32_t v[8];
        if (fread(&ao, 8, 1, in) != 1 || fread(&bo, 8, 1, in) != 1 || fread(v, 4, 8, in) != 8) return 1;
        if (only >= 0 && i != only) continue;
        const int alen = v[0], blen = v[1];
        auto pack = [&](int64_t off, int len) {
            std::vector<uint64_t> w((len + 31) / 32 + 3, 0);
            for (int j = 0; j < len; j++) w[1 + (j >> 5)] |= (uint64_t)(seq[off + j] & 3) << (2 * (j & 31));
            return w;
        };
        std::vector<uint64_t> pa = pack(ao, alen), pb = pack(bo, blen);
        const int maxtp = 2 * (std::max(alen, blen) / ts + 2);
        const int pool_cap = 1 << 20, pts_cap = 8192;
        std::vector<mi::Peb> pool(pool_cap), pts(2 * pts_cap);
        std::vector<uint16_t> tr(4 * maxtp + 8, 0);
        mi::LaResult res[64];
        std::vector<std::thread> th;
        for (int l = 0; l < 64; l++)
            th.emplace_back([&, l] {
                emu::lane = l;
                emu::g = &ctx;   // the wave context is per thread
                mi::WaveCount wc;
                if (nch == 4)
                    res[l] = mi::local_alignment<4>(sp, pa.data() + 1, alen, pb.data() + 1, blen, v[2], v[3], v[4],
                                                    v[5], -INT32_MAX, INT32_MAX, pool.data(), pool_cap, pts.data(),
                                                    pts_cap, tr.data() + maxtp, tr.data() + 3 * maxtp, maxtp, &g_abort, wc);
                else
                    res[l] = mi::local_alignment<1>(sp, pa.data() + 1, alen, pb.data() + 1, blen, v[2], v[3], v[4],
                                                    v[5], -INT32_MAX, INT32_MAX, pool.data(), pool_cap, pts.data(),
                                                    pts_cap, tr.data() + maxtp, tr.data() + 3 * maxtp, maxtp, &g_abort, wc);
            });
        for (auto& t : th) t.join();
        const mi::LaResult& r = res[0];
        int32_t f[13] = {r.abpos, r.bbpos, r.aepos, r.bepos, r.diffs, r.atlen, r.b_abpos, r.b_bbpos,
                         r.b_aepos, r.b_bepos, r.diffs, r.btlen, r.status};
        if (r.status) { for (int q = 0; q < 12; q++) f[q] = 0; }
        fwrite(f, 4, 13, out);
        if (!r.status) {
            fwrite(tr.data() + maxtp + r.afirst, 2, r.atlen, out);
            fwrite(tr.data() + 3 * maxtp + r.bfirst, 2, r.btlen, out);
        }
        fflush(out);
    }
    fclose(out);
    return 0;
}
