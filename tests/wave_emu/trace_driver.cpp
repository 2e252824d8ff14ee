// Runs the library's Compute_Trace_PTS batch entry (mi_trace_pts_host) over
// a la_vectors task file and writes the words of oracle/trace_vectors.c's
// result layout.  The A paths come from the oracle's Local_Alignment.
// Built twice: on the CPU through hip_cpu.hpp (test infrastructure) and,
// unchanged, against the GPU library.  TEST INFRASTRUCTURE ONLY.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/marvl_align.h"
#include "../../oracle/dal_oracle.h"

static void rd(void* p, size_t n, FILE* f)
{
    if (fread(p, 1, n, f) != n) {
        fprintf(stderr, "trace_driver: short task file\n");
        exit(1);
    }
}

int main(int argc, char** argv)
{
    if (argc != 3) return 1;
    FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
    if (!in || !out) return 1;
    int32_t n, ts;
    double corr;
    float freq[4];
    int64_t nb;
    rd(&n, 4, in);
    rd(&corr, 8, in);
    rd(&ts, 4, in);
    rd(freq, 16, in);
    rd(&nb, 8, in);
    std::vector<char> seq(nb + 16);
    rd(seq.data(), nb, in);
    or_spec* sp = (or_spec*)malloc(sizeof(or_spec));
    or_spec_init(sp, corr, ts, freq);
    or_work* w = or_work_new();
    std::vector<uint64_t> packed(1, 0);
    std::vector<uint16_t> points;
    std::vector<MiTraceTask> tasks;
    std::vector<int> status(n, 0), atlen(n, 0);
    auto pack = [&](int64_t off, int len) {
        const long long w0 = (long long)packed.size();
        packed.resize(packed.size() + (len + 31) / 32 + 2, 0);
        for (int j = 0; j < len; j++) packed[w0 + (j >> 5)] |= (uint64_t)(seq[off + j] & 3) << (2 * (j & 31));
        return w0;
    };
    for (int i = 0; i < n; i++) {
        int64_t ao, bo;
        int32_t v[8];
        rd(&ao, 8, in);
        rd(&bo, 8, in);
        rd(v, 32, in);
        or_path pa, pb;
        const uint16_t *ta, *tb;
        if (or_local_alignment(sp, w, seq.data() + ao, v[0], seq.data() + bo, v[1], v[2], v[3], v[4], v[5], v[6], v[7],
                               &pa, &ta, &pb, &tb)) {
            status[i] = 3;
            continue;
        }
        MiTraceTask t;
        t.aword = pack(ao, v[0]);
        t.bword = pack(bo, v[1]);
        t.alen = v[0];
        t.blen = v[1];
        t.abpos = pa.abpos;
        t.bbpos = pa.bbpos;
        t.aepos = pa.aepos;
        t.bepos = pa.bepos;
        t.tlen = pa.tlen;
        t.points = (long long)points.size();
        t.selfie = 0;
        points.insert(points.end(), ta, ta + pa.tlen);
        atlen[i] = pa.tlen;
        tasks.push_back(t);
    }
    const int nt = (int)tasks.size();
    long long cap = 16;
    for (auto& t : tasks) cap += (t.aepos - t.abpos) + (t.bepos - t.bbpos) + 16;
    std::vector<std::vector<MiTraceResult>> res(3, std::vector<MiTraceResult>(nt));
    std::vector<std::vector<int>> tr(3, std::vector<int>(cap));
    for (int m = 0; m < 3; m++) {
        const int rc = mi_trace_pts_host(ts, m - 1, packed.data(), (long long)packed.size(), points.data(),
                                         (long long)points.size(), tasks.data(), nt, res[m].data(), tr[m].data(), cap);
        if (rc) {
            fprintf(stderr, "mi_trace_pts_host: %d (%s)\n", rc, mi_last_error());
            return 2;
        }
    }
    int k = 0;
    for (int i = 0; i < n; i++) {
        const int32_t head[2] = {status[i], atlen[i]};
        fwrite(head, 4, 2, out);
        if (status[i]) continue;
        for (int m = 0; m < 3; m++) {
            const MiTraceResult& r = res[m][k];
            if (r.status) {
                fprintf(stderr, "task %d mode %d: status %d\n", i, m - 1, r.status);
                return 3;
            }
            const int32_t h2[2] = {r.diffs, r.ntrace};
            fwrite(h2, 4, 2, out);
            fwrite(tr[m].data() + r.trace, 4, r.ntrace, out);
        }
        k++;
    }
    fclose(out);
    return 0;
}
