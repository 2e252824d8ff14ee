/* Oracle: CPU restatement of Compute_Trace_PTS (dalign/align.c:5286-5397)
 * and its per-segment O(NP) aligner iter_np (dalign/align.c:4556-4916).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU trace kernel.  Pinned
 * against the reference's own Compute_Trace_PTS through the vectors of
 * tests/golden/trace/ (made by oracle/trace_vectors.c linked with the
 * reference align.c) and, for small cases, against a unit-cost edit-distance
 * DP (the segment distance iter_np returns is the Levenshtein distance).
 *
 * Row layout of the wave arrays follows the reference's allocation (rows of
 * trace_spacing + nmax + 3 ints, column 0 at offset nmax + 1, forward rows
 * for d = -2..dmax then back-pointer rows), so any access at a row edge
 * reads the same neighbour the reference reads.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "dal_oracle.h"

typedef struct {
    int **PVF, **PHF;   /* rows d = -2 .. dmax */
    int *Stop;          /* next output slot */
    const char *Aabs, *Babs;
} np_wave;

/* one segment: A[0..M) against B[0..N); returns its difference count and
 * appends the indel positions of the trace to w->Stop (dalign/align.c:4556-4916) */
static int np_segment(const char *A, int M, const char *B, int N, np_wave *w, int mode)
{
    int **PVF = w->PVF, **PHF = w->PHF;
    const int del = M - N;
    int D, low, hgh, posl = -INT32_MAX, posh = INT32_MAX;

    if (del >= 0) {
        low = 0;
        hgh = del;
    } else {
        low = del;
        hgh = 0;
    }
    if (w->Aabs == w->Babs) {   /* same sequence: stay off the main diagonal */
        if (B == A) return -1;
        if (B < A) posl = (int)(B - A) + 1;
        else posh = (int)(B - A) - 1;
    }
    int *F2, *F1 = PVF[-2], *F0 = PVF[-1], *HF;
    for (int k = low - 1; k <= hgh + 1; k++) F1[k] = F0[k] = -2;
    F0[0] = -1;
    low += 1;
    hgh -= 1;

    for (D = 0;; D++) {
        F2 = F1;
        F1 = F0;
        F0 = PVF[D];
        HF = PHF[D];
        if ((D & 1) == 0) {
            if (low > posl) low -= 1;
            if (hgh < posh) hgh += 1;
        }
        F0[hgh + 1] = F0[low - 1] = -2;

        /* furthest point on diagonal k from the three predecessors; HF records
         * which one (0: F1[k]+1, mdir/pdir as given); then extend the snake */
#define NP_STEP(MDIR, PDIR)                                                          \
    do {                                                                             \
        const int ac = F1[k] + 1;                                                    \
        if (ac < am) {                                                               \
            if (ap < am) { HF[k] = (MDIR); j = am; }                                 \
            else { HF[k] = (PDIR); j = ap; }                                         \
        } else if (ap < ac) { HF[k] = 0; j = ac; }                                   \
        else { HF[k] = (PDIR); j = ap; }                                             \
        const int lim = (N < i) ? N : i;                                             \
        while (j < lim && B[j] == a[j]) j += 1;                                      \
        F0[k] = j;                                                                   \
    } while (0)

        int k, i, j, am, ap;
        const char *a;
        j = -2;
        a = A + hgh;
        i = M - hgh;
        for (k = hgh; k > del; k--) {
            ap = j + 1;
            am = F2[k - 1];
            NP_STEP(-1, 4);
            a -= 1;
            i += 1;
        }
        j = -2;
        a = A + low;
        i = M - low;
        for (k = low; k < del; k++) {
            ap = F2[k + 1] + 1;
            am = j;
            NP_STEP(2, 1);
            a += 1;
            i -= 1;
        }
        ap = F0[del + 1] + 1;
        am = j;
        NP_STEP(2, 4);
#undef NP_STEP
        if (F0[del] >= N) break;
    }

    /* trace back from (D, del) and reverse the back-pointers into forward links */
    {
        int k, h, m, e, c;
        const int ap0 = (int)(w->Aabs - A) - 1;
        const int bp0 = (int)(B - w->Babs) + 1;
        PHF[0][0] = 3;
        c = N;
        k = del;
        e = PHF[D][k];
        PHF[D][k] = 3;
        while (e != 3) {
            h = k + e;
            if (e > 1) h -= 3;
            else if (e == 0) D -= 1;
            else D -= 2;

            if (mode == 1 && h < k) {           /* UPPERMOST: prefer the upper indel */
                const char *a = A + k;
                m = (k < 0) ? -k : 0;
                if (PVF[D][h] <= c) c = PVF[D][h] - 1;
                while (c >= m && a[c] == B[c]) c -= 1;
                if (e == -1) {
                    if (c <= PVF[D + 2][k + 1]) { e = 4; h = k + 1; D = D + 2; }
                    else if (c == PVF[D + 1][k]) { e = 0; h = k; D = D + 1; }
                    else PVF[D][h] = c + 1;
                } else {
                    m = (k == del) ? D : D - 2;
                    if (c <= PVF[m][k + 1]) { e = (k == del) ? 4 : 1; h = k + 1; D = m; }
                    else if (c == PVF[D - 1][k]) { e = 0; h = k; D = D - 1; }
                    else PVF[D][h] = c + 1;
                }
            } else if (mode == -1 && h > k) {   /* LOWERMOST: prefer the lower indel */
                const char *a = A + k;
                m = (k < 0) ? -k : 0;
                if (PVF[D][h] < c) c = PVF[D][h];
                while (c >= m && a[c] == B[c]) c -= 1;
                if (e == 1) {
                    if (c < PVF[D + 2][k - 1]) { e = 2; h = k - 1; D = D + 2; }
                    else if (c == PVF[D + 1][k]) { e = 0; h = k; D = D + 1; }
                    else PVF[D][h] = c--;
                } else {
                    m = (k == del) ? D : D - 2;
                    if (c < PVF[m][k - 1]) { e = (k == del) ? 2 : -1; h = k - 1; D = m; }
                    else if (c == PVF[D - 1][k]) { e = 0; h = k; D = D - 1; }
                    else PVF[D][h] = c--;
                }
            }
            m = PHF[D][h];
            PHF[D][h] = e;
            e = m;
            k = h;
        }

        k = D = 0;
        e = PHF[D][k];
        while (e != 3) {
            h = k - e;
            c = PVF[D][k];
            if (e > 1) h += 3;
            else if (e == 0) D += 1;
            else D += 2;
            if (h > k) *w->Stop++ = bp0 + c;
            else if (h < k) *w->Stop++ = ap0 - (c + k);
            k = h;
            e = PHF[D][h];
        }
    }
    return D + abs(del);
}

int or_trace_pts(const char *aseq, const char *bseq, const or_path *path, const uint16_t *points,
                 int trace_spacing, int mode, int *out, int64_t cap, int *ntrace, int *diffs)
{
    const int tlen = path->tlen;
    const int M = path->aepos - path->abpos, N = path->bepos - path->bbpos;
    int dmax = 0, nmax = 0;
    for (int d = 1; d < tlen; d += 2) {
        if (points[d - 1] > dmax) dmax = points[d - 1];
        if (points[d] > nmax) nmax = points[d];
    }
    if (tlen <= 1) nmax = N;
    (void)M;

    /* rows of s ints, column 0 at offset nmax + 1 (dalign/align.c:5330-5345) */
    const int64_t s = trace_spacing + nmax + 3;
    const int64_t rows = 2 * (int64_t)(dmax + 3);
    int *mem = (int *)calloc((size_t)(rows * s + 2 * s), sizeof(int));
    int **pv = (int **)malloc(sizeof(int *) * (size_t)(dmax + 3));
    int **ph = (int **)malloc(sizeof(int *) * (size_t)(dmax + 3));
    if (mem == NULL || pv == NULL || ph == NULL) return -2;
    np_wave w;
    w.PVF = pv + 2;
    w.PHF = ph + 2;
    w.PVF[-2] = mem + s + (nmax + 1);
    for (int d = -1; d <= dmax; d++) w.PVF[d] = w.PVF[d - 1] + s;
    w.PHF[-2] = w.PVF[dmax] + s;
    for (int d = -1; d <= dmax; d++) w.PHF[d] = w.PHF[d - 1] + s;

    /* the trace of a segment never holds more positions than its two lengths */
    int *tmp = (int *)malloc(sizeof(int) * (size_t)(path->aepos - path->abpos + path->bepos - path->bbpos + 16));
    w.Stop = tmp;
    w.Aabs = aseq;
    w.Babs = bseq;
    int total = 0, ab = path->abpos, bb = path->bbpos, ae = (ab / trace_spacing) * trace_spacing, be, rc = 0;
    for (int i = 1; i < tlen - 2; i += 2) {
        ae += trace_spacing;
        be = bb + points[i];
        const int d = np_segment(aseq + ab, ae - ab, bseq + bb, be - bb, &w, mode);
        if (d < 0) { rc = -1; goto done; }
        total += d;
        ab = ae;
        bb = be;
    }
    {
        const int d = np_segment(aseq + ab, path->aepos - ab, bseq + bb, path->bepos - bb, &w, mode);
        if (d < 0) { rc = -1; goto done; }
        total += d;
    }
    {
        const int64_t n = w.Stop - tmp;
        if (n > cap) { rc = -3; goto done; }
        memcpy(out, tmp, sizeof(int) * (size_t)n);
        *ntrace = (int)n;
        *diffs = total;
    }
done:
    free(tmp);
    free(mem);
    free(pv);
    free(ph);
    return rc;
}

int64_t or_trace_tasks(double ave_corr, int tspace, const float freq[4], const char *seq, int ntasks,
                       const int64_t *aoff, const int64_t *boff, const int32_t *par, int32_t *res, int64_t rcap)
{
    or_spec *sp = (or_spec *)malloc(sizeof(or_spec));
    or_work *w = or_work_new();
    int64_t o = 0;
    or_spec_init(sp, ave_corr, tspace, freq);
    for (int i = 0; i < ntasks; i++) {
        const int32_t *v = par + 8 * i;
        or_path pa, pb;
        const uint16_t *ta = NULL, *tb = NULL;
        const int rc = or_local_alignment(sp, w, seq + aoff[i], v[0], seq + boff[i], v[1], v[2], v[3], v[4], v[5],
                                          v[6], v[7], &pa, &ta, &pb, &tb);
        if (o + 2 > rcap) { o = -1; break; }
        res[o++] = rc ? 3 : 0;
        res[o++] = rc ? 0 : pa.tlen;
        if (rc) continue;
        uint16_t *pts = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(pa.tlen + 1));
        memcpy(pts, ta, sizeof(uint16_t) * (size_t)pa.tlen);
        for (int mode = -1; mode <= 1; mode++) {
            int ntr = 0, diffs = 0;
            if (o + 2 > rcap) { o = -1; break; }
            const int r = or_trace_pts(seq + aoff[i], seq + boff[i], &pa, pts, tspace, mode, res + o + 2,
                                       rcap - o - 2, &ntr, &diffs);
            if (r) { o = -1; break; }
            res[o] = diffs;
            res[o + 1] = ntr;
            o += 2 + ntr;
        }
        free(pts);
        if (o < 0) break;
    }
    free(sp);
    or_work_free(w);
    return o;
}
