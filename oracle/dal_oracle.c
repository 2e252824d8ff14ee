/* Oracle: plain-C restatement of MARVEL daligner's overlap path.
 *
 * TEST INFRASTRUCTURE ONLY (see dal_oracle.h): the checker for the HIP path,
 * never part of the product.  Every routine cites the reference lines it
 * restates; the data layout and control flow are this file's own.
 */
#define _GNU_SOURCE
#include "dal_oracle.h"

#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

/* ------------------------------------------------------------------------ */
/* Constants of the adaptive wave (dalign/align.c:145-159).                 */
#define W_PATH_LEN 60
#define W_PATH_TOP (1ULL << 60)
#define W_PATH_INT (W_PATH_TOP - 1)
#define W_TRIM_LEN 15
#define W_TRIM_MASK 0x7fff
#define W_TRIM_MLAG 200
#define W_WAVE_LAG 30

static void *xrealloc(void *p, size_t n)
{
    void *q = realloc(p, n ? n : 1);
    if (q == NULL) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        exit(1);
    }
    return q;
}

/* Bias table of New_Align_Spec (dalign/align.c:161-162). */
static const double bias_factor[10] = {.690, .690, .690, .690, .780, .850, .900, .933, .966, 1.000};

/* New_Align_Spec + set_table (dalign/align.c:196-279).  The trim tables are
 * filled iteratively: for a 15-bit column history x (first column = MSB),
 * score[x] is the running sum after all 15 columns (+mscore per match bit,
 * -dscore per mismatch bit) and table[x] = score[x] - max(prefix sums before
 * each column, starting at 0). */
void or_spec_init(or_spec *sp, double ave_corr, int tspace, const float freq[4])
{
    double match;
    int bias, x, i;

    sp->ave_corr = ave_corr;
    sp->tspace = tspace;
    for (i = 0; i < 4; i++) sp->freq[i] = freq[i];

    match = (float)(freq[0] + freq[3]);
    if (match > .5) match = 1. - match;
    bias = (int)((match + .025) * 20. - 1.);
    if (match < .2) bias = 3;

    sp->ave_path = (int)(W_PATH_LEN * (1. - bias_factor[bias] * (1. - ave_corr)));
    sp->mscore = (int)(1000 * bias_factor[bias] * (1. - ave_corr));
    sp->dscore = 1000 - sp->mscore;

    for (x = 0; x <= W_TRIM_MASK; x++) {
        int s = 0, mx = 0;
        for (i = W_TRIM_LEN - 1; i >= 0; i--) {
            if (s > mx) mx = s;
            s += ((x >> i) & 1) ? sp->mscore : -sp->dscore;
        }
        sp->table[x] = (int16_t)(s - mx);
        sp->score[x] = (int16_t)s;
    }
}

/* ------------------------------------------------------------------------ */
/* Working storage.  Band vectors are indexed directly by diagonal k through
 * an offset chosen per call so every legal diagonal of the (alen, blen) edit
 * graph, plus a guard, has a slot.                                          */

typedef struct {
    int up;     /* previous pebble on the path (-1 at the root) */
    int diag;
    int dif;
    int mark;   /* trace-point coordinate (A: x, B: y) */
} pebble;

struct or_work {
    int cap, off;
    int *V, *M, *HA, *HB, *NA, *NB;
    uint64_t *T;
    pebble *peb;
    int npeb, pcap;
    uint16_t *tr;        /* trace workspace: [rev A | fwd A][rev B | fwd B] */
    int trcap;
    int64_t cells, snake;
    int64_t nwaves, maxw, hist[9];   /* wave-width statistics: widths 1-64,65-128,...,>512 */
};

or_work *or_work_new(void)
{
    or_work *w = (or_work *)calloc(1, sizeof(or_work));
    if (w == NULL) exit(1);
    return w;
}

void or_work_free(or_work *w)
{
    if (w == NULL) return;
    free(w->V); free(w->M); free(w->HA); free(w->HB); free(w->NA); free(w->NB);
    free(w->T); free(w->peb); free(w->tr);
    free(w);
}

void or_work_counters(const or_work *w, int64_t *cells, int64_t *snake)
{
    *cells = w->cells;
    *snake = w->snake;
}

static inline void wstat(or_work *w, int width)
{
    int b = (width - 1) >> 6;
    w->nwaves += 1;
    if (width > w->maxw) w->maxw = width;
    w->hist[b > 8 ? 8 : b] += 1;
}

void or_work_widths(const or_work *w, int64_t *nwaves, int64_t *maxw, int64_t hist[9])
{
    *nwaves = w->nwaves;
    *maxw = w->maxw;
    for (int i = 0; i < 9; i++) hist[i] = w->hist[i];
}

#define BAND_GUARD 256

static void band_reserve(or_work *w, int alen, int blen)
{
    int need = alen + blen + 2 * BAND_GUARD + 1;
    if (need > w->cap) {
        w->cap = need + need / 4;
        w->V = xrealloc(w->V, sizeof(int) * w->cap);
        w->M = xrealloc(w->M, sizeof(int) * w->cap);
        w->HA = xrealloc(w->HA, sizeof(int) * w->cap);
        w->HB = xrealloc(w->HB, sizeof(int) * w->cap);
        w->NA = xrealloc(w->NA, sizeof(int) * w->cap);
        w->NB = xrealloc(w->NB, sizeof(int) * w->cap);
        w->T = xrealloc(w->T, sizeof(uint64_t) * w->cap);
    }
    w->off = blen + BAND_GUARD;
}

static int peb_push(or_work *w, int up, int diag, int dif, int mark)
{
    if (w->npeb >= w->pcap) {
        w->pcap = w->pcap * 2 + 4096;
        w->peb = xrealloc(w->peb, sizeof(pebble) * w->pcap);
    }
    pebble *p = w->peb + w->npeb;
    p->up = up;
    p->diag = diag;
    p->dif = dif;
    p->mark = mark;
    return w->npeb++;
}

/* Reverse the singly linked pebble chain ending at `tail` in place and
 * return its root (the walk at dalign/align.c:984-991). */
static int chain_reverse(pebble *P, int tail)
{
    int prev = -1, h = tail, nxt;
    while (h >= 0) {
        nxt = P[h].up;
        P[h].up = prev;
        prev = h;
        h = nxt;
    }
    return prev;
}

/* State of one pass; fields follow the variables of forward_wave /
 * reverse_wave (dalign/align.c:369-1110, :1114-1916). */
typedef struct {
    int besta, besty;
    int trima, trimy, trimd, trimha, trimhb;
    int morea, morey, mored, moreha, morehb, morem;
    int lasta, more, aclip, bclip;
    int low, hgh, dif;
} pass_state;

#define BAD_INPUT 3

/* Snake along diagonal k from row y. dir=+1 compares B[y] with A[y+k]
 * (forward); dir=-1 compares B[y-1] with A[y+k-1] (reverse).  Returns the
 * final row and sets *stop to 1 at a B sentinel, 2 at an A sentinel, 0 at a
 * mismatch.  *steps = number of matches consumed. */
static inline int snake(const char *A, const char *B, int k, int y, int dir, int *stop, int *steps)
{
    int y0 = y;
    (void)dir;
    if (dir > 0) {
        for (;;) {
            int c = B[y];
            if (c == 4) { *stop = 1; break; }
            int d = A[y + k];
            if (c != d) { *stop = (d == 4) ? 2 : 0; break; }
            y += 1;
        }
        *steps = y - y0;
    } else {
        for (;;) {
            int c = B[y - 1];
            if (c == 4) { *stop = 1; break; }
            int d = A[y + k - 1];
            if (c != d) { *stop = (d == 4) ? 2 : 0; break; }
            y -= 1;
        }
        *steps = y0 - y;
    }
    return y;
}

/* Apply `s` match columns to the (m, b) path-quality pair exactly as the
 * per-base loop of dalign/align.c:799-823 does. */
static inline void apply_matches(int *m, uint64_t *b, int s)
{
    int mm = *m;
    uint64_t bb = *b;
    for (int i = 0; i < s; i++) {
        if ((bb & W_PATH_TOP) == 0) mm += 1;
        bb = (bb << 1) | 1;
    }
    *m = mm;
    *b = bb;
}

/* forward_wave (dalign/align.c:369-1110). Returns 0 or BAD_INPUT. */
static int fwd_pass(const or_spec *sp, or_work *w, const char *A, int alen, const char *B, int blen,
                    int comp, int *mind, int maxd, int mida, int minp, int maxp,
                    or_path *ap, uint16_t *atr, or_path *bp, uint16_t *btr)
{
    const int TS = sp->tspace;
    const int boff = comp ? blen % TS : 0;
    const int lo_lim = -w->off, hi_lim = w->cap - w->off - 1;
    int *V = w->V + w->off, *M = w->M + w->off, *HA = w->HA + w->off, *HB = w->HB + w->off;
    int *NA = w->NA + w->off, *NB = w->NB + w->off;
    uint64_t *T = w->T + w->off;
    pass_state s;
    int k;

    s.hgh = maxd;
    s.low = *mind;
    s.dif = 0;
    s.more = 1;
    s.aclip = INT32_MAX;
    s.bclip = -INT32_MAX;
    s.besta = s.trima = s.morea = s.lasta = mida;
    s.besty = s.trimy = s.morey = (mida - s.hgh) >> 1;
    s.trimd = s.mored = 0;
    s.trimha = s.moreha = 0;
    s.trimhb = s.morehb = 1;
    s.morem = -1;
    w->npeb = 0;

    if (s.low - 1 < lo_lim || s.hgh + 1 > hi_lim) return BAD_INPUT;

    /* wave 0 (dalign/align.c:455-587) */
    for (k = s.hgh; k >= s.low; k--) {
        int y = (mida - k) >> 1, stop, st;
        int na = ((y + k) / TS) * TS;
        int ha = peb_push(w, -1, k, 0, na);
        na += TS;
        int nb = ((y + (TS - boff)) / TS - 1) * TS + boff;
        int hb = peb_push(w, -1, k, 0, nb);
        nb += TS;

        if (y < 0 || y > blen || y + k < 0 || y + k > alen) return BAD_INPUT;
        y = snake(A, B, k, y, +1, &stop, &st);
        w->snake += st + 1;
        if (stop == 1) { s.more = 0; if (s.bclip < k) s.bclip = k; }
        else if (stop == 2) { s.more = 0; s.aclip = k; }
        int c = (y << 1) + k;

        while (y + k >= na) { ha = peb_push(w, ha, k, 0, na); na += TS; }
        while (y >= nb) { hb = peb_push(w, hb, k, 0, nb); nb += TS; }

        if (c > s.besta) {
            s.besta = s.trima = s.lasta = c;
            s.besty = s.trimy = y;
            s.trimha = ha;
            s.trimhb = hb;
        }
        V[k] = c; T[k] = W_PATH_INT; M[k] = W_PATH_LEN;
        HA[k] = ha; HB[k] = hb; NA[k] = na; NB[k] = nb;
    }
    w->cells += s.hgh - s.low + 1;

    for (;;) {
        /* clip bookkeeping after a wave (dalign/align.c:589-625, :912-944) */
        if (s.more == 0) {
            if (B[s.besty] != 4 && A[s.besta - s.besty] != 4) s.more = 1;
            if (s.hgh >= s.aclip) {
                s.hgh = s.aclip - 1;
                if (s.morem <= M[s.aclip]) {
                    s.morem = M[s.aclip];
                    s.morea = V[s.aclip];
                    s.morey = (s.morea - s.aclip) / 2;
                    s.mored = s.dif;
                    s.moreha = HA[s.aclip];
                    s.morehb = HB[s.aclip];
                }
            }
            if (s.low <= s.bclip) {
                s.low = s.bclip + 1;
                if (s.morem <= M[s.bclip]) {
                    s.morem = M[s.bclip];
                    s.morea = V[s.bclip];
                    s.morey = (s.morea - s.bclip) / 2;
                    s.mored = s.dif;
                    s.moreha = HA[s.bclip];
                    s.morehb = HB[s.bclip];
                }
            }
            s.aclip = INT32_MAX;
            s.bclip = -INT32_MAX;
        }
        if (s.dif > 0) {
            /* WAVE_LAG trim of the band ends (dalign/align.c:946-955) */
            int n = s.besta - W_WAVE_LAG;
            while (s.hgh >= s.low) {
                if (V[s.hgh] < n) s.hgh -= 1;
                else {
                    while (V[s.low] < n) s.low += 1;
                    break;
                }
            }
        }
        if (!(s.more && s.lasta >= s.besta - W_TRIM_MLAG)) break;
        if (s.hgh < s.low) return BAD_INPUT;   /* empty band: undefined in the reference */

        /* next wave (dalign/align.c:634-910) */
        int am, ac, apv, ua, ub, n;
        uint64_t t;
        s.low -= 1;
        s.hgh += 1;
        if (s.low - 1 < lo_lim || s.hgh + 1 > hi_lim) return BAD_INPUT;
        if (s.low >= minp) { NA[s.low] = NA[s.low + 1]; NB[s.low] = NB[s.low + 1]; V[s.low] = -1; }
        else s.low += 1;
        if (s.hgh <= maxp) { NA[s.hgh] = NA[s.hgh - 1]; NB[s.hgh] = NB[s.hgh - 1]; V[s.hgh] = am = -1; }
        else am = V[--s.hgh];
        s.dif += 1;
        ac = V[s.hgh + 1] = V[s.low - 1] = -1;
        t = W_PATH_INT;
        n = W_PATH_LEN;
        ua = ub = -1;
        w->cells += s.hgh - s.low + 1;
        wstat(w, s.hgh - s.low + 1);

        for (k = s.hgh; k >= s.low; k--) {
            int c, m, ha, hb, y, stop, st;
            uint64_t b;
            apv = ac;
            ac = am;
            am = V[k - 1];
            if (ac < am && am >= apv) { c = am + 1; m = M[k - 1]; b = T[k - 1]; ha = HA[k - 1]; hb = HB[k - 1]; }
            else if (ac < am || ac < apv) { c = apv + 1; m = n; b = t; ha = ua; hb = ub; }
            else { c = ac + 2; m = M[k]; b = T[k]; ha = HA[k]; hb = HB[k]; }
            if (ha < 0 || hb < 0) return BAD_INPUT;

            if ((b & W_PATH_TOP) != 0) m -= 1;
            b <<= 1;
            y = (c - k) >> 1;
            if (y < 0 || y > blen || y + k < 0 || y + k > alen) return BAD_INPUT;
            y = snake(A, B, k, y, +1, &stop, &st);
            w->snake += st + 1;
            apply_matches(&m, &b, st);
            if (stop == 1) { s.more = 0; if (s.bclip < k) s.bclip = k; }
            else if (stop == 2) { s.more = 0; s.aclip = k; }
            c = (y << 1) + k;

            while (y + k >= NA[k]) {
                if (w->peb[ha].mark < NA[k]) ha = peb_push(w, ha, k, s.dif, NA[k]);
                NA[k] += TS;
            }
            while (y >= NB[k]) {
                if (w->peb[hb].mark < NB[k]) hb = peb_push(w, hb, k, s.dif, NB[k]);
                NB[k] += TS;
            }

            if (c > s.besta) {
                s.besta = c;
                s.besty = y;
                if (m >= sp->ave_path) {
                    s.lasta = c;
                    if (sp->table[b & W_TRIM_MASK] >= 0 &&
                        sp->table[(b >> W_TRIM_LEN) & W_TRIM_MASK] + sp->score[b & W_TRIM_MASK] >= 0) {
                        s.trima = c; s.trimy = y; s.trimd = s.dif; s.trimha = ha; s.trimhb = hb;
                    }
                }
            }
            t = T[k]; n = M[k]; ua = HA[k]; ub = HB[k];
            V[k] = c; T[k] = b; M[k] = m; HA[k] = ha; HB[k] = hb;
        }
    }

    /* trace assembly (dalign/align.c:970-1107) */
    {
        int trimx, atlen = 0, btlen = 0, h, kk, b, e, a, d;
        pebble *P = w->peb;
        if (s.morem >= 0) {
            trimx = s.morea - s.morey;
            s.trimy = s.morey; s.trimd = s.mored; s.trimha = s.moreha; s.trimhb = s.morehb;
        } else
            trimx = s.trima - s.trimy;

        h = chain_reverse(P, s.trimha);
        kk = P[h].diag;
        b = (mida - kk) / 2;
        e = 0;
        for (h = P[h].up; h >= 0; h = P[h].up) {
            kk = P[h].diag;
            a = P[h].mark - kk;
            d = P[h].dif;
            atr[atlen++] = (uint16_t)(d - e);
            atr[atlen++] = (uint16_t)(a - b);
            b = a;
            e = d;
        }
        if (b + kk != trimx) {
            atr[atlen++] = (uint16_t)(s.trimd - e);
            atr[atlen++] = (uint16_t)(s.trimy - b);
        } else if (b != s.trimy && atlen > 0) {
            atr[atlen - 1] = (uint16_t)(atr[atlen - 1] + (s.trimy - b));
            atr[atlen - 2] = (uint16_t)(atr[atlen - 2] + (s.trimd - e));
        }

        h = chain_reverse(P, s.trimhb);
        kk = P[h].diag;
        b = (mida + kk) / 2;
        e = 0;
        *mind = kk;
        for (h = P[h].up; h >= 0; h = P[h].up) {
            kk = P[h].diag;
            a = P[h].mark + kk;
            d = P[h].dif;
            btr[btlen++] = (uint16_t)(d - e);
            btr[btlen++] = (uint16_t)(a - b);
            b = a;
            e = d;
        }
        if (b - kk != s.trimy) {
            btr[btlen++] = (uint16_t)(s.trimd - e);
            btr[btlen++] = (uint16_t)(trimx - b);
        } else if (b != trimx && btlen > 0) {
            btr[btlen - 1] = (uint16_t)(btr[btlen - 1] + (trimx - b));
            btr[btlen - 2] = (uint16_t)(btr[btlen - 2] + (s.trimd - e));
        }

        ap->aepos = trimx;
        ap->bepos = s.trimy;
        ap->diffs = s.trimd;
        ap->tlen = atlen;
        if (comp) {
            bp->abpos = blen - ap->bepos;
            bp->bbpos = alen - ap->aepos;
        } else {
            bp->aepos = ap->bepos;
            bp->bepos = ap->aepos;
        }
        bp->diffs = s.trimd;
        bp->tlen = btlen;
    }
    return 0;
}

/* reverse_wave (dalign/align.c:1114-1916).  atr/btr point at the first slot of
 * the forward traces; this pass prepends in front of them.  On return
 * *a_first / *b_first give the index (<= 0) of the first trace entry. */
static int rev_pass(const or_spec *sp, or_work *w, const char *A0, int alen, const char *B0, int blen,
                    int comp, int mind, int maxd, int mida, int minp, int maxp,
                    or_path *ap, uint16_t *atr, int *a_first, or_path *bp, uint16_t *btr, int *b_first)
{
    const int TS = sp->tspace;
    const int boff = comp ? blen % TS : 0;
    const char *A = A0 - 1, *B = B0 - 1;   /* index y reads the base before row y */
    const int lo_lim = -w->off, hi_lim = w->cap - w->off - 1;
    int *V = w->V + w->off, *M = w->M + w->off, *HA = w->HA + w->off, *HB = w->HB + w->off;
    int *NA = w->NA + w->off, *NB = w->NB + w->off;
    uint64_t *T = w->T + w->off;
    pass_state s;
    int k;

    s.hgh = maxd;
    s.low = mind;
    s.dif = 0;
    s.more = 1;
    s.aclip = -INT32_MAX;
    s.bclip = INT32_MAX;
    s.besta = s.trima = s.morea = s.lasta = mida;
    s.besty = s.trimy = s.morey = (mida - s.hgh) >> 1;
    s.trimd = s.mored = 0;
    s.trimha = s.moreha = 0;
    s.trimhb = s.morehb = 1;
    s.morem = -1;
    w->npeb = 0;

    if (s.low - 1 < lo_lim || s.hgh + 1 > hi_lim) return BAD_INPUT;

    /* wave 0 (dalign/align.c:1193-1322) */
    for (k = s.low; k <= s.hgh; k++) {
        int y = (mida - k) >> 1, stop, st;
        int na = ((y + k + TS - 1) / TS - 1) * TS;
        int ha = peb_push(w, -1, k, 0, y + k);
        int nb = ((y + (TS - boff) - 1) / TS - 1) * TS + boff;
        int hb = peb_push(w, -1, k, 0, y);

        if (y < 0 || y > blen || y + k < 0 || y + k > alen) return BAD_INPUT;
        /* forward-style snake on the shifted views: compare B[y] / A[y+k] and step down */
        int y0 = y;
        for (;;) {
            int c = B[y];
            if (c == 4) { stop = 1; break; }
            int d = A[y + k];
            if (c != d) { stop = (d == 4) ? 2 : 0; break; }
            y -= 1;
        }
        st = y0 - y;
        w->snake += st + 1;
        if (stop == 1) { s.more = 0; if (s.bclip > k) s.bclip = k; }
        else if (stop == 2) { s.more = 0; s.aclip = k; }
        int c = (y << 1) + k;

        while (y + k <= na) { ha = peb_push(w, ha, k, 0, na); na -= TS; }
        while (y <= nb) { hb = peb_push(w, hb, k, 0, nb); nb -= TS; }

        if (c < s.besta) {
            s.besta = s.trima = s.lasta = c;
            s.besty = s.trimy = y;
            s.trimha = ha;
            s.trimhb = hb;
        }
        V[k] = c; T[k] = W_PATH_INT; M[k] = W_PATH_LEN;
        HA[k] = ha; HB[k] = hb; NA[k] = na; NB[k] = nb;
    }
    w->cells += s.hgh - s.low + 1;

    for (;;) {
        if (s.more == 0) {
            if (B[s.besty] != 4 && A[s.besta - s.besty] != 4) s.more = 1;
            if (s.low <= s.aclip) {
                s.low = s.aclip + 1;
                if (s.morem <= M[s.aclip]) {
                    s.morem = M[s.aclip];
                    s.morea = V[s.aclip];
                    s.morey = (s.morea - s.aclip) / 2;
                    s.mored = s.dif;
                    s.moreha = HA[s.aclip];
                    s.morehb = HB[s.aclip];
                }
            }
            if (s.hgh >= s.bclip) {
                s.hgh = s.bclip - 1;
                if (s.morem <= M[s.bclip]) {
                    s.morem = M[s.bclip];
                    s.morea = V[s.bclip];
                    s.morey = (s.morea - s.bclip) / 2;
                    s.mored = s.dif;
                    s.moreha = HA[s.bclip];
                    s.morehb = HB[s.bclip];
                }
            }
            s.aclip = -INT32_MAX;
            s.bclip = INT32_MAX;
        }
        if (s.dif > 0) {
            int n = s.besta + W_WAVE_LAG;
            while (s.hgh >= s.low) {
                if (V[s.hgh] > n) s.hgh -= 1;
                else {
                    while (V[s.low] > n) s.low += 1;
                    break;
                }
            }
        }
        if (!(s.more && s.lasta <= s.besta + W_TRIM_MLAG)) break;
        if (s.hgh < s.low) return BAD_INPUT;

        int am, ac, apv, ua, ub, n;
        uint64_t t;
        s.low -= 1;
        s.hgh += 1;
        if (s.low - 1 < lo_lim || s.hgh + 1 > hi_lim) return BAD_INPUT;
        if (s.low >= minp) { NA[s.low] = NA[s.low + 1]; NB[s.low] = NB[s.low + 1]; V[s.low] = apv = INT32_MAX; }
        else apv = V[++s.low];
        if (s.hgh <= maxp) { NA[s.hgh] = NA[s.hgh - 1]; NB[s.hgh] = NB[s.hgh - 1]; V[s.hgh] = INT32_MAX; }
        else s.hgh -= 1;
        s.dif += 1;
        ac = V[s.hgh + 1] = V[s.low - 1] = INT32_MAX;
        t = W_PATH_INT;
        n = W_PATH_LEN;
        ua = ub = -1;
        w->cells += s.hgh - s.low + 1;
        wstat(w, s.hgh - s.low + 1);

        for (k = s.low; k <= s.hgh; k++) {
            int c, m, ha, hb, y, stop, st;
            uint64_t b;
            am = ac;
            ac = apv;
            apv = V[k + 1];
            if (ac > apv && apv <= am) { c = apv - 1; m = M[k + 1]; b = T[k + 1]; ha = HA[k + 1]; hb = HB[k + 1]; }
            else if (ac > apv || ac > am) { c = am - 1; m = n; b = t; ha = ua; hb = ub; }
            else { c = ac - 2; m = M[k]; b = T[k]; ha = HA[k]; hb = HB[k]; }
            if (ha < 0 || hb < 0) return BAD_INPUT;

            if ((b & W_PATH_TOP) != 0) m -= 1;
            b <<= 1;
            y = (c - k) >> 1;
            if (y < 0 || y > blen || y + k < 0 || y + k > alen) return BAD_INPUT;
            int y0 = y;
            for (;;) {
                int cc = B[y];
                if (cc == 4) { stop = 1; break; }
                int d = A[y + k];
                if (cc != d) { stop = (d == 4) ? 2 : 0; break; }
                y -= 1;
            }
            st = y0 - y;
            w->snake += st + 1;
            apply_matches(&m, &b, st);
            if (stop == 1) { s.more = 0; if (s.bclip > k) s.bclip = k; }
            else if (stop == 2) { s.more = 0; s.aclip = k; }
            c = (y << 1) + k;

            while (y + k <= NA[k]) {
                if (w->peb[ha].mark > NA[k]) ha = peb_push(w, ha, k, s.dif, NA[k]);
                NA[k] -= TS;
            }
            while (y <= NB[k]) {
                if (w->peb[hb].mark > NB[k]) hb = peb_push(w, hb, k, s.dif, NB[k]);
                NB[k] -= TS;
            }

            if (c < s.besta) {
                s.besta = c;
                s.besty = y;
                if (m >= sp->ave_path) {
                    s.lasta = c;
                    if (sp->table[b & W_TRIM_MASK] >= 0 &&
                        sp->table[(b >> W_TRIM_LEN) & W_TRIM_MASK] + sp->score[b & W_TRIM_MASK] >= 0) {
                        s.trima = c; s.trimy = y; s.trimd = s.dif; s.trimha = ha; s.trimhb = hb;
                    }
                }
            }
            t = T[k]; n = M[k]; ua = HA[k]; ub = HB[k];
            V[k] = c; T[k] = b; M[k] = m; HA[k] = ha; HB[k] = hb;
        }
    }

    /* trace assembly (dalign/align.c:1709-1901) */
    {
        int trimx, atlen = 0, btlen = 0, h, kk, b, e, a, d;
        pebble *P = w->peb;
        if (s.morem >= 0) {
            trimx = s.morea - s.morey;
            s.trimy = s.morey; s.trimd = s.mored; s.trimha = s.moreha; s.trimhb = s.morehb;
        } else
            trimx = s.trima - s.trimy;

        h = chain_reverse(P, s.trimha);
        kk = P[h].diag;
        b = P[h].mark - kk;
        e = 0;
        if ((b + kk) % TS != 0) {
            h = P[h].up;
            if (h < 0) { a = s.trimy; d = s.trimd; }
            else { kk = P[h].diag; a = P[h].mark - kk; d = P[h].dif; }
            if (ap->tlen == 0) {
                atr[--atlen] = (uint16_t)(b - a);
                atr[--atlen] = (uint16_t)(d - e);
            } else {
                atr[1] = (uint16_t)(atr[1] + (b - a));
                atr[0] = (uint16_t)(atr[0] + (d - e));
            }
            b = a;
            e = d;
        }
        if (h >= 0) {
            for (h = P[h].up; h >= 0; h = P[h].up) {
                kk = P[h].diag;
                a = P[h].mark - kk;
                atr[--atlen] = (uint16_t)(b - a);
                d = P[h].dif;
                atr[--atlen] = (uint16_t)(d - e);
                b = a;
                e = d;
            }
            if (b + kk != trimx) {
                atr[--atlen] = (uint16_t)(b - s.trimy);
                atr[--atlen] = (uint16_t)(s.trimd - e);
            } else if (b != s.trimy && (atlen < 0 || ap->tlen > 0)) {
                atr[atlen + 1] = (uint16_t)(atr[atlen + 1] + (b - s.trimy));
                atr[atlen] = (uint16_t)(atr[atlen] + (s.trimd - e));
            }
        }

        h = chain_reverse(P, s.trimhb);
        kk = P[h].diag;
        b = P[h].mark + kk;
        e = 0;
        if ((b - kk) % TS != boff) {
            h = P[h].up;
            if (h < 0) { a = trimx; d = s.trimd; }
            else { kk = P[h].diag; a = P[h].mark + kk; d = P[h].dif; }
            if (bp->tlen == 0) {
                /* the reference stores the b-advance in both slots here (dalign/align.c:1831-1832) */
                btr[--btlen] = (uint16_t)(b - a);
                btr[--btlen] = (uint16_t)(b - a);
            } else {
                btr[1] = (uint16_t)(btr[1] + (b - a));
                btr[0] = (uint16_t)(btr[0] + (d - e));
            }
            b = a;
            e = d;
        }
        if (h >= 0) {
            for (h = P[h].up; h >= 0; h = P[h].up) {
                kk = P[h].diag;
                a = P[h].mark + kk;
                btr[--btlen] = (uint16_t)(b - a);
                d = P[h].dif;
                btr[--btlen] = (uint16_t)(d - e);
                b = a;
                e = d;
            }
            if (b - kk != s.trimy) {
                btr[--btlen] = (uint16_t)(b - trimx);
                btr[--btlen] = (uint16_t)(s.trimd - e);
            } else if (b != trimx && (btlen < 0 || bp->tlen > 0)) {
                btr[btlen + 1] = (uint16_t)(btr[btlen + 1] + (b - trimx));
                btr[btlen] = (uint16_t)(btr[btlen] + (s.trimd - e));
            }
        }

        ap->abpos = trimx;
        ap->bbpos = s.trimy;
        ap->diffs += s.trimd;
        ap->tlen -= atlen;
        if (comp) {
            bp->aepos = blen - ap->bbpos;
            bp->bepos = alen - ap->abpos;
        } else {
            bp->abpos = ap->bbpos;
            bp->bbpos = ap->abpos;
        }
        bp->diffs += s.trimd;
        bp->tlen -= btlen;
        *a_first = atlen;
        *b_first = btlen;
    }
    return 0;
}

/* Local_Alignment (dalign/align.c:1922-2060). */
int or_local_alignment(const or_spec *sp, or_work *w, const char *aseq, int alen,
                       const char *bseq, int blen, int comp, int low, int hgh,
                       int anti, int lbord, int hbord, or_path *apath,
                       const uint16_t **atrace, or_path *bpath, const uint16_t **btrace)
{
    int minp, maxp, maxtp, rc, af, bf;
    int selfie = (aseq == bseq);
    uint16_t *atr, *btr;

    maxtp = 2 * (((alen < blen) ? blen : alen) / sp->tspace + 2);
    if (4 * maxtp > w->trcap) {
        w->trcap = 4 * maxtp + 1024;
        w->tr = xrealloc(w->tr, sizeof(uint16_t) * w->trcap);
    }
    atr = w->tr + maxtp;
    btr = atr + 2 * maxtp;
    band_reserve(w, alen, blen);
    w->cells = w->snake = 0;

    if (lbord < 0) minp = (selfie && low >= 0) ? 1 : -INT32_MAX;
    else minp = low - lbord;
    if (hbord < 0) maxp = (selfie && hgh <= 0) ? -1 : INT32_MAX;
    else maxp = hgh + hbord;

    memset(apath, 0, sizeof(*apath));
    memset(bpath, 0, sizeof(*bpath));
    rc = fwd_pass(sp, w, aseq, alen, bseq, blen, comp, &low, hgh, anti, minp, maxp, apath, atr, bpath, btr);
    if (rc) return rc;
    rc = rev_pass(sp, w, aseq, alen, bseq, blen, comp, low, low, anti, minp, maxp, apath, atr, &af, bpath, btr, &bf);
    if (rc) return rc;
    atr += af;
    btr += bf;

    if (comp) {   /* B-view trace runs the other way (dalign/align.c:2004-2023) */
        int i = bpath->tlen - 2, j = 0;
        while (j < i) {
            uint16_t p;
            p = btr[i]; btr[i] = btr[j]; btr[j] = p;
            p = btr[i + 1]; btr[i + 1] = btr[j + 1]; btr[j + 1] = p;
            i -= 2;
            j += 2;
        }
    }
    *atrace = atr;
    *btrace = btr;
    return 0;
}

/* ======================================================================== */
/* DB block loading (db/DB.c:360-585 Open_DB_Block, :1349-1410            */
/* Read_All_Sequences; stub/index layouts db/DB.h:164-249).                */

static int read_stub_block(const char *dbfile, int part, int *ufirst, int *ulast)
{
    FILE *f = fopen(dbfile, "r");
    int nfiles, nblocks, i, last = 0;
    long long size;
    char a[4096], b[4096];
    if (f == NULL) return -1;
    if (fscanf(f, "files = %9d\n", &nfiles) != 1) { fclose(f); return -1; }
    for (i = 0; i < nfiles; i++)
        if (fscanf(f, "  %9d %4095s %4095s\n", &last, a, b) != 3) { fclose(f); return -1; }
    if (fscanf(f, "blocks = %9d\n", &nblocks) != 1) {
        fclose(f);
        if (part != 0) return -1;
        *ufirst = 0;
        *ulast = -1;
        return 0;
    }
    if (fscanf(f, "size = %9lld\n", &size) != 1 || part > nblocks) { fclose(f); return -1; }
    if (part > 0) {
        int v = 0;
        for (i = 1; i <= part; i++)
            if (fscanf(f, " %9d\n", &v) != 1) { fclose(f); return -1; }
        *ufirst = v;
        if (fscanf(f, " %9d\n", &v) != 1) { fclose(f); return -1; }
        *ulast = v;
    } else {
        *ufirst = 0;
        *ulast = -1;
    }
    fclose(f);
    return 0;
}

int or_block_load(or_block *b, const char *path)
{
    char root[4096], dir[4096], fn[8192];
    const char *slash, *dot;
    int part = 0, ufirst, ulast, i;
    FILE *idx, *bps;
    unsigned char hdr[88];

    memset(b, 0, sizeof(*b));
    slash = strrchr(path, '/');
    if (slash) {
        snprintf(dir, sizeof(dir), "%.*s", (int)(slash - path), path);
        snprintf(root, sizeof(root), "%s", slash + 1);
    } else {
        snprintf(dir, sizeof(dir), ".");
        snprintf(root, sizeof(root), "%s", path);
    }
    size_t rl = strlen(root);
    if (rl > 4 && strcmp(root + rl - 4, ".dam") == 0) root[rl - 4] = 0;
    else if (rl > 3 && strcmp(root + rl - 3, ".db") == 0) root[rl - 3] = 0;
    dot = strrchr(root, '.');
    if (dot && dot[1] != 0 && dot[1] != '-') {
        char *end;
        long p = strtol(dot + 1, &end, 10);
        if (*end == 0 && p != 0) {
            part = (int)p;
            root[dot - root] = 0;
        }
    }
    /* the stub is ROOT.db, else ROOT.dam (db/DB.c:407-421) */
    snprintf(fn, sizeof(fn), "%s/%s.db", dir, root);
    if (access(fn, F_OK) != 0) snprintf(fn, sizeof(fn), "%s/%s.dam", dir, root);
    if (read_stub_block(fn, part, &ufirst, &ulast)) return -1;
    snprintf(fn, sizeof(fn), "%s/.%s.idx", dir, root);
    idx = fopen(fn, "rb");
    if (idx == NULL) return -1;
    if (fread(hdr, 88, 1, idx) != 1) { fclose(idx); return -1; }
    int ureads;
    memcpy(&ureads, hdr + 0, 4);
    memcpy(b->freq, hdr + 4, 16);
    memcpy(&b->maxlen, hdr + 20, 4);
    memcpy(&b->totlen, hdr + 24, 8);
    if (ulast < 0) ulast = ureads;
    b->part = part;
    b->ufirst = ufirst;
    b->nreads = ulast - ufirst;
    b->rlen = (int *)xrealloc(NULL, sizeof(int) * (b->nreads + 1));
    b->boff = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (b->nreads + 1));
    int64_t *foff = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (b->nreads + 1));
    if (fseeko(idx, 88 + 32LL * ufirst, SEEK_SET) != 0) { fclose(idx); return -1; }
    int64_t tot = 0;
    int mx = 0;
    for (i = 0; i < b->nreads; i++) {
        unsigned char r[32];
        if (fread(r, 32, 1, idx) != 1) { fclose(idx); return -1; }
        memcpy(&b->rlen[i], r + 0, 4);
        memcpy(&foff[i], r + 8, 8);
        tot += b->rlen[i];
        if (b->rlen[i] > mx) mx = b->rlen[i];
    }
    fclose(idx);
    if (part > 0) {
        b->maxlen = mx;
        b->totlen = tot;
    }
    b->alloc = (char *)xrealloc(NULL, tot + b->nreads + 4);
    b->alloc[0] = 4;
    b->bases = b->alloc + 1;
    snprintf(fn, sizeof(fn), "%s/.%s.bps", dir, root);
    bps = fopen(fn, "rb");
    if (bps == NULL) return -1;
    int64_t o = 0;
    unsigned char *pk = NULL;
    size_t pkcap = 0;
    for (i = 0; i < b->nreads; i++) {
        int len = b->rlen[i], cl = (len + 3) / 4, j;
        if ((size_t)cl > pkcap) { pkcap = cl + 1024; pk = xrealloc(pk, pkcap); }
        if (fseeko(bps, foff[i], SEEK_SET) != 0 || (cl > 0 && fread(pk, cl, 1, bps) != 1)) { fclose(bps); return -1; }
        for (j = 0; j < len; j++) b->bases[o + j] = (char)((pk[j >> 2] >> (6 - 2 * (j & 3))) & 3);
        b->bases[o + len] = 4;
        b->boff[i] = o;
        o += len + 1;
    }
    b->boff[b->nreads] = o;
    free(pk);
    free(foff);
    fclose(bps);
    b->ureads = ureads;
    b->stem = (char *)xrealloc(NULL, strlen(dir) + strlen(root) + 3);
    sprintf(b->stem, "%s/.%s", dir, root);
    return 0;
}

void or_block_free(or_block *b)
{
    free(b->rlen);
    free(b->boff);
    free(b->alloc);
    free(b->stem);
    free(b->manno);
    free(b->mdata);
    memset(b, 0, sizeof(*b));
}

void or_block_complement(const or_block *src, or_block *dst)
{
    int i;
    *dst = *src;
    dst->rlen = (int *)xrealloc(NULL, sizeof(int) * (src->nreads + 1));
    dst->boff = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (src->nreads + 1));
    memcpy(dst->rlen, src->rlen, sizeof(int) * (src->nreads + 1));
    memcpy(dst->boff, src->boff, sizeof(int64_t) * (src->nreads + 1));
    dst->alloc = (char *)xrealloc(NULL, src->boff[src->nreads] + 1);
    dst->alloc[0] = 4;
    dst->bases = dst->alloc + 1;
    memcpy(dst->bases, src->bases, src->boff[src->nreads]);
    dst->freq[0] = src->freq[3];
    dst->freq[3] = src->freq[0];
    dst->freq[1] = src->freq[2];
    dst->freq[2] = src->freq[1];
    for (i = 0; i < src->nreads; i++) {
        char *s = dst->bases + dst->boff[i];
        int n = dst->rlen[i], l = 0, r = n - 1;
        while (l < r) {
            char c = s[l];
            s[l] = (char)(3 - s[r]);
            s[r] = (char)(3 - c);
            l++;
            r--;
        }
        if (l == r) s[l] = (char)(3 - s[l]);
    }
    /* the complemented mask: each read's points reversed and mapped x -> rlen - x
     * (dalign/daligner.c:519-560) */
    dst->stem = NULL;
    dst->manno = NULL;
    dst->mdata = NULL;
    if (src->manno) {
        int64_t nd = src->manno[src->nreads];
        dst->manno = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (src->nreads + 1));
        dst->mdata = (int32_t *)xrealloc(NULL, sizeof(int32_t) * (nd + 1));
        memcpy(dst->manno, src->manno, sizeof(int64_t) * (src->nreads + 1));
        for (i = 0; i < src->nreads; i++)
            for (int64_t k = src->manno[i], j = src->manno[i + 1] - 1; k < src->manno[i + 1]; k++, j--)
                dst->mdata[k] = src->rlen[i] - src->mdata[j];
    }
}

/* ======================================================================== */
/* Mask tracks of daligner -m: read_DB / Merge_Size / Merge_Tracks
 * (dalign/daligner.c:193-431), Load_Track (db/DB.c:967-1120), track_load /
 * track_load_block (lib/tracks.c:19-170), uncompress_chunks
 * (lib/compression.c:64-100).  A track comes back as anno[nreads+1] (in
 * points, as read_DB leaves it after dividing the byte offsets by 4) and
 * the block's slice of the points. */

static int read_file(const char *path, unsigned char **buf, int64_t *len)
{
    FILE *f = fopen(path, "rb");
    if (f == NULL) return -1;
    fseeko(f, 0, SEEK_END);
    *len = ftello(f);
    fseeko(f, 0, SEEK_SET);
    *buf = (unsigned char *)xrealloc(NULL, *len + 1);
    int ok = *len == 0 || fread(*buf, *len, 1, f) == 1;
    fclose(f);
    return ok ? 0 : -1;
}

/* zlib chunks: u64 compressed length, then the stream of <= 8 MiB of input */
static int unchunk(const unsigned char *in, int64_t len, unsigned char *out, int64_t olen)
{
    int64_t i = 0, o = 0;
    while (i < len) {
        uint64_t c;
        if (i + 8 > len) return -1;
        memcpy(&c, in + i, 8);
        if ((int64_t)c > len - i - 8) return -1;
        uLongf d = (uLongf)(olen - o);
        if (uncompress(out + o, &d, in + i + 8, (uLong)c) != Z_OK) return -1;
        i += 8 + (int64_t)c;
        o += (int64_t)d;
    }
    return o == olen ? 0 : -1;
}

/* MARVEL layout .ROOT.NAME.a2 / .d2.  The block's slice is
 * anno[ufirst .. ufirst+nreads] rebased to 0.  The reference trims a block
 * with ufirst > 1 twice (track_load at lib/tracks.c:151-163, then
 * track_load_block at :40-58), which reads outside the block's slice; the
 * single trim here is the intended one. */
static int load_marvel(const or_block *b, const char *name, int64_t **anno, int32_t **data, char *err, int el)
{
    char fn[8192];
    unsigned char *a = NULL, *d = NULL;
    int64_t alen, dlen;
    snprintf(fn, sizeof fn, "%s.%s.a2", b->stem, name);
    if (read_file(fn, &a, &alen)) return 1;   /* absent: the caller tries the Dazzler layout */
    uint16_t size;
    uint64_t len, clen, cdlen;
    if (alen < 64) { snprintf(err, el, "%s: short header", fn); free(a); return -1; }
    memcpy(&size, a + 2, 2);
    memcpy(&len, a + 8, 8);
    memcpy(&clen, a + 16, 8);
    memcpy(&cdlen, a + 24, 8);
    if (size != 8 || (b->part == 0 && len != (uint64_t)b->nreads) || len < (uint64_t)(b->ufirst + b->nreads) ||
        (int64_t)clen > alen - 64) {
        snprintf(err, el, "%s: invalid header", fn);
        free(a);
        return -1;
    }
    int64_t *full = (int64_t *)xrealloc(NULL, 8 * (len + 1));
    if (unchunk(a + 64, (int64_t)clen, (unsigned char *)full, 8 * (int64_t)(len + 1))) {
        snprintf(err, el, "%s: bad compressed offsets", fn);
        free(a); free(full);
        return -1;
    }
    free(a);
    snprintf(fn, sizeof fn, "%s.%s.d2", b->stem, name);
    int64_t tot = full[len];
    unsigned char *all = (unsigned char *)xrealloc(NULL, tot + 4);
    if (tot > 0 && (read_file(fn, &d, &dlen) || (int64_t)cdlen > dlen || unchunk(d, (int64_t)cdlen, all, tot))) {
        snprintf(err, el, "%s: missing or bad point data", fn);
        free(d); free(all); free(full);
        return -1;
    }
    free(d);
    int64_t lo = full[b->ufirst];
    *anno = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (b->nreads + 1));
    for (int i = 0; i <= b->nreads; i++) (*anno)[i] = (full[b->ufirst + i] - lo) / 4;
    int64_t np = (*anno)[b->nreads];
    *data = (int32_t *)xrealloc(NULL, sizeof(int32_t) * (np + 1));
    memcpy(*data, all + lo, 4 * np);
    free(all);
    free(full);
    return 0;
}

/* Dazzler layout: .ROOT.<part>.NAME.anno/.data when present, else the
 * whole-DB .ROOT.NAME.anno/.data read from the block's first read on. */
static int load_dazzler(const or_block *b, const char *name, int64_t **anno, int32_t **data, char *err, int el)
{
    char fn[8192];
    FILE *af = NULL, *df;
    int ispart = 0, tlen, size;
    if (b->part > 0) {
        snprintf(fn, sizeof fn, "%s.%d.%s.anno", b->stem, b->part, name);
        af = fopen(fn, "rb");
        ispart = af != NULL;
    }
    if (af == NULL) {
        snprintf(fn, sizeof fn, "%s.%s.anno", b->stem, name);
        af = fopen(fn, "rb");
    }
    if (af == NULL) { snprintf(err, el, "track %s does not exist", name); return -1; }
    if (fread(&tlen, 4, 1, af) != 1 || fread(&size, 4, 1, af) != 1 || size != 8 ||
        tlen != (ispart ? b->nreads : b->ureads)) {
        snprintf(err, el, "%s: junk or wrong size annotation", fn);
        fclose(af);
        return -1;
    }
    if (!ispart && b->part > 0) fseeko(af, 8LL * b->ufirst, SEEK_CUR);
    int64_t *an = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (b->nreads + 1));
    if (fread(an, 8, b->nreads + 1, af) != (size_t)(b->nreads + 1)) {
        snprintf(err, el, "%s: short annotation", fn);
        fclose(af); free(an);
        return -1;
    }
    fclose(af);
    if (ispart) snprintf(fn, sizeof fn, "%s.%d.%s.data", b->stem, b->part, name);
    else snprintf(fn, sizeof fn, "%s.%s.data", b->stem, name);
    int64_t off = an[0];
    for (int i = 0; i <= b->nreads; i++) an[i] = (an[i] - off) / 4;
    int64_t np = an[b->nreads];
    int32_t *dt = (int32_t *)xrealloc(NULL, sizeof(int32_t) * (np + 1));
    df = fopen(fn, "rb");
    if (df == NULL || fseeko(df, off, SEEK_SET) != 0 || (np > 0 && fread(dt, 4 * np, 1, df) != 1)) {
        snprintf(err, el, "%s: missing or short point data", fn);
        if (df) fclose(df);
        free(an); free(dt);
        return -1;
    }
    fclose(df);
    *anno = an;
    *data = dt;
    return 0;
}

/* A track the extraction can use: per read an even number of points, non
 * decreasing, within [0, rlen].  (Other tracks make tuple_thread read
 * outside the read, dalign/filter.c:424-452.) */
static int check_points(const or_block *b, const int64_t *anno, const int32_t *data)
{
    for (int i = 0; i < b->nreads; i++) {
        if (anno[i + 1] < anno[i] || ((anno[i + 1] - anno[i]) & 1)) return -1;
        int prev = 0;
        for (int64_t k = anno[i]; k < anno[i + 1]; k++) {
            if (data[k] < prev || data[k] > b->rlen[i]) return -1;
            prev = data[k];
        }
    }
    return 0;
}

int or_block_load_masks(or_block *b, char *const *names, int n, char *err, int errlen)
{
    int64_t **an = (int64_t **)xrealloc(NULL, sizeof(int64_t *) * (n + 1));
    int32_t **dt = (int32_t **)xrealloc(NULL, sizeof(int32_t *) * (n + 1));
    int t, rc = 0;
    for (t = 0; t < n; t++) {
        for (int u = 0; u < t; u++)
            if (strcmp(names[u], names[t]) == 0) {
                snprintf(err, errlen, "track %s given twice", names[t]);
                rc = -1;
            }
        if (rc == 0) {
            rc = load_marvel(b, names[t], an + t, dt + t, err, errlen);
            if (rc == 1) rc = load_dazzler(b, names[t], an + t, dt + t, err, errlen);
        }
        if (rc == 0 && check_points(b, an[t], dt[t])) {
            snprintf(err, errlen, "track %s: points not in increasing pairs within the reads", names[t]);
            free(an[t]); free(dt[t]);
            rc = -1;
        }
        if (rc) break;
    }
    if (rc) {
        for (int u = 0; u < t; u++) { free(an[u]); free(dt[u]); }
        free(an); free(dt);
        return -1;
    }
    if (n == 1) {
        b->manno = an[0];
        b->mdata = dt[0];
    } else if (n > 1) {
        /* union of the tracks: a sweep over every read's points in increasing
         * order, each toggling its own track in or out; a point is kept where
         * the number of tracks inside goes from or to 0.  Points at equal
         * positions may come in any order here and in the reference's heap:
         * the kept lists then differ only by empty gaps, which hold no k-mer. */
        int64_t cap = 0;
        for (t = 0; t < n; t++) cap += an[t][b->nreads];
        b->manno = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (b->nreads + 1));
        b->mdata = (int32_t *)xrealloc(NULL, sizeof(int32_t) * (cap + 1));
        int64_t *cur = (int64_t *)xrealloc(NULL, sizeof(int64_t) * n);
        int *in = (int *)xrealloc(NULL, sizeof(int) * n);
        int64_t o = 0;
        for (int r = 0; r < b->nreads; r++) {
            int level = 0;
            b->manno[r] = o;
            for (t = 0; t < n; t++) { cur[t] = an[t][r]; in[t] = 0; }
            for (;;) {
                int best = -1;
                for (t = 0; t < n; t++)
                    if (cur[t] < an[t][r + 1] && (best < 0 || dt[t][cur[t]] < dt[best][cur[best]])) best = t;
                if (best < 0) break;
                int x = dt[best][cur[best]++];
                in[best] ^= 1;
                if (in[best]) {
                    if (level++ == 0) b->mdata[o++] = x;
                } else if (--level == 0) {
                    b->mdata[o++] = x;
                }
            }
        }
        b->manno[b->nreads] = o;
        free(cur);
        free(in);
        for (t = 0; t < n; t++) { free(an[t]); free(dt[t]); }
    }
    free(an);
    free(dt);
    return 0;
}

/* ======================================================================== */
/* Options (dalign/daligner.c:637-649 defaults).                            */

void or_opts_default(or_opts *o)
{
    memset(o, 0, sizeof(*o));
    o->kmer = 14;
    o->binshift = 6;
    o->hitmin = 35;
    o->suppress = 0;
    o->minover = 2 * 1000;
    o->tspace = 100;
    o->nthreads = 4;
    o->nshift = 2;
    o->symmetric = 1;
    o->run_id = 1;
    o->ave_corr = .70;
    o->mem_limit = 0;
}

/* ======================================================================== */
/* k-mer index: tuple_thread + lex_sort + compress (dalign/filter.c:400-490, */
/* :168-377, :645-696, :698-903).  Final order (code, read, rpos).          */

typedef struct {
    uint64_t code;
    int32_t rpos, read;
} kmer_t;

/* stable LSD radix sort of 16-byte records by a 64-bit key extracted by
 * shift/width, 11-bit digits */
static void radix_kmers(kmer_t **src, kmer_t **tmp, int64_t n, int bits)
{
    const int D = 11;
    int64_t *cnt = (int64_t *)xrealloc(NULL, sizeof(int64_t) << D);
    for (int sh = 0; sh < bits; sh += D) {
        memset(cnt, 0, sizeof(int64_t) << D);
        kmer_t *s = *src, *t = *tmp;
        for (int64_t i = 0; i < n; i++) cnt[(s[i].code >> sh) & ((1 << D) - 1)]++;
        int64_t x = 0;
        for (int j = 0; j < (1 << D); j++) { int64_t c = cnt[j]; cnt[j] = x; x += c; }
        for (int64_t i = 0; i < n; i++) t[cnt[(s[i].code >> sh) & ((1 << D) - 1)]++] = s[i];
        *src = t;
        *tmp = s;
    }
    free(cnt);
}

/* -b weights (dalign/filter.c:718-732): per base ceil(-10000 log4 freq),
 * computed once, from the block of the first biased index a process
 * builds, and kept for every later block (its complement included). */
static int bias_set, LogBase[4];

/* biased_tuple_thread for one run of positions [p, q) of a read
 * (dalign/filter.c:492-633): a window grows forward until its weight reaches
 * 10000 K, sheds its oldest bases while it stays there, and yields the
 * window's code left-aligned to 2K bits when its weight exceeds
 * 10000 (K - 2).  Without a mask (masked == 0) a window that runs into the
 * read's end ends the read; with one it stops at q and may still yield a
 * k-mer ending at q. */
static int64_t biased_run(const char *s, int p, int q, int K, int masked, int read, kmer_t *list, int64_t m)
{
    const uint64_t kmask = (K >= 32) ? ~0ULL : ((1ULL << (2 * K)) - 1);
    const int lognorm = 10000 * K, logthresh = 10000 * (K - 2);
    const char *t = s + 1;
    uint64_t c = 0;
    int a = 0, k = 1, x;
    while (p < q) {
        x = s[p];
        a += LogBase[x];
        c = (c << 2) | (uint64_t)x;
        while (a < lognorm && k < K) {
            if (++p >= q) {
                if (!masked) return m;
                break;
            }
            k += 1;
            x = s[p];
            a += LogBase[x];
            c = (c << 2) | (uint64_t)x;
        }
        for (;;) {
            int u = a - LogBase[(int)t[p - k]];
            if (u < lognorm) break;
            a = u;
            k -= 1;
        }
        if (a > logthresh) {
            list[m].code = (c << (2 * K - 2 * k)) & kmask;
            list[m].read = read;
            list[m].rpos = p;
            m++;
        }
        p += 1;
        a -= LogBase[(int)s[p - k]];
    }
    return m;
}

static kmer_t *sort_kmers(const or_block *b, int K, int suppress, int biased, int64_t *nk_out)
{
    int64_t n = b->boff[b->nreads] - (int64_t)K * b->nreads, m = 0;
    uint64_t kmask = (K >= 32) ? ~0ULL : ((1ULL << (2 * K)) - 1);
    if (n <= 0) { *nk_out = 0; return NULL; }
    kmer_t *list = (kmer_t *)xrealloc(NULL, sizeof(kmer_t) * (n + 1));
    kmer_t *tmp = (kmer_t *)xrealloc(NULL, sizeof(kmer_t) * (n + 1));
    if (biased && !bias_set) {
        const double scale = -10000. / log(4.);
        for (int i = 0; i < 4; i++) LogBase[i] = (int)ceil(scale * log(b->freq[i]));
        bias_set = 1;
    }
    for (int r = 0; r < b->nreads; r++) {
        const char *s = b->bases + b->boff[r];
        uint64_t c = 0;
        int p = 0, x;
        if (biased) {
            if (b->manno == NULL) {
                m = biased_run(s, 0, b->rlen[r], K, 0, r, list, m);
                continue;
            }
            for (int64_t a = b->manno[r]; a <= b->manno[r + 1]; a += 2) {
                int q;
                p = (a == b->manno[r]) ? 0 : b->mdata[a - 1];
                q = (a == b->manno[r + 1]) ? b->rlen[r] : b->mdata[a];
                if (p + K <= q) m = biased_run(s, p, q, K, 1, r, list, m);
            }
            continue;
        }
        if (b->manno) {
            /* masked: only k-mers wholly inside a gap between masked
             * intervals, [0, pt0), [pt1, pt2), ..., [ptlast, rlen)
             * (dalign/filter.c:416-452) */
            for (int64_t a = b->manno[r]; a <= b->manno[r + 1]; a += 2) {
                int q;
                p = (a == b->manno[r]) ? 0 : b->mdata[a - 1];
                q = (a == b->manno[r + 1]) ? b->rlen[r] : b->mdata[a];
                if (p + K > q) continue;
                c = 0;
                for (x = 1; x < K; x++) c = (c << 2) | s[p++];
                while (p < q) {
                    c = ((c << 2) | s[p]) & kmask;
                    list[m].code = c;
                    list[m].read = r;
                    list[m].rpos = p++;
                    m++;
                }
            }
            continue;
        }
        for (x = 1; x < K; x++) c = (c << 2) | s[p++];
        while ((x = s[p]) != 4) {
            c = ((c << 2) | x) & kmask;
            list[m].code = c;
            list[m].read = r;
            list[m].rpos = p++;
            m++;
        }
    }
    radix_kmers(&list, &tmp, m, 2 * K);
    list[m].code = 1ULL << (2 * K);
    if (suppress > 0) {   /* compsize/compress: drop groups of size >= t */
        int64_t i = 0, o = 0;
        while (i < m) {
            int64_t p = i++;
            while (i < m && list[i].code == list[p].code) i++;
            if (i - p < suppress)
                while (p < i) tmp[o++] = list[p++];
        }
        kmer_t *z = list; list = tmp; tmp = z;
        m = o;
        list[m].code = 1ULL << (2 * K);
    }
    free(tmp);
    *nk_out = m;
    return list;
}

/* ======================================================================== */
/* Seeds: count_thread / limit / merge_thread / pair sort                  */
/* (dalign/filter.c:948-1231, :2097-2267).                                  */

typedef struct {
    int32_t diag, apos, aread, bread;
} seed_t;

#define MAXGRAM 10000

/* B entries of a code group that pair with A entry a in self mode
 * (dalign/filter.c:984-1001, :1101-1146): reads below a's, and with -I
 * also a's own read -- at positions below a's, or wholly against the
 * complement.  mode 1: self; 2: self -I; 3: self -I, complement. */
static inline int pairs_below(int mode, const kmer_t *b, const kmer_t *a)
{
    if (mode == 3) return b->read <= a->read;
    if (mode == 2) return b->read < a->read || (b->read == a->read && b->rpos < a->rpos);
    return b->read < a->read;
}

static seed_t *make_seeds(const or_opts *o, kmer_t *as, int64_t na, kmer_t *bs, int64_t nb,
                          int self, int same_index, int64_t *nseed, int *limit_out)
{
    int64_t *hist = (int64_t *)calloc(MAXGRAM, sizeof(int64_t));
    int64_t ia = 0, ib = 0, limit, nh = 0, total = 0;
    /* pass 1: group sizes (the merge-join over equal codes) */
    while (ia < na && ib < nb) {
        if (as[ia].code < bs[ib].code) { ia++; continue; }
        if (as[ia].code > bs[ib].code) { ib++; continue; }
        uint64_t c = as[ia].code;
        int64_t ja = ia, jb = ib, ct = 0;
        while (ia < na && as[ia].code == c) ia++;
        while (ib < nb && bs[ib].code == c) ib++;
        if (self) {
            int64_t q = jb;
            for (int64_t a = ja; a < ia; a++) {
                while (q < ib && pairs_below(self, &bs[q], &as[a])) q++;
                ct += q - jb;
            }
        } else
            ct = (ia - ja) * (ib - jb);
        if (ct < MAXGRAM) hist[ct] += 1;
        total += ct;
    }
    /* cap from the memory limit (dalign/filter.c:2107-2174) */
    {
        uint64_t M16 = o->mem_limit / 16;
        int64_t avail;
        if (same_index || (int64_t)M16 > na + 2 * nb)
            avail = (int64_t)((M16 - (uint64_t)na) / 2);
        else
            avail = (int64_t)(M16 - (uint64_t)(na + nb));
        avail = (int64_t)(avail * .98);
        int64_t tom = 0;
        int j;
        for (j = 0; j < MAXGRAM; j++) {
            tom += j * hist[j];
            if (tom > avail) break;
        }
        limit = j;
        if (o->mem_limit == 0) limit = INT32_MAX;
        if (limit == 1) {
            fprintf(stderr, "oracle: insufficient memory limit\n");
            exit(1);
        }
        if (limit > MAXGRAM) nh = total;
        else
            for (j = 1; j < limit; j++) nh += j * hist[j];
    }
    *limit_out = (int)limit;
    free(hist);
    seed_t *h = (seed_t *)xrealloc(NULL, sizeof(seed_t) * (nh + 1));
    int64_t n = 0;
    ia = ib = 0;
    while (ia < na && ib < nb) {
        if (as[ia].code < bs[ib].code) { ia++; continue; }
        if (as[ia].code > bs[ib].code) { ib++; continue; }
        uint64_t c = as[ia].code;
        int64_t ja = ia, jb = ib, ct = 0;
        while (ia < na && as[ia].code == c) ia++;
        while (ib < nb && bs[ib].code == c) ib++;
        if (self) {
            int64_t q = jb;
            for (int64_t a = ja; a < ia; a++) {
                while (q < ib && pairs_below(self, &bs[q], &as[a])) q++;
                ct += q - jb;
            }
        } else
            ct = (ia - ja) * (ib - jb);
        if (ct >= limit) continue;
        int64_t q = jb;
        for (int64_t a = ja; a < ia; a++) {
            int64_t e = ib;
            if (self) {
                while (q < ib && pairs_below(self, &bs[q], &as[a])) q++;
                e = q;
            }
            for (int64_t bb = jb; bb < e; bb++) {
                h[n].bread = bs[bb].read;
                h[n].aread = as[a].read;
                h[n].apos = as[a].rpos;
                h[n].diag = as[a].rpos - bs[bb].rpos;
                n++;
            }
        }
    }
    /* stable LSD by apos, then aread, then bread (dalign/filter.c:2030-2054) */
    seed_t *t = (seed_t *)xrealloc(NULL, sizeof(seed_t) * (n + 1));
    int64_t *cnt = (int64_t *)xrealloc(NULL, sizeof(int64_t) << 16);
    for (int f = 0; f < 3; f++)
        for (int sh = 0; sh < 32; sh += 16) {
            int32_t mxv = 0;
            for (int64_t i = 0; i < n; i++) {
                int32_t v = f == 0 ? h[i].apos : f == 1 ? h[i].aread : h[i].bread;
                if (v > mxv) mxv = v;
            }
            if (sh > 0 && (mxv >> sh) == 0) continue;
            memset(cnt, 0, sizeof(int64_t) << 16);
            for (int64_t i = 0; i < n; i++) {
                uint32_t v = (uint32_t)(f == 0 ? h[i].apos : f == 1 ? h[i].aread : h[i].bread);
                cnt[(v >> sh) & 0xffff]++;
            }
            int64_t x = 0;
            for (int j = 0; j < 65536; j++) { int64_t c = cnt[j]; cnt[j] = x; x += c; }
            for (int64_t i = 0; i < n; i++) {
                uint32_t v = (uint32_t)(f == 0 ? h[i].apos : f == 1 ? h[i].aread : h[i].bread);
                t[cnt[(v >> sh) & 0xffff]++] = h[i];
            }
            seed_t *z = h; h = t; t = z;
        }
    free(cnt);
    free(t);
    h[n].aread = h[n].bread = h[n].diag = 0x7fffffff;
    h[n].apos = 0;
    *nseed = n;
    return h;
}

/* ======================================================================== */
/* Redundancy handling (dalign/filter.c:1249-1566).                          */

typedef struct {
    int tlen, diffs, abpos, bbpos, aepos, bepos;
    int64_t toff;   /* offset into the per-pair u16 trace pool */
} mpath;

typedef struct {
    uint16_t *t;
    int64_t n, cap;
} u16pool;

static int64_t pool_reserve(u16pool *p, int64_t len)
{
    if (p->n + len > p->cap) {
        p->cap = (p->n + len) * 2 + 4096;
        p->t = xrealloc(p->t, sizeof(uint16_t) * p->cap);
    }
    int64_t o = p->n;
    p->n += len;
    return o;
}

/* Entwine: minimum |b-offset| between two paths over their shared A trace
 * points; *where = last A coordinate with distance 0.  -1 when they share no
 * interior trace point. */
static int entwine(const mpath *jp, const mpath *kp, const u16pool *P, int TS, int *where)
{
    const uint16_t *jt = P->t + jp->toff, *kt = P->t + kp->toff;
    int mn = 10000, den = 0, y2 = jp->bbpos, b2 = kp->bbpos, ac, i, j, k, ae;
    int jseg = jp->abpos / TS, kseg = kp->abpos / TS;
    if (jseg < kseg) {
        ac = kseg * TS;
        j = 1 + 2 * (kseg - jseg);
        k = 1;
        for (i = 1; i < j; i += 2) y2 += jt[i];
    } else {
        ac = jseg * TS;
        k = 1 + 2 * (jseg - kseg);
        j = 1;
        for (i = 1; i < k; i += 2) b2 += kt[i];
    }
    ae = jp->aepos < kp->aepos ? jp->aepos : kp->aepos;
    for (;;) {
        ac += TS;
        if (ac >= ae) break;
        y2 += jt[j];
        b2 += kt[k];
        j += 2;
        k += 2;
        i = abs(y2 - b2);
        if (i <= mn) {
            mn = i;
            if (i == 0) *where = ac;
        }
        den += 1;
    }
    return den == 0 ? -1 : mn;
}

/* Fusion: p1 up to A coordinate `at`, then p2 from `at` on. */
static void fusion(mpath *p1, int at, const mpath *p2, u16pool *P, int TS)
{
    int k1 = 2 * ((at / TS) - (p1->abpos / TS));
    int k2 = 2 * ((at / TS) - (p2->abpos / TS));
    int len = k1 + (p2->tlen - k2), diff = 0, n = 0, k;
    int64_t o = pool_reserve(P, len);
    uint16_t *dst = P->t + o;
    const uint16_t *s1 = P->t + p1->toff, *s2 = P->t + p2->toff;
    for (k = 0; k < k1; k += 2) {
        dst[n++] = s1[k];
        dst[n++] = s1[k + 1];
        diff += s1[k];
    }
    for (k = k2; k < p2->tlen; k += 2) {
        dst[n++] = s2[k];
        dst[n++] = s2[k + 1];
        diff += s2[k];
    }
    p1->aepos = p2->aepos;
    p1->bepos = p2->bepos;
    p1->diffs = diff;
    p1->toff = o;
    p1->tlen = n;
}

static int handle_redundancies(mpath *am, int n, mpath *bm, u16pool *P, int TS, int comp)
{
    int no = 0, j, k, aw = 0, bw = 0;
    for (j = 1; j < n; j++) {
        mpath *jp = am + j;
        for (k = no; k >= 0; k--) {
            mpath *kp = am + k;
            if (jp->abpos < kp->abpos) {
                if (!(kp->abpos <= jp->aepos && kp->bbpos <= jp->bepos)) continue;
                if (entwine(jp, kp, P, TS, &aw) != 0) continue;
                if (kp->aepos > jp->aepos) {
                    if (bm) {
                        if (comp) {
                            if (entwine(bm + k, bm + j, P, TS, &bw) != 0) continue;
                            fusion(jp, aw, kp, P, TS);
                            am[k] = *jp;
                            fusion(bm + k, bw, bm + j, P, TS);
                        } else {
                            if (entwine(bm + j, bm + k, P, TS, &bw) != 0) continue;
                            fusion(jp, aw, kp, P, TS);
                            am[k] = *jp;
                            fusion(bm + j, bw, bm + k, P, TS);
                            bm[k] = bm[j];
                        }
                    } else {
                        fusion(jp, aw, kp, P, TS);
                        am[k] = *jp;
                    }
                } else {
                    am[k] = *jp;
                    if (bm) bm[k] = bm[j];
                }
                break;
            } else {
                if (!(jp->abpos <= kp->aepos && jp->bbpos <= kp->bepos)) continue;
                if (entwine(kp, jp, P, TS, &aw) != 0) continue;
                if (kp->abpos == jp->abpos) {
                    if (kp->aepos < jp->aepos) {
                        am[k] = *jp;
                        if (bm) bm[k] = bm[j];
                    }
                } else if (jp->aepos > kp->aepos) {
                    if (bm) {
                        if (comp) {
                            if (entwine(bm + j, bm + k, P, TS, &bw) != 0) continue;
                            fusion(kp, aw, jp, P, TS);
                            fusion(bm + j, bw, bm + k, P, TS);
                            bm[k] = bm[j];
                        } else {
                            if (entwine(bm + k, bm + j, P, TS, &bw) != 0) continue;
                            fusion(kp, aw, jp, P, TS);
                            fusion(bm + k, bw, bm + j, P, TS);
                        }
                    } else
                        fusion(kp, aw, jp, P, TS);
                }
                break;
            }
        }
        if (k < 0) {
            no += 1;
            am[no] = *jp;
            if (bm) bm[no] = bm[j];
        }
    }
    return no + 1;
}

/* ======================================================================== */
/* Overlap record set and writer (dalign/align.c:3345-3360, :5704-5921).    */

static void ovl_add(or_ovlset *s, const mpath *p, const uint16_t *tr, uint32_t flags, int aread,
                    int bread, int no_trace)
{
    if (s->n >= s->cap) {
        s->cap = s->cap * 2 + 1024;
        s->ovl = xrealloc(s->ovl, sizeof(or_ovl) * s->cap);
    }
    or_ovl *o = s->ovl + s->n++;
    o->path.tlen = no_trace ? 0 : p->tlen;
    o->path.diffs = p->diffs;
    o->path.abpos = p->abpos;
    o->path.bbpos = p->bbpos;
    o->path.aepos = p->aepos;
    o->path.bepos = p->bepos;
    o->flags = flags;
    o->aread = aread;
    o->bread = bread;
    o->toff = s->tn;
    if (no_trace) return;
    int64_t nb = (int64_t)p->tlen * s->tbytes;
    if (s->tn + nb > s->tcap) {
        s->tcap = (s->tn + nb) * 2 + 65536;
        s->tr = xrealloc(s->tr, s->tcap);
    }
    for (int i = 0; i < p->tlen; i++) {
        if (s->tbytes == 1) s->tr[s->tn + i] = (uint8_t)tr[i];   /* Compress_TraceTo8 */
        else memcpy(s->tr + s->tn + 2 * i, &tr[i], 2);
    }
    s->tn += nb;
}

static int ovl_cmp(const or_ovl *l, const or_ovl *r)
{
    if (l->aread != r->aread) return l->aread < r->aread ? -1 : 1;
    if (l->bread != r->bread) return l->bread < r->bread ? -1 : 1;
    if ((l->flags & 1) != (r->flags & 1)) return (l->flags & 1) > (r->flags & 1) ? 1 : -1;
    if (l->path.abpos != r->path.abpos) return l->path.abpos < r->path.abpos ? -1 : 1;
    return 0;
}

static void msort_ovl(or_ovl *a, or_ovl *t, int64_t n)
{
    if (n < 2) return;
    int64_t h = n / 2;
    msort_ovl(a, t, h);
    msort_ovl(a + h, t, n - h);
    int64_t i = 0, j = h, o = 0;
    while (i < h && j < n) t[o++] = (ovl_cmp(a + j, a + i) < 0) ? a[j++] : a[i++];
    while (i < h) t[o++] = a[i++];
    while (j < n) t[o++] = a[j++];
    memcpy(a, t, sizeof(or_ovl) * n);
}

void or_ovlset_sort(or_ovlset *s)
{
    or_ovl *t = (or_ovl *)xrealloc(NULL, sizeof(or_ovl) * (s->n + 1));
    msort_ovl(s->ovl, t, s->n);
    free(t);
}

int or_write_las(const char *path, const or_ovlset *s, int64_t lo, int64_t hi, int tspace)
{
    FILE *f = fopen(path, "wb");
    if (f == NULL) return -1;
    int64_t n = hi - lo;
    fwrite(&n, 8, 1, f);
    fwrite(&tspace, 4, 1, f);
    for (int64_t i = lo; i < hi; i++) {
        const or_ovl *o = s->ovl + i;
        int32_t rec[10] = {o->path.tlen, o->path.diffs, o->path.abpos, o->path.bbpos,
                           o->path.aepos, o->path.bepos, (int32_t)o->flags, o->aread, o->bread, 0};
        fwrite(rec, 40, 1, f);
        if (o->path.tlen > 0) fwrite(s->tr + o->toff, s->tbytes, o->path.tlen, f);
    }
    fclose(f);
    return 0;
}

void or_ovlset_free(or_ovlset *s)
{
    free(s->ovl);
    free(s->tr);
    memset(s, 0, sizeof(*s));
}

/* ======================================================================== */
/* report_thread (dalign/filter.c:1617-1989) over the -j chunks of one pass. */

typedef struct {
    int w, mind, maxd;
    int *F;   /* [1 pad][score w][lastp w][lasta w][1 pad] */
} band_arrays;

static inline int *B_score(band_arrays *b) { return b->F + 1 - b->mind; }
static inline int *B_lastp(band_arrays *b) { return b->F + 1 + b->w - b->mind; }
static inline int *B_lasta(band_arrays *b) { return b->F + 1 + 2 * b->w - b->mind; }

static void diagonal_span(const mpath *p, const uint16_t *pts, int TS, int binshift, int *mind, int *maxd)
{
    int low, hgh, dd, i, tlen = p->tlen - 2;
    dd = p->abpos - p->bbpos;
    low = hgh = dd;
    dd = p->aepos - p->bepos;
    if (dd < low) low = dd;
    else if (dd > hgh) hgh = dd;
    dd = (p->abpos / TS) * TS - p->bbpos;
    for (i = 1; i < tlen; i += 2) {
        dd += TS - pts[i];
        if (dd < low) low = dd;
        else if (dd > hgh) hgh = dd;
    }
    *mind = (low >> binshift) - 1;
    *maxd = (hgh >> binshift) + 1;
}

static int64_t pair_key(const seed_t *h) { return ((int64_t)h->bread << 32) | (uint32_t)h->aread; }

static int or_daligner_pass(const or_opts *o, const or_spec *sp, or_block *A, or_block *B, int self, int comp,
                     kmer_t *as, int64_t na, kmer_t *bs, int64_t nb, int same_index,
                     or_ovlset *out, or_stats *st, or_work *w)
{
    const int TS = o->tspace, K = o->kmer, BS = o->binshift;
    int64_t nh;
    int limit;
    if (na == 0 || nb == 0) return 0;
    seed_t *hits = make_seeds(o, as, na, bs, nb, self ? (o->identity ? (comp ? 3 : 2) : 1) : 0, same_index, &nh,
                              &limit);
    st->nhits += nh;
    st->limit = limit;
    if (nh == 0) { free(hits); return 0; }

    const int minhit = (o->hitmin - 1) / K + 1;
    const int maxdiag = A->maxlen >> BS, mindiag = (-B->maxlen) >> BS;
    band_arrays ba;
    ba.w = maxdiag - mindiag + 1;
    ba.mind = mindiag;
    ba.maxd = maxdiag;
    ba.F = (int *)calloc(3 * ba.w + 2, sizeof(int));
    int *score = B_score(&ba), *lastp = B_lastp(&ba), *lasta = B_lasta(&ba);
    const int afirst = A->ufirst, bfirst = B->ufirst;

    mpath *am = NULL, *bm = NULL;
    int amax = 0, bmax = 0;
    u16pool pool = {0};

    int64_t *chunk = (int64_t *)xrealloc(NULL, sizeof(int64_t) * (o->nthreads + 1));
    chunk[0] = 0;
    for (int i = 1; i < o->nthreads; i++) {
        int64_t p = (nh * i) >> o->nshift;
        if (p > 0) {
            int d = hits[p - 1].bread;
            while (hits[p].bread == d) p++;
        }
        chunk[i] = p;
    }
    chunk[o->nthreads] = nh;

    for (int th = 0; th < o->nthreads; th++) {
        int64_t nidx = chunk[th], eidx = chunk[th + 1] - minhit;
        int64_t cpair = pair_key(hits + nidx), npair = cpair;
        memset(ba.F, 0, sizeof(int) * (3 * ba.w + 2));
        for (; nidx < eidx; cpair = npair) {
            if (pair_key(hits + nidx + minhit - 1) != cpair) {
                nidx += 1;
                while ((npair = pair_key(hits + nidx)) == cpair) nidx += 1;
                continue;
            }
            int ar = hits[nidx].aread, br = hits[nidx].bread;
            int alen = A->rlen[ar], blen = B->rlen[br];
            if (alen < o->hgap_min && blen < o->hgap_min) {
                nidx += 1;
                while ((npair = pair_key(hits + nidx)) == cpair) nidx += 1;
                continue;
            }
            st->npairs += 1;
            const char *aseq = A->bases + A->boff[ar], *bseq = B->bases + B->boff[br];
            int doA = (alen >= o->hgap_min);
            int doB = (o->symmetric && blen >= o->hgap_min && (ar != br || !self || !comp));
            int novla = 0, novlb = 0, amark2 = 0;
            int64_t sidx, lidx, h2, f;
            pool.n = 0;
            for (sidx = nidx; pair_key(hits + nidx) == cpair; nidx = h2) {
                int amark = amark2 + 50000;
                amark2 = amark - 10000;
                int apos;
                h2 = lidx = nidx;
                do {
                    apos = hits[nidx].apos;
                    npair = pair_key(hits + (++nidx));
                    if (apos <= amark2) h2 = nidx;
                } while (npair == cpair && apos <= amark);
                if (nidx - lidx < minhit) continue;

                for (f = lidx; f < nidx; f++) {
                    int ap = hits[f].apos, dg = hits[f].diag >> BS;
                    if (ap - lastp[dg] >= K) score[dg] += K;
                    else score[dg] += ap - lastp[dg];
                    lastp[dg] = ap;
                }
                for (f = lidx; f < nidx; f++) {
                    int ap = hits[f].apos, dg = hits[f].diag, bp = ap - dg;
                    dg >>= BS;
                    if (dg == maxdiag || dg == mindiag) st->edge_events += 1;
                    if (!(ap > lasta[dg] && (score[dg] + score[dg + 1] >= o->hitmin ||
                                             score[dg] + score[dg - 1] >= o->hitmin)))
                        continue;
                    or_path pa, pb;
                    const uint16_t *ta, *tb;
                    st->nseeds += 1;
                    int rc = or_local_alignment(sp, w, aseq, alen, bseq, blen, comp, ap - bp, ap - bp,
                                                ap + bp, -1, -1, &pa, &ta, &pb, &tb);
                    if (rc) {
                        fprintf(stderr, "oracle: undefined wave state at pair (%d,%d) seed (%d,%d)\n",
                                ar, br, ap, bp);
                        exit(3);
                    }
                    int64_t c1, c2;
                    or_work_counters(w, &c1, &c2);
                    st->cells += c1;
                    st->snake += c2;
                    st->naligned += 1;
                    {
                        mpath mp = {pa.tlen, pa.diffs, pa.abpos, pa.bbpos, pa.aepos, pa.bepos, 0};
                        int lo, hi, ae = pa.aepos, d;
                        diagonal_span(&mp, ta, TS, BS, &lo, &hi);
                        if (dg < lo) lo = dg;
                        else if (dg > hi) hi = dg;
                        if (lo < mindiag || hi > maxdiag) st->edge_events += 1;
                        for (d = lo; d <= hi; d++)
                            if (ae > lasta[d]) lasta[d] = ae;
                    }
                    if ((pa.aepos - pa.abpos) + (pa.bepos - pa.bbpos) >= o->minover) {
                        if (doA) {
                            if (novla >= amax) { amax = amax * 2 + 64; am = xrealloc(am, sizeof(mpath) * amax); }
                            mpath *m = am + novla++;
                            m->tlen = pa.tlen; m->diffs = pa.diffs; m->abpos = pa.abpos; m->bbpos = pa.bbpos;
                            m->aepos = pa.aepos; m->bepos = pa.bepos;
                            m->toff = pool_reserve(&pool, pa.tlen);
                            memcpy(pool.t + m->toff, ta, sizeof(uint16_t) * pa.tlen);
                        }
                        if (doB) {
                            if (novlb >= bmax) { bmax = bmax * 2 + 64; bm = xrealloc(bm, sizeof(mpath) * bmax); }
                            mpath *m = bm + novlb++;
                            m->tlen = pb.tlen; m->diffs = pb.diffs; m->abpos = pb.abpos; m->bbpos = pb.bbpos;
                            m->aepos = pb.aepos; m->bepos = pb.bepos;
                            m->toff = pool_reserve(&pool, pb.tlen);
                            memcpy(pool.t + m->toff, tb, sizeof(uint16_t) * pb.tlen);
                        }
                    }
                }
                for (f = lidx; f < nidx; f++) {
                    int dg = hits[f].diag >> BS;
                    score[dg] = lastp[dg] = 0;
                }
            }
            for (f = sidx; f < nidx; f++) {
                int dg = hits[f].diag >> BS, d;
                for (d = dg; d <= maxdiag; d++) {
                    if (lasta[d] == 0) break;
                    lasta[d] = 0;
                }
                for (d = dg - 1; d >= mindiag; d--) {
                    if (lasta[d] == 0) break;
                    lasta[d] = 0;
                }
            }
            if (novla > 1) {
                if (novlb > 1) novla = novlb = handle_redundancies(am, novla, bm, &pool, TS, comp);
                else novla = handle_redundancies(am, novla, NULL, &pool, TS, comp);
            } else if (novlb > 1)
                novlb = handle_redundancies(bm, novlb, NULL, &pool, TS, comp);
            for (int i = 0; i < novla; i++)
                ovl_add(out, am + i, pool.t + am[i].toff, (uint32_t)comp, ar + afirst, br + bfirst, o->no_trace);
            for (int i = 0; i < novlb; i++)
                ovl_add(out, bm + i, pool.t + bm[i].toff, (uint32_t)comp, br + bfirst, ar + afirst, o->no_trace);
            st->nrecords += novla + novlb;
        }
    }
    free(chunk);
    free(pool.t);
    free(am);
    free(bm);
    free(ba.F);
    free(hits);
    return 0;
}

/* daligner.c:929-1029: both strands of one block pair. */
int or_daligner_pair(const or_opts *o, const or_spec *sp, or_block *A, or_block *B, int same_file,
                     or_ovlset *out, or_stats *st)
{
    int64_t na, nb;
    or_work *w = or_work_new();
    out->tbytes = (o->tspace <= 125) ? 1 : 2;
    kmer_t *as = sort_kmers(A, o->kmer, o->suppress, o->biased, &na);
    st->nkmers_a = na;
    if (same_file) {
        or_daligner_pass(o, sp, A, A, 1, 0, as, na, as, na, 1, out, st, w);
        or_block cB;
        or_block_complement(A, &cB);
        kmer_t *bs = sort_kmers(&cB, o->kmer, o->suppress, o->biased, &nb);
        st->nkmers_b = nb;
        or_daligner_pass(o, sp, A, &cB, 1, 1, as, na, bs, nb, 0, out, st, w);
        free(bs);
        or_block_free(&cB);
    } else {
        kmer_t *bs = sort_kmers(B, o->kmer, o->suppress, o->biased, &nb);
        st->nkmers_b = nb;
        or_daligner_pass(o, sp, A, B, 0, 0, as, na, bs, nb, 0, out, st, w);
        free(bs);
        or_block cB;
        or_block_complement(B, &cB);
        bs = sort_kmers(&cB, o->kmer, o->suppress, o->biased, &nb);
        or_daligner_pass(o, sp, A, &cB, 0, 1, as, na, bs, nb, 0, out, st, w);
        free(bs);
        or_block_free(&cB);
    }
    free(as);
    if (o->verbose > 1) {
        fprintf(stderr, "waves %lld maxw %lld hist", (long long)w->nwaves, (long long)w->maxw);
        for (int i = 0; i < 9; i++) fprintf(stderr, " %lld", (long long)w->hist[i]);
        fprintf(stderr, "\n");
    }
    or_work_free(w);
    return 0;
}

/* ======================================================================== */
/* Flat task-array driver for tests: same task / result layout as
 * oracle/la_vectors.c.  Returns the number of tasks whose wave reached a
 * state the reference leaves undefined (their result status is 3). */
int or_la_tasks(double ave_corr, int tspace, const float freq[4], const char *seq, int ntasks,
                const int64_t *aoff, const int64_t *boff, const int32_t *par /* ntasks x 8 */,
                int32_t *res /* ntasks x 13: 12 fields + status */, uint16_t *traces, int64_t *toff /* ntasks+1 */,
                int64_t tcap)
{
    or_spec *sp = (or_spec *)malloc(sizeof(or_spec));
    or_work *w = or_work_new();
    int bad = 0;
    int64_t t = 0;
    or_spec_init(sp, ave_corr, tspace, freq);
    for (int i = 0; i < ntasks; i++) {
        const int32_t *v = par + 8 * i;
        or_path pa, pb;
        const uint16_t *ta = NULL, *tb = NULL;
        int rc = or_local_alignment(sp, w, seq + aoff[i], v[0], seq + boff[i], v[1], v[2], v[3], v[4], v[5], v[6],
                                    v[7], &pa, &ta, &pb, &tb);
        int32_t *r = res + 13 * i;
        toff[i] = t;
        if (rc) {
            memset(r, 0, sizeof(int32_t) * 13);
            r[12] = rc;
            bad += 1;
            continue;
        }
        int32_t vals[13] = {pa.abpos, pa.bbpos, pa.aepos, pa.bepos, pa.diffs, pa.tlen,
                            pb.abpos, pb.bbpos, pb.aepos, pb.bepos, pb.diffs, pb.tlen, 0};
        memcpy(r, vals, sizeof vals);
        if (t + pa.tlen + pb.tlen > tcap) {
            free(sp);
            or_work_free(w);
            return -1;
        }
        memcpy(traces + t, ta, 2 * pa.tlen);
        t += pa.tlen;
        memcpy(traces + t, tb, 2 * pb.tlen);
        t += pb.tlen;
    }
    toff[ntasks] = t;
    free(sp);
    or_work_free(w);
    return bad;
}
