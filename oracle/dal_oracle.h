/* Oracle: CPU restatement of MARVEL daligner's overlap path (dalign/ lineage).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (marvel_amd/) never links or calls it.
 *
 * Each function cites the reference file:line it restates.  Parity of every
 * function here is pinned against the compiled reference (oracle/_ref, built
 * from /root/reference by oracle/Makefile) through the golden vectors in
 * tests/golden/ and the byte-identical .las checks in tests/.
 */
#ifndef DAL_ORACLE_H
#define DAL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Alignment parameters (dalign/align.c:196-279). */
typedef struct {
    double ave_corr;
    int tspace;
    float freq[4];
    int ave_path;
    int mscore, dscore;
    int16_t score[32768];
    int16_t table[32768];
} or_spec;

/* Path record, field order of dalign/align.h:76-83 (trace kept separately). */
typedef struct {
    int tlen, diffs, abpos, bbpos, aepos, bepos;
} or_path;

typedef struct or_work or_work;

void or_spec_init(or_spec *sp, double ave_corr, int tspace, const float freq[4]);
or_work *or_work_new(void);
void or_work_free(or_work *w);

/* Local_Alignment (dalign/align.c:1922-2060).  aseq/bseq point at the first
 * base of each read inside a buffer that holds the value 4 immediately before
 * and after the read.  On return apath and bpath hold the A- and B-view paths and
 * atrace/btrace point into w (valid until the next call).  Returns 0, or a
 * nonzero code when an input drives the wave into a state the reference leaves
 * undefined (empty band, pebble -1). */
int or_local_alignment(const or_spec *sp, or_work *w, const char *aseq, int alen,
                       const char *bseq, int blen, int comp, int low, int hgh,
                       int anti, int lbord, int hbord, or_path *apath,
                       const uint16_t **atrace, or_path *bpath,
                       const uint16_t **btrace);

/* Work counters of the last or_local_alignment: wave widths summed and snake
 * compare steps, forward + reverse (the W_align units). */
void or_work_counters(const or_work *w, int64_t *cells, int64_t *snake);
/* Wave-width statistics accumulated over the life of w. */
void or_work_widths(const or_work *w, int64_t *nwaves, int64_t *maxw, int64_t hist[9]);

/* ---- DB block (db/DB.c:360-585, :1349-1410) ---------------------------- */
typedef struct {
    int nreads, ufirst, part, maxlen;
    int64_t totlen;
    float freq[4];
    int *rlen;          /* [nreads] */
    int64_t *boff;      /* [nreads+1] offsets of each read in bases */
    char *bases;        /* bases[-1] == 4, read i at bases+boff[i], 4 after each */
    char *alloc;
    int ureads;         /* reads of the whole DB */
    char *stem;         /* "DIR/.ROOT", the prefix of the DB's hidden files */
    /* soft mask (daligner -m): read i's points at mdata[manno[i] .. manno[i+1]),
     * pairs (begin, end) of masked intervals; NULL when unmasked */
    int64_t *manno;
    int32_t *mdata;
} or_block;

int or_block_load(or_block *b, const char *path);   /* "DIR/ROOT.<part>" or "ROOT" */
/* Loads the -m tracks names[0..n) of the block (Dazzler .anno/.data or MARVEL
 * .a2/.d2) and, for n > 1, merges them into one (dalign/daligner.c:373-431).
 * Returns 0, or -1 with a message in err. */
int or_block_load_masks(or_block *b, char *const *names, int n, char *err, int errlen);
void or_block_free(or_block *b);
/* reverse-complemented copy (dalign/daligner.c:456-573, inplace == 0) */
void or_block_complement(const or_block *src, or_block *dst);

/* ---- daligner options (dalign/daligner.c:615-836) ---------------------- */
typedef struct {
    int kmer, binshift, hitmin, suppress, minover, tspace, nthreads, nshift;
    int hgap_min, symmetric, run_id, no_trace, verbose;
    int biased;   /* -b: k-mers of variable length weighted by the base frequencies */
    int identity, only_identity;   /* -I: seeds of a read with itself too; -O: keep only those overlaps */
    double ave_corr;
    uint64_t mem_limit;
} or_opts;

void or_opts_default(or_opts *o);

/* One overlap record plus its (already 8-bit or 16-bit) trace bytes. */
typedef struct {
    or_path path;
    uint32_t flags;
    int aread, bread;
    int64_t toff;        /* offset of the trace bytes in the record set */
} or_ovl;

typedef struct {
    or_ovl *ovl; int64_t n, cap;
    uint8_t *tr; int64_t tn, tcap;
    int tbytes;
} or_ovlset;

/* Statistics of one Match_Filter pass. */
typedef struct {
    int64_t nkmers_a, nkmers_b, nhits, npairs, nseeds, naligned, nrecords;
    int64_t cells, snake;         /* aligner work (W_align units) */
    int limit;
    int64_t edge_events;          /* accesses of band arrays beyond [mindiag,maxdiag] */
} or_stats;

/* Full daligner on block A vs block B (self when same_file).  Appends records
 * to `out` in the reference's thread-buffer insertion order (both strands).
 * Returns 0 on success. */
int or_daligner_pair(const or_opts *o, const or_spec *sp, or_block *A, or_block *B,
                     int same_file, or_ovlset *out, or_stats *st);

/* Stable sort by (aread, bread, COMP, abpos) (dalign/align.c:5786-5813). */
void or_ovlset_sort(or_ovlset *s);
/* Write .las (dalign/align.c:3345-3350, :5909-5921) records [lo,hi). */
int or_write_las(const char *path, const or_ovlset *s, int64_t lo, int64_t hi, int tspace);
void or_ovlset_free(or_ovlset *s);

/* Compute_Trace_PTS (dalign/align.c:5286-5397, iter_np :4556-4916) on the
 * A-view path with its uint16 trace points; mode -1 LOWERMOST, 0 GREEDIEST,
 * 1 UPPERMOST.  Writes the int trace (indel positions) to out.  Returns 0,
 * -1 for a self comparison on diagonal 0, -3 when cap is too small. */
int or_trace_pts(const char *aseq, const char *bseq, const or_path *path, const uint16_t *points,
                 int trace_spacing, int mode, int *out, int64_t cap, int *ntrace, int *diffs);

/* Task driver in the layout of oracle/trace_vectors.c: per task
 * {status, a_tlen, 3 x {diffs, ntrace, trace[ntrace]}} appended to res
 * (int32 words).  Returns the number of words written, or -1 when rcap is
 * too small. */
int64_t or_trace_tasks(double ave_corr, int tspace, const float freq[4], const char *seq, int ntasks,
                       const int64_t *aoff, const int64_t *boff, const int32_t *par, int32_t *res, int64_t rcap);

#ifdef __cplusplus
}
#endif
#endif
