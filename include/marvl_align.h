/* libmarvl_align: MI355X (gfx950) implementation of MARVEL daligner's
 * overlap hot path behind a plain C ABI.
 *
 * Two groups of entry points:
 *   1. the names and signatures of MARVEL's dalign/align.h that the daligner
 *      path and the downstream tools bind (Local_Alignment and friends), so a
 *      C host links this library in place of marvl-align;
 *   2. batched entry points with no reference counterpart (mi_*), used by the
 *      GPU daligner driver, the tests and the benchmark.
 * Every function returns 0 / non-NULL on success.  Errors of the mi_* calls
 * are negative codes, with a message from mi_last_error().  The align.h
 * drop-ins keep the reference's batch-mode behaviour: they print a message
 * and exit(1) (db/DB.h:32-36).
 */
#ifndef MARVL_ALIGN_H
#define MARVL_ALIGN_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference-compatible types (dalign/align.h:76-117, :280-336) ----- */

typedef struct {
    void *trace;
    int tlen;
    int diffs;
    int abpos, bbpos;
    int aepos, bepos;
} Path;

typedef struct {
    Path *path;
    uint32_t flags;
    char *aseq;
    char *bseq;
    int alen;
    int blen;
} Alignment;

typedef struct {
    Path path;
    uint32_t flags;
    int aread;
    int bread;
} Overlap;

typedef void Work_Data;
typedef void Align_Spec;

#define COMP(x) ((x)&0x1)
#define COMP_FLAG 0x1
#define TRACE_XOVR 125

/* New_Align_Spec: replaces dalign/align.c:214-279 (decl. dalign/align.h:160).
 * Same integer parameters (ave_path, match/mismatch scores, 15-bit trim tables). */
Align_Spec *New_Align_Spec(double ave_corr, int trace_space, float *freq, int nthreads, int symmetric,
                           int only_identity, int no_trace_points);
/* dalign/align.h:162-166 */
void Free_Align_Spec(Align_Spec *spec);
int Trace_Spacing(Align_Spec *spec);
double Average_Correlation(Align_Spec *spec);
float *Base_Frequencies(Align_Spec *spec);

/* New_Work_Data / Free_Work_Data: replace dalign/align.c:64-130 (dalign/align.h:131-133). */
Work_Data *New_Work_Data(void);
void Free_Work_Data(Work_Data *work);

/* Local_Alignment: replaces dalign/align.c:1922-2060 (dalign/align.h:195).
 * Runs on the GPU (one wave64).  Same contract: align->path receives the
 * A-view path, the returned Path is the B-view path, both traces (uint16
 * pairs) live in `work` until the next call.  aseq/bseq must point into
 * buffers holding the value 4 right before and right after each read. */
Path *Local_Alignment(Alignment *align, Work_Data *work, Align_Spec *spec, int low, int hgh, int anti, int lbord,
                      int hbord);

/* Compute_Trace_PTS: replaces dalign/align.c:5286-5397 (dalign/align.h:224).
 * Runs on the GPU.  Same contract: align->path (A view, uint16 trace
 * points) gets an int trace of indel positions that lives in `work` until
 * the next call, tlen = its length, diffs = the summed segment distances.
 * mode: LOWERMOST -1, GREEDIEST 0, UPPERMOST 1.  `debug` is ignored. */
#define LOWERMOST -1
#define GREEDIEST 0
#define UPPERMOST 1
int Compute_Trace_PTS(Alignment *align, Work_Data *work, int trace_spacing, int mode, int debug);

/* Overlap record helpers: replace dalign/align.c:3328-3360, :3419-3449
 * (dalign/align.h:309-318); host code, byte-compatible .las records. */
int Read_Overlap(FILE *input, Overlap *ovl);
int Read_Trace(FILE *input, Overlap *ovl, int tbytes);
void Write_Overlap(FILE *output, Overlap *ovl, int tbytes);
void Compress_TraceTo8(Overlap *ovl);
void Decompress_TraceTo16(Overlap *ovl);
int Check_Trace_Points(Overlap *ovl, int tspace, int verbose, char *fname);
/* Print_Overlap: replaces dalign/align.c:3372-3417 (dalign/align.h:313); same text. */
void Print_Overlap(FILE *output, Overlap *ovl, int tbytes, int indent);

/* Complement_Seq: replaces dalign/align.c:3530-3546 (dalign/align.h:119): reverse
 * complement of a read of bases 0..3, in place.
 * Flip_Alignment: replaces dalign/align.c:3451-3522 (dalign/align.h:230-234): the
 * alignment with A and B exchanged; full = 1 also converts an int trace (as
 * Compute_Trace_PTS leaves it). */
void Complement_Seq(char *a, int n);
void Flip_Alignment(Alignment *align, int full);

/* Per-thread overlap buffers of an Align_Spec (dalign/align.h:324-336). */
typedef struct {
    int tbytes;
    uint64_t tmax;
    uint64_t ttop;
    void *trace;
    int no_trace;
    int omax;
    int otop;
    Overlap *ovls;
} Overlap_IO_Buffer;

/* Replace dalign/align.c:5630-5813 (dalign/align.h:338-351). */
Overlap_IO_Buffer *CreateOverlapBuffer(int nthreads, int tbytes, int no_trace);
int AddOverlapToBuffer(Overlap_IO_Buffer *iobuf, Overlap *ovl, int tbytes);
Overlap_IO_Buffer *OVL_IO_Buffer(Align_Spec *espec);
int Num_Threads(Align_Spec *espec);
int Only_Identity(Align_Spec *espec);
int Symmetric(Align_Spec *espec);
/* Write_Overlap_Buffer: replaces dalign/align.c:5815-6043.  Stable sort of
 * all buffered records by (aread, bread, COMP, abpos), then one .las file
 * (self or not symmetric) or two split at lastRead, named
 * [RRR_BBBBB/]A.a.B.b.las as the reference names them.  The sort and the
 * write run on a background thread, overlapping the caller's next block
 * pair; the call returns once the previous write is done.  The next write,
 * Free_Align_Spec, mi_flush_overlap_writes and process exit wait for it (a
 * failed write ends the process with status 1, as in the reference). */
void Write_Overlap_Buffer(Align_Spec *spec, int RUN_ID, char *ablock, char *bblock, int lastRead);
/* Reset_Overlap_Buffer: replaces dalign/align.c:6045-6056. */
void Reset_Overlap_Buffer(Align_Spec *spec);
/* Waits for the .las files handed to the background writer: 0, or 1 when a
 * write failed.  (No reference counterpart: its writes are synchronous.) */
int mi_flush_overlap_writes(void);

/* ---- batched entry points (no reference counterpart) ------------------- */

#define MI_ERR_ARG -1
#define MI_ERR_DEVICE -2
#define MI_ERR_UNDEF -3
#define MI_ERR_TIMEOUT -4   /* a kernel outran the watchdog deadline (MI_WATCHDOG_S) and was stopped */

/* Alignment parameters as New_Align_Spec takes them. */
typedef struct {
    double ave_corr;
    int trace_space;
    float freq[4];
} MiAlignParams;

/* One Local_Alignment task.  Reads are 2-bit packed in one uint64 buffer:
 * base i of a read at bits 2*(i&31) of word (aword + (i>>5)); one zero guard
 * word before aword and two after the read's last word. */
typedef struct {
    long long aword, bword;
    int alen, blen;
    int comp;             /* B read is the reverse complement (COMP flag) */
    int low, hgh, anti;   /* Local_Alignment's low, hgh, anti */
    int lbord, hbord;     /* Local_Alignment's lbord, hbord */
    int selfie;           /* A and B are the same sequence object (aseq == bseq) */
} MiLaTask;

typedef struct {
    int status;                             /* 0 ok, 3: reference-undefined wave state */
    int abpos, bbpos, aepos, bepos, diffs;  /* A-view path */
    int b_abpos, b_bbpos, b_aepos, b_bepos; /* B-view path (same diffs) */
    int atlen, btlen;                       /* trace lengths (uint16 values) */
    long long atrace, btrace;               /* offsets of the traces in the output trace array */
    long long cells, snake;                 /* work counters: wave widths, snake compare steps */
} MiLaResult;

/* Runs every task on the GPU; traces are appended to `traces` (capacity
 * ntrace uint16 values) in task order, A then B. */
int mi_la_batch_host(const MiAlignParams *prm, const uint64_t *packed, long long npacked, const MiLaTask *tasks,
                     int ntasks, MiLaResult *results, uint16_t *traces, long long ntrace);

/* The same batch on device-resident data (no reference counterpart: the
 * batched entry for callers that keep reads and tasks in HBM).  d_packed,
 * d_tasks, d_order (optional queue order, NULL = task order), d_results,
 * d_traces and d_trace_used are device pointers.  Each completed task's A
 * trace and then its B trace are appended to the trace arena d_traces
 * (capacity trace_cap uint16 values) at the offset *d_trace_used held, which
 * the call advances; results[i].atrace / btrace are arena offsets.  A task
 * that does not fit gets status 6; tasks wider than the widest window keep
 * status 1 (window) or 2 (pool).  work (host, optional, 8 values): executed
 * band cells and snake compare steps, the aligner kernel time of each of the
 * 4 window levels in microseconds (HIP events on the launch stream), and the
 * number of task re-runs.  The library keeps its scratch between calls. */
int mi_la_batch_device(const MiAlignParams *prm, const uint64_t *d_packed, long long npacked, const MiLaTask *d_tasks,
                       int ntasks, const int *d_order, MiLaResult *d_results, uint16_t *d_traces, long long trace_cap,
                       unsigned long long *d_trace_used, long long *work);

/* mi_la_batch_device without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  All scratch -- the
 * alignment tables, counters, retry lists and per-wave pebble pools -- lives
 * in one file-scope device array of 64 MiB per GPU, held until the process
 * ends; the caller passes inputs and outputs as for mi_la_batch_device (task
 * order only).  At most 1048576 tasks per call (more: MI_ERR_ARG); fewer
 * waves run than in mi_la_batch_device, so it is slower.  Every call on one
 * GPU shares the array: calls that overlap on one GPU race. */
int mi_la_batch_device_noalloc(const MiAlignParams *prm, const uint64_t *d_packed, long long npacked,
                               const MiLaTask *d_tasks, int ntasks, MiLaResult *d_results, uint16_t *d_traces,
                               long long trace_cap, unsigned long long *d_trace_used);

/* Synthetic batched-aligner workload, generated in device memory (no
 * reference counterpart: the input of the batched Local_Alignment benchmark).
 * Read pairs: template length log-uniform in [lmin, lmax], pair error
 * uniform in [emin_ppm, emax_ppm] parts per million split half into each
 * read with insertion : deletion : substitution = 9 : 4 : 2, seed point on
 * the true diagonal moved along B by up to +-offmax, COMP on half of them.
 * Every decision is an integer hash of (seed, task, position), so the task
 * set does not depend on the launch. */
typedef struct {
    unsigned long long seed;
    int lmin, lmax;
    int emin_ppm, emax_ppm;
    int offmax;
} MiGenTaskParams;

/* Writes ntasks tasks to d_tasks and their 2-bit packed reads (A then B per
 * task, laid out as MiLaTask describes) to d_packed (device pointers).
 * *npacked receives the words needed; with d_packed == NULL only that is
 * computed. */
int mi_gen_la_tasks(const MiGenTaskParams *g, int ntasks, uint64_t *d_packed, long long packed_cap,
                    long long *npacked, MiLaTask *d_tasks);

/* One Compute_Trace_PTS task: the A-view path of an alignment (its uint16
 * trace points at `points` in the points array) between two 2-bit packed
 * reads laid out as for MiLaTask (the B read already complemented when the
 * alignment is COMP). */
typedef struct {
    long long aword, bword;
    int alen, blen;
    int abpos, bbpos, aepos, bepos;
    int tlen;             /* number of uint16 trace values (pairs) */
    long long points;     /* offset of the trace values in the points array */
    int selfie;           /* A and B are the same sequence object */
} MiTraceTask;

typedef struct {
    int status;           /* 0 ok; 2: a segment of a self comparison starts on diagonal 0 */
    int diffs;            /* summed segment distances */
    int ntrace;           /* ints in the trace */
    long long trace;      /* offset of the trace in the output array */
} MiTraceResult;

/* Compute_Trace_PTS on a batch of paths (one wave per path, one lane per
 * trace-point segment); mode LOWERMOST (-1), GREEDIEST (0) or UPPERMOST (1).
 * Traces are appended to `trace` (capacity ntrace ints) in task order. */
int mi_trace_pts_host(int trace_spacing, int mode, const uint64_t *packed, long long npacked, const uint16_t *points,
                      long long npoints, const MiTraceTask *tasks, int ntasks, MiTraceResult *results, int *trace,
                      long long ntrace);

/* Device memory for callers of the device-resident entry points (no
 * reference counterpart): allocation (NULL on failure), release, blocking
 * copies host -> device (to_device = 1) or device -> host (0), and fills. */
void *mi_device_alloc(long long bytes);
int mi_device_free(void *p);
int mi_device_copy(void *dst, const void *src, long long bytes, int to_device);
int mi_device_memset(void *p, int value, long long bytes);

/* Launch shapes of the GPU kernels (no reference counterpart; results do
 * not depend on them): waves per CU of the report kernel's 64-diagonal
 * window (4..32, default 32) and of its 128-diagonal window (2..32, default
 * 16), the pebble pool of the 64-diagonal window as a power of two (12..18,
 * default 15), waves per CU of the batched aligner's 64-diagonal window
 * (4..32, default 32) and of its 128- and 256-diagonal windows (2..32,
 * default 16).  A 0 field keeps the default; NULL restores all defaults.
 * The tuner in marvel_amd/tune.py picks them from a calibration run. */
typedef struct {
    int report_waves_per_cu;
    int report_pool_log2;
    int report_l1_waves_per_cu;
    int batch_waves_per_cu;
    int batch_wide_waves_per_cu;
} MiTuning;
int mi_set_tuning(const MiTuning *t);
int mi_get_tuning(MiTuning *t);
/* The current device's architecture name (e.g. gfx950) and compute units. */
int mi_device_info(char *name, int cap, int *cus);

/* Raises (1) or lowers (0) the stop flag every long-running kernel of the
 * library polls (no reference counterpart).  Kernels in flight and launched
 * while it is raised stop at their next poll; their calls return
 * MI_ERR_TIMEOUT because tasks were left unrun. */
int mi_abort_device_work(int raise);

/* Alignment parameters New_Align_Spec derives (ave_path, mscore, dscore). */
int mi_align_params(const MiAlignParams *prm, int *ave_path, int *mscore, int *dscore);

/* Last error message of this thread. */
const char *mi_last_error(void);

/* Number of GPUs visible (0 when none); never initialises more than the runtime. */
int mi_device_count(void);

/* Device self-tests of the wave64 building blocks (0: persistent task loop,
 * 1: lane scans / shuffles / ballots, 2: 2-bit snake matcher); the caller
 * checks `out` against its own CPU computation. */
int mi_selftest(int which, const void *in, long long nin, const void *in2, long long nin2, const void *q, int nq,
                int arg, void *out, long long nout);

#ifdef __cplusplus
}
#endif
#endif
