// Internal helpers shared by the HIP translation units of libmarvl_align.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/marvl_align.h"
#include "la_wave.hpp"

namespace mi {

// Host-derived alignment constants (dalign/align.c:214-279) and their device copy.
struct HostSpec {
    double ave_corr;
    int tspace;
    float freq[4];
    int ave_path, mscore, dscore;
    int16_t table[32768];
    int16_t score[32768];
};

struct DevSpecHost {
    HostSpec host;
    DevSpec spec;   // table/score point at device memory
};

// Align_Spec of this library: the alignment constants plus the per-thread
// overlap buffers the filter appends to (dalign/align.c:52-62, :260-276).
struct SpecObj {
    HostSpec hs;
    MiAlignParams prm;
    int nthreads, symmetric, only_identity, no_trace;
    Overlap_IO_Buffer* iobuf;   // [nthreads]
    // double buffering for the background .las writer: the buffers a write
    // is still reading become the spare set at Reset_Overlap_Buffer, and the
    // spare set (already written) becomes the active one
    Overlap_IO_Buffer* spare = nullptr;   // [nthreads], made at the first reset after a write
    bool written = false;                 // the active buffers were handed to the writer
};
// Waits for the background .las writer; the process exits with status 1
// when a write failed (the reference exits inside Write_Overlap_Buffer).
void wait_overlap_writes();

void spec_build(HostSpec* s, double ave_corr, int tspace, const float* freq);
struct OutRec;
// Appends n records (their trace bytes at tr + toff) to an overlap buffer in
// one piece; the behaviour of n AddOverlapToBuffer calls.  Returns 0.
int mi_append_records(Overlap_IO_Buffer* b, const OutRec* recs, long long n, const uint8_t* tr, long long trbytes,
                      int tbytes);
// Device tables for these parameters (cached per parameter set; lives until exit).
DevSpecHost* mi_spec_device(const MiAlignParams* prm);
int mi_num_cus();
// launch shapes chosen by mi_set_tuning (0 fields: built-in defaults)
const MiTuning& mi_tuning();
void mi_set_error(const char* msg);
// Watchdog: device pointer of a host-pinned flag every long-running kernel
// polls; mi_wait_watchdog waits for stream `st` and, past the deadline
// (MI_WATCHDOG_S seconds, default 120), raises the flag so the kernel's waves
// exit, then reports the timeout.  The flag is cleared before returning.
const int* mi_abort_flag_device();
int mi_wait_watchdog(hipStream_t st, const char* what);

}  // namespace mi
