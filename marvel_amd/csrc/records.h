// Output record of the report kernel (host and device view).
#pragma once
#include <stdint.h>

namespace mi {

struct OutRec {   // one .las record plus its sort tie-breaker
    int tlen, diffs, abpos, bbpos, aepos, bepos;
    uint32_t flags;
    int aread, bread;
    int local;        // position in the pair's kept list (ties of the final sort)
    int job;          // job that wrote the record (-1: slot never written)
    long long toff;   // offset of the trace bytes in the output trace arena
    int gen;          // launch that wrote it (a job's records count from its last successful launch)
    int pad;
};

}  // namespace mi
