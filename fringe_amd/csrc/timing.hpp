// timing.hpp -- optional per-stage GPU timing of the tile engine.  When
// enabled, each stage launch is bracketed by two HIP events recorded on the
// stream the stage runs on; elapsed times are summed when read.
#pragma once
#include <hip/hip_runtime.h>

namespace fringe {

enum Stage { STAGE_PACK = 0, STAGE_GRAM = 1, STAGE_SOLVE = 2, STAGE_NMAP_PREP = 3, STAGE_NMAP_KS = 4,
             STAGE_RESOLVE = 5, STAGE_COUNT = 6 };

void timing_begin(int stage, hipStream_t s);
void timing_end(int stage, hipStream_t s);

struct StageScope {
    int st;
    hipStream_t s;
    StageScope(int stage, hipStream_t stream) : st(stage), s(stream) { timing_begin(st, s); }
    ~StageScope() { timing_end(st, s); }
};

}  // namespace fringe
