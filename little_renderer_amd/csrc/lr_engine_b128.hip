// lr_engine_b128.hip - the engine's kernels for work-groups of 128 lanes (lr_engine_k.h); one
// file per work-group size, so that the variants the autotuner picks from build in parallel.
#include "lr_engine_k.h"

void lr_eng_fill_b128(LrEngTable &t) { fill_table<128>(t); }
