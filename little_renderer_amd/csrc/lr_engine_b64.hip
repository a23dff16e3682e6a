// lr_engine_b64.hip - the engine's kernels for work-groups of 64 lanes (lr_engine_k.h); one
// file per work-group size, so that the variants the autotuner picks from build in parallel.
#include "lr_engine_k.h"

void lr_eng_fill_b64(LrEngTable &t) { fill_table<64>(t); }
