// Host entry points of the fused single-kernel step.
#pragma once

#include "bt_internal.h"

namespace ibt {

// Launch shape of the fused step: 60 output columns per 64-lane wave,
// `chunk` output rows per wave.
struct FusedShape {
    int chunk = 32;
};

int fused_run(inmsom_bt_ctx *c, double tau, int nstep, int nsteps, int mode, double *ms);
// Copies the fused state back into the canonical FP64 state (no-op when
// the fused path does not hold the state).
int fused_flush(inmsom_bt_ctx *c);
void fused_release(inmsom_bt_ctx *c);
int fused_range_check(inmsom_bt_ctx *c, long long *bad);
int fused_volume(inmsom_bt_ctx *c, double *vol);
int fused_coarsen(inmsom_bt_ctx *c, int factor, double *u, double *v, double *ssh);
int fused_set_shape(inmsom_bt_ctx *c, int chunk);
int fused_set_substeps(inmsom_bt_ctx *c, int k);
// K substeps per launch as a K-wave pipeline (FP64 and FP32; 0 =
// automatic, 1 = off) and the stages the next launches use (0: none)
int fused_set_pipeline(inmsom_bt_ctx *c, int k);
int pipe_stages(const inmsom_bt_ctx *c, int mode);
int subs_per_launch(const inmsom_bt_ctx *c, int mode);

}  // namespace ibt
