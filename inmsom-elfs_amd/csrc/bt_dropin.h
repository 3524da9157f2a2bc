// Host entry points of the fused drop-in of barotropic_dynamics.
#pragma once

#include <string>

#include "bt_internal.h"

namespace ibt {

// Whether the fused path can take a call with this ksw4 on this context's
// grid (see bt_dropin.hip); *why says why not.  ksw4 > 0 is covered: the
// biharmonic term comes from a pre-pass of reference-order kernels per
// substep (lit_diff4_terms).
bool dropin_eligible(inmsom_bt_ctx *c, int ksw4, std::string *why);
// One call of barotropic_dynamics (Function/vel_ssh.f90:379-679) on device
// arrays: the state (u, up, v, vp, ssh, sshp) is replaced by the state
// after substep nstep, acc (ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady) receives
// the averages.  Queued on the context's stream.
// scratch: 14 planes of the substeps' buffers (dropin_scratch_doubles), or
// null for library-owned buffers allocated at the first such call.
int dropin_call(inmsom_bt_ctx *c, double tau, int nstep, double *const st[6], double *const acc[4],
                const double *rhsx, const double *rhsy, const double *wflux, const double *mu, const double *rdis,
                double *scratch = nullptr, int ksw4 = 0, const double *mu4 = nullptr);
// The grid analysis (eligibility, mask bits, row tables, launch shape),
// done once per context at inmsom_bt_init.
int dropin_setup(inmsom_bt_ctx *c);
size_t dropin_scratch_doubles(const inmsom_bt_ctx *c);
// The fixed scratch array of the entry point that allocates nothing: its
// address on the current device, looked up once per device at set-up
// (inmsom_bt_init), and (noalloc_scratch) that address and the array's
// size in doubles (null before set-up).
int noalloc_scratch_setup(int device);
double *noalloc_scratch(int device, size_t *doubles);
void dropin_release(inmsom_bt_ctx *c);
// The launch the fused drop-in uses: rows per wave and whether metrics come
// from row tables (after the first call; 0 / false before).
void dropin_shape(const inmsom_bt_ctx *c, int *rows_per_wave, bool *row_metrics);

}  // namespace ibt
