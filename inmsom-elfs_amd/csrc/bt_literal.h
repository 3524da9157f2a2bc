// Host entry points of the stage-by-stage (reference-order) path.
#pragma once

#include "bt_internal.h"

namespace ibt {

// c->scal layout: [0, 64) reduction scratch, [64, 76) depth pointer table
constexpr int DEP_TABLE_OFFSET = 64;
constexpr int SCAL_DOUBLES = 128;

void lit_copy(inmsom_bt_ctx *c, double *dst, const double *src);
void lit_hh_init(inmsom_bt_ctx *c, const double *ssh, const double *sshp);
void lit_barotropic_dynamics(inmsom_bt_ctx *c, double tau, int nstep, int ksw4, const double *RHSx,
                             const double *RHSy, const double *wflux, const double *mu,
                             const double *mu4, const double *rdis);
void lit_forcing(inmsom_bt_ctx *c);
// periodic wrap of a field (no-op without periodicity) and hh_update
// (Function/depth.f90:108-170) of (hqn, hun, hvn, hhn) from the surface sh
void lit_cyclize(inmsom_bt_ctx *c, double *f);
void lit_hh_update(inmsom_bt_ctx *c, const double *sh, double *hqn, double *hun, double *hvn, double *hhn);
// uv_diff4 (Function/vel_ssh.f90:187-317) of a substep with level-n surface
// sh and filtered velocities up, vp on a grid that is not periodic, as its
// two terms per component: t[2..5] = px, qx, py, qy on lcu / lcv interior
// points (0 elsewhere), which the caller adds to its uv_diff2 result as
// RHSx_dif + px + qx, RHSy_dif - py + qy.  t[0], t[1] receive the first
// sweep's fx, fy.  Two launches; all six planes are n doubles.  rowmet:
// the metrics as a 9 x ny row table (dx dy dxt dyt dxh dyh dxb dyb rlh_s)
// when they are uniform along x on columns 2..nx-1, with mb the packed mask
// bits (MB_*) of every cell; null for the grid's per-cell planes.
void lit_diff4_terms(inmsom_bt_ctx *c, const double *up, const double *vp, const double *sh, const double *mu4,
                     double *const t[6], const double *rowmet = nullptr, const uint8_t *mb = nullptr);

}  // namespace ibt
