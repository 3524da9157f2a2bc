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
// one level of stress_components, uv_trans + uv_diff2 (vorticity into
// vort), wind stress on U/V faces and the quadratic bottom friction
void lit_stress(inmsom_bt_ctx *c, const double *u, const double *v, double *str_t, double *str_s);
void lit_trans_diff2(inmsom_bt_ctx *c, const double *u, const double *v, double *vort, const double *hq,
                     const double *hu, const double *hv, const double *hh, const double *mu, const double *str_t,
                     const double *str_s, double *radv, double *rady, double *rdfx, double *rdfy);
void lit_surface_stress(inmsom_bt_ctx *c, const double *taux, const double *tauy, double *sx, double *sy);
void lit_bottom_friction(inmsom_bt_ctx *c, const double *uu, const double *vv, const double *hhu,
                         const double *hhv, const double *hhq, double *bf);

}  // namespace ibt
