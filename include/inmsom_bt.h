/*
 * inmsom_bt: MI355X (gfx950) implementation of INMSOM's explicit
 * leap-frog barotropic (external-mode shallow-water) step.
 *
 * C ABI.  Every 2-D field is a column-major (m fastest) array of nx*ny
 * values with the reference's bounds bnd_x1..bnd_x2 x bnd_y1..bnd_y2 on a
 * single process (Control/init_arrays.f90:44-70): nx, ny include two halo
 * cells per side and the computed interior is m = 3..nx-2, n = 3..ny-2
 * (1-based).  Host arrays are passed as plain pointers; all device memory
 * is owned by the context.  Every function returns an INMSOM_BT_* status.
 *
 * Calls on one context must come from one host thread at a time; the
 * context owns one HIP stream and all work is ordered on it.
 */
#ifndef INMSOM_BT_H
#define INMSOM_BT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define INMSOM_BT_OK 0
#define INMSOM_BT_EINVAL (-1)  /* bad argument: null pointer, size or flag */
#define INMSOM_BT_EDEVICE (-2) /* HIP runtime error (message in inmsom_bt_last_error) */
#define INMSOM_BT_ERANGE (-3)  /* state failed the range check (NaN or out of bounds) */
#define INMSOM_BT_ENOMEM (-4)  /* device allocation failed */
#define INMSOM_BT_ESTATE (-5)  /* call not valid in the context's current state */

/* Arithmetic of the fused step (inmsom_bt_run). */
#define INMSOM_BT_FP64_EXACT 0 /* FP64, the reference's operation order, no fused multiply-add */
#define INMSOM_BT_FP64 1       /* FP64 with reciprocal metrics and fused multiply-add */
#define INMSOM_BT_FP32 2       /* FP32 stencils; FP64 filtered level and FP64 Robert-Asselin filter */

typedef struct inmsom_bt_ctx inmsom_bt_ctx;

/* Static grid: the reference's module globals that barotropic_dynamics
 * reads (Modules/mod_basin_grid.f90:133-190, Inc/basinpar.fi).
 * Masks: either the seven real(4) masks of gridcon, or (when lu is NULL)
 * the integer mask lbasins (0 = ocean) from which they are built exactly
 * as Service/grid_construction.f90:33-133 does. */
typedef struct inmsom_bt_grid {
    int nx, ny;
    int periodicity_x, periodicity_y; /* basinpar.fi:13-14 */
    int full_free_surface;            /* basinpar.fi:46 */
    double free_fall_acc;             /* basinpar.fi:77 (FreeFallAcc) */
    double ref_den;                   /* basinpar.fi:76 (RefDen) */
    double time_smooth;               /* basinpar.fi:3 */
    const int *lbasins;
    const float *lu, *luu, *luh, *lcu, *lcv, *llu, *llv;
    const double *dx, *dy, *dxt, *dyt, *dxh, *dyh, *dxb, *dyb;
    const double *rlh_s;    /* Coriolis parameter on H points */
    const double *hhq_rest; /* depth at rest on T points */
} inmsom_bt_grid;

/* Per-run forcing of the 2-D model step (Function/fluxes.f90 inputs and
 * the barotropic_dynamics arrays that do not change between steps). */
typedef struct inmsom_bt_forcing {
    const double *taux, *tauy; /* T-point wind stress [Pa] */
    const double *wflux;       /* surface water flux, barotropic_dynamics arg wflux */
    const double *mu, *mu4;    /* 2-D lateral viscosities, args mu, mu4 */
    const double *rdis;        /* Rayleigh friction, arg rdis (r_diss) */
    double cb_nl, ebottom;     /* quadratic drag, basinpar.fi:66-68 */
} inmsom_bt_forcing;

/* Last error message of the calling thread ("" if none). */
const char *inmsom_bt_last_error(void);

/* Builds the seven C-grid masks from lbasins on the host, as
 * Service/grid_construction.f90:33-133 (one process). */
int inmsom_bt_gridcon(int nx, int ny, int periodicity_x, int periodicity_y,
                      const int *lbasins, float *lu, float *luu, float *luh,
                      float *lcu, float *lcv, float *llu, float *llv);

/* Rayleigh friction r_diss from the bottom slope on the host,
 * Control/init_pars.f90:459-481 (r_fric: basinpar.fi:70). */
int inmsom_bt_rayleigh(const inmsom_bt_grid *grid, double r_fric, double *r_diss);

/* Uploads the static grid to device 'device' and returns a context.
 * Replaces the module initialisation the reference performs in
 * Control/init_pars.f90 (ocean_model_parameters/ocinicond) for the
 * fields the barotropic step reads. */
int inmsom_bt_init(const inmsom_bt_grid *grid, int device, inmsom_bt_ctx **out);
int inmsom_bt_finalize(inmsom_bt_ctx *ctx);

/* Drop-in for barotropic_dynamics (Function/vel_ssh.f90:379-679).
 * Arguments keep the reference's order and meaning; arrays are host
 * column-major nx*ny.  In: *_e state, RHSx/RHSy, wflux, mu, mu4, rdis.
 * Out: *_e (state after nstep substeps), ubrtr_i, vbrtr_i, ssh4gradx,
 * ssh4grady.  The reference's scratch outputs (vort, str_t2d, str_s2d,
 * fx, fy, RHS*_adv, RHS*_dif) are recomputed before any later use by its
 * caller (Control/dyncall.f90:39,65,72-84) and are not returned.  The
 * context's external-mode depths are re-initialised from ssh_e/sshp_e at
 * the end, as hh_init does (:668-675); read them with
 * inmsom_bt_get_depths.
 * Path: one fused kernel launch per substep (bit for bit the reference)
 * when full_free_surface = 1, the grid is not periodic and the two halo
 * cells of every side are land (with ksw4 > 0 the biharmonic term of each
 * substep comes from two reference-order sweeps before its launch);
 * otherwise one kernel per reference sweep (also bit for bit).
 * inmsom_bt_set_dropin_path chooses. */
int inmsom_bt_barotropic_dynamics(inmsom_bt_ctx *ctx, double tau, int nstep, int ksw4,
                                  double *ubrtr_e, double *ubrtrp_e, double *vbrtr_e,
                                  double *vbrtrp_e, double *ssh_e, double *sshp_e,
                                  double *ubrtr_i, double *vbrtr_i, double *ssh4gradx,
                                  double *ssh4grady, const double *RHSx, const double *RHSy,
                                  const double *wflux, const double *mu, const double *mu4,
                                  const double *rdis);

/* As inmsom_bt_barotropic_dynamics (Function/vel_ssh.f90:379-679, called
 * at Control/dyncall.f90:137-164) with every array a device address on the
 * context's GPU (nx*ny doubles, ready: synchronise the producing stream
 * first): no host copies.  The context's own state, depths and averages
 * (inmsom_bt_get_state / _get_depths / _get_means) are not changed.
 * Returns when the call has completed. */
int inmsom_bt_barotropic_dynamics_dev(inmsom_bt_ctx *ctx, double tau, int nstep, int ksw4,
                                      double *ubrtr_e, double *ubrtrp_e, double *vbrtr_e,
                                      double *vbrtrp_e, double *ssh_e, double *sshp_e,
                                      double *ubrtr_i, double *vbrtr_i, double *ssh4gradx,
                                      double *ssh4grady, const double *RHSx, const double *RHSy,
                                      const double *wflux, const double *mu, const double *mu4,
                                      const double *rdis);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  As inmsom_bt_barotropic_dynamics_dev with ksw4 = 0 (so
 * no mu4), through the fused kernel only, with its scratch (the substeps'
 * two state sets and carried depths, 14 planes of nx*ny doubles) in a fixed
 * 384 MiB device array the library holds on every GPU it runs on until the
 * process ends: grids above 3,595,117 cells (nx*ny) fail with
 * INMSOM_BT_EINVAL, calls the fused kernel does not cover with
 * INMSOM_BT_ESTATE.  Calls no hipMalloc or hipFree; the work is queued on
 * the context's stream and the entry returns without waiting for it
 * (inmsom_bt_synchronize waits).  All calls on one GPU share the scratch:
 * calls that overlap on one GPU race. */
int inmsom_bt_barotropic_dynamics_noalloc(inmsom_bt_ctx *ctx, double tau, int nstep, double *ubrtr_e,
                                          double *ubrtrp_e, double *vbrtr_e, double *vbrtrp_e, double *ssh_e,
                                          double *sshp_e, double *ubrtr_i, double *vbrtr_i, double *ssh4gradx,
                                          double *ssh4grady, const double *RHSx, const double *RHSy,
                                          const double *wflux, const double *mu, const double *rdis);

/* Path of the barotropic_dynamics entries: 0 = fused where the call allows
 * it (default), 1 = one kernel per reference sweep, 2 = fused (calls it
 * does not cover fail with INMSOM_BT_ESTATE).  Results are the same bit
 * for bit.  get: the path (1 or 2) a call with this ksw4 takes. */
int inmsom_bt_set_dropin_path(inmsom_bt_ctx *ctx, int path);
int inmsom_bt_get_dropin_path(inmsom_bt_ctx *ctx, int ksw4, int *path);

/* ncalls device-resident calls of barotropic_dynamics on the context's
 * state (inmsom_bt_set_state / _get_state) with the RHS of the current
 * state's wind stress and bottom drag, the context's water flux, mu, mu4
 * and Rayleigh coefficient; *ms = their HIP-event time. */
int inmsom_bt_dropin_timed(inmsom_bt_ctx *ctx, double tau, int nstep, int ksw4, int ncalls, double *ms);

/* The twelve external-mode depths hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e,
 * hhun_e, hhv_e, hhvp_e, hhvn_e, hhh_e, hhhp_e, hhhn_e
 * (Modules/mod_basin_grid.f90:179-190) as the context holds them. */
int inmsom_bt_get_depths(inmsom_bt_ctx *ctx, double *const h[12]);

/* ---- device-resident 2-D model ------------------------------------- */

int inmsom_bt_set_forcing(inmsom_bt_ctx *ctx, const inmsom_bt_forcing *f);
/* Uploads (u, up, v, vp, ssh, sshp) = (ubrtr_e, ubrtrp_e, vbrtr_e,
 * vbrtrp_e, ssh_e, sshp_e) and re-initialises the depths (hh_init). */
int inmsom_bt_set_state(inmsom_bt_ctx *ctx, const double *u, const double *up,
                        const double *v, const double *vp, const double *ssh,
                        const double *sshp);
int inmsom_bt_get_state(inmsom_bt_ctx *ctx, double *u, double *up, double *v, double *vp,
                        double *ssh, double *sshp);
/* No counterpart in the reference (its restarts go through files): a copy
 * of the context's state kept on the device, and its restore, which is
 * inmsom_bt_set_state with the saved arrays (depths re-initialised)
 * without host copies; INMSOM_BT_ESTATE if nothing was saved.  For callers
 * that rerun from one state, e.g. perturbation runs or timing. */
int inmsom_bt_save_state(inmsom_bt_ctx *ctx);
int inmsom_bt_restore_state(inmsom_bt_ctx *ctx);
/* ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady of the last barotropic_dynamics
 * call on the context's state. */
int inmsom_bt_get_means(inmsom_bt_ctx *ctx, double *ubrtr_i, double *vbrtr_i,
                        double *ssh4gradx, double *ssh4grady);

/* nsteps outer steps of the 2-D model, each the sequence
 * Control/dyncall.f90:52-164 runs for a single-level flow: wind stress
 * (Function/fluxes.f90:29-37), quadratic bottom drag from the current
 * external-mode state (:226-272), RHS assembly (dyncall.f90:114-126 with
 * the 3-D terms zero) and barotropic_dynamics(tau, nstep, ksw4), all on
 * the device in the reference's operation order; barotropic_dynamics takes
 * the path inmsom_bt_barotropic_dynamics takes (inmsom_bt_set_dropin_path). */
int inmsom_bt_model_steps(inmsom_bt_ctx *ctx, double tau, int nstep, int ksw4, int nsteps);

/* nsteps outer steps with ksw4 = 0, each through nstep launches of one
 * fused kernel (one per state-advancing substep of tau/nstep).  The
 * state after each outer step is that of
 * inmsom_bt_model_steps(ctx, tau, nstep, 0, 1): the look-ahead substeps
 * nstep+1..2*nstep, which the reference runs only to form the averages
 * ubrtr_i/vbrtr_i/ssh4grad*, are not run; the forcing is formed at the
 * first substep and frozen over the call, as the reference forms it once
 * per outer step.  Requires uniform metrics (all eight metric arrays one
 * constant) and wind, water flux and Coriolis parameter uniform along x,
 * mu = constant * lu (checked at the first call).  mode is an
 * INMSOM_BT_FP* value; the first call in a mode converts the state. */
int inmsom_bt_run(inmsom_bt_ctx *ctx, double tau, int nstep, int nsteps, int mode);

/* Same as inmsom_bt_run, timed on the context's stream with HIP events:
 * *ms receives the elapsed time of the nsteps*nstep launches. */
int inmsom_bt_run_timed(inmsom_bt_ctx *ctx, double tau, int nstep, int nsteps, int mode,
                        double *ms);

/* Launch shape of the fused step: output rows marched by one 64-lane
 * wave (0 = default: 32 for the FP64 modes, 20 for FP32).  Results do
 * not depend on it. */
int inmsom_bt_set_launch(inmsom_bt_ctx *ctx, int rows_per_wave);
int inmsom_bt_get_launch(inmsom_bt_ctx *ctx, int *rows_per_wave);

/* Substeps of one call that the fused step advances per launch: 2 (with
 * nstep >= 2 each pair of substeps is one march over the rows, the second
 * substep taking the first's rows from registers; an odd last substep runs
 * alone), 1 (one launch per substep), or 0 (default: 2 in FP64 and FP32
 * from 2^20 cells on, 1 in FP64_EXACT and on smaller grids).  Results do
 * not depend on it. */
int inmsom_bt_set_substeps_per_launch(inmsom_bt_ctx *ctx, int substeps);
/* inmsom_bt_run in FP64 and FP32: stages = K in 2..4 runs K consecutive
 * substeps per launch (across outer steps) as a pipeline of K waves per
 * workgroup, one substep per wave, the rows handed over in LDS; 1 = off
 * (inmsom_bt_set_substeps_per_launch applies); 0 = automatic (default:
 * 2 stages from 2^20 cells on while the substeps per launch are automatic
 * too).  Results agree with one launch per substep to rounding (the order
 * of contracted multiply-adds may differ). */
int inmsom_bt_set_pipeline(inmsom_bt_ctx *ctx, int stages);
/* the pipeline stages the next inmsom_bt_run in this mode uses (0: none) */
int inmsom_bt_get_pipeline(inmsom_bt_ctx *ctx, int mode, int *stages);
/* the substeps per launch the fused step uses in `mode` (1 or 2, or the
 * pipeline's stages when it is on) */
int inmsom_bt_get_substeps_per_launch(inmsom_bt_ctx *ctx, int mode, int *substeps);

/* Range check of Control/dyncall.f90:452-516 on the device state:
 * *bad receives the number of wet cells with NaN, |ssh| >= 10 or
 * |u|,|v| >= 100.  Returns INMSOM_BT_ERANGE when *bad > 0. */
int inmsom_bt_range_check(inmsom_bt_ctx *ctx, long long *bad);

/* Volume anomaly sum(ssh * dx * dy) over wet interior cells, FP64,
 * summed in a fixed order (deterministic). */
int inmsom_bt_volume(inmsom_bt_ctx *ctx, double *volume);

int inmsom_bt_synchronize(inmsom_bt_ctx *ctx);

/* ---- batched ensembles --------------------------------------------- */

/* Many independent closed basins of different sizes advanced together:
 * every substep of every member is one persistent launch over a work list
 * of (member, strip, rows) items, and the launches of one outer step are
 * replayed from a HIP graph.  Members need what inmsom_bt_run needs
 * (uniform metrics, forcing uniform along x, mu = constant * lu) and must
 * share g, rho0, time_smooth and the drag parameters.  grids[k] and
 * forcing[k] describe member k (forcing[k].rdis must be set).  State per
 * member follows inmsom_bt_set_state/get_state. */
typedef struct inmsom_bt_ens inmsom_bt_ens;

int inmsom_bt_ens_create(int nmembers, const inmsom_bt_grid *grids,
                         const inmsom_bt_forcing *forcing, int mode, int device,
                         inmsom_bt_ens **out);
int inmsom_bt_ens_destroy(inmsom_bt_ens *ens);
/* member count, wet interior cells and array cells summed over members */
int inmsom_bt_ens_size(inmsom_bt_ens *ens, int *nmembers, long long *wet_cells, long long *cells);
int inmsom_bt_ens_set_state(inmsom_bt_ens *ens, int member, const double *u, const double *up,
                            const double *v, const double *vp, const double *ssh,
                            const double *sshp);
int inmsom_bt_ens_get_state(inmsom_bt_ens *ens, int member, double *u, double *up, double *v,
                            double *vp, double *ssh, double *sshp);
/* launch shape: output rows per wave (0 = default) and waves per XCD of
 * the persistent grid (0 = default); results do not depend on it */
int inmsom_bt_ens_set_launch(inmsom_bt_ens *ens, int rows_per_wave, int waves_per_xcd);
/* substeps per launch, as inmsom_bt_set_substeps_per_launch (0 = by mode) */
int inmsom_bt_ens_set_substeps_per_launch(inmsom_bt_ens *ens, int substeps);
int inmsom_bt_ens_get_substeps_per_launch(inmsom_bt_ens *ens, int *substeps);
/* substep pairs (FP64 and FP32) as a two-wave pipeline, as
 * inmsom_bt_set_pipeline: 2 = on, 1 = off, 0 = by mode (default: on);
 * waves_per_xcd of inmsom_bt_ens_set_launch then counts two-wave
 * workgroups.  get: the stages the next run uses (0: none) */
int inmsom_bt_ens_set_pipeline(inmsom_bt_ens *ens, int stages);
int inmsom_bt_ens_get_pipeline(inmsom_bt_ens *ens, int *stages);
/* nsteps outer steps of tau, nstep substeps each; use_graph = 1 replays a
 * captured graph per outer step, 0 launches directly */
int inmsom_bt_ens_run(inmsom_bt_ens *ens, double tau, int nstep, int nsteps, int use_graph);
int inmsom_bt_ens_run_timed(inmsom_bt_ens *ens, double tau, int nstep, int nsteps, int use_graph,
                            double *ms);
/* device-side snapshot of all members' state and its restore (used by the
 * launch-shape autotuner) */
int inmsom_bt_ens_save(inmsom_bt_ens *ens);
int inmsom_bt_ens_restore(inmsom_bt_ens *ens);
/* range check over all members (INMSOM_BT_ERANGE when *bad > 0) */
int inmsom_bt_ens_range_check(inmsom_bt_ens *ens, long long *bad);
/* FP64 volume anomaly of one member (fixed-order sum) */
int inmsom_bt_ens_volume(inmsom_bt_ens *ens, int member, double *volume);


/* The 2-D -> 3-D coupling that follows barotropic_dynamics in the model
 * step, Control/dyncall.f90:169-231 (host arrays in and out, like
 * inmsom_bt_barotropic_dynamics; 3-D arrays are (nz, ny, nx), i.e. the
 * reference's column-major (m, n, k)).  The caller runs vertical_velocity
 * between the two calls, as the reference does.
 *
 * inmsom_bt_remove_barotropic: uu(:,:,k) -= uu2d on lcu and
 * vv(:,:,k) -= vv2d on lcv, k = 1..nz, m = 2..nx-1, n = 2..ny-1
 * (dyncall.f90:169-184). */
int inmsom_bt_remove_barotropic(inmsom_bt_ctx *ctx, int nz, double *uu, double *vv,
                                const double *uu2d, const double *vv2d);
/* inmsom_bt_internal_update (dyncall.f90:188-231): uu2d = ubrtr_i / hhu and
 * pgrx -= (slpr(m+1) - slpr(m)) * hhu * dyh / RefDen on lcu (v likewise);
 * ssh_internal (Function/vel_ssh.f90:682-751): ssh_new from the continuity
 * equation with the time-mean transports ubrtr_i, vbrtr_i and wflux,
 * hh_update of hhn = {hhqn, hhun, hhvn, hhhn} when full_free_surface > 0,
 * and the Robert-Asselin step of (ssh_i, sshp_i) with time_smooth / 2;
 * then uu(:,:,k) += uu2d, vv(:,:,k) += vv2d.  uu2d, vv2d, pgrx, pgry,
 * ssh_i, sshp_i, hhn and uu, vv are read and written (cells the
 * reference does not write keep their values). */
int inmsom_bt_internal_update(inmsom_bt_ctx *ctx, double tau, int nz, const double *ubrtr_i,
                              const double *vbrtr_i, const double *hhu, const double *hhv,
                              const double *slpr, const double *wflux, double *uu2d, double *vv2d,
                              double *pgrx, double *pgry, double *ssh_i, double *sshp_i,
                              double *const hhn[4], double *uu, double *vv);
/* The same two steps on arrays already in device memory (device addresses
 * on the context's GPU, same shapes and meanings): no host copies, for a
 * caller whose 3-D fields live in HBM.  Work runs on the library's stream,
 * which does not wait for other streams: the arrays must be ready when
 * the call is made.  Both calls return after the work has finished. */
int inmsom_bt_remove_barotropic_dev(inmsom_bt_ctx *ctx, int nz, double *uu, double *vv,
                                    const double *uu2d, const double *vv2d);
int inmsom_bt_internal_update_dev(inmsom_bt_ctx *ctx, double tau, int nz, const double *ubrtr_i,
                                  const double *vbrtr_i, const double *hhu, const double *hhv,
                                  const double *slpr, const double *wflux, double *uu2d, double *vv2d,
                                  double *pgrx, double *pgry, double *ssh_i, double *sshp_i,
                                  double *const hhn[4], double *uu, double *vv);
/* The forcing of the barotropic step from the 3-D model, Control/
 * dyncall.f90:37-130 with ksw_vert = ksw_lat = ksw_lat4 = 0 (host arrays
 * in and out, like inmsom_bt_barotropic_dynamics): stress_components of
 * the level n-1 velocities uup, vvp (3-D) and uup2d, vvp2d (2-D),
 * sea_surface_fluxes from taux, tauy on T points, sea_bottom_fluxes from
 * the bottom level k = nz of uu, vv, uup, vvp (drag parameters of
 * inmsom_bt_set_forcing), uv_trans of (uu, vv) and (uu2d, vv2d) and
 * uv_diff2 with amuv and amuv2d, on the 3-D model's depths hhq, hhu, hhv,
 * hhh (Function/vel_ssh.f90:2-184); pgx, pgy (3-D, or both NULL) are what
 * pressure_gradients adds to the 3-D RHS; depth_ave with the layer
 * thicknesses dz[nz] (Function/depth.f90:223-254); and the assembly of
 * :109-126 into rhsx2d, rhsy2d, which are read and written (cells the
 * reference does not write keep their values).  3-D arrays are
 * (nz, ny, nx), i.e. the reference's column-major (m, n, k). */
int inmsom_bt_forcing_3d(inmsom_bt_ctx *ctx, int nz, const double *dz, const double *uu, const double *vv,
                         const double *uup, const double *vvp, const double *amuv, const double *uu2d,
                         const double *vv2d, const double *uup2d, const double *vvp2d, const double *amuv2d,
                         const double *hhq, const double *hhu, const double *hhv, const double *hhh,
                         const double *taux, const double *tauy, const double *slpr, const double *pgx,
                         const double *pgy, double *rhsx2d, double *rhsy2d);
/* The same forcing on arrays already in device memory (device addresses on
 * the context's GPU, same shapes and meanings; dz stays a host array of
 * nz values): no host copies, for a caller whose 3-D fields live in HBM.
 * rhsx2d, rhsy2d are updated in place.  Work runs on the library's
 * stream, which does not wait for other streams: the arrays must be ready
 * when the call is made; the call returns after the work has finished.
 * Control/dyncall.f90:37-130 as above. */
int inmsom_bt_forcing_3d_dev(inmsom_bt_ctx *ctx, int nz, const double *dz, const double *uu, const double *vv,
                             const double *uup, const double *vvp, const double *amuv, const double *uu2d,
                             const double *vv2d, const double *uup2d, const double *vvp2d, const double *amuv2d,
                             const double *hhq, const double *hhu, const double *hhv, const double *hhh,
                             const double *taux, const double *tauy, const double *slpr, const double *pgx,
                             const double *pgy, double *rhsx2d, double *rhsy2d);
/* Timing of its device work: nrep repetitions on the fields the last
 * inmsom_bt_forcing_3d call with nz levels left on the device (no host
 * copies); *ms = HIP-event time of all repetitions.  No counterpart in the
 * reference (a measurement entry beside inmsom_bt_run_timed). */
int inmsom_bt_forcing_3d_timed(inmsom_bt_ctx *ctx, int nz, int nrep, double *ms);

/* DIOGIN direct-access records (Service/input_output_data.f90): field
 * (nx, ny, nz) column-major (first index fastest, as the reference's
 * arrays), lu (nx, ny) the ocean mask; the subdomain [nxb, nxe] x [nyb, nye]
 * of levels nzb..nze goes to records (nfield-1)*(nze-nzb+1)+1 .. of the
 * file path/fname (joined as fulfname does, :4-114).  Records hold float32
 * (wdstd/rdstd, :117-327) or float64 (wdstd8/rdstd8, :329-533) values with
 * no record markers.  The writer stores land points (|lu| < 0.5) as
 * undef = -1e32 and leaves them 0 in field; the reader sets them to 0.
 * The file is created if missing and never truncated (status='unknown').
 * Errors return INMSOM_BT_EINVAL (the reference stops the run). */
int inmsom_bt_wdstd(const char *path, const char *fname, int nfield, float *field, const float *lu, int nx, int ny,
                    int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze);
int inmsom_bt_rdstd(const char *path, const char *fname, int nfield, float *field, const float *lu, int nx, int ny,
                    int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze);
int inmsom_bt_wdstd8(const char *path, const char *fname, int nfield, double *field, const float *lu, int nx, int ny,
                     int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze);
int inmsom_bt_rdstd8(const char *path, const char *fname, int nfield, double *field, const float *lu, int nx, int ny,
                     int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze);
/* GrADS descriptor of a DIOGIN file, ctl_file_write
 * (Service/rw_ctl_file.f90:1-138): writes fname with its extension
 * replaced by .ctl, naming fname's .dat in DSET; same arguments and text
 * as the reference's writer (x0/y0/z0: first value and step hx/hy/hz when
 * the type is 0, all values when 1; ht the time step in seconds, >= 60;
 * yr_type 0 adds OPTIONS 365_DAY_CALENDAR). */
int inmsom_bt_ctl_write(const char *fname, float undef, int nx, int ny, int nz, int nt, int xtype, const double *x0,
                        double hx, int ytype, const double *y0, double hy, int ztype, const double *z0, double hz,
                        int yr_type, int year0, int month0, int day0, int hour0, int minute0, float ht,
                        const char *title, const char *varname);

/* Snapshot of the external-mode state in the reference's output formats
 * (Control/output.f90:115-154, 155-190, 230-260): path/ssh.dat (lu,
 * T points m = 3..nx-2, n = 3..ny-2), path/ubrtr.dat (llu, m = 2..nx-2)
 * and path/vbrtr.dat (llv, n = 2..ny-2), float32 DIOGIN records (land =
 * undef -1e32), record nrec of each, and their GrADS descriptors (.ctl,
 * as ctl_file_write: t-dimension nrec, linear axes from the first T
 * point xt0, yt0 with steps dxst, dyst, the velocity axes shifted half a
 * step, calendar/year/month/day/hour/minute of the first record, tstep
 * seconds between records).  The fields are packed on the device and
 * copied once into pinned memory; *ms (if not NULL) = HIP-event time of
 * the packing and that copy.  No counterpart as one routine in the
 * reference (local_output writes the 3-D model's fields this way). */
int inmsom_bt_write_snapshot(inmsom_bt_ctx *ctx, const char *path, int nrec, double xt0, double dxst, double yt0,
                             double dyst, int calendar, int year, int month, int day, int hour, int minute,
                             float tstep, double *ms);

/* baroclinic_dynamics, Function/vel_ssh.f90:794-1027 (with factor8,
 * Service/oc_algorithms.f90:494-524), with the reference's arguments and
 * the module arrays it reads: uu, uup, vv, vvp, RHSx, RHSy, RHSx_tran,
 * RHSy_tran (nz, ny, nx); ww and nu (the reference's anzu) (nz+1, ny, nx);
 * pgrx, pgry, rdis (r_diss), tx_surf, ty_surf, tx_bot, ty_bot (ny, nx);
 * the vertical grid dz[nz], hzt[nz+1]; the 3-D model's depths hhu, hhup,
 * hhun, hhv, hhvp, hhvn, hhh (ny, nx).  Coriolis, metrics, masks and
 * time_smooth are the context's.  uu, uup, vv, vvp are updated in place
 * (host arrays).  nz in 2..128. */
int inmsom_bt_baroclinic_dynamics(inmsom_bt_ctx *ctx, double tau, int nz, double *uu, double *uup, double *vv,
                                  double *vvp, const double *ww, const double *nu, const double *rhsx,
                                  const double *rhsy, const double *rhsx_tran, const double *rhsy_tran,
                                  const double *pgrx, const double *pgry, const double *rdis, const double *tx_surf,
                                  const double *ty_surf, const double *tx_bot, const double *ty_bot,
                                  const double *dz, const double *hzt, const double *hhu, const double *hhup,
                                  const double *hhun, const double *hhv, const double *hhvp, const double *hhvn,
                                  const double *hhh);
/* The same on device addresses (dz, hzt stay host arrays); returns after
 * the work has finished. */
int inmsom_bt_baroclinic_dynamics_dev(inmsom_bt_ctx *ctx, double tau, int nz, double *uu, double *uup, double *vv,
                                      double *vvp, const double *ww, const double *nu, const double *rhsx,
                                      const double *rhsy, const double *rhsx_tran, const double *rhsy_tran,
                                      const double *pgrx, const double *pgry, const double *rdis,
                                      const double *tx_surf, const double *ty_surf, const double *tx_bot,
                                      const double *ty_bot, const double *dz, const double *hzt, const double *hhu,
                                      const double *hhup, const double *hhun, const double *hhv,
                                      const double *hhvp, const double *hhvn, const double *hhh);
/* Timing: nrep repetitions of the device work on the fields the last
 * inmsom_bt_baroclinic_dynamics call with nz levels left on the device
 * (each repetition advances them); *ms = HIP-event time of all.  No
 * counterpart in the reference. */
int inmsom_bt_baroclinic_timed(inmsom_bt_ctx *ctx, double tau, int nz, int nrep, double *ms);

/* vertical_velocity, Function/vel_ssh.f90:1030-1076: w (nz+1, ny, nx) of
 * the 3-D model from the continuity equation with u, v (nz, ny, nx), the
 * depths hu, hv, the water flux wflux (ny, nx; times full_free_surface)
 * and dz[nz], integrated from w(nz+1) = 0 on lu interior columns, then
 * wrapped on periodic boundaries.  w is read and written (cells the
 * reference does not compute keep their values).  Host arrays; the _dev
 * variant takes device addresses (dz a host array) and returns after the
 * work has finished. */
int inmsom_bt_vertical_velocity(inmsom_bt_ctx *ctx, int nz, const double *dz, const double *u, const double *v,
                                double *w, const double *hu, const double *hv, const double *wflux);
int inmsom_bt_vertical_velocity_dev(inmsom_bt_ctx *ctx, int nz, const double *dz, const double *u, const double *v,
                                    double *w, const double *hu, const double *hv, const double *wflux);
/* Timing: nrep repetitions of the device work on the fields the last
 * inmsom_bt_vertical_velocity call with nz levels staged; *ms = HIP-event
 * time of all.  No counterpart in the reference. */
int inmsom_bt_vertical_velocity_timed(inmsom_bt_ctx *ctx, int nz, int nrep, double *ms);

/* pressure_gradients, Function/vel_ssh.f90:320-376: the hydrostatic
 * pressure of the density den (nz, ny, nx) integrated down the sigma
 * levels z[nz] between the two T columns of each lcu (x) and lcv (y)
 * interior point, its gradient (times hhu dyh, hhv dxh, FreeFallAcc /
 * RefDen) subtracted from rhsx, rhsy (nz, ny, nx) in place; hhq, hhu, hhv
 * (ny, nx) are the 3-D model's depths (the reference reads z and them from
 * Modules/mod_basin_grid.f90).  Cells off lcu (lcv) keep their values.
 * Called on zeroed rhsx, rhsy it yields the pgx, pgy that
 * inmsom_bt_forcing_3d adds.  Host arrays; the _dev variant takes device
 * addresses (z a host array) and returns after the work has finished. */
int inmsom_bt_pressure_gradients(inmsom_bt_ctx *ctx, int nz, const double *z, const double *den, const double *hhq,
                                 const double *hhu, const double *hhv, double *rhsx, double *rhsy);
int inmsom_bt_pressure_gradients_dev(inmsom_bt_ctx *ctx, int nz, const double *z, const double *den,
                                     const double *hhq, const double *hhu, const double *hhv, double *rhsx,
                                     double *rhsy);
/* Timing: nrep repetitions of the device work on the fields the last
 * inmsom_bt_pressure_gradients call with nz levels staged (each repetition
 * subtracts again); *ms = HIP-event time of all.  No counterpart in the
 * reference. */
int inmsom_bt_pressure_gradients_timed(inmsom_bt_ctx *ctx, int nz, int nrep, double *ms);

/* The density definition of the 3-D step, Control/dyncall.f90:437-449,
 * with the equation of state of Service/oc_algorithms.f90:138-161 (Brydon
 * et al. 1999): on lu points of m = 2..nx-1, n = 2..ny-1 and every level,
 * den = density(tt, ss, FreeFallAcc * RefDen * hhq * z(k)) and den_pot =
 * density(tt, ss, 0), deviations from 1020 kg/m^3; tt, ss, den, den_pot
 * (nz, ny, nx), hhq (ny, nx), z[nz].  den, den_pot are read and written
 * (cells the reference does not write keep their values); den is what
 * inmsom_bt_pressure_gradients reads.  Host arrays; the _dev variant takes
 * device addresses (z a host array) and returns after the work has
 * finished. */
int inmsom_bt_density(inmsom_bt_ctx *ctx, int nz, const double *z, const double *tt, const double *ss,
                      const double *hhq, double *den, double *den_pot);
int inmsom_bt_density_dev(inmsom_bt_ctx *ctx, int nz, const double *z, const double *tt, const double *ss,
                          const double *hhq, double *den, double *den_pot);
/* Timing: nrep repetitions of the device work on the fields the last
 * inmsom_bt_density call with nz levels staged; *ms = HIP-event time of
 * all.  No counterpart in the reference. */
int inmsom_bt_density_timed(inmsom_bt_ctx *ctx, int nz, int nrep, double *ms);

/* stress_components, Function/mixing.f90:115-168, of the 3-D velocities u,
 * v (nz, ny, nx), as Control/dyncall.f90:36 calls it for the Smagorinsky
 * viscosity: on lu points of the interior str_t = dy/dx (u/dyh - u_w/dyh_w)
 * - dx/dy (v/dxh - v_s/dxh_s), on luu points str_s = dxb/dyb (u_n/dxt_n -
 * u/dxt) + dyb/dxb (v_e/dyt_e - v/dyt), every level, then both wrapped on
 * periodic boundaries.  str_t, str_s (nz, ny, nx) are read and written
 * (cells the reference does not write keep their values).  Host arrays;
 * the _dev variant takes device addresses and returns after the work has
 * finished; _timed repeats the device work on the fields the last
 * host-array call staged (*ms = HIP-event time of all; no counterpart in
 * the reference). */
int inmsom_bt_stress_components(inmsom_bt_ctx *ctx, int nz, const double *u, const double *v, double *str_t,
                                double *str_s);
int inmsom_bt_stress_components_dev(inmsom_bt_ctx *ctx, int nz, const double *u, const double *v,
                                    double *str_t, double *str_s);
int inmsom_bt_stress_components_timed(inmsom_bt_ctx *ctx, int nz, int nrep, double *ms);

/* The lateral viscosity the 3-D step hands the barotropic step when
 * ksw_lat > 0, Control/dyncall.f90:40-43: smagorinsky_coeff
 * (Function/mixing.f90:171-219) and depth_ave(amuv, amuv2d, lu, 0)
 * (Function/depth.f90:223-254).  On lu points of the interior and every
 * level, mu = (0.4 / pi)^2 dx dy sqrt(str_t^2 + mean of str_s^2 over the
 * four H corners of the T cell), amuv = visc2 + mu, amts = dift + 0.1 mu
 * (visc2, dift: lvisc_2, ldiff_ts of the reference); both wrapped on
 * periodic boundaries; then amuv2d = sum_k amuv dz on lu points of
 * m = 2..nx-1, n = 2..ny-1 -- the mu argument of barotropic_dynamics.
 * str_t, str_s (nz, ny, nx) from stress_components; amts, amuv (nz, ny, nx)
 * and amuv2d (ny, nx) are read and written (cells the reference does not
 * write keep their values).  Host arrays; the _dev variant takes device
 * addresses (dz a host array) and returns after the work has finished. */
int inmsom_bt_smagorinsky(inmsom_bt_ctx *ctx, int nz, double visc2, double dift, const double *str_t,
                          const double *str_s, const double *dz, double *amts, double *amuv, double *amuv2d);
int inmsom_bt_smagorinsky_dev(inmsom_bt_ctx *ctx, int nz, double visc2, double dift, const double *str_t,
                              const double *str_s, const double *dz, double *amts, double *amuv,
                              double *amuv2d);
/* Timing: nrep repetitions of the device work on the fields the last
 * inmsom_bt_smagorinsky call with nz levels staged; *ms = HIP-event time
 * of all.  No counterpart in the reference. */
int inmsom_bt_smagorinsky_timed(inmsom_bt_ctx *ctx, int nz, int nrep, double *ms);

/* Timing of the coupling's device work: nrep repetitions of the kernels of
 * inmsom_bt_remove_barotropic then inmsom_bt_internal_update on the fields
 * the last inmsom_bt_internal_update call left on the device (it must have
 * had at least nz levels; INMSOM_BT_ESTATE otherwise), without the host
 * copies; *ms = HIP-event time of all repetitions.  Each repetition
 * advances those device fields; nothing is copied back.  No counterpart in
 * the reference (a measurement entry beside inmsom_bt_run_timed). */
int inmsom_bt_coupling_timed(inmsom_bt_ctx *ctx, double tau, int nz, int nrep, double *ms);

/* Block average of the current state onto a grid `factor` times coarser
 * (what a coarse model of the same ocean sees; used to build the learned
 * closure's training targets and to judge coarse runs).  Interior blocks of
 * factor x factor cells; outputs are ((ny-4)/factor+4, (nx-4)/factor+4)
 * arrays with zero halos: ssh = mean over the block's wet cells, u = mean
 * of the fine U faces on the block's east edge, v = mean of the fine V
 * faces on its north edge.  No counterpart in the reference. */
int inmsom_bt_coarsen_state(inmsom_bt_ctx *ctx, int factor, double *u, double *v, double *ssh);

/* ------------------------------------------------------------------------
 * Learned lateral-momentum closure for coarse grids.  No counterpart in the
 * reference (its lateral momentum closure is the constant/Smagorinsky
 * viscosity of Function/mixing.f90:1-146); the plain viscosity stays the
 * fallback.  A 4-layer 3x3 CNN, 4 -> 32 -> 32 -> 32 -> 2 channels with ReLU
 * after the first three layers, BF16 weights and activations, FP32
 * accumulation on the matrix cores.  Inputs per grid point: u, v, ssh and
 * hhq_rest of the basin's (ny, nx) arrays, shifted and scaled; outputs: the
 * u and v tendencies [m/s^2] on U and V points.
 *
 * Parameter vector (float): per layer l = 0..3 the weights W[32][kpad_l]
 * then the bias b[32], with cin_0 = 8 (channels 4..7 zero), cin_l = 32,
 * kpad_l = 9 cin_l rounded up to 16, and k = t cin_l + ci where tap
 * t = 3 (dy + 1) + (dx + 1) reads input (y + dy, x + dx); zero padding.
 * Layer 3 uses output channels 0 and 1 only.
 *
 * The object also trains itself on the device: Adam (beta1 0.9, beta2 0.999,
 * eps 1e-8) on FP32 master weights, on the masked mean-square error of the
 * two tendencies over a batch of `batch` images of ny x nx. */
typedef struct inmsom_bt_cnn inmsom_bt_cnn;

int inmsom_bt_cnn_create(int batch, int ny, int nx, int device, inmsom_bt_cnn **out);
int inmsom_bt_cnn_destroy(inmsom_bt_cnn *net);
int inmsom_bt_cnn_num_params(inmsom_bt_cnn *net, long long *n);
/* x_c = (field_c - in_shift[c]) * in_scale[c] for (u, v, ssh, depth);
 * tendency = output * out_scale */
int inmsom_bt_cnn_set_norm(inmsom_bt_cnn *net, const float *in_shift, const float *in_scale,
                           const float *out_scale);
/* what the network's outputs mean: INMSOM_BT_CNN_UV (default) the u and v
 * tendencies (outputs 0, 1 times out_scale[0], out_scale[1]);
 * INMSOM_BT_CNN_PSI a stream-function tendency psi on H points (output 0
 * times out_scale[0]) whose differences are the tendencies,
 * du(m, n) = -(psi(m, n) - psi(m, n-1)), dv(m, n) = psi(m, n) - psi(m-1, n):
 * divergence-free on the C-grid (the grid step is folded into out_scale).
 * Applies to training (loss) and to the closure in the step. */
#define INMSOM_BT_CNN_UV 0
#define INMSOM_BT_CNN_PSI 1
int inmsom_bt_cnn_set_output(inmsom_bt_cnn *net, int kind);
/* The closure's inference in the step: 2 runs the four layers in one pass
 * per tile (the halo rings recomputed: fewer launches, more products), 1
 * one pass per layer (the first two fused), 0 (default) the one pass for
 * batches of at most 4096 tiles of 8 x 32 pixels.  The tendencies are the
 * same bits either way.  No counterpart in the reference. */
int inmsom_bt_cnn_set_fused(inmsom_bt_cnn *net, int mode);
/* set (and reset the Adam moments) / read the FP32 master parameters */
int inmsom_bt_cnn_set_params(inmsom_bt_cnn *net, const float *params);
int inmsom_bt_cnn_get_params(inmsom_bt_cnn *net, float *params);
/* gradient of the last inmsom_bt_cnn_loss_grad / _train_step */
int inmsom_bt_cnn_get_grad(inmsom_bt_cnn *net, float *grad);
/* one batch: x [batch][ny][nx][8] network inputs (rounded to BF16),
 * target [batch][ny][nx][2] (may be NULL), mask [batch][ny][nx] with bit 0
 * = U point counted in the loss, bit 1 = V point (may be NULL) */
int inmsom_bt_cnn_set_batch(inmsom_bt_cnn *net, const float *x, const float *target,
                            const uint8_t *mask);
/* forward of the batch: out [batch][ny][nx][2] (may be NULL) */
int inmsom_bt_cnn_forward(inmsom_bt_cnn *net, float *out);
/* hidden activation of the last forward (layer 0..2): [batch][ny][nx][32] */
int inmsom_bt_cnn_get_activation(inmsom_bt_cnn *net, int layer, float *out);
/* loss and gradient of the batch without an update */
int inmsom_bt_cnn_loss_grad(inmsom_bt_cnn *net, double *loss);
/* keep dL/dz of every layer of the next backward passes (on = 1: four
 * batch-sized BF16 buffers; 0 frees them), and read layer `layer`'s
 * (0..3; z_3 = the outputs, z_0..2 before the ReLU) as float, [B, H, W,
 * 32] channels-last.  No reference counterpart: the network's per-layer
 * parity tests read them */
int inmsom_bt_cnn_keep_backprop(inmsom_bt_cnn *net, int on);
int inmsom_bt_cnn_get_backprop(inmsom_bt_cnn *net, int layer, float *out);
/* device-resident training set of nsamples images and one Adam step on the
 * samples idx[0..batch-1]; *loss is the batch loss before the step */
int inmsom_bt_cnn_dataset_alloc(inmsom_bt_cnn *net, int nsamples);
int inmsom_bt_cnn_dataset_set(inmsom_bt_cnn *net, int i, const float *x, const float *target,
                              const uint8_t *mask);
int inmsom_bt_cnn_train_step(inmsom_bt_cnn *net, const int *idx, float lr, double *loss);
/* loss of the samples idx[0..batch-1] without a gradient or an update */
int inmsom_bt_cnn_eval(inmsom_bt_cnn *net, const int *idx, double *loss);
int inmsom_bt_cnn_synchronize(inmsom_bt_cnn *net);

/* Attach a closure to an ensemble (NULL detaches): after every outer step
 * the network runs on every member (batched BF16-MFMA inference over the
 * uneven basins, inside the per-step graph) and u += tau * tendency on U/V
 * points, except for members whose closure-tendency L2 norm exceeds
 * `ratio` times that of their plain viscous tendency nu * Laplacian(u, v):
 * those keep the plain step (per step).  The network must outlive the
 * attachment; its weights are read at every step. */
int inmsom_bt_ens_set_closure(inmsom_bt_ens *ens, inmsom_bt_cnn *net, double ratio);
/* the closure after every `every`-th outer step (1 = every step, the
 * default; 2, 4 or 8), adding `every` steps' worth of its tendency
 * (u += every * tau * tendency); the count restarts at set_closure,
 * set_state and restore */
int inmsom_bt_ens_set_closure_interval(inmsom_bt_ens *ens, int every);
/* per member: steps that fell back to plain viscosity, and the closure and
 * viscous tendency norms of the last step (either pointer may be NULL) */
int inmsom_bt_ens_closure_stats(inmsom_bt_ens *ens, int *nfallback, double *norms);

#ifdef __cplusplus
}
#endif
#endif
