/*
 * CPU restatement of INMSOM's explicit leap-frog barotropic step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never
 * calls into it.
 *
 * Conventions (identical to the reference on one MPI rank):
 *   - every 2-D field is a column-major (m fastest) array of nx*ny
 *     doubles (masks: floats), indexed with the reference's 1-based
 *     (m,n); bnd_x1 = bnd_y1 = 1, bnd_x2 = nx, bnd_y2 = ny
 *     (Control/init_arrays.f90:44-70 with p_size = (1,1));
 *   - the computed interior is m = 3..nx-2, n = 3..ny-2
 *     (nx_start = mmm = 3, nx_end = mm = nx-2, Inc/basinpar.fi:7-10);
 *   - arithmetic is FP64 and follows the reference's evaluation order
 *     term by term; the library is compiled with -ffp-contract=off so
 *     that no multiply-add is fused.
 */
#ifndef BT_ORACLE_H
#define BT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny;                  /* array extents incl. the 2-cell halo */
    int periodicity_x, periodicity_y;
    int full_free_surface;
    double free_fall_acc;        /* FreeFallAcc */
    double ref_den;              /* RefDen */
    double time_smooth;          /* Robert-Asselin coefficient */
    /* real(4) masks, Service/grid_construction.f90:33-112 */
    const float *lu, *luu, *luh, *lcu, *lcv, *llu, *llv;
    /* metrics, Modules/mod_basin_grid.f90:161-164 */
    const double *dx, *dy, *dxt, *dyt, *dxh, *dyh, *dxb, *dyb;
    const double *rlh_s;         /* Coriolis parameter on H points */
    const double *hhq_rest;      /* depth at rest on T points */
} bto_grid;

/* The twelve external-mode depth arrays (Modules/mod_basin_grid.f90:179-190):
 * index 0..11 = hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e,
 *               hhv_e, hhvp_e, hhvn_e, hhh_e, hhhp_e, hhhn_e           */
typedef struct { double *h[12]; } bto_depths;

/* Service/grid_construction.f90:33-112: the seven C-grid masks from the
 * integer land mask (0 = ocean).  Outputs are zero-filled first. */
void bto_gridcon(int nx, int ny, int periodicity_x, int periodicity_y,
                 const int *lbasins, float *lu, float *luu, float *luh,
                 float *lcu, float *lcv, float *llu, float *llv);

/* Control/init_pars.f90:459-481: Rayleigh friction from bottom slope. */
void bto_rayleigh_init(const bto_grid *g, double r_fric, double *r_diss);

/* Function/depth.f90:2-105 */
void bto_hh_init(const bto_grid *g, bto_depths *d, const double *ssh,
                 const double *sshp);
/* Function/depth.f90:108-170 */
void bto_hh_update(const bto_grid *g, double *hqn, double *hun, double *hvn,
                   double *hhn, const double *sh);
/* Function/depth.f90:174-220 */
void bto_hh_shift(const bto_grid *g, bto_depths *d, int nstep);

/* Function/mixing.f90:115-168 (nlev = 1) */
void bto_stress_components(const bto_grid *g, const double *u, const double *v,
                           double *str_t, double *str_s);
/* Function/vel_ssh.f90:2-114 (nlev = 1) */
void bto_uv_trans(const bto_grid *g, const double *u, const double *v,
                  double *vort, const double *hq, const double *hu,
                  const double *hv, const double *hh, double *rhsx,
                  double *rhsy);
/* Function/vel_ssh.f90:117-184 (nlev = 1) */
void bto_uv_diff2(const bto_grid *g, const double *mu, const double *str_t,
                  const double *str_s, const double *hq, const double *hu,
                  const double *hv, const double *hh, double *rhsx,
                  double *rhsy);
/* Function/vel_ssh.f90:187-317 (nlev = 1) */
void bto_uv_diff4(const bto_grid *g, const double *mu, double *str_t,
                  double *str_s, double *fx, double *fy, const double *hq,
                  const double *hu, const double *hv, const double *hh,
                  double *rhsx, double *rhsy);

/* Function/vel_ssh.f90:379-679.  Argument order follows the reference. */
void bto_barotropic_dynamics(const bto_grid *g, bto_depths *d, double tau,
                             int nstep, int ksw4,
                             double *ubrtr_e, double *ubrtrp_e,
                             double *vbrtr_e, double *vbrtrp_e,
                             double *ssh_e, double *sshp_e,
                             double *ubrtr_i, double *vbrtr_i,
                             double *ssh4gradx, double *ssh4grady,
                             const double *rhsx, const double *rhsy,
                             const double *wflux, const double *mu,
                             const double *mu4, double *vort, double *str_t2d,
                             double *str_s2d, double *fx, double *fy,
                             const double *rdis, double *rhsx_adv,
                             double *rhsy_adv, double *rhsx_dif,
                             double *rhsy_dif);

/* Function/fluxes.f90:29-37: wind stress averaged to U/V faces
 * (the ksw_ssbc = 1,2 branch). */
void bto_surface_stress(const bto_grid *g, const double *taux,
                        const double *tauy, double *surf_x, double *surf_y);

/* Function/fluxes.f90:226-272 with type_fric = 2, applied to a
 * single-level (2-D) flow: uu/vv are the current velocities, uup/vvp the
 * previous ones, hhu/hhv/hhq the depths the drag is evaluated with. */
void bto_bottom_stress(const bto_grid *g, double cb_nl, double ebottom,
                       const double *uu, const double *vv, const double *uup,
                       const double *vvp, const double *hhu, const double *hhv,
                       const double *hhq, double *bottom_friction,
                       double *bot_x, double *bot_y);

/* Control/dyncall.f90:110-130 restricted to a 2-D flow: every 3-D
 * contribution (dispersion, transport, sea-level pressure) is zero, so
 * RHSx2d = (surf + bot) * dxt * dyh on lcu and likewise on lcv. */
void bto_rhs_assembly(const bto_grid *g, const double *surf_x,
                      const double *surf_y, const double *bot_x,
                      const double *bot_y, double *rhsx2d, double *rhsy2d);

/* One 2-D model step, the sequence Control/dyncall.f90:52-164 executes
 * for a single-level flow: surface stress, bottom drag from the current
 * external-mode state and depths, RHS assembly, then
 * barotropic_dynamics(tau, nstep, ksw4).  All scratch arrays live in
 * `work` (at least 16*nx*ny doubles, zero-filled by the caller once). */
void bto_model_step(const bto_grid *g, bto_depths *d, double tau, int nstep,
                    int ksw4, double cb_nl, double ebottom, double *ubrtr_e,
                    double *ubrtrp_e, double *vbrtr_e, double *vbrtrp_e,
                    double *ssh_e, double *sshp_e, double *ubrtr_i,
                    double *vbrtr_i, double *ssh4gradx, double *ssh4grady,
                    const double *taux, const double *tauy,
                    const double *wflux, const double *mu, const double *mu4,
                    const double *rdis, double *work);

/* Control/dyncall.f90:37-130 (ksw_vert = ksw_lat = ksw_lat4 = 0): the 2-D
 * forcing RHSx2d, RHSy2d of the barotropic step from the 3-D and 2-D
 * velocities (stress components, surface and bottom stresses, uv_trans
 * and uv_diff2 at nz levels and in 2-D, the caller's 3-D pressure-gradient
 * terms or NULL, depth_ave with dz, assembly of :109-126).  3-D arrays are
 * (nz, ny, nx); rhsx2d, rhsy2d are read and written. */
void bto_forcing_3d(const bto_grid *g, int nz, const double *dz, double cb_nl, double ebottom,
                    const double *uu, const double *vv, const double *uup, const double *vvp,
                    const double *amuv, const double *uu2d, const double *vv2d,
                    const double *uup2d, const double *vvp2d, const double *amuv2d,
                    const double *hhq, const double *hhu, const double *hhv, const double *hhh,
                    const double *taux, const double *tauy, const double *slpr,
                    const double *pgx, const double *pgy, double *rhsx2d, double *rhsy2d);

/* Control/dyncall.f90:169-184: the barotropic part removed from the 3-D
 * velocity, uu(:,:,k) -= uu2d on lcu and vv(:,:,k) -= vv2d on lcv for
 * k = 1..nz over m = 2..nx-1, n = 2..ny-1; uu, vv are (nz, ny, nx). */
void bto_remove_barotropic(const bto_grid *g, int nz, double *uu, double *vv,
                           const double *uu2d, const double *vv2d);

/* Control/dyncall.f90:188-231 without vertical_velocity: uu2d = ubrtr_i/hhu
 * and pgrx -= (slpr(m+1)-slpr(m))*hhu*dyh/RefDen on lcu (v likewise on
 * lcv); ssh_internal (Function/vel_ssh.f90:682-751): continuity with the
 * time-mean transports into ssh_new, hh_update of (hhqn, hhun, hhvn, hhhn)
 * from it when full_free_surface > 0, and the Asselin step of
 * (ssh_i, sshp_i) with time_smooth/2; then uu(:,:,k) += uu2d, vv += vv2d. */
void bto_internal_update(const bto_grid *g, double tau, int nz,
                         const double *ubrtr_i, const double *vbrtr_i,
                         const double *hhu, const double *hhv, const double *slpr,
                         const double *wflux, double *uu2d, double *vv2d,
                         double *pgrx, double *pgry, double *ssh_i, double *sshp_i,
                         double *hhqn, double *hhun, double *hhvn, double *hhhn,
                         double *uu, double *vv);

/* Function/vel_ssh.f90:794-1027 (baroclinic_dynamics) with factor8
 * (Service/oc_algorithms.f90:494-524): 3-D arrays (nz, ny, nx), ww and nu
 * (nz+1, ny, nx), dz[nz], hzt[nz+1]; uu, uup, vv, vvp read and written.
 * nz >= 2. */
void bto_baroclinic_dynamics(const bto_grid *g, int nz, double tau, const double *dz, const double *hzt,
                             double *uu, double *uup, double *vv, double *vvp, const double *ww,
                             const double *nu, const double *rhsx, const double *rhsy,
                             const double *rhsx_tran, const double *rhsy_tran, const double *pgrx,
                             const double *pgry, const double *rdis, const double *tx_surf,
                             const double *ty_surf, const double *tx_bot, const double *ty_bot,
                             const double *hhu, const double *hhup, const double *hhun, const double *hhv,
                             const double *hhvp, const double *hhvn, const double *hhh);

/* Function/vel_ssh.f90:1030-1076 (vertical_velocity): u, v (nz, ny, nx),
 * w (nz+1, ny, nx) read and written, dz[nz]; hu, hv, wflux (ny, nx). */
/* Function/mixing.f90:171-219 + Function/depth.f90:223-254 (kbcl = 0) */
void bto_smagorinsky(const bto_grid *g, int nz, double visc2, double dift, const double *str_t,
                     const double *str_s, const double *dz, double *amts, double *amuv, double *amuv2d);
void bto_vertical_velocity(const bto_grid *g, int nz, const double *dz, const double *u, const double *v, double *w,
                           const double *hu, const double *hv, const double *wflux);

/* Function/vel_ssh.f90:320-376 (pressure_gradients): den, rhsx, rhsy
 * (nz, ny, nx), rhsx/rhsy updated in place on lcu/lcv; z[nz]; hhq, hhu,
 * hhv (ny, nx). */
void bto_pressure_gradients(const bto_grid *g, int nz, const double *z, const double *den, const double *hhq,
                            const double *hhu, const double *hhv, double *rhsx, double *rhsy);

/* Control/dyncall.f90:437-449 with Service/oc_algorithms.f90:138-161: den
 * (at the level's pressure) and den_pot (at zero pressure) from tt, ss on
 * lu points of m = 2..nx-1, n = 2..ny-1; (nz, ny, nx) arrays, z[nz]. */
void bto_density(const bto_grid *g, int nz, const double *z, const double *tt, const double *ss, const double *hhq,
                 double *den, double *den_pot);

/* Wet-cell volume anomaly sum(eta * dx * dy) over lu interior cells,
 * summed row by row in a fixed order. */
double bto_volume(const bto_grid *g, const double *ssh);

/* OpenMP threads the oracle's loops run on. */
int bto_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
