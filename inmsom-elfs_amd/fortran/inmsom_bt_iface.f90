! ISO_C_BINDING interface of libinmsom_bt (include/inmsom_bt.h).
!
! A Fortran host uses this module in place of the reference's
! Function/vel_ssh.f90 barotropic part:
!   * bt_gpu_init(...)            uploads the static grid once
!                                 (the module globals of Modules/mod_basin_grid.f90);
!   * barotropic_dynamics_gpu(...) has the argument list of the reference's
!                                 barotropic_dynamics (Function/vel_ssh.f90:379-406)
!                                 and runs it on the GPU;
!   * remove_barotropic_gpu(...) and internal_update_gpu(...) replace the
!                                 2-D -> 3-D coupling loops around vertical_velocity
!                                 (Control/dyncall.f90:169-231, with ssh_internal);
!   * the inmsom_bt_* interfaces expose the whole C ABI.
module inmsom_bt
use iso_c_binding
implicit none

integer(c_int), parameter :: INMSOM_BT_OK = 0, INMSOM_BT_EINVAL = -1, INMSOM_BT_EDEVICE = -2, &
                             INMSOM_BT_ERANGE = -3, INMSOM_BT_ENOMEM = -4, INMSOM_BT_ESTATE = -5
integer(c_int), parameter :: INMSOM_BT_FP64_EXACT = 0, INMSOM_BT_FP64 = 1, INMSOM_BT_FP32 = 2

! struct inmsom_bt_grid
type, bind(C) :: bt_grid
  integer(c_int) :: nx, ny, periodicity_x, periodicity_y, full_free_surface
  real(c_double) :: free_fall_acc, ref_den, time_smooth
  type(c_ptr) :: lbasins
  type(c_ptr) :: lu, luu, luh, lcu, lcv, llu, llv
  type(c_ptr) :: dx, dy, dxt, dyt, dxh, dyh, dxb, dyb
  type(c_ptr) :: rlh_s, hhq_rest
end type bt_grid

! struct inmsom_bt_forcing
type, bind(C) :: bt_forcing
  type(c_ptr) :: taux, tauy, wflux, mu, mu4, rdis
  real(c_double) :: cb_nl, ebottom
end type bt_forcing

! the context barotropic_dynamics_gpu works on
type(c_ptr), save :: bt_ctx = c_null_ptr

interface
  function inmsom_bt_last_error() bind(C, name="inmsom_bt_last_error")
    import :: c_ptr
    type(c_ptr) :: inmsom_bt_last_error
  end function
  integer(c_int) function inmsom_bt_gridcon(nx, ny, px, py, lbasins, lu, luu, luh, lcu, lcv, llu, llv) &
      bind(C, name="inmsom_bt_gridcon")
    import :: c_int, c_ptr
    integer(c_int), value :: nx, ny, px, py
    type(c_ptr), value :: lbasins, lu, luu, luh, lcu, lcv, llu, llv
  end function
  integer(c_int) function inmsom_bt_rayleigh(grid, r_fric, r_diss) bind(C, name="inmsom_bt_rayleigh")
    import :: c_int, c_double, c_ptr, bt_grid
    type(bt_grid), intent(in) :: grid
    real(c_double), value :: r_fric
    type(c_ptr), value :: r_diss
  end function
  integer(c_int) function inmsom_bt_init(grid, device, ctx) bind(C, name="inmsom_bt_init")
    import :: c_int, c_ptr, bt_grid
    type(bt_grid), intent(in) :: grid
    integer(c_int), value :: device
    type(c_ptr), intent(out) :: ctx
  end function
  integer(c_int) function inmsom_bt_finalize(ctx) bind(C, name="inmsom_bt_finalize")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function
  integer(c_int) function inmsom_bt_barotropic_dynamics(ctx, tau, nstep, ksw4, ubrtr_e, ubrtrp_e, &
      vbrtr_e, vbrtrp_e, ssh_e, sshp_e, ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady, rhsx, rhsy, wflux, &
      mu, mu4, rdis) bind(C, name="inmsom_bt_barotropic_dynamics")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nstep, ksw4
    type(c_ptr), value :: ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e, ubrtr_i, vbrtr_i, &
                          ssh4gradx, ssh4grady, rhsx, rhsy, wflux, mu, mu4, rdis
  end function
  integer(c_int) function inmsom_bt_get_depths(ctx, h) bind(C, name="inmsom_bt_get_depths")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    type(c_ptr), intent(in) :: h(12)
  end function
  integer(c_int) function inmsom_bt_set_forcing(ctx, f) bind(C, name="inmsom_bt_set_forcing")
    import :: c_int, c_ptr, bt_forcing
    type(c_ptr), value :: ctx
    type(bt_forcing), intent(in) :: f
  end function
  integer(c_int) function inmsom_bt_set_state(ctx, u, up, v, vp, ssh, sshp) bind(C, name="inmsom_bt_set_state")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx, u, up, v, vp, ssh, sshp
  end function
  integer(c_int) function inmsom_bt_get_state(ctx, u, up, v, vp, ssh, sshp) bind(C, name="inmsom_bt_get_state")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx, u, up, v, vp, ssh, sshp
  end function
  integer(c_int) function inmsom_bt_get_means(ctx, ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady) &
      bind(C, name="inmsom_bt_get_means")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx, ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady
  end function
  integer(c_int) function inmsom_bt_model_steps(ctx, tau, nstep, ksw4, nsteps) bind(C, name="inmsom_bt_model_steps")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nstep, ksw4, nsteps
  end function
  integer(c_int) function inmsom_bt_run(ctx, tau, nstep, nsteps, mode) bind(C, name="inmsom_bt_run")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nstep, nsteps, mode
  end function
  integer(c_int) function inmsom_bt_run_timed(ctx, tau, nstep, nsteps, mode, ms) bind(C, name="inmsom_bt_run_timed")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nstep, nsteps, mode
    real(c_double), intent(out) :: ms
  end function
  integer(c_int) function inmsom_bt_set_launch(ctx, rows) bind(C, name="inmsom_bt_set_launch")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: rows
  end function
  integer(c_int) function inmsom_bt_get_launch(ctx, rows) bind(C, name="inmsom_bt_get_launch")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), intent(out) :: rows
  end function
  integer(c_int) function inmsom_bt_set_substeps_per_launch(ctx, substeps) &
      bind(C, name="inmsom_bt_set_substeps_per_launch")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: substeps
  end function
  integer(c_int) function inmsom_bt_get_substeps_per_launch(ctx, mode, substeps) &
      bind(C, name="inmsom_bt_get_substeps_per_launch")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: mode
    integer(c_int), intent(out) :: substeps
  end function
  integer(c_int) function inmsom_bt_range_check(ctx, bad) bind(C, name="inmsom_bt_range_check")
    import :: c_int, c_long_long, c_ptr
    type(c_ptr), value :: ctx
    integer(c_long_long), intent(out) :: bad
  end function
  integer(c_int) function inmsom_bt_volume(ctx, vol) bind(C, name="inmsom_bt_volume")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: vol
  end function
  integer(c_int) function inmsom_bt_synchronize(ctx) bind(C, name="inmsom_bt_synchronize")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function
  integer(c_int) function inmsom_bt_remove_barotropic(ctx, nz, uu, vv, uu2d, vv2d) &
      bind(C, name="inmsom_bt_remove_barotropic")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: uu, vv, uu2d, vv2d
  end function
  integer(c_int) function inmsom_bt_internal_update(ctx, tau, nz, ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux, &
      uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i, hhn, uu, vv) bind(C, name="inmsom_bt_internal_update")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nz
    type(c_ptr), value :: ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux, uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i
    type(c_ptr), intent(in) :: hhn(4)
    type(c_ptr), value :: uu, vv
  end function
  ! the same on device addresses (e.g. from hipMalloc or OpenMP use_device_addr)
  integer(c_int) function inmsom_bt_remove_barotropic_dev(ctx, nz, uu, vv, uu2d, vv2d) &
      bind(C, name="inmsom_bt_remove_barotropic_dev")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: uu, vv, uu2d, vv2d
  end function
  integer(c_int) function inmsom_bt_internal_update_dev(ctx, tau, nz, ubrtr_i, vbrtr_i, hhu, hhv, slpr, &
      wflux, uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i, hhn, uu, vv) bind(C, name="inmsom_bt_internal_update_dev")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nz
    type(c_ptr), value :: ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux, uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i
    type(c_ptr), intent(in) :: hhn(4)
    type(c_ptr), value :: uu, vv
  end function
  integer(c_int) function inmsom_bt_coupling_timed(ctx, tau, nz, nrep, ms) bind(C, name="inmsom_bt_coupling_timed")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nz, nrep
    real(c_double), intent(out) :: ms
  end function
  ! forcing of the barotropic step from the 3-D model (Control/dyncall.f90:37-130)
  integer(c_int) function inmsom_bt_forcing_3d(ctx, nz, dz, uu, vv, uup, vvp, amuv, uu2d, vv2d, uup2d, &
      vvp2d, amuv2d, hhq, hhu, hhv, hhh, taux, tauy, slpr, pgx, pgy, rhsx2d, rhsy2d) &
      bind(C, name="inmsom_bt_forcing_3d")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: dz, uu, vv, uup, vvp, amuv, uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh
    type(c_ptr), value :: taux, tauy, slpr, pgx, pgy, rhsx2d, rhsy2d
  end function
  integer(c_int) function inmsom_bt_forcing_3d_dev(ctx, nz, dz, uu, vv, uup, vvp, amuv, uu2d, vv2d, uup2d, &
      vvp2d, amuv2d, hhq, hhu, hhv, hhh, taux, tauy, slpr, pgx, pgy, rhsx2d, rhsy2d) &
      bind(C, name="inmsom_bt_forcing_3d_dev")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: dz, uu, vv, uup, vvp, amuv, uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh
    type(c_ptr), value :: taux, tauy, slpr, pgx, pgy, rhsx2d, rhsy2d
  end function
  integer(c_int) function inmsom_bt_forcing_3d_timed(ctx, nz, nrep, ms) bind(C, name="inmsom_bt_forcing_3d_timed")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz, nrep
    real(c_double), intent(out) :: ms
  end function
  ! baroclinic_dynamics (Function/vel_ssh.f90:794-1027)
  integer(c_int) function inmsom_bt_baroclinic_dynamics(ctx, tau, nz, uu, uup, vv, vvp, ww, nu, rhsx, rhsy, &
      rhsx_tran, rhsy_tran, pgrx, pgry, rdis, tx_surf, ty_surf, tx_bot, ty_bot, dz, hzt, hhu, hhup, hhun, &
      hhv, hhvp, hhvn, hhh) bind(C, name="inmsom_bt_baroclinic_dynamics")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nz
    type(c_ptr), value :: uu, uup, vv, vvp, ww, nu, rhsx, rhsy, rhsx_tran, rhsy_tran, pgrx, pgry, rdis
    type(c_ptr), value :: tx_surf, ty_surf, tx_bot, ty_bot, dz, hzt, hhu, hhup, hhun, hhv, hhvp, hhvn, hhh
  end function
  integer(c_int) function inmsom_bt_baroclinic_dynamics_dev(ctx, tau, nz, uu, uup, vv, vvp, ww, nu, rhsx, &
      rhsy, rhsx_tran, rhsy_tran, pgrx, pgry, rdis, tx_surf, ty_surf, tx_bot, ty_bot, dz, hzt, hhu, hhup, &
      hhun, hhv, hhvp, hhvn, hhh) bind(C, name="inmsom_bt_baroclinic_dynamics_dev")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: tau
    integer(c_int), value :: nz
    type(c_ptr), value :: uu, uup, vv, vvp, ww, nu, rhsx, rhsy, rhsx_tran, rhsy_tran, pgrx, pgry, rdis
    type(c_ptr), value :: tx_surf, ty_surf, tx_bot, ty_bot, dz, hzt, hhu, hhup, hhun, hhv, hhvp, hhvn, hhh
  end function
  ! vertical_velocity (Function/vel_ssh.f90:1030-1076)
  integer(c_int) function inmsom_bt_vertical_velocity(ctx, nz, dz, u, v, w, hu, hv, wflux) &
      bind(C, name="inmsom_bt_vertical_velocity")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: dz, u, v, w, hu, hv, wflux
  end function
  integer(c_int) function inmsom_bt_vertical_velocity_dev(ctx, nz, dz, u, v, w, hu, hv, wflux) &
      bind(C, name="inmsom_bt_vertical_velocity_dev")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: dz, u, v, w, hu, hv, wflux
  end function
  ! pressure_gradients (Function/vel_ssh.f90:320-376)
  integer(c_int) function inmsom_bt_pressure_gradients(ctx, nz, z, den, hhq, hhu, hhv, rhsx, rhsy) &
      bind(C, name="inmsom_bt_pressure_gradients")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: z, den, hhq, hhu, hhv, rhsx, rhsy
  end function
  integer(c_int) function inmsom_bt_pressure_gradients_dev(ctx, nz, z, den, hhq, hhu, hhv, rhsx, rhsy) &
      bind(C, name="inmsom_bt_pressure_gradients_dev")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: z, den, hhq, hhu, hhv, rhsx, rhsy
  end function
  ! the density definition (Control/dyncall.f90:437-449)
  integer(c_int) function inmsom_bt_density(ctx, nz, z, tt, ss, hhq, den, den_pot) &
      bind(C, name="inmsom_bt_density")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: z, tt, ss, hhq, den, den_pot
  end function
  integer(c_int) function inmsom_bt_density_dev(ctx, nz, z, tt, ss, hhq, den, den_pot) &
      bind(C, name="inmsom_bt_density_dev")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: z, tt, ss, hhq, den, den_pot
  end function
  ! stress_components of the 3-D velocities (Function/mixing.f90:115-168)
  integer(c_int) function inmsom_bt_stress_components(ctx, nz, u, v, str_t, str_s) &
      bind(C, name="inmsom_bt_stress_components")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: u, v, str_t, str_s
  end function
  integer(c_int) function inmsom_bt_stress_components_dev(ctx, nz, u, v, str_t, str_s) &
      bind(C, name="inmsom_bt_stress_components_dev")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    type(c_ptr), value :: u, v, str_t, str_s
  end function
  ! smagorinsky_coeff + depth_ave (Control/dyncall.f90:40-43)
  integer(c_int) function inmsom_bt_smagorinsky(ctx, nz, visc2, dift, str_t, str_s, dz, amts, amuv, amuv2d) &
      bind(C, name="inmsom_bt_smagorinsky")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    real(c_double), value :: visc2, dift
    type(c_ptr), value :: str_t, str_s, dz, amts, amuv, amuv2d
  end function
  integer(c_int) function inmsom_bt_smagorinsky_dev(ctx, nz, visc2, dift, str_t, str_s, dz, amts, amuv, amuv2d) &
      bind(C, name="inmsom_bt_smagorinsky_dev")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: nz
    real(c_double), value :: visc2, dift
    type(c_ptr), value :: str_t, str_s, dz, amts, amuv, amuv2d
  end function
  function c_strlen(s) bind(C, name="strlen")
    import :: c_ptr, c_size_t
    type(c_ptr), value :: s
    integer(c_size_t) :: c_strlen
  end function
end interface

contains

! Stops with the library's message when status is not INMSOM_BT_OK
! (the reference stops the run on a failed range check, dyncall.f90:452-516).
subroutine bt_check(status, where)
  integer(c_int), intent(in) :: status
  character(*), intent(in) :: where
  character(kind=c_char), pointer :: msg(:)
  type(c_ptr) :: p
  integer :: n
  if (status == INMSOM_BT_OK) return
  p = inmsom_bt_last_error()
  n = int(c_strlen(p))
  call c_f_pointer(p, msg, [max(n, 1)])
  write(*, '(a,a,i0,a,*(a))') trim(where), ' failed, status ', status, ': ', msg(1:n)
  error stop 1
end subroutine bt_check

! The reference's barotropic_dynamics argument list (Function/vel_ssh.f90:379-439)
! on the context bt_ctx.  The scratch arguments vort .. RHSy_dif are not
! written: the reference's caller recomputes them before any use
! (Control/dyncall.f90:39,65,72-84).
subroutine barotropic_dynamics_gpu(tau, nstep, ksw4, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, &
                                   sshp_e, ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady, RHSx, RHSy, wflux, &
                                   mu, mu4, vort, str_t2d, str_s2d, fx, fy, rdis, RHSx_adv, RHSy_adv, &
                                   RHSx_dif, RHSy_dif)
  real(8), intent(in) :: tau
  integer, intent(in) :: nstep, ksw4
  real(8), target, intent(inout) :: ubrtr_e(*), ubrtrp_e(*), vbrtr_e(*), vbrtrp_e(*), ssh_e(*), sshp_e(*)
  real(8), target, intent(inout) :: ubrtr_i(*), vbrtr_i(*), ssh4gradx(*), ssh4grady(*)
  real(8), target, intent(in) :: RHSx(*), RHSy(*), wflux(*), mu(*), mu4(*), rdis(*)
  real(8), intent(inout) :: vort(*), str_t2d(*), str_s2d(*), fx(*), fy(*)
  real(8), intent(inout) :: RHSx_adv(*), RHSy_adv(*), RHSx_dif(*), RHSy_dif(*)
  call bt_check(inmsom_bt_barotropic_dynamics(bt_ctx, real(tau, c_double), int(nstep, c_int), &
                int(ksw4, c_int), c_loc(ubrtr_e), c_loc(ubrtrp_e), c_loc(vbrtr_e), c_loc(vbrtrp_e), &
                c_loc(ssh_e), c_loc(sshp_e), c_loc(ubrtr_i), c_loc(vbrtr_i), c_loc(ssh4gradx), &
                c_loc(ssh4grady), c_loc(RHSx), c_loc(RHSy), c_loc(wflux), c_loc(mu), c_loc(mu4), &
                c_loc(rdis)), 'barotropic_dynamics_gpu')
end subroutine barotropic_dynamics_gpu

! Control/dyncall.f90:169-184 on the context bt_ctx: the barotropic part
! uu2d, vv2d taken out of the 3-D velocity uu, vv (nz levels) before
! vertical_velocity.
subroutine remove_barotropic_gpu(nz, uu, vv, uu2d, vv2d)
  integer, intent(in) :: nz
  real(8), target, intent(inout) :: uu(*), vv(*)
  real(8), target, intent(in) :: uu2d(*), vv2d(*)
  call bt_check(inmsom_bt_remove_barotropic(bt_ctx, int(nz, c_int), c_loc(uu), c_loc(vv), c_loc(uu2d), &
                c_loc(vv2d)), 'remove_barotropic_gpu')
end subroutine remove_barotropic_gpu

! Control/dyncall.f90:188-231 on the context bt_ctx, after vertical_velocity:
! the time-mean 2-D velocity and the sea-level-pressure term of pgrx, pgry,
! ssh_internal (Function/vel_ssh.f90:682-751, which also updates the 3-D
! model's depths hhqn, hhun, hhvn, hhhn when full_free_surface > 0) and the
! barotropic part put back into uu, vv.
subroutine internal_update_gpu(tau, nz, ubrtr_i, vbrtr_i, hhu, hhv, slpr, wf_tot, uu2d, vv2d, pgrx, pgry, &
                               ssh_i, sshp_i, hhqn, hhun, hhvn, hhhn, uu, vv)
  real(8), intent(in) :: tau
  integer, intent(in) :: nz
  real(8), target, intent(in) :: ubrtr_i(*), vbrtr_i(*), hhu(*), hhv(*), slpr(*), wf_tot(*)
  real(8), target, intent(inout) :: uu2d(*), vv2d(*), pgrx(*), pgry(*), ssh_i(*), sshp_i(*)
  real(8), target, intent(inout) :: hhqn(*), hhun(*), hhvn(*), hhhn(*), uu(*), vv(*)
  type(c_ptr) :: hhn(4)
  hhn = [c_loc(hhqn), c_loc(hhun), c_loc(hhvn), c_loc(hhhn)]
  call bt_check(inmsom_bt_internal_update(bt_ctx, real(tau, c_double), int(nz, c_int), c_loc(ubrtr_i), &
                c_loc(vbrtr_i), c_loc(hhu), c_loc(hhv), c_loc(slpr), c_loc(wf_tot), c_loc(uu2d), c_loc(vv2d), &
                c_loc(pgrx), c_loc(pgry), c_loc(ssh_i), c_loc(sshp_i), hhn, c_loc(uu), c_loc(vv)), &
                'internal_update_gpu')
end subroutine internal_update_gpu

! Control/dyncall.f90:37-130 (ksw_vert = ksw_lat = ksw_lat4 = 0) on the
! context bt_ctx: RHSx2d, RHSy2d from the 3-D model's level-n velocities
! uu, vv, level n-1 uup, vvp, viscosity amuv (nz levels of thickness dz),
! the 2-D velocities and viscosity, the depths hhq .. hhh, wind stress,
! sea-level pressure and, when given, the terms pgx, pgy that
! pressure_gradients adds to the 3-D RHS.
subroutine forcing_3d_gpu(nz, dz, uu, vv, uup, vvp, amuv, uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, &
                          hhh, taux, tauy, slpr, RHSx2d, RHSy2d, pgx, pgy)
  integer, intent(in) :: nz
  real(8), target, intent(in) :: dz(*), uu(*), vv(*), uup(*), vvp(*), amuv(*)
  real(8), target, intent(in) :: uu2d(*), vv2d(*), uup2d(*), vvp2d(*), amuv2d(*), hhq(*), hhu(*), hhv(*)
  real(8), target, intent(in) :: hhh(*), taux(*), tauy(*), slpr(*)
  real(8), target, intent(inout) :: RHSx2d(*), RHSy2d(*)
  real(8), target, intent(in), optional :: pgx(*), pgy(*)
  type(c_ptr) :: px, py
  px = c_null_ptr
  py = c_null_ptr
  if (present(pgx)) px = c_loc(pgx)
  if (present(pgy)) py = c_loc(pgy)
  call bt_check(inmsom_bt_forcing_3d(bt_ctx, int(nz, c_int), c_loc(dz), c_loc(uu), c_loc(vv), c_loc(uup), &
                c_loc(vvp), c_loc(amuv), c_loc(uu2d), c_loc(vv2d), c_loc(uup2d), c_loc(vvp2d), c_loc(amuv2d), &
                c_loc(hhq), c_loc(hhu), c_loc(hhv), c_loc(hhh), c_loc(taux), c_loc(tauy), c_loc(slpr), px, py, &
                c_loc(RHSx2d), c_loc(RHSy2d)), 'forcing_3d_gpu')
end subroutine forcing_3d_gpu

! Function/vel_ssh.f90:794-1027 on the context bt_ctx with the reference's
! argument list, followed by the module arrays it reads: the vertical grid
! dz(nz), hzt(nz+1) and the 3-D model's depths hhu .. hhh.  uu, uup, vv,
! vvp are updated in place.
subroutine baroclinic_dynamics_gpu(tau, nz, uu, uup, vv, vvp, ww, anzu, RHSx, RHSy, RHSx_tran, RHSy_tran, &
                                   pgrx, pgry, r_diss, tx_surf, ty_surf, tx_bot, ty_bot, dz, hzt, hhu, hhup, &
                                   hhun, hhv, hhvp, hhvn, hhh)
  real(8), intent(in) :: tau
  integer, intent(in) :: nz
  real(8), target, intent(inout) :: uu(*), uup(*), vv(*), vvp(*)
  real(8), target, intent(in) :: ww(*), anzu(*), RHSx(*), RHSy(*), RHSx_tran(*), RHSy_tran(*)
  real(8), target, intent(in) :: pgrx(*), pgry(*), r_diss(*), tx_surf(*), ty_surf(*), tx_bot(*), ty_bot(*)
  real(8), target, intent(in) :: dz(*), hzt(*), hhu(*), hhup(*), hhun(*), hhv(*), hhvp(*), hhvn(*), hhh(*)
  call bt_check(inmsom_bt_baroclinic_dynamics(bt_ctx, real(tau, c_double), int(nz, c_int), c_loc(uu), &
                c_loc(uup), c_loc(vv), c_loc(vvp), c_loc(ww), c_loc(anzu), c_loc(RHSx), c_loc(RHSy), &
                c_loc(RHSx_tran), c_loc(RHSy_tran), c_loc(pgrx), c_loc(pgry), c_loc(r_diss), c_loc(tx_surf), &
                c_loc(ty_surf), c_loc(tx_bot), c_loc(ty_bot), c_loc(dz), c_loc(hzt), c_loc(hhu), c_loc(hhup), &
                c_loc(hhun), c_loc(hhv), c_loc(hhvp), c_loc(hhvn), c_loc(hhh)), 'baroclinic_dynamics_gpu')
end subroutine baroclinic_dynamics_gpu

! Function/vel_ssh.f90:1030-1076 on the context bt_ctx, the reference's
! arguments followed by the vertical grid dz(nz) it reads; w (nz+1
! levels) is read and written.
subroutine vertical_velocity_gpu(nz, u, v, w, hu, hv, wflux, dz)
  integer, intent(in) :: nz
  real(8), target, intent(in) :: u(*), v(*), hu(*), hv(*), wflux(*), dz(*)
  real(8), target, intent(inout) :: w(*)
  call bt_check(inmsom_bt_vertical_velocity(bt_ctx, int(nz, c_int), c_loc(dz), c_loc(u), c_loc(v), c_loc(w), &
                c_loc(hu), c_loc(hv), c_loc(wflux)), 'vertical_velocity_gpu')
end subroutine vertical_velocity_gpu

! Function/vel_ssh.f90:320-376 on the context bt_ctx, the reference's
! arguments followed by what it reads from Modules/mod_basin_grid.f90: the
! sigma levels z(nz) and the depths hhq, hhu, hhv.  RHSx, RHSy are read and
! written.
subroutine pressure_gradients_gpu(nz, den, RHSx, RHSy, z, hhq, hhu, hhv)
  integer, intent(in) :: nz
  real(8), target, intent(in) :: den(*), z(*), hhq(*), hhu(*), hhv(*)
  real(8), target, intent(inout) :: RHSx(*), RHSy(*)
  call bt_check(inmsom_bt_pressure_gradients(bt_ctx, int(nz, c_int), c_loc(z), c_loc(den), c_loc(hhq), &
                c_loc(hhu), c_loc(hhv), c_loc(RHSx), c_loc(RHSy)), 'pressure_gradients_gpu')
end subroutine pressure_gradients_gpu

! The density definition of Control/dyncall.f90:437-449 (the loop over lu
! points calling Service/oc_algorithms.f90:138-161) on the context bt_ctx:
! den at the level's pressure and den_pot at zero pressure from tt, ss, with
! the sigma levels z(nz) and the depth hhq.  den, den_pot are read and
! written.
subroutine density_gpu(nz, tt, ss, den, den_pot, z, hhq)
  integer, intent(in) :: nz
  real(8), target, intent(in) :: tt(*), ss(*), z(*), hhq(*)
  real(8), target, intent(inout) :: den(*), den_pot(*)
  call bt_check(inmsom_bt_density(bt_ctx, int(nz, c_int), c_loc(z), c_loc(tt), c_loc(ss), c_loc(hhq), &
                c_loc(den), c_loc(den_pot)), 'density_gpu')
end subroutine density_gpu

! Function/mixing.f90:115-168 on the context bt_ctx for nz levels, the
! reference's arguments (its nlev is nz); str_t, str_s are read and written.
subroutine stress_components_gpu(u, v, str_t, str_s, nz)
  integer, intent(in) :: nz
  real(8), target, intent(in) :: u(*), v(*)
  real(8), target, intent(inout) :: str_t(*), str_s(*)
  call bt_check(inmsom_bt_stress_components(bt_ctx, int(nz, c_int), c_loc(u), c_loc(v), c_loc(str_t), &
                c_loc(str_s)), 'stress_components_gpu')
end subroutine stress_components_gpu

! Control/dyncall.f90:40-43 on the context bt_ctx: smagorinsky_coeff
! (Function/mixing.f90:171-219) with the reference's arguments, then
! depth_ave(amuv, amuv2d, lu, 0) (Function/depth.f90:223-254) with the
! vertical grid dz(nz) it reads.  amts, amuv, amuv2d are read and written;
! amuv2d is the mu argument of barotropic_dynamics_gpu.
subroutine smagorinsky_gpu(nz, visc2, dift, str_t, str_s, amts, amuv, amuv2d, dz)
  integer, intent(in) :: nz
  real(8), intent(in) :: visc2, dift
  real(8), target, intent(in) :: str_t(*), str_s(*), dz(*)
  real(8), target, intent(inout) :: amts(*), amuv(*), amuv2d(*)
  call bt_check(inmsom_bt_smagorinsky(bt_ctx, int(nz, c_int), real(visc2, c_double), real(dift, c_double), &
                c_loc(str_t), c_loc(str_s), c_loc(dz), c_loc(amts), c_loc(amuv), c_loc(amuv2d)), 'smagorinsky_gpu')
end subroutine smagorinsky_gpu

end module inmsom_bt
