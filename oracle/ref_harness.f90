! Test harness that drives the REFERENCE Fortran hot path on synthetic
! inputs.  TEST INFRASTRUCTURE ONLY: it links the reference's unmodified
! sources (built by oracle/Makefile into oracle/_ref/) and writes the
! vectors that pin the C restatement in oracle/bt_oracle.c.
!
! usage: ref_harness <case.bin> <out.bin>
!
! case.bin (little-endian stream, written by inmsom_bt.cases.write_case):
!   int32 magic ('BTC1'), version, nx, ny, nsteps, nstep, ksw4, mode, nout
!   int32 out_steps(nout)
!   real8 tau, cb_nl, ebottom
!   int32 lbasins(nx,ny)
!   real8 hhq_rest, dx, dy, dxt, dyt, dxh, dyh, dxb, dyb, rlh_s, taux,
!         tauy, wflux, mu, mu4, rdis, ssh, sshp, u, up, v, vp   (nx,ny each)
!
! The output starts with the 7 masks lu, luu, luh, lcu, lcv, llu, llv
! (as real8 (nx,ny)) built by the reference's gridcon.
! mode 0: time loop.  Each step is the 2-D path of Control/dyncall.f90:
!   sea_surface_fluxes (ksw_ssbc=1) and sea_bottom_fluxes on the
!   external-mode state, RHS assembly (:114-126 with every 3-D term
!   zero), then barotropic_dynamics(tau, nstep, ksw4).  After each step
!   listed in out_steps it writes, as real8 (nx,ny): ssh_e, sshp_e, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e,
!   ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady and the 12 hh*_e arrays.
! mode 1: one call of each helper routine on the case's state arrays
!   (see the write sequence at the end of this file).
! mode 2: the 2-D -> 3-D coupling after one barotropic step.  One step as
!   in mode 0 gives ubrtr_i, vbrtr_i and pgrx/pgry (the ssh4grad outputs);
!   synthetic 3-D velocities and internal-mode fields are set (below), all
!   inputs are written, then the sequence of Control/dyncall.f90:169-231
!   without vertical_velocity runs -- the removal and re-add loops and the
!   uu2d / pgrx update restated from dyncall (they sit inside
!   ocean_model_step), ssh_internal called as the reference has it -- and
!   uu, vv, uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i, hhqn, hhun, hhvn, hhhn
!   are written.
! mode 3: the forcing of the barotropic step, Control/dyncall.f90:37-130
!   with ksw_vert = ksw_lat = ksw_lat4 = 0, on synthetic 3-D fields at the
!   compiled nz: the inputs are written (dz, uu, vv, uup, vvp, amuv, the
!   2-D fields, the 3-D model depths, slpr, 3-D pressure-gradient terms
!   that stand for what pressure_gradients adds to RHSx3d, and the
!   initial RHSx2d, RHSy2d), the reference's stress_components,
!   sea_surface_fluxes, sea_bottom_fluxes, uv_trans, uv_diff2 and depth_ave
!   run in dyncall's order, the assembly loop of :109-126 is restated
!   (it sits inside ocean_model_step), and RHSx2d, RHSy2d are written.
! mode 4: baroclinic_dynamics (Function/vel_ssh.f90:794-1027) on synthetic
!   3-D fields at the compiled nz (the reference's own module depths from
!   hh_init, a non-uniform vertical grid hzt, dz): the inputs are written,
!   the routine is called, and uu, uup, vv, vvp are written.
! mode 5: vertical_velocity (Function/vel_ssh.f90:1030-1077) on synthetic
!   3-D velocities, the depths of hh_init and the case's wflux: the inputs
!   (dz, uu, vv, the initial ww, hhu, hhv, wflux) and the resulting ww are
!   written.
! mode 6: pressure_gradients (Function/vel_ssh.f90:320-376) on a synthetic
!   density and sigma grid at the compiled nz, the depths of hh_init and
!   synthetic 3-D right-hand sides: the inputs (z, den, RHSx3d, RHSy3d,
!   hhq, hhu, hhv) and the updated RHSx3d, RHSy3d are written.
! mode 7: the density definition (Control/dyncall.f90:437-449, restated
!   here as it is inline in the reference's caller) with the reference's
!   own density function (Service/oc_algorithms.f90:138-161) on synthetic
!   tt, ss, a non-uniform sigma grid and the depths of hh_init: the inputs
!   (z, tt, ss, the initial den, den_pot, hhq) and den, den_pot are written.
! mode 9: the time loop of mode 0 without output, timed by the reference's
!   own start_timer / end_timer (mpi_wtime, Modules/mod_mpi_parallel_tools.f90)
!   around the nsteps steps: after the masks it writes int32 nsteps, int32
!   OpenMP threads, real8 seconds.
! mode 10: the Smagorinsky viscosity of Control/dyncall.f90:40-43 (ksw_lat >
!   0): the reference's stress_components (Function/mixing.f90:115-168) of
!   synthetic 3-D velocities at the compiled nz, then smagorinsky_coeff
!   (:171-219) with lvisc_2 = 1e3, ldiff_ts = 5e2 and depth_ave(amuv, amuv2d,
!   lu, 0) (Function/depth.f90:223-254) from initial amts, amuv, amuv2d the
!   routines must leave where they do not write: the inputs (dz, the
!   velocities uup, vvp, their stress_t, stress_s, amts, amuv, amuv2d) and
!   the outputs (amts, amuv, amuv2d) are written.
! mode 8: the reference's grid set-up basinpar (Service/basin_parameters.f90:
!   191-300 with grid_parameters_spherical / _carthesian, :764-949) for the
!   compiled Inc/basinpar.fi, on zeroed metric arrays: dx, dy, dxt, dyt, dxh,
!   dyh, dxb, dyb and rlh_s are written (the case's fields are not used).
program ref_harness
use omp_lib
use main_basin_pars
use mpi_parallel_tools
use basin_grid
use ocean_variables
implicit none

character(512) :: case_path, out_path, mask_path
integer :: ierr, magic, version, cnx, cny, nsteps, nstep, ksw4, mode, nout
integer :: step, k, m, n, io
integer, allocatable :: out_steps(:), lb(:,:)
real(8) :: tau, cb_nl_in, ebottom_in, tsec
real(8), allocatable :: wflux(:,:), mu(:,:), mu4(:,:), rdis(:,:)
real(8), allocatable :: ssh4gx(:,:), ssh4gy(:,:), vort(:,:), strt(:,:), strs(:,:)
real(8), allocatable :: fxa(:,:), fya(:,:), radv(:,:), rady(:,:), rdfx(:,:), rdfy(:,:)
real(8), allocatable :: ax(:,:), ay(:,:)
real(8), allocatable :: pgx3(:,:,:), pgy3(:,:,:)
real(8), external :: density

call get_command_argument(1, case_path)
call get_command_argument(2, out_path)

call mpi_array_boundary_definition
call model_grid_allocate
call ocean_variables_allocate
call oceanbc_arrays_allocate

open(21, file=trim(case_path), access='stream', form='unformatted', status='old', iostat=io)
if (io /= 0) then
  print *, 'cannot open case file'
  call mpi_abort(cart_comm, 5, ierr)
endif
read(21) magic, version, cnx, cny, nsteps, nstep, ksw4, mode, nout
if (cnx /= nx .or. cny /= ny) then
  print *, 'case dims', cnx, cny, 'differ from the compiled nx, ny', nx, ny
  call mpi_abort(cart_comm, 6, ierr)
endif
allocate(out_steps(max(nout,1)), lb(nx,ny))
if (nout > 0) read(21) out_steps(1:nout)
read(21) tau, cb_nl_in, ebottom_in
read(21) lb

allocate(wflux(nx,ny), mu(nx,ny), mu4(nx,ny), rdis(nx,ny), ssh4gx(nx,ny), ssh4gy(nx,ny))
allocate(vort(nx,ny), strt(nx,ny), strs(nx,ny), fxa(nx,ny), fya(nx,ny))
allocate(radv(nx,ny), rady(nx,ny), rdfx(nx,ny), rdfy(nx,ny), ax(nx,ny), ay(nx,ny))
vort = 0.0d0; strt = 0.0d0; strs = 0.0d0; fxa = 0.0d0; fya = 0.0d0
radv = 0.0d0; rady = 0.0d0; rdfx = 0.0d0; rdfy = 0.0d0; ssh4gx = 0.0d0; ssh4gy = 0.0d0

read(21) hhq_rest, dx, dy, dxt, dyt, dxh, dyh, dxb, dyb, rlh_s, taux, tauy, &
         wflux, mu, mu4, rdis, ssh_e, sshp_e, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e
close(21)

! masks through the reference's own gridcon, from a mask file in its format
mask_path = trim(out_path)//'.mask.txt'
open(22, file=trim(mask_path), status='replace')
write(22, '(a)') 'synthetic mask (0 - ocean, 1 - land)'
do n = ny, 1, -1
  write(22, '(*(i1))') (lb(m,n), m = 1, nx)
enddo
close(22)
call gridcon(trim(mask_path))

open(23, file=trim(out_path), access='stream', form='unformatted', status='replace')
call write_masks()

if (mode == 0 .or. mode == 9) then
  ksw_ssbc = 1
  ksw_lat4 = ksw4
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  r_diss = rdis
  k = 1
  if (mode == 9) call start_timer(tsec)
  do step = 1, nsteps
    ! the 2-D flow is its own bottom layer; depths are the external-mode ones
    uu(:,:,nz)  = ubrtr_e
    vv(:,:,nz)  = vbrtr_e
    uup(:,:,nz) = ubrtrp_e
    vvp(:,:,nz) = vbrtrp_e
    hhu = hhu_e
    hhv = hhv_e
    hhq = hhq_e
    call sea_surface_fluxes
    call sea_bottom_fluxes
    RHSx2d = 0.0d0
    RHSy2d = 0.0d0
    do n = ny_start, ny_end
      do m = nx_start, nx_end
        if (lcu(m,n) > 0.5) then
          RHSx2d(m,n) = RHSx2d(m,n) + (surf_stress_x(m,n) + bot_stress_x(m,n))*dxt(m,n)*dyh(m,n)
        endif
        if (lcv(m,n) > 0.5) then
          RHSy2d(m,n) = RHSy2d(m,n) + (surf_stress_y(m,n) + bot_stress_y(m,n))*dyt(m,n)*dxh(m,n)
        endif
      enddo
    enddo
    call barotropic_dynamics(tau, nstep, ksw4, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, &
                             ssh_e, sshp_e, ubrtr_i, vbrtr_i, ssh4gx, ssh4gy,       &
                             RHSx2d, RHSy2d, wflux, mu, mu4, vort, strt, strs,      &
                             fxa, fya, r_diss, radv, rady, rdfx, rdfy)
    if (k <= nout) then
      if (out_steps(k) == step) then
        write(23) ssh_e, sshp_e, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, &
                  ubrtr_i, vbrtr_i, ssh4gx, ssh4gy
        write(23) hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
                  hhh_e, hhhp_e, hhhn_e
        k = k + 1
      endif
    endif
  enddo
  if (mode == 9) then
    call end_timer(tsec)
    write(23) nsteps, omp_get_max_threads(), tsec
  endif
else if (mode == 2) then
  ksw_ssbc = 1
  ksw_lat4 = ksw4
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  r_diss = rdis
  uu(:,:,nz)  = ubrtr_e
  vv(:,:,nz)  = vbrtr_e
  uup(:,:,nz) = ubrtrp_e
  vvp(:,:,nz) = vbrtrp_e
  hhu = hhu_e
  hhv = hhv_e
  hhq = hhq_e
  call sea_surface_fluxes
  call sea_bottom_fluxes
  RHSx2d = 0.0d0
  RHSy2d = 0.0d0
  do n = ny_start, ny_end
    do m = nx_start, nx_end
      if (lcu(m,n) > 0.5) RHSx2d(m,n) = (surf_stress_x(m,n) + bot_stress_x(m,n))*dxt(m,n)*dyh(m,n)
      if (lcv(m,n) > 0.5) RHSy2d(m,n) = (surf_stress_y(m,n) + bot_stress_y(m,n))*dyt(m,n)*dxh(m,n)
    enddo
  enddo
  call barotropic_dynamics(tau, nstep, ksw4, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, &
                           ssh_e, sshp_e, ubrtr_i, vbrtr_i, ssh4gx, ssh4gy,       &
                           RHSx2d, RHSy2d, wflux, mu, mu4, vort, strt, strs,      &
                           fxa, fya, r_diss, radv, rady, rdfx, rdfy)
  ! synthetic coupling inputs
  do k = 1, nz
    uu(:,:,k) = ubrtrp_e*(1.0d0 + 0.05d0*k) + 0.01d0*k
    vv(:,:,k) = vbrtrp_e*(1.0d0 - 0.03d0*k) - 0.02d0*k
  enddo
  uu2d = ubrtrp_e
  vv2d = vbrtrp_e
  hhu = hhu_e
  hhv = hhv_e
  slpr = 1.0d4*sshp_e + 50.0d0*taux
  pgrx = ssh4gx
  pgry = ssh4gy
  ssh_i = sshp_e
  sshp_i = ssh_e
  hhqn = 0.0d0; hhun = 0.0d0; hhvn = 0.0d0; hhhn = 0.0d0
  write(23) ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux, uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i
  write(23) uu, vv
  ! Control/dyncall.f90:169-184
  do n = ny_start-1, ny_end+1
    do m = nx_start-1, nx_end+1
      if (lcu(m,n) > 0.5) uu(m,n,1:nz) = uu(m,n,1:nz) - uu2d(m,n)
      if (lcv(m,n) > 0.5) vv(m,n,1:nz) = vv(m,n,1:nz) - vv2d(m,n)
    enddo
  enddo
  ! Control/dyncall.f90:188-206
  do n = ny_start-1, ny_end+1
    do m = nx_start-1, nx_end+1
      if (lcu(m,n) > 0.5) then
        uu2d(m,n) = ubrtr_i(m,n)/hhu(m,n)
        pgrx(m,n) = pgrx(m,n) - (slpr(m+1,n) - slpr(m,n))*hhu(m,n)*dyh(m,n)/RefDen
      endif
      if (lcv(m,n) > 0.5) then
        vv2d(m,n) = vbrtr_i(m,n)/hhv(m,n)
        pgry(m,n) = pgry(m,n) - (slpr(m,n+1) - slpr(m,n))*hhv(m,n)*dxh(m,n)/RefDen
      endif
    enddo
  enddo
  call ssh_internal(tau, ssh_i, sshp_i, ubrtr_i, vbrtr_i, wflux)
  ! Control/dyncall.f90:216-231
  do n = ny_start-1, ny_end+1
    do m = nx_start-1, nx_end+1
      if (lcu(m,n) > 0.5) uu(m,n,1:nz) = uu(m,n,1:nz) + uu2d(m,n)
      if (lcv(m,n) > 0.5) vv(m,n,1:nz) = vv(m,n,1:nz) + vv2d(m,n)
    enddo
  enddo
  write(23) uu, vv
  write(23) uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i, hhqn, hhun, hhvn, hhhn
else if (mode == 3) then
  ksw_ssbc = 1
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  hhq = hhq_e; hhu = hhu_e; hhv = hhv_e; hhh = hhh_e
  allocate(pgx3(nx,ny,nz), pgy3(nx,ny,nz))
  do k = 1, nz
    dz(k) = (0.5d0 + k)/(0.5d0*nz + 0.5d0*nz*(nz + 1))
    uu(:,:,k)  = (ubrtr_e*(1.0d0 + 0.05d0*k) + 0.01d0*k)*lcu
    vv(:,:,k)  = (vbrtr_e*(1.0d0 - 0.03d0*k) - 0.02d0*k)*lcv
    uup(:,:,k) = (ubrtrp_e*(1.0d0 + 0.04d0*k) + 0.005d0*k)*lcu
    vvp(:,:,k) = (vbrtrp_e*(1.0d0 - 0.02d0*k) - 0.01d0*k)*lcv
    amuv(:,:,k) = mu*(1.0d0 + 0.1d0*k)
    pgx3(:,:,k) = 1.0d-2*(sshp_e - ssh_e)*k
    pgy3(:,:,k) = -2.0d-2*(ssh_e + 0.1d0)*k
  enddo
  uu2d = ubrtr_e; vv2d = vbrtr_e; uup2d = ubrtrp_e; vvp2d = vbrtrp_e
  amuv2d = mu
  slpr = 1.0d4*sshp_e + 50.0d0*taux
  RHSx2d = 7.0d0; RHSy2d = -3.0d0
  write(23) dz
  write(23) uu, vv, uup, vvp, amuv, pgx3, pgy3
  write(23) uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh, slpr, RHSx2d, RHSy2d
  ! Control/dyncall.f90:38-39
  call stress_components(uup, vvp, stress_t, stress_s, nz)
  call stress_components(uup2d, vvp2d, stress_t2d, stress_s2d, 1)
  ! :53-59
  call sea_surface_fluxes
  call sea_bottom_fluxes
  ! :65-81
  RHSx3d = 0.0d0; RHSy3d = 0.0d0
  RHSx3d_tran = 0.0d0; RHSy3d_tran = 0.0d0
  RHSx2d_diff_disp = 0.0d0; RHSy2d_diff_disp = 0.0d0
  call uv_trans(uu, vv, r_vort, hhq, hhu, hhv, hhh, RHSx3d_tran, RHSy3d_tran, nz)
  call uv_trans(uu2d, vv2d, r_vort2d, hhq, hhu, hhv, hhh, RHSx2d_tran_disp, RHSy2d_tran_disp, 1)
  call uv_diff2(amuv, stress_t, stress_s, hhq, hhu, hhv, hhh, RHSx3d, RHSy3d, nz)
  call uv_diff2(amuv2d, stress_t2d, stress_s2d, hhq, hhu, hhv, hhh, RHSx2d_diff_disp, RHSy2d_diff_disp, 1)
  ! :92-98, what pressure_gradients adds
  RHSx3d = RHSx3d + pgx3
  RHSy3d = RHSy3d + pgy3
  ! :100-104
  call depth_ave(RHSx3d, RHSx2d, llu, 0)
  call depth_ave(RHSy3d, RHSy2d, llv, 0)
  call depth_ave(RHSx3d_tran, RHSx2d_tran, llu, 0)
  call depth_ave(RHSy3d_tran, RHSy2d_tran, llv, 0)
  ! :107-126
  do n = ny_start, ny_end
    do m = nx_start, nx_end
      if (lcu(m,n) > 0.5) then
        RHSx2d(m,n) = RHSx2d(m,n) - RHSx2d_diff_disp(m,n) + RHSx2d_tran(m,n) - RHSx2d_tran_disp(m,n) &
                    + (surf_stress_x(m,n) + bot_stress_x(m,n))*dxt(m,n)*dyh(m,n)                   &
                    - (slpr(m+1,n) - slpr(m,n))*hhu(m,n)*dyh(m,n)/RefDen
      endif
      if (lcv(m,n) > 0.5) then
        RHSy2d(m,n) = RHSy2d(m,n) - RHSy2d_diff_disp(m,n) + RHSy2d_tran(m,n) - RHSy2d_tran_disp(m,n) &
                    + (surf_stress_y(m,n) + bot_stress_y(m,n))*dyt(m,n)*dxh(m,n)                   &
                    - (slpr(m,n+1) - slpr(m,n))*hhv(m,n)*dxh(m,n)/RefDen
      endif
    enddo
  enddo
  write(23) RHSx2d, RHSy2d
else if (mode == 4) then
  ! baroclinic_dynamics (Function/vel_ssh.f90:794-1027) on synthetic 3-D
  ! fields at the compiled nz, the depths of hh_init, a non-uniform
  ! vertical grid; inputs written, then the updated uu, uup, vv, vvp
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  hhu = hhu_e; hhup = hhup_e; hhun = hhun_e
  hhv = hhv_e; hhvp = hhvp_e; hhvn = hhvn_e; hhh = hhh_e
  do k = 1, nz
    dz(k) = (0.5d0 + k)/(0.5d0*nz + 0.5d0*nz*(nz + 1))
    uu(:,:,k)  = (ubrtr_e*(1.0d0 + 0.05d0*k) + 0.01d0*k)*lcu
    vv(:,:,k)  = (vbrtr_e*(1.0d0 - 0.03d0*k) - 0.02d0*k)*lcv
    uup(:,:,k) = (ubrtrp_e*(1.0d0 + 0.04d0*k) + 0.005d0*k)*lcu
    vvp(:,:,k) = (vbrtrp_e*(1.0d0 - 0.02d0*k) - 0.01d0*k)*lcv
    RHSx3d(:,:,k) = 1.0d2*(ubrtr_e + 0.1d0*k)*dxt*dyh*1.0d-3
    RHSy3d(:,:,k) = -1.0d2*(vbrtr_e - 0.05d0*k)*dxh*dyt*1.0d-3
    RHSx3d_tran(:,:,k) = 3.0d1*(ssh_e - 0.02d0*k)*dxt*dyh*1.0d-3
    RHSy3d_tran(:,:,k) = 2.0d1*(sshp_e + 0.03d0*k)*dxh*dyt*1.0d-3
  enddo
  do k = 1, nz + 1
    hzt(k) = (0.6d0 + 0.05d0*k)/nz
    ww(:,:,k) = 1.0d-4*(ssh_e - 0.5d0*sshp_e + 0.01d0*k)
    anzu(:,:,k) = 1.0d-3*(1.0d0 + 0.2d0*k) + 1.0d-2*abs(ubrtr_e)
  enddo
  pgrx = 5.0d1*(sshp_e - ssh_e)*dxt*dyh*1.0d-3
  pgry = -4.0d1*(sshp_e + ssh_e)*dxh*dyt*1.0d-3
  r_diss = rdis
  surf_stress_x = 1.0d-4*taux
  surf_stress_y = 1.0d-4*(taux - 0.3d0)
  bot_stress_x = -1.0d-5*ubrtr_e
  bot_stress_y = -1.0d-5*vbrtr_e
  write(23) dz, hzt
  write(23) uu, uup, vv, vvp, RHSx3d, RHSy3d, RHSx3d_tran, RHSy3d_tran
  write(23) ww, anzu
  write(23) pgrx, pgry, r_diss, surf_stress_x, surf_stress_y, bot_stress_x, bot_stress_y, &
            hhu, hhup, hhun, hhv, hhvp, hhvn, hhh
  call baroclinic_dynamics(tau, uu, uup, vv, vvp, ww, anzu, RHSx3d, RHSy3d, RHSx3d_tran, RHSy3d_tran, &
                           pgrx, pgry, r_diss, surf_stress_x, surf_stress_y, bot_stress_x, bot_stress_y)
  write(23) uu, uup, vv, vvp
else if (mode == 5) then
  ! vertical_velocity (Function/vel_ssh.f90:1030-1077) on synthetic 3-D
  ! velocities at the compiled nz, the depths of hh_init and the case's
  ! wflux; ww starts at values the routine must leave where it does not
  ! write
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  do k = 1, nz
    dz(k) = (0.5d0 + k)/(0.5d0*nz + 0.5d0*nz*(nz + 1))
    uu(:,:,k) = (ubrtr_e*(1.0d0 + 0.05d0*k) + 0.01d0*k)*lcu
    vv(:,:,k) = (vbrtr_e*(1.0d0 - 0.03d0*k) - 0.02d0*k)*lcv
  enddo
  do k = 1, nz + 1
    ww(:,:,k) = 7.0d0 + 0.1d0*k + 1.0d-3*ssh_e
  enddo
  write(23) dz
  write(23) uu, vv, ww
  write(23) hhu_e, hhv_e, wflux
  call vertical_velocity(uu, vv, ww, hhu_e, hhv_e, wflux)
  write(23) ww
else if (mode == 6) then
  ! pressure_gradients (Function/vel_ssh.f90:320-376); a stratified,
  ! horizontally varying density and a non-uniform sigma grid
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  hhq = hhq_e; hhu = hhu_e; hhv = hhv_e
  do k = 1, nz
    z(k) = (dble(k) - 0.5d0)**1.3d0/(dble(nz) - 0.4d0)**1.3d0
    den(:,:,k) = 1025.0d0 + 0.4d0*k + 2.0d0*ssh_e - 0.5d0*ubrtr_e + 0.3d0*vbrtrp_e*k
    RHSx3d(:,:,k) = 1.0d2*(ubrtr_e + 0.1d0*k)*dxt*dyh*1.0d-3
    RHSy3d(:,:,k) = -1.0d2*(vbrtr_e - 0.05d0*k)*dxh*dyt*1.0d-3
  enddo
  write(23) z
  write(23) den, RHSx3d, RHSy3d
  write(23) hhq, hhu, hhv
  call pressure_gradients(den, RHSx3d, RHSy3d)
  write(23) RHSx3d, RHSy3d
else if (mode == 7) then
  ! the density definition, Control/dyncall.f90:437-449
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  hhq = hhq_e
  do k = 1, nz
    z(k) = (dble(k) - 0.5d0)**1.3d0/(dble(nz) - 0.4d0)**1.3d0
    tt(:,:,k) = 18.0d0 - 0.6d0*k + 40.0d0*ssh_e + 3.0d0*ubrtr_e
    ss(:,:,k) = 17.5d0 + 0.2d0*k - 20.0d0*sshp_e + 2.0d0*vbrtr_e
  enddo
  den = 7.0d0
  den_pot = -7.0d0
  write(23) z
  write(23) tt, ss, den, den_pot
  write(23) hhq
  do n = ny_start - 1, ny_end + 1
    do m = nx_start - 1, nx_end + 1
      if (lu(m,n) > 0.5) then
        do k = 1, nz
          den(m,n,k) = density(tt(m,n,k), ss(m,n,k), FreeFallAcc*RefDen*hhq(m,n)*z(k))
          den_pot(m,n,k) = density(tt(m,n,k), ss(m,n,k), 0.0d0)
        enddo
      endif
    enddo
  enddo
  write(23) den, den_pot
else if (mode == 10) then
  do k = 1, nz
    dz(k) = (0.5d0 + k)/(0.5d0*nz + 0.5d0*nz*(nz + 1))
    uup(:,:,k) = (ubrtr_e*(1.0d0 + 0.05d0*k) + 0.01d0*k)*lcu
    vvp(:,:,k) = (vbrtr_e*(1.0d0 - 0.03d0*k) - 0.02d0*k)*lcv
    amts(:,:,k) = 3.0d0 + 0.25d0*k + ssh_e
    amuv(:,:,k) = 5.0d0 - 0.5d0*k + sshp_e
  enddo
  amuv2d = -7.0d0 + ubrtr_e
  call stress_components(uup, vvp, stress_t, stress_s, nz)
  lvisc_2 = 1.0d3
  ldiff_ts = 5.0d2
  write(23) dz
  write(23) uup, vvp
  write(23) stress_t, stress_s, amts, amuv, amuv2d
  call smagorinsky_coeff(lvisc_2, ldiff_ts, stress_t, stress_s, amts, amuv)
  call depth_ave(amuv, amuv2d, lu, 0)
  write(23) amts, amuv, amuv2d
else if (mode == 8) then
  ! the reference's own grid set-up (Service/basin_parameters.f90, basinpar)
  ! for the compiled Inc/basinpar.fi, from zeroed metric arrays
  dx = 0.0d0; dy = 0.0d0; dxt = 0.0d0; dyt = 0.0d0
  dxh = 0.0d0; dyh = 0.0d0; dxb = 0.0d0; dyb = 0.0d0
  rlh_s = 0.0d0
  call basinpar
  write(23) dx, dy, dxt, dyt, dxh, dyh, dxb, dyb, rlh_s
else
  ! helper routines, one call each, on the case's state arrays
  ! hh_init from (ssh_e, sshp_e)
  call hh_init(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
               hhh_e, hhhp_e, hhhn_e, ssh_e, sshp_e, hhq_rest)
  write(23) hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
            hhh_e, hhhp_e, hhhn_e
  ! hh_update from ubrtr_e (used as a surface height field)
  call hh_update(hhqn_e, hhun_e, hhvn_e, hhhn_e, ubrtr_e, hhq_rest)
  write(23) hhqn_e, hhun_e, hhvn_e, hhhn_e
  ! hh_shift with nstep
  call hh_shift(hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
                hhh_e, hhhp_e, hhhn_e, nstep)
  write(23) hhq_e, hhqp_e, hhqn_e, hhu_e, hhup_e, hhun_e, hhv_e, hhvp_e, hhvn_e, &
            hhh_e, hhhp_e, hhhn_e
  ! stress_components(ubrtrp_e, vbrtrp_e)
  call stress_components(ubrtrp_e, vbrtrp_e, strt, strs, 1)
  write(23) strt, strs
  ! uv_trans(ubrtr_e, vbrtr_e) with the current depths
  call uv_trans(ubrtr_e, vbrtr_e, vort, hhq_e, hhu_e, hhv_e, hhh_e, radv, rady, 1)
  write(23) vort, radv, rady
  ! uv_diff2(mu) on the stresses above
  call uv_diff2(mu, strt, strs, hhq_e, hhu_e, hhv_e, hhh_e, rdfx, rdfy, 1)
  write(23) rdfx, rdfy
  ! uv_diff4(mu4) accumulated onto the uv_diff2 result
  call uv_diff4(mu4, strt, strs, fxa, fya, hhq_e, hhu_e, hhv_e, hhh_e, rdfx, rdfy, 1)
  write(23) fxa, fya, strt, strs, rdfx, rdfy
  ! wind stress and quadratic drag
  ksw_ssbc = 1
  uu(:,:,nz)  = ubrtr_e
  vv(:,:,nz)  = vbrtr_e
  uup(:,:,nz) = ubrtrp_e
  vvp(:,:,nz) = vbrtrp_e
  hhu = hhu_e
  hhv = hhv_e
  hhq = hhq_e
  call sea_surface_fluxes
  call sea_bottom_fluxes
  write(23) surf_stress_x, surf_stress_y, BottomFriction, bot_stress_x, bot_stress_y
endif
close(23)
call mpi_finalize(ierr)

contains

subroutine write_masks()
  ax = dble(lu);  write(23) ax
  ax = dble(luu); write(23) ax
  ax = dble(luh); write(23) ax
  ax = dble(lcu); write(23) ax
  ax = dble(lcv); write(23) ax
  ax = dble(llu); write(23) ax
  ax = dble(llv); write(23) ax
end subroutine write_masks

end program ref_harness
