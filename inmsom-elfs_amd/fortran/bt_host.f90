! Fortran host for the GPU barotropic step, in the shape of the reference's
! driver (inmsom_head.f90:1-333 with the 2-D part of Control/dyncall.f90):
! it reads the run parameters from ocean_run.par / phys_proc.par in the
! reference's positional format, uploads the grid once, runs the time loop
! with the range check after every step and reports the same timer the
! reference prints (time_barotrop).
!
! usage: inmsom_bt_host <par dir> <case.bin> <out.bin> [literal|fp64_exact|fp64|fp32] [g] [coupling.bin|-]
!                       [forcing.bin|-] [baroclinic.bin|-] [pressure.bin|-] [density.bin]
!
!   ocean_run.par  line 2 time_step [s], line 3 run_duration [days],
!                  line 9 nstep_barotrop                 (inmsom_head.f90:47-55)
!   phys_proc.par  line 6 ksw_lat4                       (init_pars.f90:17-57)
!   case.bin       grid, forcing and initial state, written by
!                  inmsom_bt.cases.write_case (layout in oracle/ref_harness.f90);
!                  its nsteps field, when positive, overrides run_duration.
!   out.bin        final ssh_e, sshp_e, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e (real8)
!   coupling.bin   optional 3-D side of the last step: int32 nz, then real8
!                  hhu, hhv, slpr, uu2d, vv2d, ssh_i, sshp_i (nx x ny) and
!                  uu, vv (nx x ny x nz).  After the last step the host runs
!                  the 2-D -> 3-D coupling (Control/dyncall.f90:169-231) with
!                  the step's ubrtr_i, vbrtr_i, pgrx, pgry and wflux, and
!                  appends uu, vv, uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i,
!                  hhqn, hhun, hhvn, hhhn to out.bin ('-': no coupling).
!   forcing.bin    optional 3-D model fields: int32 nz, real8 dz(nz), then
!                  uu, vv, uup, vvp, amuv, pgx, pgy (nx x ny x nz) and uu2d,
!                  vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh, slpr,
!                  RHSx2d, RHSy2d (nx x ny).  The host forms the forcing of
!                  the barotropic step (Control/dyncall.f90:37-130) with the
!                  case's wind stress and appends RHSx2d, RHSy2d to out.bin
!                  ('-': none).
!   baroclinic.bin optional: int32 nz, real8 tau, dz(nz), hzt(nz+1), then
!                  uu, uup, vv, vvp, RHSx, RHSy, RHSx_tran, RHSy_tran
!                  (nx x ny x nz), ww, anzu (nx x ny x (nz+1)), pgrx, pgry,
!                  r_diss, tx_surf, ty_surf, tx_bot, ty_bot, hhu, hhup, hhun,
!                  hhv, hhvp, hhvn, hhh (nx x ny).  The host runs
!                  baroclinic_dynamics (Function/vel_ssh.f90:794-1027) and
!                  appends uu, uup, vv, vvp to out.bin ('-': none).
!   pressure.bin   optional: int32 nz, real8 z(nz), then den, RHSx, RHSy
!                  (nx x ny x nz), hhq, hhu, hhv (nx x ny).  The host runs
!                  pressure_gradients (Function/vel_ssh.f90:320-376) and
!                  appends RHSx, RHSy to out.bin ('-': none).
!   density.bin    optional: int32 nz, real8 z(nz), then tt, ss, den, den_pot
!                  (nx x ny x nz), hhq (nx x ny).  The host runs the density
!                  definition (Control/dyncall.f90:437-449) and appends den,
!                  den_pot to out.bin.
!
! Mode 'literal' runs the stage-by-stage device path (any metrics, the
! biharmonic term); the fused modes need uniform metrics.
program inmsom_bt_host
use iso_c_binding
use inmsom_bt
implicit none

character(512) :: pardir, case_path, out_path, arg, cpl_path, frc_path, bcl_path, prs_path, eos_path
character(256) :: lines(64)
character(16) :: mode_name
integer :: nl, ierr, magic, version, nx, ny, nsteps_case, nstep_case, ksw4_case, cmode, nout
integer :: nstep_barotrop, ksw_lat4, nsteps, step, mode, nz
integer, allocatable, target :: lbasins(:,:), out_steps(:)
real(8) :: time_step, run_duration, tau_case, cb_nl, ebottom, g
real(8), allocatable, target :: f(:,:,:)
real(8), allocatable, target :: st(:,:,:)
real(8), allocatable, target :: rdis(:,:)
! the 'dropin' mode: the caller-side arrays of barotropic_dynamics (its RHS,
! the four averages and the scratch arguments it does not write)
real(8), allocatable, target :: rhs(:,:,:), acc(:,:,:), scr(:,:,:)
! coupling: c2 = ubrtr_i vbrtr_i pgrx pgry hhu hhv slpr uu2d vv2d ssh_i sshp_i hhqn hhun hhvn hhhn
real(8), allocatable, target :: c2(:,:,:), uu(:,:,:), vv(:,:,:)
! forcing: f3 = uu vv uup vvp amuv pgx pgy; f2 = uu2d vv2d uup2d vvp2d amuv2d hhq hhu hhv hhh slpr RHSx2d RHSy2d
real(8), allocatable, target :: f3(:,:,:,:), f2(:,:,:), dzv(:)
! baroclinic: b3 = uu uup vv vvp RHSx RHSy RHSx_tran RHSy_tran; bw = ww anzu; b2 = 14 planes
real(8), allocatable, target :: b3(:,:,:,:), bw(:,:,:,:), b2(:,:,:), bdz(:), bhzt(:)
! pressure gradients: p3 = den RHSx RHSy; p2 = hhq hhu hhv
real(8), allocatable, target :: p3(:,:,:,:), p2(:,:,:), pz(:)
! density: e3 = tt ss den den_pot; e2 = hhq
real(8), allocatable, target :: e3(:,:,:,:), e2(:,:), ez(:)
integer(c_long_long) :: bad
integer(8) :: c0, c1, crate
real(8) :: time_barotrop, vol0, vol1
type(bt_grid) :: grid
type(bt_forcing) :: forc
integer(c_int) :: rc
! field order inside f: hhq_rest dx dy dxt dyt dxh dyh dxb dyb rlh_s taux tauy wflux mu mu4
integer, parameter :: NF = 15

call get_command_argument(1, pardir)
call get_command_argument(2, case_path)
call get_command_argument(3, out_path)
mode_name = 'literal'
if (command_argument_count() >= 4) call get_command_argument(4, mode_name)
g = 9.81d0
if (command_argument_count() >= 5) then
  call get_command_argument(5, arg)
  read(arg, *) g
endif

! ---- parameters (positional text files as the reference reads them)
call read_lines(trim(pardir)//'/ocean_run.par', lines, nl)
read(lines(2), *) time_step
read(lines(3), *) run_duration
read(lines(9), *) nstep_barotrop
call read_lines(trim(pardir)//'/phys_proc.par', lines, nl)
read(lines(6), *) ksw_lat4

! ---- grid, forcing and initial state
open(21, file=trim(case_path), access='stream', form='unformatted', status='old', iostat=ierr)
if (ierr /= 0) then
  write(*,*) 'cannot open ', trim(case_path)
  error stop 5
endif
read(21) magic, version, nx, ny, nsteps_case, nstep_case, ksw4_case, cmode, nout
allocate(out_steps(max(nout, 1)), lbasins(nx, ny), f(nx, ny, NF), st(nx, ny, 6), rdis(nx, ny))
if (nout > 0) read(21) out_steps(1:nout)
read(21) tau_case, cb_nl, ebottom
read(21) lbasins
read(21) f
read(21) rdis
read(21) st            ! ssh, sshp, u, up, v, vp
close(21)

nsteps = nint(run_duration * 86400.0d0 / time_step)
if (nsteps_case > 0) nsteps = nsteps_case
select case (trim(mode_name))
case ('literal');    mode = -1
case ('dropin');     mode = -2
case ('fp64_exact'); mode = INMSOM_BT_FP64_EXACT
case ('fp64');       mode = INMSOM_BT_FP64
case ('fp32');       mode = INMSOM_BT_FP32
case default
  write(*,*) 'unknown mode ', trim(mode_name)
  error stop 2
end select

grid%nx = nx; grid%ny = ny
grid%periodicity_x = 0; grid%periodicity_y = 0
grid%full_free_surface = 1
grid%free_fall_acc = g
grid%ref_den = 1025.0d0
grid%time_smooth = 0.05d0
grid%lbasins = c_loc(lbasins)
grid%lu = c_null_ptr; grid%luu = c_null_ptr; grid%luh = c_null_ptr; grid%lcu = c_null_ptr
grid%lcv = c_null_ptr; grid%llu = c_null_ptr; grid%llv = c_null_ptr
grid%hhq_rest = c_loc(f(1,1,1))
grid%dx = c_loc(f(1,1,2));  grid%dy = c_loc(f(1,1,3))
grid%dxt = c_loc(f(1,1,4)); grid%dyt = c_loc(f(1,1,5))
grid%dxh = c_loc(f(1,1,6)); grid%dyh = c_loc(f(1,1,7))
grid%dxb = c_loc(f(1,1,8)); grid%dyb = c_loc(f(1,1,9))
grid%rlh_s = c_loc(f(1,1,10))
call bt_check(inmsom_bt_init(grid, 0_c_int, bt_ctx), 'inmsom_bt_init')

forc%taux = c_loc(f(1,1,11)); forc%tauy = c_loc(f(1,1,12)); forc%wflux = c_loc(f(1,1,13))
forc%mu = c_loc(f(1,1,14)); forc%mu4 = c_loc(f(1,1,15)); forc%rdis = c_loc(rdis)
forc%cb_nl = cb_nl; forc%ebottom = ebottom
call bt_check(inmsom_bt_set_forcing(bt_ctx, forc), 'inmsom_bt_set_forcing')
call bt_check(inmsom_bt_set_state(bt_ctx, c_loc(st(1,1,3)), c_loc(st(1,1,4)), c_loc(st(1,1,5)), &
              c_loc(st(1,1,6)), c_loc(st(1,1,1)), c_loc(st(1,1,2))), 'inmsom_bt_set_state')
call bt_check(inmsom_bt_volume(bt_ctx, vol0), 'inmsom_bt_volume')

write(*,'(a,i0,a,i0,a,f8.2,a,i0,a,i0,a,a)') ' grid ', nx, ' x ', ny, ', time step ', time_step, &
      ' s, nstep_barotrop ', nstep_barotrop, ', steps ', nsteps, ', mode ', trim(mode_name)

! ---- time loop (inmsom_head.f90:236-316, dyncall.f90:109-166)
if (mode == -2) then
  ! the caller's RHS of the 'dropin' mode: the wind stress on U and V faces
  ! as a face force divided by the reference density, fixed over the run
  allocate(rhs(nx, ny, 2), acc(nx, ny, 4), scr(nx, ny, 9))
  rhs(:,:,1) = f(:,:,11) * f(:,:,4) * f(:,:,7) / 1025.0d0
  rhs(:,:,2) = f(:,:,12) * f(:,:,6) * f(:,:,5) / 1025.0d0
  acc = 0.0d0
  scr = 0.0d0
endif
time_barotrop = 0.0d0
call system_clock(count_rate=crate)
do step = 1, nsteps
  call system_clock(c0)
  if (mode == -2) then
    ! the reference's own call (Control/dyncall.f90:137-164) on host arrays
    call barotropic_dynamics_gpu(time_step, nstep_barotrop, ksw_lat4, st(:,:,3), st(:,:,4), st(:,:,5), st(:,:,6), &
                                 st(:,:,1), st(:,:,2), acc(:,:,1), acc(:,:,2), acc(:,:,3), acc(:,:,4), &
                                 rhs(:,:,1), rhs(:,:,2), f(:,:,13), f(:,:,14), f(:,:,15), scr(:,:,1), &
                                 scr(:,:,2), scr(:,:,3), scr(:,:,4), scr(:,:,5), rdis, scr(:,:,6), &
                                 scr(:,:,7), scr(:,:,8), scr(:,:,9))
    ! (the call leaves its result in the context's state too: the range
    ! check below and the final read see it)
    rc = INMSOM_BT_OK
  else if (mode < 0) then
    rc = inmsom_bt_model_steps(bt_ctx, time_step, int(nstep_barotrop, c_int), &
                               int(ksw_lat4, c_int), 1_c_int)
  else
    rc = inmsom_bt_run(bt_ctx, time_step, int(nstep_barotrop, c_int), 1_c_int, int(mode, c_int))
  endif
  call bt_check(rc, 'barotropic step')
  call system_clock(c1)
  time_barotrop = time_barotrop + dble(c1 - c0) / dble(crate)
  ! range check after every step (dyncall.f90:452-516)
  rc = inmsom_bt_range_check(bt_ctx, bad)
  if (rc == INMSOM_BT_ERANGE) then
    write(*,'(a,i0,a,i0)') ' range check failed at step ', step, ': bad cells ', bad
    error stop 3
  endif
  call bt_check(rc, 'inmsom_bt_range_check')
enddo

call bt_check(inmsom_bt_get_state(bt_ctx, c_loc(st(1,1,3)), c_loc(st(1,1,4)), c_loc(st(1,1,5)), &
              c_loc(st(1,1,6)), c_loc(st(1,1,1)), c_loc(st(1,1,2))), 'inmsom_bt_get_state')
call bt_check(inmsom_bt_volume(bt_ctx, vol1), 'inmsom_bt_volume')
open(23, file=trim(out_path), access='stream', form='unformatted', status='replace')
write(23) st
if (mode == -2) write(23) acc
if (command_argument_count() >= 6) then
  call get_command_argument(6, cpl_path)
  if (trim(cpl_path) /= '-') then
    call coupling_step(trim(cpl_path))
    write(23) uu, vv, c2(:,:,8:9), c2(:,:,3:4), c2(:,:,10:15)
  endif
endif
if (command_argument_count() >= 7) then
  call get_command_argument(7, frc_path)
  if (trim(frc_path) /= '-') then
    call forcing_step(trim(frc_path))
    write(23) f2(:,:,11:12)
  endif
endif
if (command_argument_count() >= 8) then
  call get_command_argument(8, bcl_path)
  if (trim(bcl_path) /= '-') then
    call baroclinic_step(trim(bcl_path))
    write(23) b3(:,:,:,1:4)
  endif
endif
if (command_argument_count() >= 9) then
  call get_command_argument(9, prs_path)
  if (trim(prs_path) /= '-') then
    call pressure_step(trim(prs_path))
    write(23) p3(:,:,:,2:3)
  endif
endif
if (command_argument_count() >= 10) then
  call get_command_argument(10, eos_path)
  call density_step(trim(eos_path))
  write(23) e3(:,:,:,3:4)
endif
close(23)
write(*,'(a,es12.5)') ' Time barotropic: ', time_barotrop
write(*,'(a,es12.5)') ' volume drift (relative): ', abs(vol1 - vol0) / max(abs(vol0), 1.0d-300)
call bt_check(inmsom_bt_finalize(bt_ctx), 'inmsom_bt_finalize')

contains

! The 2-D -> 3-D coupling of Control/dyncall.f90:169-231 after the last
! step, on the time means that step left (inmsom_bt_get_means).
subroutine coupling_step(path)
  character(*), intent(in) :: path
  integer :: io
  open(24, file=path, access='stream', form='unformatted', status='old', iostat=io)
  if (io /= 0) then
    write(*,*) 'cannot open ', path
    error stop 5
  endif
  read(24) nz
  allocate(c2(nx, ny, 15), uu(nx, ny, nz), vv(nx, ny, nz))
  c2 = 0.0d0
  read(24) c2(:,:,5:11)
  read(24) uu, vv
  close(24)
  call bt_check(inmsom_bt_get_means(bt_ctx, c_loc(c2(1,1,1)), c_loc(c2(1,1,2)), c_loc(c2(1,1,3)), &
                c_loc(c2(1,1,4))), 'inmsom_bt_get_means')
  call remove_barotropic_gpu(nz, uu, vv, c2(:,:,8), c2(:,:,9))
  ! (the 3-D model's vertical_velocity runs here)
  call internal_update_gpu(time_step, nz, c2(:,:,1), c2(:,:,2), c2(:,:,5), c2(:,:,6), c2(:,:,7), f(:,:,13), &
                           c2(:,:,8), c2(:,:,9), c2(:,:,3), c2(:,:,4), c2(:,:,10), c2(:,:,11), c2(:,:,12), &
                           c2(:,:,13), c2(:,:,14), c2(:,:,15), uu, vv)
  write(*,'(a,i0,a)') ' coupling step done (', nz, ' levels)'
end subroutine coupling_step

! The forcing of the barotropic step from the 3-D model, Control/
! dyncall.f90:37-130, on the fields of `path` and the case's wind stress.
subroutine forcing_step(path)
  character(*), intent(in) :: path
  integer :: io, nzf
  open(25, file=path, access='stream', form='unformatted', status='old', iostat=io)
  if (io /= 0) then
    write(*,*) 'cannot open ', path
    error stop 5
  endif
  read(25) nzf
  allocate(dzv(nzf), f3(nx, ny, nzf, 7), f2(nx, ny, 12))
  read(25) dzv
  read(25) f3
  read(25) f2
  close(25)
  call forcing_3d_gpu(nzf, dzv, f3(:,:,:,1), f3(:,:,:,2), f3(:,:,:,3), f3(:,:,:,4), f3(:,:,:,5), f2(:,:,1), &
                      f2(:,:,2), f2(:,:,3), f2(:,:,4), f2(:,:,5), f2(:,:,6), f2(:,:,7), f2(:,:,8), f2(:,:,9), &
                      f(:,:,11), f(:,:,12), f2(:,:,10), f2(:,:,11), f2(:,:,12), f3(:,:,:,6), f3(:,:,:,7))
  write(*,'(a,i0,a)') ' forcing step done (', nzf, ' levels)'
end subroutine forcing_step

! baroclinic_dynamics (Function/vel_ssh.f90:794-1027) on the fields of `path`.
subroutine baroclinic_step(path)
  character(*), intent(in) :: path
  integer :: io, nzb
  real(8) :: tau_b
  open(26, file=path, access='stream', form='unformatted', status='old', iostat=io)
  if (io /= 0) then
    write(*,*) 'cannot open ', path
    error stop 5
  endif
  read(26) nzb
  read(26) tau_b
  allocate(bdz(nzb), bhzt(nzb + 1), b3(nx, ny, nzb, 8), bw(nx, ny, nzb + 1, 2), b2(nx, ny, 14))
  read(26) bdz, bhzt
  read(26) b3
  read(26) bw
  read(26) b2
  close(26)
  call baroclinic_dynamics_gpu(tau_b, nzb, b3(:,:,:,1), b3(:,:,:,2), b3(:,:,:,3), b3(:,:,:,4), bw(:,:,:,1), &
                               bw(:,:,:,2), b3(:,:,:,5), b3(:,:,:,6), b3(:,:,:,7), b3(:,:,:,8), b2(:,:,1), &
                               b2(:,:,2), b2(:,:,3), b2(:,:,4), b2(:,:,5), b2(:,:,6), b2(:,:,7), bdz, bhzt, &
                               b2(:,:,8), b2(:,:,9), b2(:,:,10), b2(:,:,11), b2(:,:,12), b2(:,:,13), b2(:,:,14))
  write(*,'(a,i0,a)') ' baroclinic step done (', nzb, ' levels)'
end subroutine baroclinic_step

! pressure_gradients (Function/vel_ssh.f90:320-376) on the fields of `path`.
subroutine pressure_step(path)
  character(*), intent(in) :: path
  integer :: io, nzp
  open(27, file=path, access='stream', form='unformatted', status='old', iostat=io)
  if (io /= 0) then
    write(*,*) 'cannot open ', path
    error stop 5
  endif
  read(27) nzp
  allocate(pz(nzp), p3(nx, ny, nzp, 3), p2(nx, ny, 3))
  read(27) pz
  read(27) p3
  read(27) p2
  close(27)
  call pressure_gradients_gpu(nzp, p3(:,:,:,1), p3(:,:,:,2), p3(:,:,:,3), pz, p2(:,:,1), p2(:,:,2), p2(:,:,3))
  write(*,'(a,i0,a)') ' pressure gradients done (', nzp, ' levels)'
end subroutine pressure_step

! The density definition (Control/dyncall.f90:437-449) on the fields of `path`.
subroutine density_step(path)
  character(*), intent(in) :: path
  integer :: io, nze
  open(28, file=path, access='stream', form='unformatted', status='old', iostat=io)
  if (io /= 0) then
    write(*,*) 'cannot open ', path
    error stop 5
  endif
  read(28) nze
  allocate(ez(nze), e3(nx, ny, nze, 4), e2(nx, ny))
  read(28) ez
  read(28) e3
  read(28) e2
  close(28)
  call density_gpu(nze, e3(:,:,:,1), e3(:,:,:,2), e3(:,:,:,3), e3(:,:,:,4), ez, e2)
  write(*,'(a,i0,a)') ' density done (', nze, ' levels)'
end subroutine density_step

! Service/read_write_parameters.f90:2-41: the file's lines, one per entry.
subroutine read_lines(path, buf, n)
  character(*), intent(in) :: path
  character(256), intent(out) :: buf(:)
  integer, intent(out) :: n
  integer :: io
  n = 0
  open(31, file=path, status='old', iostat=io)
  if (io /= 0) then
    write(*,*) 'cannot open ', path
    error stop 5
  endif
  do
    if (n >= size(buf)) exit
    read(31, '(a)', iostat=io) buf(n + 1)
    if (io /= 0) exit
    n = n + 1
  enddo
  close(31)
end subroutine read_lines

end program inmsom_bt_host
