! Test infrastructure (never part of the product): calls the reference's
! own ctl_file_write (Service/rw_ctl_file.f90:1-138, compiled from the
! reference's unmodified source) on the arguments of a case file, so that
! the descriptor text of inmsom_bt_ctl_write can be compared with it byte
! for byte.
!
! usage: ctl_harness <case.txt>
!   line 1 fname, line 2 title, line 3 varname (each as given, trailing
!   blanks dropped); line 4 undef nx ny nz nt xtype ytype ztype yr_type
!   year0 month0 day0 hour0 minute0 ht hx hy hz; then one line per axis:
!   all its values (LEVELS) or its first value (LINEAR), list-directed.
program ctl_harness
implicit none
character(512) :: case_path, fname, title, varname
real :: undef, ht
integer :: nx, ny, nz, nt, xtype, ytype, ztype, yr_type, year0, month0, day0, hour0, minute0
real(8) :: hx, hy, hz
real(8), allocatable :: x0(:), y0(:), z0(:)

call get_command_argument(1, case_path)
open(11, file=trim(case_path), status='old', action='read')
read(11, '(a)') fname
read(11, '(a)') title
read(11, '(a)') varname
read(11, *) undef, nx, ny, nz, nt, xtype, ytype, ztype, yr_type, year0, month0, day0, hour0, minute0, ht, hx, hy, hz
allocate(x0(nx), y0(ny), z0(nz))
! all values of a LEVELS axis, the first of a LINEAR one
if (xtype == 1) then
  read(11, *) x0
else
  read(11, *) x0(1)
endif
if (ytype == 1) then
  read(11, *) y0
else
  read(11, *) y0(1)
endif
if (ztype == 1) then
  read(11, *) z0
else
  read(11, *) z0(1)
endif
close(11)
call ctl_file_write(trim(fname), undef, nx, ny, nz, nt, xtype, x0, hx, ytype, y0, hy, ztype, z0, hz, &
                    yr_type, year0, month0, day0, hour0, minute0, ht, trim(title), trim(varname))
end program ctl_harness
