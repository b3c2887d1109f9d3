!=======================================================================
! Test infrastructure: drives the reference's immersed-boundary table
! routines (02_solver/f90/lib/ibm_body.f90, linked from oracle/_ref by
! oracle/Makefile) in the order the pre-processor calls them
! (01_pre_processor/preprocessing.py:124-135, 166-203, 221-230, 272), so
! that the case builder's restatement can be compared with them bit for bit.
!
!   ibm_tables <axes.bin> <tables.bin>
!
! axes.bin (stream, native endian): nx ny nz heat les xprd yprd zprd (int64;
! nx etc. = number of grid lines, prd 1 = periodic), then X(0:nx), XM(0:nx),
! Y(0:ny), YM(0:ny), Z(0:nz), ZM(0:nz) (float64) and the fix flags
! ufix_x(1:nx-1), lfix_x, ufix_y, lfix_y, ufix_z, lfix_z (int64).
!
! tables.bin, per component u, v, w (and t when heat = 1): nbody nintp ninner,
! fcp(1:nintp+ninner, 1:3), intpindx(1:nintp, 1:3), geomfac(1:nintp, 3, 3, 3)
! in Fortran order; then, when les = 1: nzero and iszero(1:nx-1,1:ny-1,1:nz-1).
!=======================================================================
program ibm_tables
  implicit none
  integer(8) :: nx, ny, nz, heat, les, prd(3), ncomp, c
  real(8), allocatable :: x(:), xm(:), y(:), ym(:), z(:), zm(:)
  real(8), allocatable :: xx(:, :), yy(:, :), zz(:, :)
  integer(8), allocatable :: ufx(:), lfx(:), ufy(:), lfy(:), ufz(:), lfz(:)
  integer(8), allocatable :: iszero(:, :, :)
  integer(8) :: nzero
  character(len=512) :: fin, fout
  character(len=3) :: sx, sy, sz

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  read (10) nx, ny, nz, heat, les, prd
  allocate (x(0:nx), xm(0:nx), y(0:ny), ym(0:ny), z(0:nz), zm(0:nz))
  allocate (ufx(nx - 1), lfx(nx - 1), ufy(ny - 1), lfy(ny - 1), ufz(nz - 1), lfz(nz - 1))
  read (10) x, xm, y, ym, z, zm
  read (10) ufx, lfx, ufy, lfy, ufz, lfz
  close (10)

  ! the pre-processor passes its option strings, upper case
  sx = merge('ON ', 'OFF', prd(1) == 1)
  sy = merge('ON ', 'OFF', prd(2) == 1)
  sz = merge('ON ', 'OFF', prd(3) == 1)
  allocate (xx(-1:nx + 1, 3), yy(-1:ny + 1, 3), zz(-1:nz + 1, 3))
  call geomfac_preset(nx, x, xm, trim(sx), xx)
  call geomfac_preset(ny, y, ym, trim(sy), yy)
  call geomfac_preset(nz, z, zm, trim(sz), zz)

  open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
  ncomp = 3
  if (heat == 1) ncomp = 4
  do c = 1, ncomp
    select case (c)
    case (1)
      call one(c, x, ym, zm, xx(:, 1), yy(:, 3), zz(:, 2))
    case (2)
      call one(c, xm, y, zm, xx(:, 2), yy(:, 1), zz(:, 3))
    case (3)
      call one(c, xm, ym, z, xx(:, 3), yy(:, 2), zz(:, 1))
    case (4)
      call one(c, xm, ym, zm, xx(:, 3), yy(:, 3), zz(:, 3))
    end select
  end do
  if (les == 1) then
    allocate (iszero(1:nx - 1, 1:ny - 1, 1:nz - 1))
    call find_zero_nu_sgs(nx, ny, nz, xm, ym, zm, nzero, iszero, 0.0d0)
    write (11) nzero, iszero
  end if
  close (11)

contains

  subroutine one(dir, xc, yc, zc, xp, yp, zp)
    integer(8), intent(in) :: dir
    real(8), intent(in) :: xc(0:nx), yc(0:ny), zc(0:nz)
    real(8), intent(in) :: xp(-1:nx + 1), yp(-1:ny + 1), zp(-1:nz + 1)
    integer(8), allocatable :: inout(:, :, :), fcp(:, :), ityp(:, :), indx(:, :)
    real(8), allocatable :: gf(:, :, :, :)
    integer(8) :: nbody, nintp, ninner, nb

    allocate (inout(0:nx, 0:ny, 0:nz))
    call find_inout(nx, ny, nz, xc, yc, zc, nbody, inout, 0.0d0)
    allocate (fcp(nbody, 3), ityp(nbody, 1), indx(nbody, 3))
    call findbdy_intp(dir, nx, ny, nz, nbody, inout, ufx, ufy, ufz, lfx, lfy, lfz, &
                      nintp, ninner, fcp, ityp, indx)
    ! the pre-processor trims the tables to nintp (+ ninner) rows before
    ! computing the weights (preprocessing.py:189-199)
    nb = nintp + ninner
    allocate (gf(nintp, 3, 3, 3))
    call geomfac_intp(nx, ny, nz, xp, yp, zp, nintp, nb, fcp(1:nb, :), indx(1:nintp, :), gf, 0.0d0)
    write (11) nbody, nintp, ninner
    write (11) fcp(1:nb, :), indx(1:nintp, :), gf
  end subroutine one

end program ibm_tables
