!=======================================================================
! solver_exec on MI355X: the time loop of the LESwHT solver with the time
! step on the GPU through the C ABI (include/leswht.h).
!
! Run, like the reference executable, with the solver directory (the one
! holding settings.input) as the working directory (run_solver.sh:41). It
! reads the same inputs (settings.input, boundary.input, ../output/ibmpre/*,
! the grid file, the initial field; misc_init.f90:19-786) and writes the same
! outputs: ../output/field/fld%07d at every NPRINT steps and at the end
! (solver.f90:171,185; misc_hist.f90:212-252) and the history files
! ../output/ftr/{fhist,ftemp,fcdcl,fheatflux,ftime}.dat, fcmfr.dat of a
! driven channel (ICH /= 0), the tracer files
! f{u,v,w,p}trace.dat and mgftresiduemax.dat with the reference's records and
! formats (misc_hist.f90:2-209, 255-284; slv_auxi.f90:573-840;
! poiss_z_ft.f90:49,293-297). With IREAD = 0 it makes the initial field
! (makefld) and writes it as fld0000000 first, as the reference does.
! Unlike the reference, an error ends the program with a non-zero status.
!
! The body is the circular cylinder of geometry/funcbody_example/
! funcbody_cylinder.f90 (radius 0.5), or radius LW_BODY_RADIUS if set.
!=======================================================================
program solver_exec
  use, intrinsic :: iso_c_binding
  use leswht_c
  implicit none

  type(c_ptr) :: h
  type(lw_settings) :: st
  type(lw_history) :: hs
  type(lw_step_stats) :: ss
  type(lw_body) :: body
  integer(c_int) :: rc
  integer(8) :: clk0, clk1, rate
  real(8) :: wall
  character(len=64) :: env
  character(len=68) :: benv
  integer :: elen, estat
  logical :: heat, ibm
  integer(c_int64_t) :: ntrace, ntr, nres, q
  integer(c_int64_t) :: trpts(192)
  real(c_double) :: trv(256), t_begin
  integer(c_int64_t), allocatable :: rkind(:), rkv(:), rkk(:)
  real(c_double), allocatable :: rrm(:)
  integer(c_int64_t), parameter :: maxit = 3 * 4096
  integer(c_int64_t), allocatable :: bsub(:), bicg(:), bloc(:)
  real(c_double), allocatable :: brm(:)
  integer(c_int64_t) :: sicg(3)
  real(c_double) :: srm(3)

  body%kind = 0
  body%p = 0.0d0
  body%p(1) = 0.5d0
  call get_environment_variable('LW_BODY_RADIUS', env, elen, estat)
  if (estat == 0 .and. elen > 0) read (env(1:elen), *) body%p(1)
  ! LW_BODY = "kind p1 [p2 [p3]]": 0 cylinder (radius), 1 square (half side,
  ! angle), 2 ellipse (semi-axes, angle), as lw_body in leswht.h
  call get_environment_variable('LW_BODY', env, elen, estat)
  if (estat == 0 .and. elen > 0) then
    body%p = 0.0d0
    ! the slash ends the list, so parameters not given stay 0
    benv = env(1:elen)//' /'
    read (benv, *, iostat=estat) body%kind, body%p
  end if

  rc = lw_create_from_dir('.'//c_null_char, body, h)
  call check(rc, 'reading the case')
  rc = lw_get_settings(h, st)
  call check(rc, 'settings')
  rc = lw_get_history(h, hs)
  call check(rc, 'initial state')
  heat = st%ihtrans == 1
  ibm = st%ibmon == 1
  rc = lw_trace_info(h, ntrace, ntr, trpts)
  call check(rc, 'trace cells')

  ! makefld is followed by writefield (solver.f90:50-53)
  if (st%iread /= 1) call write_field(hs%ihist)
  call open_histories(st%iread == 1)
  ! poisinit opens the solver's logs anew on every run: poiss_z_ft its
  ! residual log (poiss_z_ft.f90:49), poiss_mgrd its iteration and solve logs
  ! (poiss_mgrd.f90:710-711)
  if (st%poisson == 1) then
    allocate (bsub(maxit), bicg(maxit), bloc(3 * maxit), brm(maxit))
    open (77, file='../output/ftr/poiss_itr.dat')
    open (78, file='../output/ftr/ftrpoittr.dat')
  else if (st%poisson == 0 .or. st%poisson == 3) then
    allocate (rkind(64), rkv(64), rkk(64), rrm(64))
    open (77, file='../output/ftr/mgftresiduemax.dat')
  end if
  call system_clock(count_rate=rate)

  do while ((hs%ntime < st%nend) .and. (hs%time < st%tend))
    t_begin = hs%time
    call system_clock(clk0)
    rc = lw_step(h, 1_c_int64_t, ss)
    call check(rc, 'time step')
    call system_clock(clk1)
    wall = dble(clk1 - clk0) / dble(rate)
    rc = lw_get_history(h, hs)
    call check(rc, 'history')
    if (st%poisson == 1) then
      nres = lw_get_bicg_log(h, bsub, bicg, brm, bloc, maxit)
      do q = 1, min(nres, maxit)
        write (77, 102) bicg(q), brm(q), bloc(3 * q - 2), bloc(3 * q - 1), bloc(3 * q)
      end do
      nres = lw_get_bicg_solves(h, sicg, srm, 3_c_int64_t)
      do q = 1, min(nres, 3_c_int64_t)
        write (78, 103) t_begin, hs%dt, sicg(q), srm(q)
      end do
    else if (st%poisson == 0 .or. st%poisson == 3) then
      ! mg2d's rows, then gsor2d's (poiss_z_ft: per sweep above IMGSOR;
      ! poiss_x_ft: at the iteration limit)
      nres = lw_get_residual_rows(h, rkind, rkv, rkk, rrm, 0_c_int64_t)
      if (nres > size(rkv)) then
        deallocate (rkind, rkv, rkk, rrm)
        allocate (rkind(nres), rkv(nres), rkk(nres), rrm(nres))
      end if
      nres = lw_get_residual_rows(h, rkind, rkv, rkk, rrm, int(size(rkv), c_int64_t))
      do q = 1, min(nres, int(size(rkv), c_int64_t))
        if (rkind(q) == 0) then
          write (77, 999) rkv(q), rkk(q), rrm(q)
        else if (rkind(q) == 1) then
          write (77, 998) rkv(q), rkk(q), rrm(q)
        else if (rkind(q) == 2) then
          write (77, 997) ' mg, time = ', t_begin, rkv(q), rkk(q), rrm(q)
        else
          write (77, 996) 'SOR', rkv(q), rkk(q), rrm(q)
        end if
      end do
    end if
999 format('KV=', i4, 3x, 'KK=', i4, 3x, 'RM=', es24.16)
998 format('KV=', i3, 3x, 'KK=', i3, 3x, 'RG=', es24.16)
997 format(a13, f13.5, ' IV=', i5, ' II=', i5, ' RG=', es16.8)
996 format(a5, '  IV=', i5, '  II=', i5, '  RG=', es23.15)
102 format(i4, es15.7, 3i4)
103 format(2f13.5, i5, es15.7)

    write (*, 320) hs%ntime, hs%time, hs%dt
    write (*, 330) hs%cflmax, hs%dvmax, ss%mg_cycles_max
320 format('--------------------------', i6, '  TIME=', f10.5, '  DT=', f12.8)
330 format('  CFL=', es13.5, '  MAX DIV=', es13.5, '  MG CYCLES=', i5)

    if (hs%written == 1) then
      write (2000, 130) hs%time, hs%dt, hs%cflmax, hs%dvmax, hs%qmmax
      if (heat) write (2012, 130) hs%time, hs%tbulk(1), hs%tbulk(2), hs%tbulk_avg(1), hs%tbulk_avg(2)
      if (ibm) write (2001, 140) hs%time, hs%cd(1), hs%cd(2), hs%cd(3), &
        hs%cd_avg(1), hs%cd_avg(2), hs%cd_avg(3)
      if (heat) write (2002, 140) hs%time, hs%hflux(1), hs%hflux(2), hs%hflux(3), &
        hs%hflux_avg(1), hs%hflux_avg(2), hs%hflux_avg(3)
      if (st%ich /= 0) write (2010, 140) hs%time, hs%qflux, hs%ubulk, hs%pmiavg, &
        hs%cmfravg(1), hs%cmfravg(2), hs%cmfravg(3)
    end if
    if (ntrace > 0) then
      if (mod(hs%ntime, ntr) == 0) then
        rc = lw_get_trace(h, trv, 256_c_int64_t)
        call check(rc, 'tracer')
        write (2005, 150) hs%time, trv(1:ntrace)
        write (2006, 150) hs%time, trv(ntrace + 1:2 * ntrace)
        write (2007, 150) hs%time, trv(2 * ntrace + 1:3 * ntrace)
        write (2008, 150) hs%time, trv(3 * ntrace + 1:4 * ntrace)
      end if
    end if
    if (mod(hs%ntime, st%nprint) == 0) call write_field(hs%ihist)
    write (2011, 130) hs%time, hs%s_sgs, hs%s_rhsnlhs, hs%s_poisson, wall
  end do

  call close_histories()
  if (mod(hs%ntime, st%nprint) /= 0) call write_field(hs%ihist)
  if (st%iavg == 1) then
    rc = lw_avg_finalize(h)
    call check(rc, 'field_avg_finalize')
  end if
  rc = lw_destroy(h)

130 format(f13.5, 4es15.7)
140 format(f13.5, 6es15.7)
150 format(f15.7, 10000es15.7)

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= lw_ok) then
      write (*, '(a,a,a,i0,a,a)') 'solver_exec: ', what, ' failed (', code, '): ', lw_error_text()
      error stop 1
    end if
  end subroutine check

  subroutine write_field(ihist)
    integer(c_int64_t), intent(in) :: ihist
    character(len=26) :: tname
    write (tname, '(a,i7.7)') '../output/field/fld', ihist
    rc = lw_write_field(h, trim(tname)//c_null_char)
    call check(rc, 'writing '//trim(tname))
  end subroutine write_field

  ! ftrfiles(1): new files, or appended when the run continues a field
  subroutine open_histories(append)
    logical, intent(in) :: append
    character(len=8) :: pos
    pos = 'asis'
    if (append) pos = 'append'
    open (2000, file='../output/ftr/fhist.dat', position=trim(pos))
    open (2001, file='../output/ftr/fcdcl.dat', position=trim(pos))
    open (2002, file='../output/ftr/fheatflux.dat', position=trim(pos))
    open (2012, file='../output/ftr/ftemp.dat', position=trim(pos))
    open (2011, file='../output/ftr/ftime.dat', position=trim(pos))
    if (ntrace /= 0) then
      open (2005, file='../output/ftr/futrace.dat', position=trim(pos))
      open (2006, file='../output/ftr/fvtrace.dat', position=trim(pos))
      open (2007, file='../output/ftr/fwtrace.dat', position=trim(pos))
      open (2008, file='../output/ftr/fptrace.dat', position=trim(pos))
    end if
    if (st%ich /= 0) open (2010, file='../output/ftr/fcmfr.dat', position=trim(pos))
  end subroutine open_histories

  subroutine close_histories()
    close (2000)
    close (2001)
    close (2002)
    close (2012)
    close (2011)
    if (st%poisson /= 2) close (77)
    if (st%poisson == 1) close (78)
    if (st%ich /= 0) close (2010)
    if (ntrace /= 0) then
      close (2005)
      close (2006)
      close (2007)
      close (2008)
    end if
  end subroutine close_histories

end program solver_exec
