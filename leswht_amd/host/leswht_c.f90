!=======================================================================
! ISO_C_BINDING view of the C ABI in include/leswht.h: the structs, field
! for field in the same order, and the entry points the Fortran host calls.
!=======================================================================
module leswht_c
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: lw_ok = 0

  ! settings.input + boundary.input + ibmpre_prdic.bin (lw_settings)
  type, bind(c) :: lw_settings
    real(c_double) :: re, pr, gr, tflu_i, tsol_i, eps_ptr, udrv_i, dt_size, cflfac, resid1, &
                      wwsor, csgsts, csgshf, tend, ptb_tst, avg_tst
    integer(c_int64_t) :: grdir, ireset, iread, iavg, ipzero, nend, nprint, npriavg, npin, &
                          idtopt, nlev, nbli, ioldv, mgitr, imgsor, iles, insmdl, itemdl, idvmon, &
                          filter, ibmon, masson, imovingon, ihtrans, idecomp
    integer(c_int64_t) :: bc_ybtm, bc_ytop, bc_zbtm, bc_ztop, bc_t_ybtm, bc_t_ytop, bc_t_zbtm, &
                          bc_t_ztop, ich, iconjg, xprdic, yprdic, zprdic, iintp
    real(c_double) :: val_t_ybtm, val_t_ytop, val_t_zbtm, val_t_ztop
    integer(c_int64_t) :: poisson
  end type lw_settings

  ! per-step results (lw_step_stats)
  type, bind(c) :: lw_step_stats
    real(c_double) :: dt, time, cflmax, dvmax, qmmax, csgsts, csgshf
    integer(c_int64_t) :: ntime, mg_cycles_max, mg_cycles_sum
    real(c_double) :: ms_sgs, ms_momentum, ms_poisson, ms_energy
  end type lw_step_stats

  ! history quantities of the last step (lw_history)
  type, bind(c) :: lw_history
    integer(c_int64_t) :: ntime, ihist, written
    real(c_double) :: time, dt, cflmax, dvmax, qmmax
    real(c_double) :: tbulk(2), tbulk_avg(2)
    real(c_double) :: cd(3), cd_avg(3)
    real(c_double) :: hflux(3), hflux_avg(3)
    real(c_double) :: s_sgs, s_rhsnlhs, s_poisson, s_energy
    real(c_double) :: qflux, ubulk, pmiavg, cmfravg(3)
  end type lw_history

  ! body level set (lw_body): kind 0 = circular cylinder of radius p(1)
  type, bind(c) :: lw_body
    integer(c_int64_t) :: kind
    real(c_double) :: p(4)
  end type lw_body

  interface
    integer(c_int) function lw_create_from_dir(solver_dir, body, h) bind(c, name='lw_create_from_dir')
      import :: c_int, c_char, c_ptr, lw_body
      character(kind=c_char), dimension(*), intent(in) :: solver_dir
      type(lw_body), intent(in) :: body
      type(c_ptr), intent(out) :: h
    end function lw_create_from_dir

    integer(c_int) function lw_destroy(h) bind(c, name='lw_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function lw_destroy

    integer(c_int) function lw_step(h, nsteps, stats) bind(c, name='lw_step')
      import :: c_int, c_ptr, c_int64_t, lw_step_stats
      type(c_ptr), value :: h
      integer(c_int64_t), value :: nsteps
      type(lw_step_stats), intent(out) :: stats
    end function lw_step

    integer(c_int) function lw_get_history(h, out) bind(c, name='lw_get_history')
      import :: c_int, c_ptr, lw_history
      type(c_ptr), value :: h
      type(lw_history), intent(out) :: out
    end function lw_get_history

    ! TRACE block and tracer (misc_init.f90:47-52, misc_hist.f90:184-209)
    integer(c_int) function lw_trace_info(h, ntrace, ntr, pts) bind(c, name='lw_trace_info')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: ntrace, ntr
      integer(c_int64_t), intent(out) :: pts(*)
    end function lw_trace_info

    integer(c_int) function lw_get_trace(h, out, cap) bind(c, name='lw_get_trace')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: cap
    end function lw_get_trace

    ! the mgftresiduemax.dat rows of the last step (poiss_z_ft.f90:293-297)
    integer(c_int64_t) function lw_get_mg_residuals(h, kv, kk, rm, cap) bind(c, name='lw_get_mg_residuals')
      import :: c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: kv(*), kk(*)
      real(c_double), intent(out) :: rm(*)
      integer(c_int64_t), value :: cap
    end function lw_get_mg_residuals

    ! every mgftresiduemax.dat row of the last step in file order: kind 0
    ! mg2d's RM rows, kind 1 gsor2d's RG rows (poiss_z_ft.f90:293-297, 584)
    integer(c_int64_t) function lw_get_residual_rows(h, kind, kv, kk, val, cap) bind(c, name='lw_get_residual_rows')
      import :: c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: kind(*), kv(*), kk(*)
      real(c_double), intent(out) :: val(*)
      integer(c_int64_t), value :: cap
    end function lw_get_residual_rows

    ! the poiss_itr.dat and ftrpoittr.dat rows of the last step (poiss_mgrd.f90:185,221)
    integer(c_int64_t) function lw_get_bicg_log(h, sub, icg, rm, loc, cap) bind(c, name='lw_get_bicg_log')
      import :: c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: sub(*), icg(*), loc(*)
      real(c_double), intent(out) :: rm(*)
      integer(c_int64_t), value :: cap
    end function lw_get_bicg_log

    integer(c_int64_t) function lw_get_bicg_solves(h, icg, rm, cap) bind(c, name='lw_get_bicg_solves')
      import :: c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: icg(*)
      real(c_double), intent(out) :: rm(*)
      integer(c_int64_t), value :: cap
    end function lw_get_bicg_solves

    integer(c_int) function lw_get_settings(h, out) bind(c, name='lw_get_settings')
      import :: c_int, c_ptr, lw_settings
      type(c_ptr), value :: h
      type(lw_settings), intent(out) :: out
    end function lw_get_settings

    integer(c_int) function lw_write_field(h, path) bind(c, name='lw_write_field')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: path
    end function lw_write_field

    ! field_avg_finalize (misc_hist.f90:530-697)
    integer(c_int) function lw_avg_finalize(h) bind(c, name='lw_avg_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function lw_avg_finalize

    type(c_ptr) function lw_last_error() bind(c, name='lw_last_error')
      import :: c_ptr
    end function lw_last_error
  end interface

contains

  ! the library's last error message as a Fortran string
  function lw_error_text() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: chars(:)
    type(c_ptr) :: p
    integer :: n
    p = lw_last_error()
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, chars, [4096])
    n = 0
    do while (n < 4096)
      if (chars(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function lw_error_text

end module leswht_c
