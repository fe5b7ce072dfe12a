! lh_amd.f90 -- Fortran 2003 interface (iso_c_binding) to the MI355X solver library
! liblh_amd.so, one-to-one with include/littleharsh_amd.h.
!
! The reference host keeps its own set-up and loop; the substep body
!     call RHS0_u1/u2/u3 ; call nonlinear ; call solveU/V/U ; meanflow ; solveP ; v_corr
! (littleharsh.f90:166-188) becomes one call, status = lh_substep(h, kRK, dt).
! State arrays are the reference's own cfield storage, passed with c_loc(u1(midband)%f).
module lh_amd
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LH_OK = 0, LH_ERR_ARG = 1, LH_ERR_HIP = 2, LH_ERR_NUMERIC = 3
  integer(c_int), parameter :: LH_PREC_FP64 = 0, LH_PREC_MIXED = 1

  ! same member order and types as struct lh_config
  type, bind(C) :: lh_config
    integer(c_int) :: nx, nz, px, pz
    integer(c_int) :: ny, dsty
    real(c_double) :: re, alp, bet
    integer(c_int) :: flag_ctpress
    real(c_double) :: mpgx, qxt, ddthetavi
    type(c_ptr) :: yu, yv, dthdyu, dthdyv, dyu2i, dyv2i
    type(c_ptr) :: gridweighting, rk
    type(c_ptr) :: k1x, k2x, k1z, k2z
    integer(c_int) :: nlist_u, nlist_v
    type(c_ptr) :: list_u, list_v
    integer(c_int) :: tip_u(2), tip_v(2)
    integer(c_int) :: ntilex, ntilez
    integer(c_int) :: band_v(2)
    integer(c_int) :: precision
  end type

  interface
    integer(c_int) function lh_create(cfg, h) bind(C, name="lh_create")
      import :: c_int, c_ptr, lh_config
      type(lh_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function lh_destroy(h) bind(C, name="lh_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function lh_last_error(h) bind(C, name="lh_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lh_set_state(h, u1, u2, u3, p) bind(C, name="lh_set_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, u1, u2, u3, p
    end function
    integer(c_int) function lh_get_state(h, u1, u2, u3, p) bind(C, name="lh_get_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, u1, u2, u3, p
    end function
    integer(c_int) function lh_set_nonlinear(h, n1, n2, n3) bind(C, name="lh_set_nonlinear")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, n1, n2, n3
    end function
    integer(c_int) function lh_get_nonlinear(h, n1, n2, n3) bind(C, name="lh_get_nonlinear")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, n1, n2, n3
    end function
    ! forcing added to Nu_new every substep (a coarse channel's closure term)
    integer(c_int) function lh_set_forcing(h, f1, f2, f3) bind(C, name="lh_set_forcing")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, f1, f2, f3
    end function
    integer(c_int) function lh_get_forcing(h, f1, f2, f3) bind(C, name="lh_get_forcing")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, f1, f2, f3
    end function
    ! the learned closure's network (trained and evaluated on the GPU)
    integer(c_int) function lh_mlp_create(nin, nhid, nlayers, nout, batch, seed, m) bind(C, name="lh_mlp_create")
      import :: c_int, c_ptr, c_long_long
      integer(c_int), value :: nin, nhid, nlayers, nout, batch
      integer(c_long_long), value :: seed
      type(c_ptr), intent(out) :: m
    end function
    integer(c_int) function lh_mlp_destroy(m) bind(C, name="lh_mlp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_long_long) function lh_mlp_param_count(m) bind(C, name="lh_mlp_param_count")
      import :: c_long_long, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function lh_mlp_set_params(m, params) bind(C, name="lh_mlp_set_params")
      import :: c_int, c_ptr
      type(c_ptr), value :: m, params
    end function
    integer(c_int) function lh_mlp_get_params(m, params) bind(C, name="lh_mlp_get_params")
      import :: c_int, c_ptr
      type(c_ptr), value :: m, params
    end function
    integer(c_int) function lh_mlp_set_data(m, x, y, n) bind(C, name="lh_mlp_set_data")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: m, x, y
      integer(c_long_long), value :: n
    end function
    integer(c_int) function lh_mlp_train(m, steps, lr, seed, loss_last) bind(C, name="lh_mlp_train")
      import :: c_int, c_ptr, c_float, c_long_long
      type(c_ptr), value :: m
      integer(c_int), value :: steps
      real(c_float), value :: lr
      integer(c_long_long), value :: seed
      real(c_float), intent(out) :: loss_last
    end function
    integer(c_int) function lh_mlp_set_fused(m, on) bind(C, name="lh_mlp_set_fused")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: on
    end function
    ! the closure bound to a coarse channel (arrays: c_loc of the host's own storage)
    integer(c_int) function lh_closure_create(h, m, yc, dxc, dzc, shift_x, shift_z, sf, xmean, xscale, &
                                              ymean, yscale, c) bind(C, name="lh_closure_create")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, m, yc, sf, xmean, xscale, ymean, yscale
      real(c_double), value :: dxc, dzc, shift_x, shift_z
      type(c_ptr), intent(out) :: c
    end function
    integer(c_int) function lh_closure_create_smagorinsky(h, yc, dxc, dzc, shift_x, shift_z, cs2, c) &
        bind(C, name="lh_closure_create_smagorinsky")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, yc
      real(c_double), value :: dxc, dzc, shift_x, shift_z, cs2
      type(c_ptr), intent(out) :: c
    end function
    ! its forcing from the current state, on the GPU (call before each lh_step)
    integer(c_int) function lh_closure_apply(c) bind(C, name="lh_closure_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function lh_closure_set_mean_only(c, on) bind(C, name="lh_closure_set_mean_only")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: on
    end function
    integer(c_int) function lh_closure_destroy(c) bind(C, name="lh_closure_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function lh_substep(h, krk, dt) bind(C, name="lh_substep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: krk
      real(c_double), value :: dt
    end function
    ! the three stages of substep kRK, for parity bisection (in turn they equal lh_substep)
    integer(c_int) function lh_nonlinear(h, krk) bind(C, name="lh_nonlinear")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: krk
    end function
    integer(c_int) function lh_solve_u(h, krk, dt) bind(C, name="lh_solve_u")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: krk
      real(c_double), value :: dt
    end function
    integer(c_int) function lh_solve_p(h, krk, dt) bind(C, name="lh_solve_p")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: krk
      real(c_double), value :: dt
    end function
    integer(c_int) function lh_step(h, dt) bind(C, name="lh_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
    end function
    integer(c_int) function lh_step_cfl(h, cfl, dt_used) bind(C, name="lh_step_cfl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: cfl
      real(c_double), intent(out) :: dt_used
    end function
    integer(c_int) function lh_cfl_dt(h, cfl, dt) bind(C, name="lh_cfl_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: cfl
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function lh_get_scalars(h, mpgx, qx, dt, cflmax) bind(C, name="lh_get_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: mpgx, qx, dt, cflmax
    end function
    integer(c_int) function lh_max_div(h, err) bind(C, name="lh_max_div")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: err
    end function
    ! err_dev: a device address (for example from hipMalloc at set-up); nothing is allocated
    integer(c_int) function lh_max_div_noalloc(h, err_dev) bind(C, name="lh_max_div_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, err_dev
    end function
    integer(c_int) function lh_stats_accumulate(h) bind(C, name="lh_stats_accumulate")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lh_stats_read(h, out, nstat, reset) bind(C, name="lh_stats_read")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nstat
      integer(c_int), value :: reset
    end function
    integer(c_int) function lh_cond_stats_size(h, n) bind(C, name="lh_cond_stats_size")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: h
      integer(c_long_long), intent(out) :: n
    end function
    integer(c_int) function lh_cond_stats_read(h, out, n, reset) bind(C, name="lh_cond_stats_read")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_long_long), value :: n
      integer(c_int), value :: reset
    end function
    integer(c_int) function lh_spectra_accumulate(h) bind(C, name="lh_spectra_accumulate")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lh_spectra_read(h, out, reset) bind(C, name="lh_spectra_read")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int), value :: reset
    end function
    integer(c_int) function lh_timing_enable(h, enable) bind(C, name="lh_timing_enable")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: enable
    end function
    integer(c_int) function lh_timing_reset(h) bind(C, name="lh_timing_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lh_timing_read(h, kernel, ms, launches, bytes) bind(C, name="lh_timing_read")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: h
      integer(c_int), value :: kernel
      real(c_double), intent(out) :: ms, bytes
      integer(c_long_long), intent(out) :: launches
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! lh_last_error as a Fortran string
  function lh_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n, i
    p = lh_last_error(h)
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  end function

  ! the reference's error convention: print and stop (e.g. start.f90:149-150)
  subroutine lh_check(h, status, where)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= LH_OK) then
      write(*,*) 'lh error in ', where, ': ', lh_error_message(h), ' (status', status, ')'
      stop 1
    end if
  end subroutine

end module lh_amd
