! lceq_c.f90 -- iso_c_binding interface to liblceq.so (include/lceq.h).
!
! A Fortran host (the driver lceq_main.f90, or a reference build that links
! lceq_dropin.f instead of the reference's own update routines) calls the
! device-resident lattice through these interfaces.  Argument meaning follows
! the reference subroutines named in include/lceq.h; every function returns
! the C status (0 = ok).  uint64 arguments (seeds, pass counters) are passed
! as integer(c_int64_t).
module lceq_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LCEQ_OK = 0
  integer(c_int), parameter :: LCEQ_SWEEP_KP_OR_METRO = 0, LCEQ_SWEEP_REFERENCE = 1

  ! lattice handle of the drop-in routines (lceq_dropin.f)
  type(c_ptr), save :: lceq_handle = c_null_ptr
  ! drop-in coherence (lceq_dropin.f): with LCEQ_DROPIN_SYNC=deferred the
  ! update routines leave u and a on the device, and lceq_dropin_download
  ! copies them into COMMON when host code needs them
  logical, save :: lceq_deferred = .false.

  interface
    integer(c_int) function lceq_create(lx, ly, lz, h) bind(C, name="lceq_create")
      import :: c_int, c_ptr
      integer(c_int), value :: lx, ly, lz
      type(c_ptr) :: h
    end function
    integer(c_int) function lceq_destroy(h) bind(C, name="lceq_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function lceq_last_error() bind(C, name="lceq_last_error")
      import :: c_ptr
    end function
    integer(c_int) function lceq_sync(h) bind(C, name="lceq_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lceq_set_couplings(h, beta, beta_a, beta_2, beta_4) &
        bind(C, name="lceq_set_couplings")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: beta, beta_a, beta_2, beta_4
    end function
    integer(c_int) function lceq_set_gauge_fixing(h, lambda, alpha, limit_or, max_iter) &
        bind(C, name="lceq_set_gauge_fixing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: lambda, alpha, limit_or
      integer(c_int), value :: max_iter
    end function
    integer(c_int) function lceq_set_imode(h, imode) bind(C, name="lceq_set_imode")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: imode
    end function
    integer(c_int) function lceq_rng_philox(h, seed, pass) bind(C, name="lceq_rng_philox")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t), value :: seed, pass
    end function
    integer(c_int) function lceq_rng_measinit(h, rank, iseed) bind(C, name="lceq_rng_measinit")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: rank
      integer(c_int) :: iseed(16)
    end function
    integer(c_int64_t) function lceq_measinit_key(iseed) bind(C, name="lceq_measinit_key")
      import :: c_int, c_int64_t
      integer(c_int) :: iseed(16)
    end function
    integer(c_int) function lceq_rng_feed(h, uniforms, n) bind(C, name="lceq_rng_feed")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: uniforms(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function lceq_rng_state(h, consumed, pass) bind(C, name="lceq_rng_state")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t) :: consumed, pass
    end function
    integer(c_int) function lceq_put_links(h, u, ld) bind(C, name="lceq_put_links")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*), intent(in) :: u(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_get_links(h, u, ld) bind(C, name="lceq_get_links")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*) :: u(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_put_higgs(h, a, ld) bind(C, name="lceq_put_higgs")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*), intent(in) :: a(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_get_higgs(h, a, ld) bind(C, name="lceq_get_higgs")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*) :: a(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_put_links_fixed(h, u, ld) bind(C, name="lceq_put_links_fixed")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*), intent(in) :: u(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_get_links_fixed(h, u, ld) bind(C, name="lceq_get_links_fixed")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*) :: u(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_put_higgs_fixed(h, a, ld) bind(C, name="lceq_put_higgs_fixed")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*), intent(in) :: a(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_get_higgs_fixed(h, a, ld) bind(C, name="lceq_get_higgs_fixed")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*) :: a(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_get_staple(h, v, ld) bind(C, name="lceq_get_staple")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*) :: v(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_get_force(h, f, ld) bind(C, name="lceq_get_force")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(*) :: f(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function lceq_confsave(h, u_path, a_path, real_bytes) &
        bind(C, name="lceq_confsave")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*) :: u_path, a_path
      integer(c_int), value :: real_bytes
    end function
    integer(c_int) function lceq_oldconf(h, u_path, a_path, real_bytes) &
        bind(C, name="lceq_oldconf")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*) :: u_path, a_path
      integer(c_int), value :: real_bytes
    end function
    integer(c_int) function lceq_unitconf(h) bind(C, name="lceq_unitconf")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lceq_testconf(h) bind(C, name="lceq_testconf")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lceq_hotstart(h, seed, sigma) bind(C, name="lceq_hotstart")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), value :: seed
      real(c_double), value :: sigma
    end function
    ! staples(eo,oe,mu), f/staples.f:9
    integer(c_int) function lceq_staples(h, eo, mu) bind(C, name="lceq_staples")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: eo, mu
    end function
    ! heat_kpi(b,v,a0_old,nacch), f/heat_kpi.f:9
    integer(c_int) function lceq_heat_kpi(h, n, b, v, a0_old, nacch) bind(C, name="lceq_heat_kpi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_double), intent(in) :: b(*), a0_old(*)
      real(c_double) :: v(*)
      integer(c_int) :: nacch
    end function
    ! agauge(measflag,accept), f/agauge.f:7
    integer(c_int) function lceq_agauge(h, measflag, accept) bind(C, name="lceq_agauge")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: measflag
      real(c_double) :: accept
    end function
    integer(c_int) function lceq_link_or(h, measflag, accept) bind(C, name="lceq_link_or")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: measflag
      real(c_double) :: accept
    end function
    ! a_force(eo), f/a_force.f:7
    integer(c_int) function lceq_a_force(h, eo) bind(C, name="lceq_a_force")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: eo
    end function
    ! gauss(X,R,NH), f/gauss.f:10
    integer(c_int) function lceq_gauss(h, x, r, nh) bind(C, name="lceq_gauss")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: x(*)
      real(c_double), intent(in) :: r(*)
      integer(c_int), value :: nh
    end function
    ! ahiggs1(delta,measflag,accept), f/ahiggs1.f:7
    integer(c_int) function lceq_ahiggs1(h, delta, measflag, accept) bind(C, name="lceq_ahiggs1")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: delta, accept
      integer(c_int), value :: measflag
    end function
    ! ahiggs2(measflag,accept), f/ahiggs2.f:7
    integer(c_int) function lceq_ahiggs2(h, measflag, accept) bind(C, name="lceq_ahiggs2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: measflag
      real(c_double) :: accept
    end function
    ! ahiggs3(measflag,accept), f/ahiggs3.f:7
    integer(c_int) function lceq_ahiggs3(h, measflag, accept) bind(C, name="lceq_ahiggs3")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: measflag
      real(c_double) :: accept
    end function
    ! renorm(), f/renorm.f:6
    integer(c_int) function lceq_renorm(h) bind(C, name="lceq_renorm")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! field guard (no reference counterpart): LCEQ_ERR_FIELD on non-finite
    ! values or, for tol >= 0, links further than tol from SU(2)
    integer(c_int) function lceq_check_fields(h, tol, max_dev, nonfinite) &
        bind(C, name="lceq_check_fields")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      real(c_double), value :: tol
      real(c_double), intent(out) :: max_dev
      integer(c_int64_t), intent(out) :: nonfinite
    end function
    ! u_f_renorm(), f/u_f_renorm.f:6
    integer(c_int) function lceq_u_f_renorm(h) bind(C, name="lceq_u_f_renorm")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! or_gauge_fixing(da1), f/or_gauge_fixing1.f:5
    integer(c_int) function lceq_or_gauge_fixing(h, da1) bind(C, name="lceq_or_gauge_fixing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: da1
    end function
    ! landau_gauge(da1,iter_or1), f/landau_gauge.f:6
    integer(c_int) function lceq_landau_gauge(h, da1, iter_or1) bind(C, name="lceq_landau_gauge")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: da1
      integer(c_int) :: iter_or1
    end function
    ! correlator(meas,mstart), f/correlator.f:11
    integer(c_int) function lceq_correlator(h, corr_a, corr_u, corr_3) bind(C, name="lceq_correlator")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: corr_a(*), corr_u(*), corr_3(*)
    end function
    ! scalarcorr(meas,mstart), f/scalarcorr.f:10
    integer(c_int) function lceq_scalarcorr(h, mean, corr) bind(C, name="lceq_scalarcorr")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: mean, corr(*)
    end function
    ! polyakov(meas,mstart), f/polyakov.f:10
    integer(c_int) function lceq_polyakov(h, mean_sq, corr) bind(C, name="lceq_polyakov")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: mean_sq, corr(*)
    end function
    ! MULTALL{NN,DN,ND,DD,AD,AN}, f/matmul.f (kind 0..5; 0-based site maps)
    integer(c_int) function lceq_multall(h, kind, op1, op2, out, n, s1, s2) bind(C, name="lceq_multall")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: kind, n
      real(c_double), intent(in) :: op1(*), op2(*)
      real(c_double) :: out(*)
      type(c_ptr), value :: s1, s2
    end function
    ! COMMON /aver_loc/ plaq0, a0kin, a02, a04, a0f(3)
    integer(c_int) function lceq_observables(h, out) bind(C, name="lceq_observables")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: out(7)
    end function
    integer(c_int) function lceq_sweep(h, schedule, nsweep, n_or, delta, measflag) &
        bind(C, name="lceq_sweep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: schedule, nsweep, n_or, measflag
      real(c_double) :: delta
    end function
    integer(c_int) function lceq_history(h, out, n) bind(C, name="lceq_history")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: out(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function lceq_profile_begin(h, kernel_class, max_launches) &
        bind(C, name="lceq_profile_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kernel_class, max_launches
    end function
    integer(c_int) function lceq_profile_end(h, total_ms, launches) bind(C, name="lceq_profile_end")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: total_ms
      integer(c_int) :: launches
    end function
  end interface

contains

  ! message of the last failure (lceq_last_error) as a Fortran string
  function lceq_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(lceq_last_error(), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function

  ! abort as the reference does on fatal errors (f/confget.f:61-66)
  subroutine lceq_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= LCEQ_OK) then
      write(*, '(a,a,a,i0,a,a)') 'lceq: ', where, ' failed with status ', rc, ': ', &
          trim(lceq_error_message())
      call abort()
    end if
  end subroutine
end module lceq_c
