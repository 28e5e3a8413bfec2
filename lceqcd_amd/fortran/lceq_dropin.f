*     lceq_dropin.f -- drop-in replacements for the reference's update and
*     measurement subroutines.  A reference build links this file (and
*     lceq_c.f90, liblceq.so) instead of f/agauge.f, f/staples.f,
*     f/heat_kpi.f, f/a_force.f, f/ahiggs1-3.f, f/gauss.f, f/renorm.f,
*     f/u_f_renorm.f, f/landau_gauge.f, f/or_gauge_fixing1.f,
*     f/correlator.f, f/scalarcorr.f and f/polyakov.f.  The routines keep
*     the reference's names, argument lists and COMMON side effects and
*     INCLUDE the reference's own headers (compile with
*     -I<reference>/f/INCLUDE and -fdefault-real-8: the device path is
*     FP64).  One rank only (nproc = 1, dimproc = 1,1,1).
*
*     Uniforms: by default each routine draws exactly the uniforms the
*     reference routine would (counts of f/heat_kpi.f:33-93,
*     f/agauge.f:167-169, f/ahiggs1.f:51-83, f/ahiggs2.f:83-111,
*     f/ahiggs3.f:109-111, f/gauss.f:21-26) from the compiler's
*     random_number intrinsic -- the reference's own stream, seeded as the
*     reference seeds it -- and feeds them to the device, which consumes
*     them in the reference's order.  With LCEQ_DROPIN_RNG=philox the
*     device's counter-based streams are used instead, seeded as
*     f/measinit.f:214-217 seeds random_number on rank 0 (lceq_rng_measinit:
*     key iseed(1) + 2^32 iseed(2) of the RANLUX seeds, which are also left
*     in COMMON iseed as the reference's measinit leaves them).
*
*     Coherence: the first call uploads u and a from COMMON /linkvar/ to
*     the device lattice.  Every routine that changes a field copies it
*     back into COMMON before returning, so unmodified host code (meas2a,
*     confsave) reads current values.  Code that changes u or a in COMMON
*     between calls calls lceq_dropin_upload afterwards.  With
*     LCEQ_DROPIN_SYNC=deferred the update routines leave u and a on the
*     device (the copies back cost more than the updates); host code then
*     calls lceq_dropin_download before it reads u or a (before confsave,
*     for instance).  The gauge-fixed u_f, a_f are copied back either way.

*     create the device lattice from the COMMON blocks and upload u, a
      subroutine lceq_dropin_attach
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'parameter.inc'
      INCLUDE 'input_parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer(c_int) rc
      integer(c_int) ls(16)
      character(len=16) mode
      integer ln, st
      if (c_associated(lceq_handle)) return
      if (nproc .ne. 1) then
        write(*,*) 'lceq_dropin: one rank only, nproc =', nproc
        call abort()
      endif
      rc = lceq_create(n_lattice(1), n_lattice(2), n_lattice(3),
     &                 lceq_handle)
      call lceq_check(rc, 'lceq_create')
      call lceq_dropin_params
      call get_environment_variable('LCEQ_DROPIN_SYNC', mode, ln, st)
      lceq_deferred = st .eq. 0 .and. mode(1:8) .eq. 'deferred'
      call get_environment_variable('LCEQ_DROPIN_RNG', mode, ln, st)
      if (st .eq. 0 .and. mode(1:6) .eq. 'philox') then
        call lceq_check(lceq_rng_measinit(lceq_handle, 0, ls),
     &       'rng_measinit')
        iseed(1:16) = ls
      endif
      call lceq_dropin_upload
      end

*     feed the next n uniforms of the random_number stream to the device
*     (reference-stream mode); nothing in philox mode
      subroutine lceq_dropin_uniforms(n)
      use lceq_c
      implicit none
      integer n
      real, allocatable :: r(:)
      real(c_double), allocatable :: d(:)
      character(len=16) mode
      integer ln, st
      call get_environment_variable('LCEQ_DROPIN_RNG', mode, ln, st)
      if (st .eq. 0 .and. mode(1:6) .eq. 'philox') return
      allocate(r(n), d(n))
      call random_number(r)
      d = r
      call lceq_check(lceq_rng_feed(lceq_handle, d,
     &     int(n, c_int64_t)), 'rng_feed')
      deallocate(r, d)
      end

*     couplings and gauge-fixing parameters from COMMON /paramet/
      subroutine lceq_dropin_params
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'input_parameter.inc'
      call lceq_check(lceq_set_couplings(lceq_handle,
     &     real(beta, c_double), real(beta_a, c_double),
     &     real(beta_2, c_double), real(beta_4, c_double)),
     &     'set_couplings')
      call lceq_check(lceq_set_gauge_fixing(lceq_handle,
     &     real(lambda, c_double), real(alpha, c_double),
     &     real(limit_or, c_double), 100000000), 'set_gauge_fixing')
      call lceq_check(lceq_set_imode(lceq_handle, imode), 'set_imode')
      end

      subroutine lceq_dropin_upload
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer(c_int64_t) ld
      ld = maxarray
      call lceq_check(lceq_put_links(lceq_handle, u, ld), 'put_links')
      call lceq_check(lceq_put_higgs(lceq_handle, a, ld), 'put_higgs')
      end

*     the updated u (a) back into COMMON, unless the copies are deferred
      subroutine lceq_dropin_get_links
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      if (lceq_deferred) return
      call lceq_check(lceq_get_links(lceq_handle, u,
     &     int(maxarray, c_int64_t)), 'get_links')
      end

      subroutine lceq_dropin_get_higgs
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      if (lceq_deferred) return
      call lceq_check(lceq_get_higgs(lceq_handle, a,
     &     int(maxarray, c_int64_t)), 'get_higgs')
      end

*     u and a from the device into COMMON (needed with
*     LCEQ_DROPIN_SYNC=deferred before host code reads them)
      subroutine lceq_dropin_download
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      if (.not. c_associated(lceq_handle)) return
      call lceq_check(lceq_get_links(lceq_handle, u,
     &     int(maxarray, c_int64_t)), 'get_links')
      call lceq_check(lceq_get_higgs(lceq_handle, a,
     &     int(maxarray, c_int64_t)), 'get_higgs')
      end

      subroutine lceq_dropin_get_fixed
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      call lceq_check(lceq_get_links_fixed(lceq_handle, u_f,
     &     int(maxarray, c_int64_t)), 'get_links_fixed')
      call lceq_check(lceq_get_higgs_fixed(lceq_handle, a_f,
     &     int(maxarray, c_int64_t)), 'get_higgs_fixed')
      end

      subroutine lceq_dropin_put_fixed
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      call lceq_check(lceq_put_links_fixed(lceq_handle, u_f,
     &     int(maxarray, c_int64_t)), 'put_links_fixed')
      call lceq_check(lceq_put_higgs_fixed(lceq_handle, a_f,
     &     int(maxarray, c_int64_t)), 'put_higgs_fixed')
      end

*     /aver_loc/ from the device observables
      subroutine lceq_dropin_obs(link, higgs)
      use lceq_c
      implicit none
      INCLUDE 'measurement.inc'
      logical link, higgs
      real(8) o(7)
      call lceq_check(lceq_observables(lceq_handle, o), 'observables')
      if (link) plaq0 = o(1)
      if (higgs) then
        a0kin = o(2)
        a02 = o(3)
        a04 = o(4)
      endif
      end

*     f/agauge.f:7
      subroutine agauge(measflag, accept)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'measurement.inc'
      integer measflag
      real accept
      real(8) acc
      call lceq_dropin_attach
      call lceq_dropin_uniforms(42*ownhalfvol(3))
      call lceq_check(lceq_agauge(lceq_handle, measflag, acc), 'agauge')
      accept = acc
      plaq0 = 0.0
      if (measflag .eq. 1) call lceq_dropin_obs(.true., .false.)
      call lceq_dropin_get_links
      end

*     f/staples.f:9 -- the staple lands in COMMON /helpfields/ v
      subroutine staples(eo, oe, mu)
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer eo, oe, mu
      call lceq_dropin_attach
      call lceq_check(lceq_staples(lceq_handle, eo, mu), 'staples')
      call lceq_check(lceq_get_staple(lceq_handle, v,
     &     int(maxarray, c_int64_t)), 'get_staple')
      end

*     f/heat_kpi.f:9 on ownhalfvol(3) sites
      subroutine heat_kpi(b, v, a0_old, nacch)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      real b(ownhalfvol(3)), a0_old(ownhalfvol(3))
      complex v(2, ownhalfvol(3))
      integer nacch
      real(8) w(4, ownhalfvol(3))
      integer(c_int) n
      integer i
      call lceq_dropin_attach
      call lceq_dropin_uniforms(6*ownhalfvol(3))
      call lceq_check(lceq_heat_kpi(lceq_handle, ownhalfvol(3),
     &     real(b, c_double), w, real(a0_old, c_double), n),
     &     'heat_kpi')
      do i = 1, ownhalfvol(3)
        v(1, i) = cmplx(w(1, i), w(2, i))
        v(2, i) = cmplx(w(3, i), w(4, i))
      enddo
      nacch = n
      end

*     f/a_force.f:7 -- the force lands in COMMON /helpfields/ force
      subroutine a_force(eo)
      use lceq_c
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer eo
      call lceq_dropin_attach
      call lceq_check(lceq_a_force(lceq_handle, eo), 'a_force')
      call lceq_check(lceq_get_force(lceq_handle, force,
     &     int(maxarray, c_int64_t)), 'get_force')
      end

*     f/gauss.f:10
      subroutine gauss(x, r, nh)
      use lceq_c
      implicit none
      integer nh
      real x(nh), r(nh)
      real(8) xd(nh)
      call lceq_dropin_attach
      call lceq_dropin_uniforms(2*nh)
      call lceq_check(lceq_gauss(lceq_handle, xd, real(r, c_double),
     &     nh), 'gauss')
      x = xd
      end

*     f/ahiggs1.f:7
      subroutine ahiggs1(delta, measflag, accept)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'measurement.inc'
      real delta, accept
      integer measflag
      real(8) d, acc
      call lceq_dropin_attach
      call lceq_dropin_uniforms(8*ownhalfvol(3))
      d = delta
      call lceq_check(lceq_ahiggs1(lceq_handle, d, measflag, acc),
     &     'ahiggs1')
      delta = d
      accept = acc
      a0kin = 0.0
      a02 = 0.0
      a04 = 0.0
      if (measflag .eq. 1) call lceq_dropin_obs(.false., .true.)
      call lceq_dropin_get_higgs
      end

*     f/ahiggs2.f:7 (accept is set only when measflag = 1, f/ahiggs2.f:142)
      subroutine ahiggs2(measflag, accept)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'measurement.inc'
      integer measflag
      real accept
      real(8) acc
      call lceq_dropin_attach
      call lceq_dropin_uniforms(14*ownhalfvol(3))
      call lceq_check(lceq_ahiggs2(lceq_handle, measflag, acc),
     &     'ahiggs2')
      a0kin = 0.0
      a02 = 0.0
      a04 = 0.0
      if (measflag .eq. 1) then
        accept = acc
        call lceq_dropin_obs(.false., .true.)
      endif
      call lceq_dropin_get_higgs
      end

*     f/ahiggs3.f:7
      subroutine ahiggs3(measflag, accept)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'measurement.inc'
      integer measflag
      real accept
      real(8) acc
      call lceq_dropin_attach
      call lceq_dropin_uniforms(2*ownhalfvol(3))
      call lceq_check(lceq_ahiggs3(lceq_handle, measflag, acc),
     &     'ahiggs3')
      a0kin = 0.0
      a02 = 0.0
      a04 = 0.0
      if (measflag .eq. 1) then
        accept = acc
        call lceq_dropin_obs(.false., .true.)
      endif
      call lceq_dropin_get_higgs
      end

*     f/renorm.f:6
      subroutine renorm
      use lceq_c
      implicit none
      call lceq_dropin_attach
      call lceq_check(lceq_renorm(lceq_handle), 'renorm')
      call lceq_dropin_get_links
      end

*     f/u_f_renorm.f:6
      subroutine u_f_renorm
      use lceq_c
      implicit none
      call lceq_dropin_attach
      call lceq_dropin_put_fixed
      call lceq_check(lceq_u_f_renorm(lceq_handle), 'u_f_renorm')
      call lceq_dropin_get_fixed
      end

*     f/or_gauge_fixing1.f:5 on COMMON u_f, a_f
      subroutine or_gauge_fixing(da1)
      use lceq_c
      implicit none
      real da1
      real(8) d
      call lceq_dropin_attach
      call lceq_dropin_put_fixed
      call lceq_check(lceq_or_gauge_fixing(lceq_handle, d),
     &     'or_gauge_fixing')
      da1 = d
      call lceq_dropin_get_fixed
      end

*     f/landau_gauge.f:6
      subroutine landau_gauge(da1, iter_or1)
      use lceq_c
      implicit none
      real da1
      integer iter_or1
      real(8) d
      integer(c_int) it
      call lceq_dropin_attach
      call lceq_check(lceq_landau_gauge(lceq_handle, d, it),
     &     'landau_gauge')
      da1 = d
      iter_or1 = it
      call lceq_dropin_get_fixed
      end

*     f/correlator.f:11 -- same files and records as f/correlator.f:166-181
      subroutine correlator(meas, mstart)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'filenames.inc'
      INCLUDE 'measurement.inc'
      integer meas, mstart, z
      real(8) ca(n_lattice(3)), cu(n_lattice(3)), c3(n_lattice(3))
      call lceq_dropin_attach
      call lceq_check(lceq_correlator(lceq_handle, ca, cu, c3),
     &     'correlator')
      if (myrank .eq. root) then
        open(unit=corr_unit, file=corr_file, status='unknown',
     &       position='append')
        write(corr_unit, '(i3.3)') mstart + meas
        do z = 1, n_lattice(3)
          write(corr_unit, '(2(e23.16,1x))') ca(z), cu(z)
        enddo
        close(corr_unit)
        open(unit=checkp_unit, file=checkp_file, status='unknown',
     &       position='append')
        write(checkp_unit, '(i3.3)') mstart + meas
        do z = 1, n_lattice(3)
          write(checkp_unit, '(e23.16)') c3(z)
        enddo
        close(checkp_unit)
      endif
      call lceq_dropin_get_fixed
      end

*     f/scalarcorr.f:10 -- records of f/scalarcorr.f:106-113
      subroutine scalarcorr(meas, mstart)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'filenames.inc'
      integer meas, mstart, z
      real(8) m, c(n_lattice(3))
      call lceq_dropin_attach
      call lceq_check(lceq_scalarcorr(lceq_handle, m, c), 'scalarcorr')
      if (myrank .eq. root) then
        open(unit=corra2_unit, file=corra2_file, status='unknown',
     &       position='append')
        write(corra2_unit, '(i5.5)') mstart + meas
        write(corra2_unit, '(e23.16)') m
        do z = 1, n_lattice(3)
          write(corra2_unit, '(e23.16)') c(z)
        enddo
        close(corra2_unit)
      endif
      end

*     f/polyakov.f:10 -- records of f/polyakov.f:110-117
      subroutine polyakov(meas, mstart)
      use lceq_c
      implicit none
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'filenames.inc'
      integer meas, mstart, z
      real(8) m, c(n_lattice(3))
      call lceq_dropin_attach
      call lceq_check(lceq_polyakov(lceq_handle, m, c), 'polyakov')
      if (myrank .eq. root) then
        open(unit=polyak_unit, file=polyak_file, status='unknown',
     &       position='append')
        write(polyak_unit, '(i5.5)') mstart + meas
        write(polyak_unit, '(e23.16)') m
        do z = 1, n_lattice(3)
          write(polyak_unit, '(e23.16)') c(z)
        enddo
        close(polyak_unit)
      endif
      end
