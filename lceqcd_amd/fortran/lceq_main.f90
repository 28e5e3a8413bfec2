! lceq_main.f90 -- Fortran host driver of the device-resident lattice.
!
! Runs the schedule of the reference main program (f/main.f:12-87) through
! the C ABI (lceq_c): 100 thermalisation sweeps of ahiggs2, ahiggs3, agauge,
! renorm without measurement, then nsweep measured sweeps; every imeas sweeps
! Landau gauge fixing and the z-correlators.  The input is the reference's
! 20-line list-directed inputfile (f/paraminit.f:82-101, e.g. runs/inputfile),
! the couplings follow f/paraminit.f:150-153, and the output files carry the
! reference's names (f/measinit.f:246-273) and record formats (f/meas1.f:91,
! f/correlator.f:168-181, f/meas2a.f:83-97, f/meas1.f:70-76).
!
!   lceq_main [inputfile] [seed | ranlux] [ntherm] [realbytes] [corrflag]
!
! ntherm (default 100, as f/main.f:20) is the number of thermalisation sweeps;
! realbytes (8, or 4 for files of the as-shipped single-precision build) is
! the real size of the configuration files.  corrflag selects the
! measurement of f/main.f:55-79 (the reference fixes corr_flag = 1 in its
! source): 1 Landau gauge + the correlators (correl_, checkp_, afield_
! files), 2 polyakov + scalarcorr (polyak_, corra2_ files), 3 both.
!
! Differences from the reference, by design: one device holds the whole
! lattice (dimproc and nproc are read and ignored); the uniforms come from the
! counter-based Philox streams, seeded by default (or with the argument
! `ranlux`) as f/measinit.f:214-217 seeds the reference's random_number on rank
! 0: RANLUX level 3 from 314159, 16 seeds int(10^6 u), which also fill
! iseed(1:16) in the joblog; the streams' key is iseed(1) + 2^32 iseed(2)
! (lceq_rng_measinit).  A number as `seed` keys the streams directly.  The hot
! start draws from the same key.  init = 0 or 1 is the hot
! start (Haar links, Higgs N(0, 0.5^2)), 2 is unitconf, 4 is testconf, 3 is
! oldconf (f/confget.f:318-526) from the file f/measinit.f:246-250 names, or,
! when that lower-case name does not exist, from the upper-case name that
! f/confsave.f:249-253 writes.  The configuration is saved at the end as
! f/finalize.f:43 does.
program lceq_main
  use iso_c_binding
  use lceq_c
  implicit none

  integer :: nl(3), dimproc(3), nproc, init, rand, save, confstart
  integer :: inorm, nsweep_flag, ntime, nsweep, imeas, imode
  real(8) :: beta, x, y, lambda, alpha, limit_or, limit_it
  real(8) :: beta_a, beta_2, beta_4, pi
  real(8), parameter :: sigma = 3.1759114d0
  character(len=256) :: fin, arg
  character(len=49) :: action_file, joblog_file, corr_file, checkp_file, afield_file
  character(len=49) :: polyak_file, corra2_file
  integer :: betaint, xint, yint, isweep, meas, mstart, nmeas, iter1, iter_or, z, ntherm
  integer(c_int64_t) :: seed, nbad
  integer(c_int) :: iseed(16)
  logical :: ranlux_seed
  integer :: realbytes, corr_flag
  real(8) :: pmean
  real(8), allocatable :: pcorr(:)
  character(len=49) :: u_conf_file, a_conf_file
  type(c_ptr) :: h
  real(8) :: acc1, acc2, acc3, acco, acch, accg, da1, da, obs(7), maxdev
  real(8), allocatable :: corr_a(:), corr_u(:), corr_3(:)
  integer :: u_act = 3, u_job = 4, u_cor = 31, u_afi = 32, u_chk = 33, u_pol = 34, u_ca2 = 35

  fin = 'inputfile'
  if (command_argument_count() >= 1) call get_command_argument(1, fin)
  ranlux_seed = .true.
  iseed = 0
  seed = 0
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    if (trim(arg) /= 'ranlux') then
      read(arg, *) seed
      ranlux_seed = .false.
    end if
  end if
  ntherm = 100
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) ntherm
  end if
  realbytes = 8
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) realbytes
  end if
  corr_flag = 1
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg)
    read(arg, *) corr_flag
  end if
  if (corr_flag < 1 .or. corr_flag > 3) then
    write(*, *) 'corrflag must be 1, 2 or 3:', corr_flag
    call abort()
  end if

  ! ---- f/paraminit.f:82-101
  open(1, file=trim(fin), status='old', err=20)
  read(1, *, err=20) nl(1), nl(2), nl(3)
  read(1, *, err=20) dimproc(1), dimproc(2), dimproc(3)
  read(1, *, err=20) nproc
  read(1, *, err=20) init
  read(1, *, err=20) rand
  read(1, *, err=20) save
  read(1, *, err=20) confstart
  read(1, *, err=20) beta
  read(1, *, err=20) x
  read(1, *, err=20) y
  read(1, *, err=20) lambda
  read(1, *, err=20) alpha
  read(1, *, err=20) limit_or
  read(1, *, err=20) limit_it
  read(1, *, err=20) inorm
  read(1, *, err=20) nsweep_flag
  read(1, *, err=20) ntime
  read(1, *, err=20) nsweep
  read(1, *, err=20) imeas
  read(1, *, err=20) imode
  close(1)
  goto 30
20 write(*, *) 'break: data input ', trim(fin)
  call abort()
30 continue

  ! ---- couplings, f/paraminit.f:44-47,150-153
  pi = 2.0d0 * acos(0.0d0)
  beta_a = beta
  beta_2 = beta_a * (3 + 8 * y / beta**2 - sigma * (4 + 5 * x) / (2 * pi * beta) - &
           ((20 * x - 10 * x**2) * (log(1.5d0 * beta) + 0.09d0) + 8.7d0 + 11.6d0 * x) / &
           (2 * pi**2 * beta**2))
  beta_4 = x * beta

  call lceq_check(lceq_create(nl(1), nl(2), nl(3), h), 'lceq_create')
  call lceq_check(lceq_set_couplings(h, beta, beta_a, beta_2, beta_4), 'set_couplings')
  call lceq_check(lceq_set_gauge_fixing(h, lambda, alpha, limit_or, 10000000), 'set_gauge_fixing')
  call lceq_check(lceq_set_imode(h, imode), 'set_imode')
  if (ranlux_seed) then
    ! f/measinit.f:214-217 on rank 0
    call lceq_check(lceq_rng_measinit(h, 0, iseed), 'rng_measinit')
    seed = lceq_measinit_key(iseed)
  else
    call lceq_check(lceq_rng_philox(h, seed, 0_c_int64_t), 'rng_philox')
  end if

  ! ---- start configuration (f/confget.f:51-59)
  select case (init)
  case (0, 1)
    call lceq_check(lceq_hotstart(h, seed, 0.5d0), 'hotstart')
  case (2)
    call lceq_check(lceq_unitconf(h), 'unitconf')
  case (4)
    call lceq_check(lceq_testconf(h), 'testconf')
  case (3)
    call conf_name(u_conf_file, 'u', confstart, .false.)
    call conf_name(a_conf_file, 'a', confstart, .false.)
    if (.not. exists(u_conf_file)) call conf_name(u_conf_file, 'u', confstart, .true.)
    if (.not. exists(a_conf_file)) call conf_name(a_conf_file, 'a', confstart, .true.)
    call lceq_check(lceq_oldconf(h, trim(u_conf_file) // c_null_char, &
                                 trim(a_conf_file) // c_null_char, realbytes), 'oldconf')
  case default
    write(*, *) 'init out of range or unsupported:', init
    call abort()
  end select

  ! ---- output files, f/measinit.f:51-170,246-273
  betaint = nint(1000 * beta)
  xint = nint(10000 * x)
  yint = nint(10000 * y)
  write(action_file, 1000) 'action', nl(3), betaint, xint, yint, confstart
  write(joblog_file, 1000) 'joblog', nl(3), betaint, xint, yint, confstart
  write(corr_file, 1000) 'correl', nl(3), betaint, xint, yint, confstart
  write(afield_file, 1000) 'afield', nl(3), betaint, xint, yint, confstart
  write(checkp_file, 1000) 'checkp', nl(3), betaint, xint, yint, confstart
  write(polyak_file, 1000) 'polyak', nl(3), betaint, xint, yint, confstart
  write(corra2_file, 1000) 'corra2', nl(3), betaint, xint, yint, confstart
  call header(u_act, action_file)
  call header(u_job, joblog_file)
  open(u_job, file=joblog_file, status='old', position='append')
  if (ranlux_seed) write(u_job, '(a,16(1x,i7))') '# iseed', iseed
  write(u_job, '(a,1x,i20)') '# philox key', seed
  close(u_job)
  call header(u_cor, corr_file)
  call header(u_afi, afield_file)
  call header(u_chk, checkp_file)
  call header(u_pol, polyak_file)
  call header(u_ca2, corra2_file)
  allocate(corr_a(nl(3)), corr_u(nl(3)), corr_3(nl(3)), pcorr(nl(3)))

  ! ---- thermalisation, f/main.f:20-25
  do isweep = 1, ntherm
    call lceq_check(lceq_ahiggs2(h, 0, acc2), 'ahiggs2')
    call lceq_check(lceq_ahiggs3(h, 0, acc3), 'ahiggs3')
    call lceq_check(lceq_agauge(h, 0, acc1), 'agauge')
    call lceq_check(lceq_renorm(h), 'renorm')
  end do

  ! ---- production, f/main.f:28-83
  acco = 0
  acch = 0
  accg = 0
  meas = 0
  iter_or = 0
  da = 0
  nmeas = nsweep / imeas
  mstart = confstart / imeas
  open(u_act, file=action_file, status='old', position='append')
  do isweep = 1, nsweep
    call lceq_check(lceq_ahiggs2(h, 1, acc2), 'ahiggs2')
    call lceq_check(lceq_ahiggs3(h, 1, acc3), 'ahiggs3')
    call lceq_check(lceq_agauge(h, 1, acc1), 'agauge')
    call lceq_check(lceq_renorm(h), 'renorm')
    acco = acco + acc3
    acch = acch + acc2
    accg = accg + acc1
    ! meas1, f/meas1.f:91
    call lceq_check(lceq_observables(h, obs), 'observables')
    write(u_act, '(i5,1x,4(e23.16,1x))') isweep + confstart, obs(1), obs(2), obs(3), obs(4)
    if (mod(isweep, imeas) == 0) then
      meas = meas + 1
      ! guard: abort, as the reference does on errors, if the fields have
      ! gone non-finite or the links have left SU(2)
      call lceq_check(lceq_check_fields(h, 1d-10, maxdev, nbad), 'check_fields')
      if (corr_flag == 2) then  ! f/main.f:68-71
        call polyakov_out()
        call scalarcorr_out()
        cycle
      end if
      call lceq_check(lceq_landau_gauge(h, da1, iter1), 'landau_gauge')
      da = da + da1
      iter_or = iter_or + iter1
      ! meas2 -> correlator, f/correlator.f:166-181
      call lceq_check(lceq_correlator(h, corr_a, corr_u, corr_3), 'correlator')
      open(u_cor, file=corr_file, status='old', position='append')
      write(u_cor, '(i3.3)') mstart + meas
      do z = 1, nl(3)
        write(u_cor, '(2(e23.16,1x))') corr_a(z), corr_u(z)
      end do
      close(u_cor)
      open(u_chk, file=checkp_file, status='old', position='append')
      write(u_chk, '(i3.3)') mstart + meas
      do z = 1, nl(3)
        write(u_chk, '(e23.16)') corr_3(z)
      end do
      close(u_chk)
      ! a_f average, f/meas2a.f:39-97
      call lceq_check(lceq_observables(h, obs), 'observables')
      open(u_afi, file=afield_file, status='old', position='append')
      write(u_afi, '(i5.5,1x,3(e23.16,1x))') meas + mstart, obs(5), obs(6), obs(7)
      close(u_afi)
      if (corr_flag == 3) call polyakov_out()  ! f/main.f:73-78
      if (meas == nmeas) then
        open(u_job, file=joblog_file, status='old', position='append')
        write(u_job, '(a,1x,i20)') 'number of iteration for or. gauge fixing :', iter_or / nmeas
        write(u_job, '(a,1x,e20.10)') 'precision of the condition |d_mu A_mu|^2=0 :', da / nmeas
        close(u_job)
      end if
    end if
  end do
  close(u_act)
  ! acceptances, f/meas1.f:62-77
  open(u_job, file=joblog_file, status='old', position='append')
  write(u_job, '(a,1x,e10.5)') 'Overrelaxation acceptance :', acco / max(nsweep, 1)
  write(u_job, '(a,1x,e10.5)') 'heat-bath acceptance : ', acch / max(nsweep, 1)
  write(u_job, '(a,1x,e10.5)') 'Gauge  acceptance    : ', accg / max(nsweep, 1)
  close(u_job)
  ! ---- f/finalize.f:43, f/confsave.f:249-253
  call conf_name(u_conf_file, 'u', confstart + nsweep, .true.)
  call conf_name(a_conf_file, 'a', confstart + nsweep, .true.)
  call lceq_check(lceq_confsave(h, trim(u_conf_file) // c_null_char, &
                                trim(a_conf_file) // c_null_char, realbytes), 'confsave')
  call lceq_check(lceq_destroy(h), 'lceq_destroy')

1000 format(a6, '_s', i2.2, '_be', i5.5, '_x', i5.5, '_y', i5.5, '_u', i6.6)

contains

  ! f/polyakov.f:110-117: record number, mean(P)^2, the correlator
  subroutine polyakov_out()
    call lceq_check(lceq_polyakov(h, pmean, pcorr), 'polyakov')
    open(u_pol, file=polyak_file, status='old', position='append')
    write(u_pol, '(i5.5)') mstart + meas
    write(u_pol, '(e23.16)') pmean
    do z = 1, nl(3)
      write(u_pol, '(e23.16)') pcorr(z)
    end do
    close(u_pol)
  end subroutine

  ! f/scalarcorr.f:106-113: record number, mean(rho), the correlator
  subroutine scalarcorr_out()
    call lceq_check(lceq_scalarcorr(h, pmean, pcorr), 'scalarcorr')
    open(u_ca2, file=corra2_file, status='old', position='append')
    write(u_ca2, '(i5.5)') mstart + meas
    write(u_ca2, '(e23.16)') pmean
    do z = 1, nl(3)
      write(u_ca2, '(e23.16)') pcorr(z)
    end do
    close(u_ca2)
  end subroutine

  ! configuration file names: f/confsave.f:249-253 (saved) and
  ! f/measinit.f:246-250 (the case oldconf opens)
  subroutine conf_name(name, field, sweep, saved)
    character(len=*), intent(out) :: name
    character(len=1), intent(in) :: field
    integer, intent(in) :: sweep
    logical, intent(in) :: saved
    if (saved) then
      write(name, 1010) field, nl(3), nint(1000 * beta), nint(10000 * x), nint(10000 * y), sweep
    else
      write(name, 1020) field, nl(3), nint(1000 * beta), nint(10000 * x), nint(10000 * y), sweep
    end if
1010 format(a1, '_conf_s', i2.2, '_BE', i5.5, '_x', i5.5, '_y', i5.5, '_U', i6.6)
1020 format(a1, '_conf_s', i2.2, '_be', i5.5, '_x', i5.5, '_y', i5.5, '_u', i6.6)
  end subroutine

  logical function exists(name)
    character(len=*), intent(in) :: name
    inquire(file=trim(name), exist=exists)
  end function

  ! f/measinit.f:64-72 header lines (formats 1001-1003)
  subroutine header(unit, name)
    integer, intent(in) :: unit
    character(len=*), intent(in) :: name
    open(unit, file=name, status='unknown', position='append')
    write(unit, '(a,3(i2.2,1x))') '# ', nl(1), nl(2), nl(3)
    write(unit, '(a,3(f7.4,1x))') '# ', beta, x, y
    write(unit, '(a,i5.5)') '# ', nsweep
    close(unit)
  end subroutine
end program lceq_main
