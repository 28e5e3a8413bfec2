*     refharness.f -- test-infrastructure driver for the reference's own
*     hot-path subroutines (compiled from /root/reference/f by
*     oracle/ref/Makefile into oracle/_ref/).  It fills the reference
*     COMMON blocks for a single MPI rank (dimproc = 1,1,1), runs one
*     command and writes the results.  Used only to produce golden
*     vectors under tests/golden and to check the C restatement.
*
*     Uniform capture: the reference draws from the compiler intrinsic
*     random_number.  The harness seeds it, draws the exact number of
*     uniforms the command will consume (plus one), re-seeds with the
*     same seed and runs the command, which therefore consumes exactly
*     the captured stream.  The extra draw after the run must equal the
*     captured extra uniform; that proves the consumption count.
*
*     Command 16 times nrep sweeps of the reference's production schedule
*     (f/main.f:37-41) with the compiler's system_clock, without the
*     uniform capture, and writes the elapsed seconds and the last plaq0.
*
*     Commands 14 (confsave) and 15 (oldconf) run the reference's
*     configuration I/O in the working directory on a 1x1x1 Cartesian
*     communicator: 14 writes the confsave files for beta, x = lambda,
*     y = alpha and sweep number arg1; 15 reads the files u_in and a_in
*     and writes the fields.
*
*     Input  (stream): int32 lx,ly,lz,cmd,arg1,arg2,imode,nrep,seed
*                      real64 beta,beta_a,beta_2,beta_4,lambda,alpha,
*                             limit_or,delta
*                      real64 u(4,N,0:1,3), a(3,N,0:1), extra(N)*2
*     Output (stream): records of (int32 n, real64 x(n)).
      program refharness
      implicit none
      INCLUDE 'mpif.h'
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'parameter.inc'
      INCLUDE 'input_parameter.inc'
      INCLUDE 'pointer.inc'
      INCLUDE 'gaugefield.inc'
      INCLUDE 'measurement.inc'
      INCLUDE 'parallel_mpi_types.inc'
      INCLUDE 'filenames.inc'

      integer(4) hdr(9)
      real(8) par(8)
      real(8), allocatable :: ub(:), ab(:), ex1(:), ex2(:), uni(:)
      real(8), allocatable :: outb(:)
      real, allocatable :: bb(:), a0o(:), rr(:), xx(:)
      complex, allocatable :: vv(:,:)
      integer lx,ly,lz,cmd,arg1,arg2,nrep,iseedv,n,i,j,eo,mu,k,rep
      integer ns, nacch, iter1, nuni
      integer(8) clk0, clk1, clkrate
      integer, allocatable :: sd(:)
      real acc, delta, da1, rnext
      character(len=512) fin, fout
      integer dims3(3), ctype, rtype
      logical per3(3)

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      call MPI_INIT(ierror)
      comm = MPI_COMM_WORLD
      nproc = 1
      myrank = 0

      open(10, file=fin, access='stream', form='unformatted',
     &     status='old')
      read(10) hdr
      read(10) par
      lx = hdr(1)
      ly = hdr(2)
      lz = hdr(3)
      cmd = hdr(4)
      arg1 = hdr(5)
      arg2 = hdr(6)
      imode = hdr(7)
      nrep = hdr(8)
      iseedv = hdr(9)
      beta = par(1)
      beta_a = par(2)
      beta_2 = par(3)
      beta_4 = par(4)
      lambda = par(5)
      alpha = par(6)
      limit_or = par(7)
      delta = par(8)

*     single-rank geometry, as f/topology.f:87-96,179-180 set it up
      n_lattice(1) = lx
      n_lattice(2) = ly
      n_lattice(3) = lz
      do mu = 1, 3
        dimproc(mu) = 1
        extension(mu) = n_lattice(mu)
        lowbound(mu) = 1
        upbound(mu) = extension(mu)
      enddo
      ownvol(1) = extension(1)
      ownhalfvol(1) = extension(1)/2
      ownvol(2) = extension(2)*ownvol(1)
      ownhalfvol(2) = ownvol(2)/2
      ownvol(3) = extension(3)*ownvol(2)
      ownhalfvol(3) = ownvol(3)/2
      n = ownhalfvol(3)
      if (n + ownhalfvol(2) + 4 .gt. maxarray) then
        write(*,*) 'lattice too large for parameter.inc'
        call abort()
      endif
      do i = 1, maxarray
        normalsites(i) = i
      enddo
      call adresses

      allocate(ub(4*n*2*3), ab(3*n*2), ex1(n), ex2(n))
      read(10) ub
      read(10) ab
      read(10) ex1
      read(10) ex2
      close(10)
      k = 0
      do mu = 1, 3
        do eo = 0, 1
          do i = 1, n
            u(1,i,eo,mu) = cmplx(ub(k+1), ub(k+2))
            u(2,i,eo,mu) = cmplx(ub(k+3), ub(k+4))
            k = k + 4
          enddo
        enddo
      enddo
      k = 0
      do eo = 0, 1
        do i = 1, n
          a(1,i,eo) = ab(k+1)
          a(2,i,eo) = ab(k+2)
          a(3,i,eo) = ab(k+3)
          k = k + 3
        enddo
      enddo

*     uniforms each command consumes (reference call order)
      nuni = 0
      if (cmd .eq. 3) nuni = 42*n*nrep
      if (cmd .eq. 5) nuni = 8*n*nrep
      if (cmd .eq. 6) nuni = 14*n*nrep
      if (cmd .eq. 7) nuni = 2*n*nrep
      if (cmd .eq. 11) nuni = 58*n*nrep
      if (cmd .eq. 12) nuni = 6*n
      if (cmd .eq. 13) nuni = 2*n
      call random_seed(size=ns)
      allocate(sd(ns))
      sd = iseedv
      call random_seed(put=sd)
      allocate(uni(nuni+1))
      do i = 1, nuni+1
        call random_number(rnext)
        uni(i) = rnext
      enddo
      call random_seed(put=sd)

      open(11, file=fout, access='stream', form='unformatted',
     &     status='replace')

      if (cmd .eq. 1) then
        allocate(outb(6*n))
        do k = 1, 2
          j = 0
          do eo = 0, 1
            do mu = 1, 3
              do i = 1, n
                j = j + 1
                if (k .eq. 1) outb(j) = siteup(i,eo,mu)
                if (k .eq. 2) outb(j) = sitedn(i,eo,mu)
              enddo
            enddo
          enddo
          call wr(11, outb, 6*n)
        enddo
      else if (cmd .eq. 2) then
        call staples(arg1, 1-arg1, arg2)
        call wrv(11, n)
      else if (cmd .eq. 3) then
        do rep = 1, nrep
          call agauge(arg1, acc)
          call wrs(11, acc)
          call wrs(11, plaq0)
        enddo
        call wru(11, n)
      else if (cmd .eq. 4) then
        call a_force(arg1)
        allocate(outb(3*n))
        do i = 1, n
          do j = 1, 3
            outb(3*(i-1)+j) = force(j,i)
          enddo
        enddo
        call wr(11, outb, 3*n)
      else if (cmd .ge. 5 .and. cmd .le. 7) then
        do rep = 1, nrep
          if (cmd .eq. 5) call ahiggs1(delta, arg1, acc)
          if (cmd .eq. 6) call ahiggs2(arg1, acc)
          if (cmd .eq. 7) call ahiggs3(arg1, acc)
          call wrs(11, acc)
          call wrs(11, delta)
          call wrs(11, a02)
          call wrs(11, a04)
          call wrs(11, a0kin)
        enddo
        call wra(11, n)
      else if (cmd .eq. 8) then
        call renorm
        call wru(11, n)
      else if (cmd .eq. 9) then
*       nrep iterations of the landau_gauge loop body (f/landau_gauge.f:54-59)
        do i = 1, n
          do eo = 0, 1
            do mu = 1, 3
              u_f(1,i,eo,mu) = u(1,i,eo,mu)
              u_f(2,i,eo,mu) = u(2,i,eo,mu)
            enddo
            a_f(1,i,eo) = a(1,i,eo)
            a_f(2,i,eo) = a(2,i,eo)
            a_f(3,i,eo) = a(3,i,eo)
          enddo
        enddo
        do rep = 1, nrep
          call or_gauge_fixing(da1)
          call u_f_renorm
          call wrs(11, da1)
        enddo
        call wruf(11, n)
      else if (cmd .eq. 10) then
        call landau_gauge(da1, iter1)
        call wrs(11, da1)
        rnext = iter1
        call wrs(11, rnext)
        call wruf(11, n)
      else if (cmd .eq. 11) then
*       reference production sweep, f/main.f:37-41
        do rep = 1, nrep
          call ahiggs2(1, acc)
          call wrs(11, acc)
          call ahiggs3(1, acc)
          call wrs(11, acc)
          call agauge(1, acc)
          call wrs(11, acc)
          call renorm
          call wrs(11, plaq0)
          call wrs(11, a0kin)
          call wrs(11, a02)
          call wrs(11, a04)
        enddo
        call wru(11, n)
        call wra(11, n)
      else if (cmd .eq. 12) then
        allocate(bb(n), a0o(n), vv(2,n), outb(4*n))
        do i = 1, n
          bb(i) = ex1(i)
          a0o(i) = ex2(i)
        enddo
        call heat_kpi(bb, vv, a0o, nacch)
        do i = 1, n
          outb(4*i-3) = real(vv(1,i))
          outb(4*i-2) = aimag(vv(1,i))
          outb(4*i-1) = real(vv(2,i))
          outb(4*i) = aimag(vv(2,i))
        enddo
        call wr(11, outb, 4*n)
        rnext = nacch
        call wrs(11, rnext)
      else if (cmd .eq. 14 .or. cmd .eq. 15) then
        dims3 = 1
        per3 = .true.
        call MPI_CART_CREATE(MPI_COMM_WORLD, 3, dims3, per3, .false.,
     &                       comm, ierror)
        if (storage_size(u(1,1,0,1)) .eq. 128) then
          ctype = MPI_DOUBLE_COMPLEX
          rtype = MPI_DOUBLE_PRECISION
        else
          ctype = MPI_COMPLEX
          rtype = MPI_REAL
        endif
        call MPI_TYPE_CONTIGUOUS(nc*ownhalfvol(2), ctype, load_u_type,
     &                           ierror)
        call MPI_TYPE_COMMIT(load_u_type, ierror)
        save_u_type = load_u_type
        call MPI_TYPE_CONTIGUOUS(3*ownhalfvol(2), rtype, load_a_type,
     &                           ierror)
        call MPI_TYPE_COMMIT(load_a_type, ierror)
        save_a_type = load_a_type
        u_unit = 20
        a_unit = 21
        if (cmd .eq. 14) then
          x = lambda
          y = alpha
          confstart = arg1
          nsweep = 0
          call confsave
        else
          u_conf_file = 'u_in'
          a_conf_file = 'a_in'
          call oldconf
          call wru(11, n)
          call wra(11, n)
        endif
      else if (cmd .eq. 16) then
        call system_clock(clk0, clkrate)
        do rep = 1, nrep
          call ahiggs2(1, acc)
          call ahiggs3(1, acc)
          call agauge(1, acc)
          call renorm
        enddo
        call system_clock(clk1)
        rnext = real(clk1 - clk0) / real(clkrate)
        call wrs(11, rnext)
        call wrs(11, plaq0)
      else if (cmd .eq. 13) then
        allocate(rr(n), xx(n), outb(n))
        do i = 1, n
          rr(i) = ex1(i)
        enddo
        call gauss(xx, rr, n)
        do i = 1, n
          outb(i) = xx(i)
        enddo
        call wr(11, outb, n)
      endif

*     consumption check and the captured stream
      call random_number(rnext)
      if (nuni .gt. 0) then
        if (rnext .eq. uni(nuni+1)) then
          rnext = 1.0
        else
          rnext = 0.0
        endif
        call wrs(11, rnext)
        call wr(11, uni, nuni)
      endif
      close(11)
      call MPI_FINALIZE(ierror)
      end

      subroutine wr(iu, x, m)
      implicit none
      integer iu, m
      real(8) x(m)
      integer(4) m4
      m4 = m
      write(iu) m4
      write(iu) x
      end

      subroutine wrs(iu, s)
      implicit none
      integer iu
      real s
      real(8) x(1)
      x(1) = s
      call wr(iu, x, 1)
      end

      subroutine wru(iu, n)
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer iu, n, i, eo, mu, k
      real(8) ob(24*n)
      k = 0
      do mu = 1, 3
        do eo = 0, 1
          do i = 1, n
            ob(k+1) = real(u(1,i,eo,mu))
            ob(k+2) = aimag(u(1,i,eo,mu))
            ob(k+3) = real(u(2,i,eo,mu))
            ob(k+4) = aimag(u(2,i,eo,mu))
            k = k + 4
          enddo
        enddo
      enddo
      call wr(iu, ob, 24*n)
      end

      subroutine wruf(iu, n)
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer iu, n, i, eo, mu, k
      real(8) ob(24*n), oa(6*n)
      k = 0
      do mu = 1, 3
        do eo = 0, 1
          do i = 1, n
            ob(k+1) = real(u_f(1,i,eo,mu))
            ob(k+2) = aimag(u_f(1,i,eo,mu))
            ob(k+3) = real(u_f(2,i,eo,mu))
            ob(k+4) = aimag(u_f(2,i,eo,mu))
            k = k + 4
          enddo
        enddo
      enddo
      call wr(iu, ob, 24*n)
      k = 0
      do eo = 0, 1
        do i = 1, n
          oa(k+1) = a_f(1,i,eo)
          oa(k+2) = a_f(2,i,eo)
          oa(k+3) = a_f(3,i,eo)
          k = k + 3
        enddo
      enddo
      call wr(iu, oa, 6*n)
      end

      subroutine wra(iu, n)
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer iu, n, i, eo, k
      real(8) oa(6*n)
      k = 0
      do eo = 0, 1
        do i = 1, n
          oa(k+1) = a(1,i,eo)
          oa(k+2) = a(2,i,eo)
          oa(k+3) = a(3,i,eo)
          k = k + 3
        enddo
      enddo
      call wr(iu, oa, 6*n)
      end

      subroutine wrv(iu, n)
      implicit none
      INCLUDE 'parameter.inc'
      INCLUDE 'gaugefield.inc'
      integer iu, n, i, eo, mu, k
      real(8) ob(24*n)
      k = 0
      do mu = 1, 3
        do eo = 0, 1
          do i = 1, n
            ob(k+1) = real(v(1,i,eo,mu))
            ob(k+2) = aimag(v(1,i,eo,mu))
            ob(k+3) = real(v(2,i,eo,mu))
            ob(k+4) = aimag(v(2,i,eo,mu))
            k = k + 4
          enddo
        enddo
      enddo
      call wr(iu, ob, 24*n)
      end

*     Single-rank self messages: confsave and oldconf send each slice to
*     root with a blocking MPI_SEND before root posts the receive, and
*     this MPI does not complete a blocking send to self that way.  The
*     harness therefore defines MPI_SEND as a nonblocking send whose
*     request is freed; the MPI_RECV that follows in the same routine
*     completes it before the buffer is touched again.
      subroutine mpi_send(buf, cnt, dtype, dest, tag, comm0, ierr)
      implicit none
      integer buf(*), cnt, dtype, dest, tag, comm0, ierr, req
      call pmpi_isend(buf, cnt, dtype, dest, tag, comm0, req, ierr)
      call pmpi_request_free(req, ierr)
      end
