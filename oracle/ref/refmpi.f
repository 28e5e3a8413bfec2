*     refmpi.f -- times the reference's own sweep under MPI, with its
*     halo exchanges, on a Cartesian rank grid (test infrastructure: the
*     CPU baseline beside the GPU; built by oracle/ref/Makefile into
*     oracle/_ref/ from the reference sources under /root/reference/f).
*
*     Usage: mpiexec -n P refmpi lx ly lz px py pz nwarm nsweep
*       P = px*py*pz ranks, each holding an (lx/px) x (ly/py) x (lz/pz)
*       block.  The rank grid and the COMMON geometry are set up as
*       f/topology.f:87-180 sets them (periodic MPI_CART_CREATE with
*       reorder, neighbours from MPI_CART_SHIFT, local bounds from the
*       Cartesian coordinates); f/initialize.f:41-47 then calls
*       typeconstruct and adresses, which are called here in the same
*       order.  Fields start from a per-rank random configuration (links
*       normalised 4-vectors, Higgs components uniform in [-1,1)), the
*       compiler's random_number seeded per rank.
*     Then nwarm untimed and nsweep timed sweeps of the schedule of
*     f/main.f:37-41 (ahiggs2, ahiggs3, agauge, renorm with measflag 1),
*     between barriers.  Rank 0 prints one line:
*       refmpi P px py pz lx ly lz nsweep seconds_per_sweep plaq0
*     with the slowest rank's time (MPI_MAX).
      program refmpi
      implicit none
      INCLUDE 'mpif.h'
      INCLUDE 'parallel_parameter.inc'
      INCLUDE 'parameter.inc'
      INCLUDE 'input_parameter.inc'
      INCLUDE 'pointer.inc'
      INCLUDE 'gaugefield.inc'
      INCLUDE 'measurement.inc'

      integer nargs, iv(8), k, mu, eo, i, ns, rep, nwarm, nsw, nrank
      character(len=32) arg
      logical periods(3), reorder
      real acc, q(4), qn
      real(8) t0, t1, dt, dtmax
      integer, allocatable :: sd(:)

      call MPI_INIT(ierror)
      call MPI_COMM_SIZE(MPI_COMM_WORLD, nrank, ierror)
      nargs = command_argument_count()
      if (nargs .lt. 8) then
        write(*,*) 'usage: refmpi lx ly lz px py pz nwarm nsweep'
        call MPI_ABORT(MPI_COMM_WORLD, 1, ierror)
      endif
      do k = 1, 8
        call get_command_argument(k, arg)
        read(arg, *) iv(k)
      enddo
      nwarm = iv(7)
      nsw = iv(8)

*     couplings of the production configuration (trace convention
*     mapped onto the reference action, as the harness takes them)
      beta = 9.0
      beta_a = 0.175
      beta_2 = -0.45
      beta_4 = 0.025
      imode = 0

*     the rank grid and local geometry (f/topology.f:87-180)
      nproc = iv(4)*iv(5)*iv(6)
      if (nproc .ne. nrank) then
        write(*,*) 'px*py*pz must equal the number of ranks'
        call MPI_ABORT(MPI_COMM_WORLD, 1, ierror)
      endif
      do mu = 1, 3
        n_lattice(mu) = iv(mu)
        dimproc(mu) = iv(3+mu)
        extension(mu) = n_lattice(mu)/dimproc(mu)
        if (extension(mu)*dimproc(mu) .ne. n_lattice(mu) .or.
     &      extension(mu) .lt. 2 .or. mod(extension(mu),2) .ne. 0) then
          write(*,*) 'each local extent must be even and >= 2'
          call MPI_ABORT(MPI_COMM_WORLD, 1, ierror)
        endif
        periods(mu) = .true.
      enddo
      if (extension(1) .gt. nx .or. extension(2) .gt. ny .or.
     &    extension(3) .gt. nz) then
        write(*,*) 'local block larger than parameter.inc allows'
        call MPI_ABORT(MPI_COMM_WORLD, 1, ierror)
      endif
      ownvol(1) = extension(1)
      ownhalfvol(1) = extension(1)/2
      ownvol(2) = extension(2)*ownvol(1)
      ownhalfvol(2) = ownvol(2)/2
      ownvol(3) = extension(3)*ownvol(2)
      ownhalfvol(3) = ownvol(3)/2
      reorder = .true.
      call MPI_CART_CREATE(MPI_COMM_WORLD, 3, dimproc, periods,
     &                     reorder, comm, ierror)
      call MPI_COMM_RANK(comm, myrank, ierror)
      do mu = 1, 3
        call MPI_CART_SHIFT(comm, mu-1, 1, neigh(-mu), neigh(mu),
     &                      ierror)
      enddo
      neigh(0) = root
      call MPI_CART_COORDS(comm, myrank, 3, mycoords, ierror)
      do mu = 1, 3
        lowbound(mu) = mycoords(mu)*extension(mu) + 1
        upbound(mu) = lowbound(mu) + extension(mu) - 1
      enddo

      call typeconstruct
      do i = 1, maxarray
        normalsites(i) = i
      enddo
      call adresses

*     per-rank random start
      call random_seed(size=ns)
      allocate(sd(ns))
      sd = 12345 + 1009*myrank
      call random_seed(put=sd)
      do mu = 1, 3
        do eo = 0, 1
          do i = 1, ownhalfvol(3)
            call random_number(q)
            q = 2.0*q - 1.0
            qn = 1.0/sqrt(q(1)**2 + q(2)**2 + q(3)**2 + q(4)**2)
            u(1,i,eo,mu) = cmplx(q(1)*qn, q(2)*qn)
            u(2,i,eo,mu) = cmplx(q(3)*qn, q(4)*qn)
          enddo
        enddo
      enddo
      do eo = 0, 1
        do i = 1, ownhalfvol(3)
          call random_number(q(1:3))
          a(1,i,eo) = 2.0*q(1) - 1.0
          a(2,i,eo) = 2.0*q(2) - 1.0
          a(3,i,eo) = 2.0*q(3) - 1.0
        enddo
      enddo

      do rep = 1, nwarm
        call ahiggs2(1, acc)
        call ahiggs3(1, acc)
        call agauge(1, acc)
        call renorm
      enddo
      call MPI_BARRIER(comm, ierror)
      t0 = MPI_WTIME()
      do rep = 1, nsw
        call ahiggs2(1, acc)
        call ahiggs3(1, acc)
        call agauge(1, acc)
        call renorm
      enddo
      call MPI_BARRIER(comm, ierror)
      t1 = MPI_WTIME()
      dt = (t1 - t0)/dble(nsw)
      call MPI_REDUCE(dt, dtmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX,
     &                root, comm, ierror)
      if (myrank .eq. root) then
        write(*,'(a,7(1x,i4),1x,i6,1x,es14.6,1x,es22.14)') 'refmpi',
     &    nproc, dimproc(1), dimproc(2), dimproc(3),
     &    n_lattice(1), n_lattice(2), n_lattice(3), nsw, dtmax,
     &    dble(plaq0)
      endif
      call MPI_FINALIZE(ierror)
      end
