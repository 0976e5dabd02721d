! cpuref_driver.f90 -- TEST / MEASUREMENT INFRASTRUCTURE ONLY.
!
! Calls the reference's own SUBROUTINE DIRECT (src/direct.f:2-3, compiled from its sources by
! `make -C oracle cpuref` into oracle/_ref/) the way src/Eddy6.f:2722-2723 calls it: one z slab
! of DP per MPI rank (NZ = (NZG-2)/MYSIZE + 2, inpall.f:514), global x2 metrics CPG/CWG, the
! caller's scratch arrays for DUMMY/DUMMYT, MTYPE = MPI_DOUBLE_PRECISION.
!
! Usage: mpiexec -np P cpuref_driver <case.bin> [<out.bin>]
!   case.bin (stream, little-endian): int32 M, L, N, mode, reps, ipres, mp, lp, np;
!     real64 DELYSQ, DELZSQ; AP, AU, RP, RU (NX each); CPG, CWG (NZG each);
!     mode 0: the RHS as a full DP(NX,NY,NZG) array follows (interior read);
!     mode 1: a deterministic RHS is formed here (a smooth pattern minus its volume-weighted
!     mean, so that the Neumann problem is solvable).
!   out.bin: the assembled DP(NX,NY,NZG) after the first call (z ghost planes 0).
! Prints one line per rank-0 run: plan+first call seconds and the per-call wall seconds of
! `reps` further calls (the slowest rank of each call), timed with MPI_WTIME around DIRECT
! alone (TCLOCK of info.f:350-362 is MPI_WTIME too).
program cpuref_driver
  implicit none
  include 'mpif.h'
  integer :: ierr, rank, nprocs, m, l, n, mode, reps, ipres, ipres_arg, mp, lp, np
  integer :: nx, ny, nz, nzg, nl, i, j, k, kg, r, u, mtype, comm
  real(8) :: delysq, delzsq, t0, t1, dt, dtmax, sw, swb, wsum(2), gsum(2)
  real(8), allocatable :: ap(:), au(:), rp(:), ru(:), cpg(:), cwg(:)
  real(8), allocatable :: dp(:, :, :), rhs(:, :, :), dummy(:), dummyt(:), glob(:, :, :), buf(:)
  real(8), allocatable :: times(:)
  character(len=512) :: fin, fout
  integer(8) :: off

  call MPI_INIT(ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, nprocs, ierr)
  comm = MPI_COMM_WORLD
  mtype = MPI_DOUBLE_PRECISION
  call get_command_argument(1, fin)
  fout = ''
  if (command_argument_count() >= 2) call get_command_argument(2, fout)

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) m, l, n, mode, reps, ipres_arg, mp, lp, np
  nx = m + 2
  ny = l + 2
  nzg = n + 2
  if (mod(n, nprocs) /= 0 .or. mod(l, nprocs) /= 0) then
    if (rank == 0) write(*, '(a)') 'CPUREF error: N and L must be multiples of the rank count'
    call MPI_ABORT(MPI_COMM_WORLD, 1, ierr)
  end if
  nl = n / nprocs
  nz = nl + 2
  allocate(ap(nx), au(nx), rp(nx), ru(nx), cpg(nzg), cwg(nzg))
  read(u) delysq, delzsq
  read(u) ap, au, rp, ru, cpg, cwg
  allocate(dp(nx, ny, nz), rhs(nx, ny, nz), dummy(nx * ny * nz), dummyt(nx * ny * nz))
  rhs = 0.0d0
  if (mode == 0) then
    ! this rank's planes K = 1..NZ are global planes rank*NL + K
    allocate(buf(nx * ny))
    do k = 1, nz
      kg = rank * nl + k
      off = 1_8 + 36_8 + 16_8 + 8_8 * (4_8 * nx + 2_8 * nzg) + 8_8 * int(nx, 8) * int(ny, 8) * int(kg - 1, 8)
      read(u, pos=off) buf
      rhs(:, :, k) = reshape(buf, (/nx, ny/))
    end do
    deallocate(buf)
  else
    wsum = 0.0d0
    do k = 2, nl + 1
      kg = rank * nl + k
      do j = 2, l + 1
        do i = 2, m + 1
          rhs(i, j, k) = sin(0.37d0 * i + 0.11d0 * j) * cos(0.23d0 * kg) + 0.01d0 * mod(7 * i + 13 * j + 17 * kg, 11)
          sw = rp(i) / (ap(i) * cpg(kg))
          wsum(1) = wsum(1) + sw * rhs(i, j, k)
          wsum(2) = wsum(2) + sw
        end do
      end do
    end do
    call MPI_ALLREDUCE(wsum, gsum, 2, MPI_DOUBLE_PRECISION, MPI_SUM, MPI_COMM_WORLD, ierr)
    swb = gsum(1) / gsum(2)
    rhs(2:m + 1, 2:l + 1, 2:nl + 1) = rhs(2:m + 1, 2:l + 1, 2:nl + 1) - swb
  end if
  close(u)

  ! first call: DIRECT builds and SAVEs its plan (direct.f:121-585), then solves
  dp = rhs
  ipres = ipres_arg
  call MPI_BARRIER(MPI_COMM_WORLD, ierr)
  t0 = MPI_WTIME()
  call DIRECT(dp, dummy, dummyt, delysq, delzsq, ap, au, cpg, cwg, rp, ru, nx, ny, nz, nzg, nprocs, rank, &
              comm, mtype, ipres, mp, lp, np)
  t1 = MPI_WTIME() - t0
  call MPI_REDUCE(t1, dtmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, 0, MPI_COMM_WORLD, ierr)
  if (rank == 0) write(*, '(a,i0,a,es12.5)') 'CPUREF ranks=', nprocs, ' first_s=', dtmax

  if (len_trim(fout) > 0) then
    ! gather the interior planes to rank 0 (equal counts per rank) and write DP(NX,NY,NZG)
    allocate(buf(nx * ny * nl))
    buf = reshape(dp(:, :, 2:nl + 1), (/nx * ny * nl/))
    if (rank == 0) then
      allocate(glob(nx, ny, nzg))
      glob = 0.0d0
    else
      allocate(glob(1, 1, 1))
    end if
    block
      real(8), allocatable :: all(:)
      if (rank == 0) then
        allocate(all(nx * ny * nl * nprocs))
      else
        allocate(all(1))
      end if
      call MPI_GATHER(buf, nx * ny * nl, MPI_DOUBLE_PRECISION, all, nx * ny * nl, MPI_DOUBLE_PRECISION, 0, &
                      MPI_COMM_WORLD, ierr)
      if (rank == 0) then
        glob(:, :, 2:n + 1) = reshape(all, (/nx, ny, n/))
        open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
        write(u) glob
        close(u)
      end if
    end block
    deallocate(buf)
  end if

  allocate(times(max(reps, 1)))
  do r = 1, reps
    dp = rhs
    ipres = ipres_arg
    call MPI_BARRIER(MPI_COMM_WORLD, ierr)
    t0 = MPI_WTIME()
    call DIRECT(dp, dummy, dummyt, delysq, delzsq, ap, au, cpg, cwg, rp, ru, nx, ny, nz, nzg, nprocs, rank, &
                comm, mtype, ipres, mp, lp, np)
    dt = MPI_WTIME() - t0
    call MPI_REDUCE(dt, dtmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, 0, MPI_COMM_WORLD, ierr)
    times(r) = dtmax
  end do
  if (rank == 0 .and. reps > 0) then
    write(*, '(a,i0,a)', advance='no') 'CPUREF ranks=', nprocs, ' call_s='
    do r = 1, reps
      write(*, '(es12.5,1x)', advance='no') times(r)
    end do
    write(*, *)
  end if
  call MPI_FINALIZE(ierr)
end program cpuref_driver
