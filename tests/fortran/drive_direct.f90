! drive_direct.f90 -- calls the drop-in DIRECT exactly as the reference's time loop does
! (src/Eddy6.f:2722-2723: DP, two scratch arrays, DELYSQ, DELZSQ, the metric arrays, the
! dimensions, the MPI arguments, IPRES and the BC flags).
!
! Usage: drive_direct <input.bin> <output.bin>
! input (stream, native endian): int32 nx, ny, nz, mp, lp, np, ipres; float64 delysq, delzsq;
! float64 ap(nx), au(nx), rp(nx), ru(nx), cp(nz), cw(nz), dp(nx*ny*nz).
! output: float64 dp(nx*ny*nz) after the call.
program drive_direct
  implicit none
  integer :: nx, ny, nz, mp, lp, np, ipres, mysize, myrank, comm, mtype, nzg
  real(8) :: delysq, delzsq
  real(8), allocatable :: ap(:), au(:), rp(:), ru(:), cp(:), cw(:)
  real(8), allocatable :: dp(:,:,:), ub(:), vb(:)
  character(len=1024) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nx, ny, nz, mp, lp, np, ipres
  read(u) delysq, delzsq
  allocate(ap(nx), au(nx), rp(nx), ru(nx), cp(nz), cw(nz), dp(nx, ny, nz))
  allocate(ub(nx*ny*nz), vb(nx*ny*nz))
  read(u) ap, au, rp, ru, cp, cw, dp
  close(u)
  mysize = 1
  myrank = 0
  comm = 0
  mtype = 0
  nzg = nz
  ub = 0.0d0
  vb = 0.0d0
  call DIRECT(dp, ub, vb, delysq, delzsq, ap, au, cp, cw, rp, ru, &
              nx, ny, nz, nzg, mysize, myrank, comm, mtype, ipres, mp, lp, np)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) dp
  close(u)
end program drive_direct
