! direct_hip.f90 -- drop-in SUBROUTINE DIRECT for the reference's projection path.
!
! Same name and 23-argument list as the reference's SUBROUTINE DIRECT (src/direct.f:2-3,
! declarations :11-23), called unchanged from src/Eddy6.f:2722-2723. Link this object in
! place of direct.o; the solve runs on the GPU through the C-ABI entry point eddy_direct
! (include/eddyproj.h). REAL is 8 bytes in the reference (-r8, Makefile:49), so the shim
! declares real(c_double) explicitly; INTEGER is 4 bytes.
!
! Errors: the reference prints FISHPACK IERROR to unit 6 and continues (direct.f:545, 872);
! this shim prints the library status and message the same way and continues, unless the
! environment variable EDDY_DIRECT_STRICT is set, in which case it stops.
subroutine DIRECT(DP, DUMMY, DUMMYT, DELYSQ, DELZSQ, AP, AU, CP, CW, RP, RU, &
                  NX, NY, NZ, NZG, MYSIZE, MYRANK, MPI_COMM_EDDY, MTYPE, IPRES, MP, LP, NP)
  use, intrinsic :: iso_c_binding, only: c_int, c_double, c_char, c_ptr, c_f_pointer, c_null_char
  implicit none
  integer(c_int) :: NX, NY, NZ, NZG, MYSIZE, MYRANK, MPI_COMM_EDDY, MTYPE, IPRES, MP, LP, NP
  real(c_double) :: DP(NX, NY, NZ), DUMMY(*), DUMMYT(*)
  real(c_double) :: DELYSQ, DELZSQ
  real(c_double) :: AP(*), AU(*), CP(*), CW(*), RP(*), RU(*)

  interface
    integer(c_int) function eddy_direct(dp, dummy, dummyt, delysq, delzsq, ap, au, cp, cw, rp, ru, &
                                        nx, ny, nz, nzg, mysize, myrank, comm, mtype, ipres, mp, lp, np) &
                                        bind(C, name="eddy_direct")
      import :: c_int, c_double
      real(c_double) :: dp(*), dummy(*), dummyt(*)
      real(c_double) :: delysq, delzsq
      real(c_double) :: ap(*), au(*), cp(*), cw(*), rp(*), ru(*)
      integer(c_int) :: nx, ny, nz, nzg, mysize, myrank, comm, mtype, ipres, mp, lp, np
    end function eddy_direct
    function eddy_last_error() bind(C, name="eddy_last_error")
      import :: c_ptr
      type(c_ptr) :: eddy_last_error
    end function eddy_last_error
  end interface

  integer(c_int) :: ierr
  character(len=256) :: strict
  integer :: slen, sstat

  ierr = eddy_direct(DP, DUMMY, DUMMYT, DELYSQ, DELZSQ, AP, AU, CP, CW, RP, RU, &
                     NX, NY, NZ, NZG, MYSIZE, MYRANK, MPI_COMM_EDDY, MTYPE, IPRES, MP, LP, NP)
  if (ierr /= 0) then
    write(6, '(A,I0,A,A)') ' DIRECT (HIP): IERROR = ', ierr, '  ', trim(c_message(eddy_last_error()))
    call get_environment_variable('EDDY_DIRECT_STRICT', strict, slen, sstat)
    if (sstat == 0 .and. slen > 0) stop 'DIRECT (HIP) failed'
  end if

contains

  function c_message(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=200) :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: i
    s = ' '
    call c_f_pointer(p, chars, [200])
    do i = 1, 200
      if (chars(i) == c_null_char) exit
      s(i:i) = chars(i)
    end do
  end function c_message

end subroutine DIRECT
