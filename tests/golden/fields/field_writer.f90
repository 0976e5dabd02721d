! Writes the field-file fixtures of tests/test_fieldio.py with a Fortran runtime, in the
! record layout of the reference's IOSCALAR and IOVECTOR (src/io.f:8-200): sequential
! unformatted records, INTEGER*4, REAL*8 (the reference builds with -r8).
!   ioscalar_dir1.res   [NX NY NZG 1], NZG planes, [NSTEP], [TIME], [DTM1 GRAV]
!   ioscalar_dir2.res   [NX NY NZG 1 TIME DTM1 GRAV], NZG planes
!   iovector_dir1.res   [NX NY NZG 3], NZG records of the U, V, W planes
!   iovector_dir2.res   [NX NY NZG 3 TIME], NZG records of the U, V, W planes
! NZG = 2*(5-2)+2 = 8: the global file of two ranks with NZ = 5 each.
! Values: P(I,J,K) = I + 10 J + 100 K + 0.5, U = P + 0.25, V = -P, W = 2 P.
program field_writer
  implicit none
  integer, parameter :: nx = 6, ny = 5, nzg = 8, nstep = 42
  real(8), parameter :: time = 1.25d0, dtm1 = 1.0d-3, grav = 9.81d0
  real(8) :: p(nx, ny, nzg), u(nx, ny, nzg), v(nx, ny, nzg), w(nx, ny, nzg)
  integer :: i, j, k

  do k = 1, nzg
    do j = 1, ny
      do i = 1, nx
        p(i, j, k) = dble(i) + 10.0d0 * dble(j) + 100.0d0 * dble(k) + 0.5d0
      end do
    end do
  end do
  u = p + 0.25d0
  v = -p
  w = 2.0d0 * p

  open(19, file='ioscalar_dir1.res', form='unformatted', status='replace')
  write(19) nx, ny, nzg, 1
  do k = 1, nzg
    write(19) ((p(i, j, k), i = 1, nx), j = 1, ny)
  end do
  write(19) nstep
  write(19) time
  write(19) dtm1, grav
  close(19)

  open(19, file='ioscalar_dir2.res', form='unformatted', status='replace')
  write(19) nx, ny, nzg, 1, time, dtm1, grav
  do k = 1, nzg
    write(19) ((p(i, j, k), i = 1, nx), j = 1, ny)
  end do
  close(19)

  open(19, file='iovector_dir1.res', form='unformatted', status='replace')
  write(19) nx, ny, nzg, 3
  do k = 1, nzg
    write(19) ((u(i, j, k), i = 1, nx), j = 1, ny), ((v(i, j, k), i = 1, nx), j = 1, ny), &
              ((w(i, j, k), i = 1, nx), j = 1, ny)
  end do
  close(19)

  open(19, file='iovector_dir2.res', form='unformatted', status='replace')
  write(19) nx, ny, nzg, 3, time
  do k = 1, nzg
    write(19) ((u(i, j, k), i = 1, nx), j = 1, ny), ((v(i, j, k), i = 1, nx), j = 1, ny), &
              ((w(i, j, k), i = 1, nx), j = 1, ny)
  end do
  close(19)
end program field_writer
