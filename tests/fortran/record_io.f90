! Fortran side of the file-format checks (no GPU): unformatted sequential
! records as output.f90 / input.f90 write and read them, and the
! '(I6,20E12.4)' lines of ener_time.data (spectral_dns_dealiased.f90:136).
!   record_io w path n : write(10) a, a(i) = cmplx(i, -i/2), complex(sp)
!   record_io r path n : read(10) a and check it; prints "ok" or "bad"
!   record_io e path   : appends ener_time lines for fixed values
program record_io
  implicit none
  integer, parameter :: sp = kind(1.0)
  character(len=256) :: mode, path, arg
  integer :: n, i
  complex(sp), allocatable :: a(:)
  real(sp) :: v(4, 5)
  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  if (mode == 'e') then
    v(:, 1) = [0.0_sp, 1.0e-3_sp, 1.5_sp, -0.25_sp]
    v(:, 2) = [0.123456_sp, 9.99951e-1_sp, 1234.5678_sp, -6.02e-23_sp]
    v(:, 3) = [12.0_sp, 0.002_sp, 1.4541583_sp, 3.3e7_sp]
    v(:, 4) = [-1.0_sp, 5.0e-5_sp, 99999.0_sp, 0.5_sp]
    v(:, 5) = [7.25_sp, -3.0e-9_sp, 0.1_sp, 2.0_sp]
    open(25, file=trim(path), status='replace')
    do i = 1, 5
      write(25, '(I6,20E12.4)') 10 * i - 7, v(:, i)
    end do
    close(25)
    stop
  end if
  call get_command_argument(3, arg)
  read(arg, *) n
  allocate(a(n))
  if (mode == 'w') then
    do i = 1, n
      a(i) = cmplx(real(i, sp), -0.5_sp * i, sp)
    end do
    open(10, file=trim(path), status='replace', form='unformatted')
    write(10) a
    close(10)
  else
    open(10, file=trim(path), status='old', form='unformatted')
    read(10) a
    close(10)
    do i = 1, n
      if (a(i) /= cmplx(real(i, sp), -0.5_sp * i, sp)) then
        write(*, '(a)') 'bad'
        stop
      end if
    end do
    write(*, '(a)') 'ok'
  end if
end program record_io
