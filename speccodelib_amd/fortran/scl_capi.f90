! iso_c_binding interfaces to the C ABI of libscl (include/scl.h).
! Arrays are passed as assumed-type, assumed-size dummies, i.e. by address,
! in the caller's precision; the plan fixes how the library reads them.
module scl_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: SCL_SP = 0, SCL_DP = 1

  interface
    function scl_last_error() bind(C, name='scl_last_error') result(p)
      import
      type(c_ptr) :: p
    end function

    integer(c_int32_t) function scl_fftwplan3d_prec(ix, iy, iz, prec) bind(C, name='scl_fftwplan3d_prec')
      import
      integer(c_int32_t), value :: ix, iy, iz, prec
    end function

    integer(c_int32_t) function scl_destroyplan3d() bind(C, name='scl_destroyplan3d')
      import
    end function

    integer(c_int32_t) function scl_wavenumber3d(prec, kx, ky, kz, k2, lx1, ly, lz) &
        bind(C, name='scl_wavenumber3d')
      import
      integer(c_int32_t), value :: prec, lx1, ly, lz
      type(*), dimension(*) :: kx, ky, kz, k2
    end function

    integer(c_int32_t) function scl_convec_dns(plan, vx, vy, vz, wx, wy, wz, lx1, ly, lz, nx, ny, nz) &
        bind(C, name='scl_convec_dns')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz, wx, wy, wz
      integer(c_int32_t), value :: lx1, ly, lz, nx, ny, nz
    end function

    integer(c_int32_t) function scl_force(plan, wx, wy, wz, vx, vy, vz, k2, lx1, ly, lz, fkmax, eps) &
        bind(C, name='scl_force')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: wx, wy, wz, vx, vy, vz, k2
      integer(c_int32_t), value :: lx1, ly, lz
      real(c_double), value :: fkmax, eps
    end function

    integer(c_int32_t) function scl_force_heli(plan, fx, fy, fz, vx, vy, vz, wx, wy, wz, k2, lx1, ly, lz, &
        fkmax, eps, eta) bind(C, name='scl_force_heli')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: fx, fy, fz, vx, vy, vz, wx, wy, wz, k2
      integer(c_int32_t), value :: lx1, ly, lz
      real(c_double), value :: fkmax, eps, eta
    end function

    integer(c_int32_t) function scl_symmetrize3d(plan, c, lx1, ly, lz) bind(C, name='scl_symmetrize3d')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: c
      integer(c_int32_t), value :: lx1, ly, lz
    end function

    integer(c_int32_t) function scl_dealiasing(plan, c, k2, lx1, ly, lz) bind(C, name='scl_dealiasing')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: c, k2
      integer(c_int32_t), value :: lx1, ly, lz
    end function

    integer(c_int32_t) function scl_projection(plan, vx, vy, vz, kx, ky, kz, lx1, ly, lz) &
        bind(C, name='scl_projection')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz, kx, ky, kz
      integer(c_int32_t), value :: lx1, ly, lz
    end function

    integer(c_int32_t) function scl_ratesdudt(plan, vx, vy, vz, nx_, ny_, nz_, forced, fkmax, eps, eta) &
        bind(C, name='scl_ratesdudt')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz, nx_, ny_, nz_
      integer(c_int32_t), value :: forced
      real(c_double), value :: fkmax, eps, eta
    end function

    integer(c_int32_t) function scl_initialize3d(prec, vx, vy, vz, nx, ny, nz, iseed, akp, u0, kcut) &
        bind(C, name='scl_initialize3d')
      import
      integer(c_int32_t), value :: prec, nx, ny, nz, iseed
      type(*), dimension(*) :: vx, vy, vz
      real(c_double), value :: akp, u0, kcut
    end function

    integer(c_int32_t) function scl_rk4step(plan, vx, vy, vz, nu, dt, forced, fkmax, eps, eta) &
        bind(C, name='scl_rk4step')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz
      real(c_double), value :: nu, dt, fkmax, eps, eta
      integer(c_int32_t), value :: forced
    end function
  end interface

contains

  ! Status check mirroring the reference's STOP on fatal conditions.
  subroutine scl_check(status, where)
    integer(c_int32_t), intent(in) :: status
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status == 0) return
    call c_f_pointer(scl_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(3a,i0,2a)') 'libscl: ', where, ' failed (status ', status, '): ', transfer(msg(1:n), repeat(' ', n))
    error stop 'libscl error'
  end subroutine scl_check

end module scl_capi
