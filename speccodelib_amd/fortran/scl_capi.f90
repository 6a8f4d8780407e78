! iso_c_binding interfaces to the C ABI of libscl (include/scl.h).
! Arrays are passed as assumed-type, assumed-size dummies, i.e. by address,
! in the caller's precision; the plan fixes how the library reads them.
module scl_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: SCL_SP = 0, SCL_DP = 1

  ! struct scl_dns_params of scl.h (same member order; bind(C) gives the C
  ! layout and padding)
  type, bind(C) :: scl_dns_params
    real(c_double) :: nu = 0, dt = 0
    integer(c_int32_t) :: forced = 0
    real(c_double) :: fkmax = 0, eps = 0, eta = 0
    integer(c_int32_t) :: dealias_rhs = 0, use_graph = 0, sgs = 0
    real(c_double) :: cs2 = 0, delta = 0, delta_test = 0, k_test = 0
    integer(c_int32_t) :: iupdate = 10, scalar = 0
    real(c_double) :: grad = 0, schmidt = 1
    integer(c_int32_t) :: stepper = 0
    real(c_double) :: beta = 0.5, dtw = 0
    integer(c_int32_t) :: ieout = 10
    real(c_double) :: umax = 5.196152422706632d0, dt_first = 0, kcut = 0
  end type scl_dns_params

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

    integer(c_int32_t) function scl_initialize_heli(prec, ux, uy, uz, nx, ny, nz, iseed, hrel, akp, u0) &
        bind(C, name='scl_initialize_heli')
      import
      integer(c_int32_t), value :: prec, nx, ny, nz, iseed
      type(*), dimension(*) :: ux, uy, uz
      real(c_double), value :: hrel, akp, u0
    end function

    integer(c_int32_t) function scl_rk4step(plan, vx, vy, vz, nu, dt, forced, fkmax, eps, eta) &
        bind(C, name='scl_rk4step')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz
      real(c_double), value :: nu, dt, fkmax, eps, eta
      integer(c_int32_t), value :: forced
    end function

    integer(c_int32_t) function scl_smag(plan, vx, vy, vz, t11, t12, t13, t22, t23, t33, kx, ky, kz, &
        lx1, ly, lz, nx, ny, nz, delta) bind(C, name='scl_smag')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz, t11, t12, t13, t22, t23, t33, kx, ky, kz
      integer(c_int32_t), value :: lx1, ly, lz, nx, ny, nz
      real(c_double), value :: delta
    end function

    integer(c_int32_t) function scl_dynsmag(plan, ux, uy, uz, t11, t12, t13, t22, t23, t33, kx, ky, kz, g_test, &
        lx1, ly, lz, nx, ny, nz, lx, delta, delta_test, mm, lm, cs2, update) bind(C, name='scl_dynsmag')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: ux, uy, uz, t11, t12, t13, t22, t23, t33, kx, ky, kz, g_test
      integer(c_int32_t), value :: lx1, ly, lz, nx, ny, nz, lx, update
      real(c_double), value :: delta, delta_test
      real(c_double) :: mm, lm, cs2
    end function

    integer(c_int32_t) function scl_convecphi_dns(plan, vx, vy, vz, phi, wphi, kx, ky, kz, lx1, ly, lz, &
        nx, ny, nz) bind(C, name='scl_convecphi_dns')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz, phi, wphi, kx, ky, kz
      integer(c_int32_t), value :: lx1, ly, lz, nx, ny, nz
    end function

    integer(c_int32_t) function scl_padd(prec, u, ub, lx1, ly, lz, lxb1, lyb, lzb) bind(C, name='scl_padd')
      import
      integer(c_int32_t), value :: prec, lx1, ly, lz, lxb1, lyb, lzb
      type(*), dimension(*) :: u, ub
    end function

    integer(c_int32_t) function scl_unpadd(prec, ub, u, lx1, ly, lz, lxb1, lyb, lzb) bind(C, name='scl_unpadd')
      import
      integer(c_int32_t), value :: prec, lx1, ly, lz, lxb1, lyb, lzb
      type(*), dimension(*) :: ub, u
    end function

    integer(c_int32_t) function scl_convec_padded(plan, vx, vy, vz, wx, wy, wz, lx1, ly, lz, lxb1, lyb, lzb, &
        nxb, nyb, nzb) bind(C, name='scl_convec_padded')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz, wx, wy, wz
      integer(c_int32_t), value :: lx1, ly, lz, lxb1, lyb, lzb, nxb, nyb, nzb
    end function

    integer(c_int32_t) function scl_smag_dealias(plan, vx, vy, vz, t11, t12, t13, t22, t23, t33, kx, ky, kz, &
        lx1, ly, lz, lxb1, lyb, lzb, nxb, nyb, nzb, delta) bind(C, name='scl_smag_dealias')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: vx, vy, vz, t11, t12, t13, t22, t23, t33, kx, ky, kz
      integer(c_int32_t), value :: lx1, ly, lz, lxb1, lyb, lzb, nxb, nyb, nzb
      real(c_double), value :: delta
    end function

    integer(c_int32_t) function scl_skewness(plan, ux, kx, s, nx, ny, nz) bind(C, name='scl_skewness')
      import
      type(c_ptr), value :: plan
      type(*), dimension(*) :: ux, kx
      real(c_double) :: s
      integer(c_int32_t), value :: nx, ny, nz
    end function

    ! explicit plan handles and the device-resident solver
    integer(c_int32_t) function scl_plan3d_create(nx, ny, nz, prec, plan) bind(C, name='scl_plan3d_create')
      import
      integer(c_int32_t), value :: nx, ny, nz, prec
      type(c_ptr) :: plan
    end function

    integer(c_int32_t) function scl_plan3d_destroy(plan) bind(C, name='scl_plan3d_destroy')
      import
      type(c_ptr), value :: plan
    end function

    integer(c_int32_t) function scl_dns_create(plan, params, dns) bind(C, name='scl_dns_create')
      import
      type(c_ptr), value :: plan
      type(scl_dns_params), intent(in) :: params
      type(c_ptr) :: dns
    end function

    integer(c_int32_t) function scl_dns_destroy(dns) bind(C, name='scl_dns_destroy')
      import
      type(c_ptr), value :: dns
    end function

    integer(c_int32_t) function scl_dns_set_velocity(dns, vx, vy, vz) bind(C, name='scl_dns_set_velocity')
      import
      type(c_ptr), value :: dns
      type(*), dimension(*) :: vx, vy, vz
    end function

    integer(c_int32_t) function scl_dns_get_velocity(dns, vx, vy, vz) bind(C, name='scl_dns_get_velocity')
      import
      type(c_ptr), value :: dns
      type(*), dimension(*) :: vx, vy, vz
    end function

    integer(c_int32_t) function scl_dns_set_scalar(dns, phi) bind(C, name='scl_dns_set_scalar')
      import
      type(c_ptr), value :: dns
      type(*), dimension(*) :: phi
    end function

    integer(c_int32_t) function scl_dns_get_scalar(dns, phi) bind(C, name='scl_dns_get_scalar')
      import
      type(c_ptr), value :: dns
      type(*), dimension(*) :: phi
    end function

    integer(c_int32_t) function scl_dns_step(dns, nsteps) bind(C, name='scl_dns_step')
      import
      type(c_ptr), value :: dns
      integer(c_int32_t), value :: nsteps
    end function

    integer(c_int32_t) function scl_dns_sync(dns) bind(C, name='scl_dns_sync')
      import
      type(c_ptr), value :: dns
    end function

    integer(c_int32_t) function scl_dns_spectrum(dns, ek, nbins) bind(C, name='scl_dns_spectrum')
      import
      type(c_ptr), value :: dns
      real(c_double), dimension(*) :: ek
      integer(c_int32_t), value :: nbins
    end function

    ! binds an FP64 solver to the external scl_derivs (derivs(x, y, dydx, ny)
    ! of rk5.f90:3-10), which a Fortran RK driver then takes as its derivs
    integer(c_int32_t) function scl_derivs_bind(dns) bind(C, name='scl_derivs_bind')
      import
      type(c_ptr), value :: dns
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
