! The reference's module and routine names, argument lists and array
! layouts, with bodies that call libscl through module scl_capi.  A program
! written against the reference (dns-deform.f90, spectral_les.f90,
! spectral_dns_scalar.f90 and the drivers of the same family) links these
! instead of fftwplan.f90, symmetrize.f90, wavenumber.f90, initialize.f90,
! convec_dns.f90, force.f90, force_heli.f90, dealiasing.f90,
! projection.f90, smag.f90, dynsmag.f90, convecphi_dns.f90,
! skewness.f90 and initialize_heli.f90, and FFTW 2.x: its rfftwnd_f77_one_complex_to_real /
! _real_to_complex calls resolve to the library's own entry points of those
! names.
!
! The kind sp comes from the caller's module mconstant (constant_sp.f90 or
! constant_dp.f90 in the reference); the plan made by fftwplan3d takes its
! precision from it.

! fftwplan.f90:1-31 (module mfftwplan3d)
module mfftwplan3d
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  ! Kept for source compatibility: the library holds the plan; these are
  ! only tokens passed back to rfftwnd_f77_one_*.
  integer*8 :: R2C3D = 1, C2R3D = -1
contains
  integer(c_int32_t) function plan_prec()
    if (kind(1.0_sp) == c_float) then
      plan_prec = SCL_SP
    else
      plan_prec = SCL_DP
    end if
  end function plan_prec

  subroutine fftwplan3d(ix, iy, iz)
    integer :: ix, iy, iz
    call scl_check(scl_fftwplan3d_prec(ix, iy, iz, plan_prec()), 'fftwplan3d')
  end subroutine fftwplan3d

  ! FFTW_ESTIMATE planning in the reference; the library has one plan kind
  subroutine fftwplan3de(ix, iy, iz)
    integer :: ix, iy, iz
    call scl_check(scl_fftwplan3d_prec(ix, iy, iz, plan_prec()), 'fftwplan3de')
  end subroutine fftwplan3de

  subroutine destroyplan3d
    call scl_check(scl_destroyplan3d(), 'destroyplan3d')
  end subroutine destroyplan3d
end module mfftwplan3d

! wavenumber.f90:1-11, 89-129 (module mwavenumber, the low-memory 3D form)
module mwavenumber
  use mconstant
  use mfftwplan3d, only: plan_prec
  use scl_capi
  implicit none
  interface wavenumber
    module procedure wavenumberlowmem3d, wavenumberhighmem3d
  end interface wavenumber
contains
  subroutine wavenumberlowmem3d(kx, ky, kz, k2, lx1, ly, lz)
    integer, intent(in) :: lx1, ly, lz
    real(sp), dimension(lx1), intent(out) :: kx
    real(sp), dimension(ly), intent(out) :: ky
    real(sp), dimension(lz), intent(out) :: kz
    real(sp), dimension(lx1, ly, lz), intent(out) :: k2
    call scl_check(scl_wavenumber3d(plan_prec(), kx, ky, kz, k2, lx1, ly, lz), 'wavenumber')
  end subroutine wavenumberlowmem3d

  ! wavenumber.f90:131-158: the same wavenumbers as 3D arrays (the LES
  ! drivers' form, spectral_les.f90:91,99)
  subroutine wavenumberhighmem3d(kx, ky, kz, k2, lx1, ly, lz)
    integer, intent(in) :: lx1, ly, lz
    real(sp), dimension(lx1, ly, lz), intent(out) :: kx, ky, kz, k2
    real(sp) :: ax(lx1), ay(ly), az(lz)
    integer :: jj, kk
    call scl_check(scl_wavenumber3d(plan_prec(), ax, ay, az, k2, lx1, ly, lz), 'wavenumber')
    do kk = 1, lz
      do jj = 1, ly
        kx(:, jj, kk) = ax
        ky(:, jj, kk) = ay(jj)
        kz(:, jj, kk) = az(kk)
      end do
    end do
  end subroutine wavenumberhighmem3d
end module mwavenumber

! symmetrize.f90:1-6, 32-64 (module msymmetrize, the 3D form)
module msymmetrize
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  interface symmetrize
    module procedure symmetrize3d
  end interface symmetrize
contains
  subroutine symmetrize3d(c, lx1, ly, lz)
    integer, intent(in) :: lx1, ly, lz
    complex(sp), dimension(lx1, ly, lz), intent(inout) :: c
    call scl_check(scl_symmetrize3d(c_null_ptr, c, lx1, ly, lz), 'symmetrize')
  end subroutine symmetrize3d
end module msymmetrize

! initialize.f90:1-10, 373-463 (module minitialize; the complex projection
! of :328-332 is used, see the library's initialize3d)
module minitialize
  use mconstant
  use mfftwplan3d, only: plan_prec
  use scl_capi
  implicit none
  interface initialize
    module procedure initialize3dvellowmemdealiased
  end interface initialize
contains
  subroutine initialize3dvellowmemdealiased(vx, vy, vz, kx, ky, kz, k2, iseed, lx1, ly, lz, akp, u0, kcut)
    integer, intent(in) :: lx1, ly, lz, iseed
    real(sp), intent(in) :: akp, u0, kcut
    complex(sp), dimension(lx1, ly, lz), intent(inout) :: vx, vy, vz
    real(sp), dimension(lx1, ly, lz), intent(in) :: k2
    real(sp), dimension(lx1), intent(in) :: kx
    real(sp), dimension(ly), intent(in) :: ky
    real(sp), dimension(lz), intent(in) :: kz
    call scl_check(scl_initialize3d(plan_prec(), vx, vy, vz, 2 * (lx1 - 1), ly, lz, iseed, &
                                    real(akp, c_double), real(u0, c_double), real(kcut, c_double)), 'initialize')
  end subroutine initialize3dvellowmemdealiased
end module minitialize

! convec_dns.f90:1-55
subroutine convec_dns(vx, vy, vz, wx, wy, wz, lx1, ly, lz, nx, ny, nz)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, nx, ny, nz
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz
  complex(sp), dimension(lx1, ly, lz), intent(inout) :: wx, wy, wz
  call scl_check(scl_convec_dns(c_null_ptr, vx, vy, vz, wx, wy, wz, lx1, ly, lz, nx, ny, nz), 'convec_dns')
end subroutine convec_dns

! force.f90:3-31
subroutine force(wx, wy, wz, vx, vy, vz, k2, lx1, ly, lz, fkmax, eps)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz
  real(sp), intent(in) :: fkmax, eps
  complex(sp), dimension(lx1, ly, lz), intent(inout) :: wx, wy, wz
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz
  real(sp), dimension(lx1, ly, lz), intent(in) :: k2
  call scl_check(scl_force(c_null_ptr, wx, wy, wz, vx, vy, vz, k2, lx1, ly, lz, &
                           real(fkmax, c_double), real(eps, c_double)), 'force')
end subroutine force

! force_heli.f90:1-57
subroutine force_heli(fx, fy, fz, vx, vy, vz, wx, wy, wz, k2, lx1, ly, lz, fkmax, eps, eta)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz
  real(sp), intent(in) :: fkmax, eps, eta
  complex(sp), dimension(lx1, ly, lz), intent(out) :: fx, fy, fz
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz, wx, wy, wz
  real(sp), dimension(lx1, ly, lz), intent(in) :: k2
  call scl_check(scl_force_heli(c_null_ptr, fx, fy, fz, vx, vy, vz, wx, wy, wz, k2, lx1, ly, lz, &
                                real(fkmax, c_double), real(eps, c_double), real(eta, c_double)), 'force_heli')
end subroutine force_heli

! dealiasing.f90:2-37
subroutine dealiasing(c, k2, lx1, ly, lz)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz
  complex(sp), dimension(lx1, ly, lz), intent(inout) :: c
  real(sp), dimension(lx1, ly, lz), intent(in) :: k2
  call scl_check(scl_dealiasing(c_null_ptr, c, k2, lx1, ly, lz), 'dealiasing')
end subroutine dealiasing

! projection.f90:2-31
subroutine projection(vx, vy, vz, kx, ky, kz, lx1, ly, lz)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz
  real(sp), dimension(lx1), intent(in) :: kx
  real(sp), dimension(ly), intent(in) :: ky
  real(sp), dimension(lz), intent(in) :: kz
  complex(sp), dimension(lx1, ly, lz), intent(inout) :: vx, vy, vz
  call scl_check(scl_projection(c_null_ptr, vx, vy, vz, kx, ky, kz, lx1, ly, lz), 'projection')
end subroutine projection

! The RHS of dns-deform.f90:259-288 (an internal procedure there) as an
! external routine: w <- N(v); iforce 0 none, 1 force(), 2 force_heli().
subroutine ratesdudt_dns(vx, vy, vz, wx, wy, wz, lx1, ly, lz, iforce, fkmax, eps, eta)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, iforce
  real(sp), intent(in) :: fkmax, eps, eta
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz
  complex(sp), dimension(lx1, ly, lz), intent(out) :: wx, wy, wz
  call scl_check(scl_ratesdudt(c_null_ptr, vx, vy, vz, wx, wy, wz, iforce, real(fkmax, c_double), &
                               real(eps, c_double), real(eta, c_double)), 'ratesdudt')
end subroutine ratesdudt_dns

! The Lawson IF-RK4 step with 2/3 truncation of dns-deform.f90:160-237 (inline
! there) as an external routine.
subroutine rk4step(vx, vy, vz, lx1, ly, lz, rnu, dt, iforce, fkmax, eps, eta)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, iforce
  real(sp), intent(in) :: rnu, dt, fkmax, eps, eta
  complex(sp), dimension(lx1, ly, lz), intent(inout) :: vx, vy, vz
  call scl_check(scl_rk4step(c_null_ptr, vx, vy, vz, real(rnu, c_double), real(dt, c_double), iforce, &
                             real(fkmax, c_double), real(eps, c_double), real(eta, c_double)), 'rk4step')
end subroutine rk4step

! initialize_heli.f90:1-109 (the start of spectral_heli_dns.f90)
subroutine initialize_heli(ux, uy, uz, kx, ky, kz, k2, lx1, ly, lz, iseed, hrel, akp, u0)
  use iso_c_binding
  use mconstant
  use mfftwplan3d, only: plan_prec
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, iseed
  real(sp), intent(in) :: akp, u0, hrel
  complex(sp), dimension(lx1, ly, lz), intent(out) :: ux, uy, uz
  real(sp), dimension(lx1, ly, lz), intent(in) :: k2
  real(sp), dimension(lx1), intent(in) :: kx
  real(sp), dimension(ly), intent(in) :: ky
  real(sp), dimension(lz), intent(in) :: kz
  call scl_check(scl_initialize_heli(plan_prec(), ux, uy, uz, 2 * (lx1 - 1), ly, lz, iseed, real(hrel, c_double), &
                                     real(akp, c_double), real(u0, c_double)), 'initialize_heli')
end subroutine initialize_heli

! dynsmag.f90:1-236.  The logical update and the real(sp) mm, lm, cs2 of the
! reference become the C ABI's int32 and doubles here; with update false
! cs2 is read (the static branch, :196-233) and mm, lm are left as they are.
subroutine dynsmag(ux, uy, uz, t11, t12, t13, t22, t23, t33, kx, ky, kz, g_test, &
                   lx1, ly, lz, nx, ny, nz, lx, delta, delta_test, mm, lm, cs2, update)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, nx, ny, nz, lx
  real(sp), intent(in) :: delta, delta_test
  logical, intent(in) :: update
  real(sp), intent(inout) :: cs2, mm, lm
  complex(sp), dimension(lx1, ly, lz), intent(in) :: ux, uy, uz
  real(sp), dimension(lx1, ly, lz), intent(in) :: kx, ky, kz, g_test
  complex(sp), dimension(lx1, ly, lz), intent(out) :: t11, t12, t13, t22, t23, t33
  real(c_double) :: dmm, dlm, dcs2
  integer(c_int32_t) :: iu
  dmm = mm
  dlm = lm
  dcs2 = cs2
  iu = 0
  if (update) iu = 1
  call scl_check(scl_dynsmag(c_null_ptr, ux, uy, uz, t11, t12, t13, t22, t23, t33, kx, ky, kz, g_test, &
                             lx1, ly, lz, nx, ny, nz, lx, real(delta, c_double), real(delta_test, c_double), &
                             dmm, dlm, dcs2, iu), 'dynsmag')
  if (update) then
    mm = real(dmm, sp)
    lm = real(dlm, sp)
    cs2 = real(dcs2, sp)
  end if
end subroutine dynsmag

! convecphi_dns.f90:1-49: wphi = i k . FFT(u phi)/N^3 (the reference's sign)
subroutine convecphi_dns(vx, vy, vz, phi, wphi, kx, ky, kz, lx1, ly, lz, nx, ny, nz)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, nx, ny, nz
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz, phi
  complex(sp), dimension(lx1, ly, lz), intent(inout) :: wphi
  real(sp), dimension(lx1), intent(in) :: kx
  real(sp), dimension(ly), intent(in) :: ky
  real(sp), dimension(lz), intent(in) :: kz
  call scl_check(scl_convecphi_dns(c_null_ptr, vx, vy, vz, phi, wphi, kx, ky, kz, lx1, ly, lz, nx, ny, nz), &
                 'convecphi_dns')
end subroutine convecphi_dns

! skewness.f90:1-48
subroutine skewness(ux, kx, s, nx, ny, nz)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: nx, ny, nz
  complex(sp), dimension(nx / 2 + 1, ny, nz), intent(in) :: ux
  real(sp), dimension(nx / 2 + 1), intent(in) :: kx
  real(sp), intent(out) :: s
  real(c_double) :: ds
  call scl_check(scl_skewness(c_null_ptr, ux, kx, ds, nx, ny, nz), 'skewness')
  s = real(ds, sp)
end subroutine skewness
