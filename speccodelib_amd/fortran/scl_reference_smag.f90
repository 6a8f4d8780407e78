! smag of smag.f90 (the unpadded products) in an object of its own: a
! program built with smag_dealias.f90's smag (3/2-rule products) links
! libscl_fortran_padded.a first and this one is never pulled in.
! smag.f90:1-78 (c1 = 0.026 inside the library)
subroutine smag(vx, vy, vz, t11, t12, t13, t22, t23, t33, kx, ky, kz, lx1, ly, lz, nx, ny, nz, delta)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, nx, ny, nz
  real(sp), intent(in) :: delta
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz
  complex(sp), dimension(lx1, ly, lz), intent(out) :: t11, t12, t13, t22, t23, t33
  real(sp), dimension(lx1), intent(in) :: kx
  real(sp), dimension(ly), intent(in) :: ky
  real(sp), dimension(lz), intent(in) :: kz
  call scl_check(scl_smag(c_null_ptr, vx, vy, vz, t11, t12, t13, t22, t23, t33, kx, ky, kz, lx1, ly, lz, &
                          nx, ny, nz, real(delta, c_double)), 'smag')
end subroutine smag
