! The 3/2-rule routines with the reference's names and argument lists:
! padd (padd.f90:1-14), unpadd (unpadd.f90:1-13), the padded convec that
! spectral_heli_dynsmag.f90:229 calls, and smag of smag_dealias.f90:1-75
! (its own argument list; it replaces smag.f90's smag, so these are built
! into libscl_fortran_padded.a, linked before libscl_fortran.a by programs
! that use the 3/2 rule).  The products run on a plan of 2n points per axis
! that the library makes from the plan of fftwplan3d: the modes kept by
! unpadd are those of the reference's 3n/2 grid for band-limited fields.
subroutine padd(u, ub, lx1, ly, lz, lxb1, lyb, lzb)
  use iso_c_binding
  use mconstant
  use mfftwplan3d, only: plan_prec
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, lxb1, lyb, lzb
  complex(sp), dimension(lx1, ly, lz), intent(in) :: u
  complex(sp), dimension(lxb1, lyb, lzb), intent(out) :: ub
  call scl_check(scl_padd(plan_prec(), u, ub, lx1, ly, lz, lxb1, lyb, lzb), 'padd')
end subroutine padd

subroutine unpadd(ub, u, lx1, ly, lz, lxb1, lyb, lzb)
  use iso_c_binding
  use mconstant
  use mfftwplan3d, only: plan_prec
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, lxb1, lyb, lzb
  complex(sp), dimension(lxb1, lyb, lzb), intent(in) :: ub
  complex(sp), dimension(lx1, ly, lz), intent(out) :: u
  call scl_check(scl_unpadd(plan_prec(), ub, u, lx1, ly, lz, lxb1, lyb, lzb), 'unpadd')
end subroutine unpadd

! w <- the unaliased Lamb vector v x w (spectral_heli_dynsmag.f90:229)
subroutine convec(vx, vy, vz, wx, wy, wz, lx1, ly, lz, lxb1, lyb, lzb, nxb, nyb, nzb)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, lxb1, lyb, lzb, nxb, nyb, nzb
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz
  complex(sp), dimension(lx1, ly, lz), intent(inout) :: wx, wy, wz
  call scl_check(scl_convec_padded(c_null_ptr, vx, vy, vz, wx, wy, wz, lx1, ly, lz, lxb1, lyb, lzb, &
                                   nxb, nyb, nzb), 'convec (padded)')
end subroutine convec

! smag_dealias.f90:1-75 (c1 = 0.026; its strain as written there)
subroutine smag(vx, vy, vz, tau11, tau12, tau13, tau22, tau23, tau33, kx, ky, kz, lx1, ly, lz, &
                lxb1, lyb, lzb, nxb, nyb, nzb, delta)
  use iso_c_binding
  use mconstant
  use scl_capi
  implicit none
  integer, intent(in) :: lx1, ly, lz, lxb1, lyb, lzb, nxb, nyb, nzb
  real(sp), intent(in) :: delta
  complex(sp), dimension(lx1, ly, lz), intent(in) :: vx, vy, vz
  complex(sp), dimension(lx1, ly, lz), intent(out) :: tau11, tau12, tau13, tau22, tau23, tau33
  real(sp), dimension(lx1, ly, lz), intent(in) :: kx, ky, kz
  call scl_check(scl_smag_dealias(c_null_ptr, vx, vy, vz, tau11, tau12, tau13, tau22, tau23, tau33, kx, ky, kz, &
                                  lx1, ly, lz, lxb1, lyb, lzb, nxb, nyb, nzb, real(delta, c_double)), &
                 'smag (smag_dealias)')
end subroutine smag
