! Precision module for building the library's Fortran module set and its
! test drivers on their own: the reference's module name and entities
! (module mconstant of constant_sp.f90), single-precision sp.
module mconstant
  implicit none
  integer, parameter :: sp = kind(1.0)
  integer, parameter :: dp = kind(1.0d0)
  real(sp), parameter :: mytiny = tiny(1.0_sp)
  complex(sp), parameter :: eye = (0.0_sp, 1.0_sp)
  real(sp), parameter :: pi = 3.1415926536_sp
end module mconstant
