! The 3/2-rule routines by their reference names (padd, unpadd, the padded
! convec of spectral_heli_dynsmag.f90:229, smag of smag_dealias.f90), linked
! from libscl_fortran_padded.a before libscl_fortran.a, on two seeded
! fields of the library's initial condition.
!   usage: padded_gpu nx outfile
! Writes to outfile as a stream of complex(sp) (lx1,ly,lz) arrays: vx, vy,
! vz, wx, wy, wz (inputs), the padded Lamb vector (3), tau11..tau33 of
! smag (6), then unpadd(padd(vx)) and the (lxb1,lyb,lzb) array padd(vx).
program padded_gpu
  use iso_c_binding
  use mconstant
  use mwavenumber
  use mfftwplan3d
  use minitialize
  implicit none
  integer :: nx, lx, lx1, ly, lz, nxb, lxb1
  real(sp) :: kcut, delta
  complex(sp), allocatable, dimension(:,:,:) :: vx, vy, vz, wx, wy, wz, lx_, ly_, lz_, u2
  complex(sp), allocatable, dimension(:,:,:) :: t11, t12, t13, t22, t23, t33, ub
  real(sp), allocatable :: kx(:), ky(:), kz(:), k2(:,:,:), kx3(:,:,:), ky3(:,:,:), kz3(:,:,:), k23(:,:,:)
  character(len=256) :: arg, outfile
  integer :: u

  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, outfile)
  lx = nx / 2; lx1 = lx + 1; ly = nx; lz = nx
  nxb = 3 * nx / 2; lxb1 = nxb / 2 + 1
  kcut = 2.0 * lx / 3
  delta = 2.0_sp * pi / nx
  call fftwplan3d(nx, nx, nx)
  allocate(kx(lx1), ky(ly), kz(lz), k2(lx1, ly, lz))
  allocate(kx3(lx1,ly,lz), ky3(lx1,ly,lz), kz3(lx1,ly,lz), k23(lx1,ly,lz))
  allocate(vx(lx1,ly,lz), vy(lx1,ly,lz), vz(lx1,ly,lz), wx(lx1,ly,lz), wy(lx1,ly,lz), wz(lx1,ly,lz))
  allocate(lx_(lx1,ly,lz), ly_(lx1,ly,lz), lz_(lx1,ly,lz), u2(lx1,ly,lz), ub(lxb1,nxb,nxb))
  allocate(t11(lx1,ly,lz), t12(lx1,ly,lz), t13(lx1,ly,lz), t22(lx1,ly,lz), t23(lx1,ly,lz), t33(lx1,ly,lz))
  call wavenumber(kx, ky, kz, k2, lx1, ly, lz)
  call wavenumberhighmem3d(kx3, ky3, kz3, k23, lx1, ly, lz)
  call initialize(vx, vy, vz, kx, ky, kz, k2, 1, lx1, ly, lz, 4.0_sp, 1.0_sp, kcut)
  call initialize(wx, wy, wz, kx, ky, kz, k2, -7, lx1, ly, lz, 3.0_sp, 1.0_sp, kcut)
  lx_ = wx; ly_ = wy; lz_ = wz
  call convec(vx, vy, vz, lx_, ly_, lz_, lx1, ly, lz, lxb1, nxb, nxb, nxb, nxb, nxb)
  call smag(vx, vy, vz, t11, t12, t13, t22, t23, t33, kx3, ky3, kz3, lx1, ly, lz, lxb1, nxb, nxb, &
            nxb, nxb, nxb, delta)
  call padd(vx, ub, lx1, ly, lz, lxb1, nxb, nxb)
  call unpadd(ub, u2, lx1, ly, lz, lxb1, nxb, nxb)
  open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(u) vx, vy, vz, wx, wy, wz, lx_, ly_, lz_, t11, t12, t13, t22, t23, t33, u2, ub
  close(u)
  call destroyplan3d
  write(*, '(a,i0)') 'padded_gpu done: nx=', nx
end program padded_gpu
