! Forced-DNS IF-RK4 run through the reference-named routines of libscl's
! Fortran module set, in the dns-deform.f90 structure (plan, wavenumbers,
! k2e, seeded initial field, RK4 loop with an RHS built from convec_dns,
! force, symmetrize and projection, 2/3 cut).
!   usage: dns_deform_gpu nx nsteps mode outfile
!   mode 1: RHS from convec_dns/force/symmetrize/projection, RK4 stages here
!   mode 2: rk4step (the whole step in one call)
!   mode 3: as 1, with u x w built here from rfftwnd_f77_one_* transforms
! Writes vx, vy, vz (complex(sp), (lx1,ly,lz) each) to outfile as a stream.
program dns_deform_gpu
  use mconstant
  use mwavenumber
  use mfftwplan3d
  use msymmetrize
  use minitialize
  implicit none
  integer :: nx, nsteps, mode, lx, lx1, ly, lz, istep, iseed
  real(sp) :: rnu, dt, eps, fkmax, akp, u0, kcut, kcut2
  complex(sp), allocatable, dimension(:,:,:) :: vx, vy, vz, wx, wy, wz, kux, kuy, kuz, dux, duy, duz
  real(sp), allocatable :: kx(:), ky(:), kz(:), k2(:,:,:), k2e(:,:,:)
  character(len=256) :: arg, outfile

  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) nsteps
  call get_command_argument(3, arg); read(arg, *) mode
  call get_command_argument(4, outfile)
  rnu = 0.01_sp; dt = 0.002_sp; eps = 0.1_sp; fkmax = 2.0_sp; akp = 4.0_sp; u0 = 1.0_sp; iseed = 1

  lx = nx / 2; lx1 = lx + 1; ly = nx; lz = nx
  kcut = 2.0 * lx / 3; kcut2 = kcut * kcut
  call fftwplan3d(nx, nx, nx)
  allocate(kx(lx1), ky(ly), kz(lz), k2(lx1, ly, lz), k2e(lx1, ly, lz))
  allocate(vx(lx1,ly,lz), vy(lx1,ly,lz), vz(lx1,ly,lz), wx(lx1,ly,lz), wy(lx1,ly,lz), wz(lx1,ly,lz))
  allocate(kux(lx1,ly,lz), kuy(lx1,ly,lz), kuz(lx1,ly,lz), dux(lx1,ly,lz), duy(lx1,ly,lz), duz(lx1,ly,lz))
  call wavenumber(kx, ky, kz, k2, lx1, ly, lz)
  k2e = exp(-rnu * k2 * dt / 2)
  call initialize(vx, vy, vz, kx, ky, kz, k2, iseed, lx1, ly, lz, akp, u0, kcut)

  do istep = 1, nsteps
    if (mode == 2) then
      call rk4step(vx, vy, vz, lx1, ly, lz, rnu, dt, 1, fkmax, eps, 0.0_sp)
      cycle
    end if
    call rhs(vx, vy, vz, wx, wy, wz)
    kux = dt * wx; kuy = dt * wy; kuz = dt * wz
    dux = kux; duy = kuy; duz = kuz
    kux = k2e * (vx + kux / 2); kuy = k2e * (vy + kuy / 2); kuz = k2e * (vz + kuz / 2)
    call rhs(kux, kuy, kuz, wx, wy, wz)
    kux = dt * wx / k2e; kuy = dt * wy / k2e; kuz = dt * wz / k2e
    dux = dux + 2 * kux; duy = duy + 2 * kuy; duz = duz + 2 * kuz
    kux = k2e * (vx + kux / 2); kuy = k2e * (vy + kuy / 2); kuz = k2e * (vz + kuz / 2)
    call rhs(kux, kuy, kuz, wx, wy, wz)
    kux = dt * wx / k2e; kuy = dt * wy / k2e; kuz = dt * wz / k2e
    dux = dux + 2 * kux; duy = duy + 2 * kuy; duz = duz + 2 * kuz
    kux = k2e * k2e * (vx + kux); kuy = k2e * k2e * (vy + kuy); kuz = k2e * k2e * (vz + kuz)
    call rhs(kux, kuy, kuz, wx, wy, wz)
    kux = dt * wx / k2e / k2e; kuy = dt * wy / k2e / k2e; kuz = dt * wz / k2e / k2e
    dux = dux + kux; duy = duy + kuy; duz = duz + kuz
    vx = k2e * k2e * (vx + dux / 6); vy = k2e * k2e * (vy + duy / 6); vz = k2e * k2e * (vz + duz / 6)
    where (k2 >= kcut2)
      vx = 0; vy = 0; vz = 0
    end where
  end do

  open(11, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(11) vx, vy, vz
  close(11)
  call destroyplan3d
  write(*, '(a,i0,a,i0,a,i0)') 'dns_deform_gpu done: nx=', nx, ' steps=', nsteps, ' mode=', mode

contains

  subroutine rhs(lvx, lvy, lvz, lwx, lwy, lwz)
    complex(sp), dimension(:,:,:) :: lvx, lvy, lvz, lwx, lwy, lwz
    integer :: ii, jj, kk
    do kk = 1, lz
      do jj = 1, ly
        do ii = 1, lx1
          lwx(ii,jj,kk) = eye * ky(jj) * lvz(ii,jj,kk) - eye * kz(kk) * lvy(ii,jj,kk)
          lwy(ii,jj,kk) = eye * kz(kk) * lvx(ii,jj,kk) - eye * kx(ii) * lvz(ii,jj,kk)
          lwz(ii,jj,kk) = eye * kx(ii) * lvy(ii,jj,kk) - eye * ky(jj) * lvx(ii,jj,kk)
        end do
      end do
    end do
    if (mode == 3) then
      call lamb_fftw(lvx, lvy, lvz, lwx, lwy, lwz)
    else
      call convec_dns(lvx, lvy, lvz, lwx, lwy, lwz, lx1, ly, lz, nx, nx, nx)
    end if
    call force(lwx, lwy, lwz, lvx, lvy, lvz, k2, lx1, ly, lz, fkmax, eps)
    call symmetrize(lwx, lx1, ly, lz)
    call symmetrize(lwy, lx1, ly, lz)
    call symmetrize(lwz, lx1, ly, lz)
    call projection(lwx, lwy, lwz, kx, ky, kz, lx1, ly, lz)
  end subroutine rhs

  ! u x w on FFTW's in-place packed layout, with the FFTW 2.x entry names
  subroutine lamb_fftw(lvx, lvy, lvz, lwx, lwy, lwz)
    complex(sp), dimension(:,:,:) :: lvx, lvy, lvz, lwx, lwy, lwz
    complex(sp), allocatable, dimension(:,:,:) :: ax, ay, az, px, py, pz
    real(sp) :: c, ignore
    allocate(ax, source=lvx)
    allocate(ay, source=lvy)
    allocate(az, source=lvz)
    c = 1.0_sp / real(nx, sp) ** 3
    call rfftwnd_f77_one_complex_to_real(C2R3D, ax, ignore)
    call rfftwnd_f77_one_complex_to_real(C2R3D, ay, ignore)
    call rfftwnd_f77_one_complex_to_real(C2R3D, az, ignore)
    call rfftwnd_f77_one_complex_to_real(C2R3D, lwx, ignore)
    call rfftwnd_f77_one_complex_to_real(C2R3D, lwy, ignore)
    call rfftwnd_f77_one_complex_to_real(C2R3D, lwz, ignore)
    ! real parts hold the even x points, imaginary parts the odd ones
    px = cmplx(real(ay) * real(lwz) - real(az) * real(lwy), aimag(ay) * aimag(lwz) - aimag(az) * aimag(lwy), sp) * c
    py = cmplx(real(az) * real(lwx) - real(ax) * real(lwz), aimag(az) * aimag(lwx) - aimag(ax) * aimag(lwz), sp) * c
    pz = cmplx(real(ax) * real(lwy) - real(ay) * real(lwx), aimag(ax) * aimag(lwy) - aimag(ay) * aimag(lwx), sp) * c
    call rfftwnd_f77_one_real_to_complex(R2C3D, px, ignore)
    call rfftwnd_f77_one_real_to_complex(R2C3D, py, ignore)
    call rfftwnd_f77_one_real_to_complex(R2C3D, pz, ignore)
    lwx = px
    lwy = py
    lwz = pz
  end subroutine lamb_fftw

end program dns_deform_gpu
