! Forced-DNS IF-RK4 run through the reference-named routines of libscl's
! Fortran module set, in the dns-deform.f90 structure (plan, wavenumbers,
! k2e, seeded initial field, RK4 loop with an RHS built from convec_dns,
! force, symmetrize and projection, 2/3 cut).
!   usage: dns_deform_gpu nx nsteps mode outfile
!   mode 1: RHS from convec_dns/force/symmetrize/projection, RK4 stages here
!   mode 2: rk4step (the whole step in one call)
!   mode 3: as 1, with u x w built here from rfftwnd_f77_one_* transforms
!   mode 4: LES + passive scalar, the RHS of spectral_les.f90:205-238 with
!           dynsmag (coefficient updated at the first RHS, then static) and
!           the scalar of spectral_dns_scalar.f90 through convecphi_dns
!   mode 5: LES with smag (smag.f90's fixed c1), no scalar
!   mode 6: the library's derivs (scl_derivs, bound to an FP64 solver)
!           stepped by a classical RK4 driver written here against the
!           derivs(x, y, dydx, ny) interface of rk5.f90:3-10
!   mode 7: initialize_heli (seed -9, hrel 0.6, akp 4, u0 1) only
! Writes vx, vy, vz (complex(sp), (lx1,ly,lz) each; mode 4 also phi; mode 6
! complex(dp)) to outfile as a stream.
program dns_deform_gpu
  use iso_c_binding
  use mconstant
  use mwavenumber
  use mfftwplan3d
  use msymmetrize
  use minitialize
  use scl_capi
  implicit none
  integer :: nx, nsteps, mode, lx, lx1, ly, lz, istep, iseed
  real(sp) :: rnu, dt, eps, fkmax, akp, u0, kcut, kcut2
  complex(sp), allocatable, dimension(:,:,:) :: vx, vy, vz, wx, wy, wz, kux, kuy, kuz, dux, duy, duz
  complex(sp), allocatable, dimension(:,:,:) :: phi, wphi, kphi, dphi, t11, t12, t13, t22, t23, t33
  real(sp), allocatable :: kx(:), ky(:), kz(:), k2(:,:,:), k2e(:,:,:), k2p(:,:,:)
  real(sp), allocatable :: kx3(:,:,:), ky3(:,:,:), kz3(:,:,:), k23(:,:,:), g_test(:,:,:)
  real(sp) :: delta, delta_test, k_test, mm, lm, cs2, sc, grad
  logical :: update
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

  if (mode == 4 .or. mode == 5) then
    call les_run()
    call destroyplan3d
    write(*, '(a,i0,a,i0,a,i0)') 'dns_deform_gpu done: nx=', nx, ' steps=', nsteps, ' mode=', mode
    stop
  end if
  if (mode == 6) then
    call derivs_run()
    call destroyplan3d
    write(*, '(a,i0,a,i0,a,i0)') 'dns_deform_gpu done: nx=', nx, ' steps=', nsteps, ' mode=', mode
    stop
  end if
  if (mode == 7) then
    ! the helical start of spectral_heli_dns.f90:92 (seed -9, hrel 0.6)
    call initialize_heli(vx, vy, vz, kx, ky, kz, k2, lx1, ly, lz, -9, 0.6_sp, akp, u0)
    open(11, file=trim(outfile), access='stream', form='unformatted', status='replace')
    write(11) vx, vy, vz
    close(11)
    call destroyplan3d
    write(*, '(a,i0,a,i0)') 'dns_deform_gpu done: nx=', nx, ' mode=', mode
    stop
  end if

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

  ! LES (+ scalar) run: the set-up of spectral_les.f90:80-109 (delta =
  ! pi/kcut, k_test = kcut/2, delta_test = 2 delta, sharp g_test, 3D
  ! wavenumbers) and IF-RK4 stages as in the DNS modes.
  subroutine les_run()
    real(sp), allocatable :: e2(:,:,:)
    allocate(kx3(lx1,ly,lz), ky3(lx1,ly,lz), kz3(lx1,ly,lz), k23(lx1,ly,lz), g_test(lx1,ly,lz), k2p(lx1,ly,lz))
    allocate(phi(lx1,ly,lz), wphi(lx1,ly,lz), kphi(lx1,ly,lz), dphi(lx1,ly,lz))
    allocate(t11(lx1,ly,lz), t12(lx1,ly,lz), t13(lx1,ly,lz), t22(lx1,ly,lz), t23(lx1,ly,lz), t33(lx1,ly,lz))
    call wavenumber(kx3, ky3, kz3, k23, lx1, ly, lz)
    delta = pi / kcut
    k_test = kcut / 2
    delta_test = 2 * delta
    where (k23 <= k_test * k_test)
      g_test = 1
    elsewhere
      g_test = 0
    end where
    rnu = 1.0e-3_sp; sc = 0.7_sp; grad = 1.0_sp; cs2 = 0.0289_sp; mm = 0; lm = 0
    update = mode == 4
    k2p = exp(-rnu / sc * k2 * dt / 2)
    k2e = exp(-rnu * k2 * dt / 2)
    phi = 0
    allocate(e2(lx1,ly,lz))
    e2 = k2e * k2e
    do istep = 1, nsteps
      call rhs_les(vx, vy, vz, phi, wx, wy, wz, wphi)
      kux = dt * wx; kuy = dt * wy; kuz = dt * wz; kphi = dt * wphi
      dux = kux; duy = kuy; duz = kuz; dphi = kphi
      kux = k2e * (vx + kux / 2); kuy = k2e * (vy + kuy / 2); kuz = k2e * (vz + kuz / 2)
      kphi = k2p * (phi + kphi / 2)
      call rhs_les(kux, kuy, kuz, kphi, wx, wy, wz, wphi)
      kux = dt * wx / k2e; kuy = dt * wy / k2e; kuz = dt * wz / k2e; kphi = dt * wphi / k2p
      dux = dux + 2 * kux; duy = duy + 2 * kuy; duz = duz + 2 * kuz; dphi = dphi + 2 * kphi
      kux = k2e * (vx + kux / 2); kuy = k2e * (vy + kuy / 2); kuz = k2e * (vz + kuz / 2)
      kphi = k2p * (phi + kphi / 2)
      call rhs_les(kux, kuy, kuz, kphi, wx, wy, wz, wphi)
      kux = dt * wx / k2e; kuy = dt * wy / k2e; kuz = dt * wz / k2e; kphi = dt * wphi / k2p
      dux = dux + 2 * kux; duy = duy + 2 * kuy; duz = duz + 2 * kuz; dphi = dphi + 2 * kphi
      kux = e2 * (vx + kux); kuy = e2 * (vy + kuy); kuz = e2 * (vz + kuz); kphi = k2p * k2p * (phi + kphi)
      call rhs_les(kux, kuy, kuz, kphi, wx, wy, wz, wphi)
      kux = dt * wx / e2; kuy = dt * wy / e2; kuz = dt * wz / e2; kphi = dt * wphi / (k2p * k2p)
      dux = dux + kux; duy = duy + kuy; duz = duz + kuz; dphi = dphi + kphi
      vx = e2 * (vx + dux / 6); vy = e2 * (vy + duy / 6); vz = e2 * (vz + duz / 6)
      phi = k2p * k2p * (phi + dphi / 6)
      where (k2 >= kcut2)
        vx = 0; vy = 0; vz = 0; phi = 0
      end where
    end do
    write(*, '(a,es14.6,a,es14.6,a,es14.6)') 'cs2 ', cs2, ' mm ', mm, ' lm ', lm
    open(11, file=trim(outfile), access='stream', form='unformatted', status='replace')
    if (mode == 4) then
      write(11) vx, vy, vz, phi
    else
      write(11) vx, vy, vz
    end if
    close(11)
  end subroutine les_run

  ! the LES momentum RHS of spectral_les.f90:205-238 (vorticity, SGS stress
  ! by dynsmag or smag, convec_dns, forcing, -i k_j tau_ij, projection),
  ! symmetrized, and (mode 4) the scalar RHS -convecphi_dns - grad u_z
  ! (spectral_dns_scalar.f90:236,245)
  subroutine rhs_les(lvx, lvy, lvz, lphi, lwx, lwy, lwz, lwphi)
    complex(sp), dimension(:,:,:) :: lvx, lvy, lvz, lphi, lwx, lwy, lwz, lwphi
    lwx = eye * (ky3 * lvz - kz3 * lvy)
    lwy = eye * (kz3 * lvx - kx3 * lvz)
    lwz = eye * (kx3 * lvy - ky3 * lvx)
    if (mode == 4) then
      call dynsmag(lvx, lvy, lvz, t11, t12, t13, t22, t23, t33, kx3, ky3, kz3, g_test, &
                   lx1, ly, lz, nx, nx, nx, lx, delta, delta_test, mm, lm, cs2, update)
      update = .false.
    else
      call smag(lvx, lvy, lvz, t11, t12, t13, t22, t23, t33, kx, ky, kz, lx1, ly, lz, nx, nx, nx, delta)
    end if
    call convec_dns(lvx, lvy, lvz, lwx, lwy, lwz, lx1, ly, lz, nx, nx, nx)
    call force(lwx, lwy, lwz, lvx, lvy, lvz, k2, lx1, ly, lz, fkmax, eps)
    lwx = lwx - eye * (kx3 * t11 + ky3 * t12 + kz3 * t13)
    lwy = lwy - eye * (kx3 * t12 + ky3 * t22 + kz3 * t23)
    lwz = lwz - eye * (kx3 * t13 + ky3 * t23 + kz3 * t33)
    call symmetrize(lwx, lx1, ly, lz)
    call symmetrize(lwy, lx1, ly, lz)
    call symmetrize(lwz, lx1, ly, lz)
    call projection(lwx, lwy, lwz, kx, ky, kz, lx1, ly, lz)
    if (mode == 4) then
      call convecphi_dns(lvx, lvy, lvz, lphi, lwphi, kx, ky, kz, lx1, ly, lz, nx, nx, nx)
      lwphi = -lwphi - grad * lvz
      call symmetrize(lwphi, lx1, ly, lz)
    else
      lwphi = 0
    end if
  end subroutine rhs_les

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

  ! derivs form: an FP64 plan and solver (forced DNS, the driver's nu, dt,
  ! forcing) bound to scl_derivs; a classical RK4 of dy/dt = N(y) - nu k^2 y
  ! over the three complex(dp) fields seen as ny real(dp)
  subroutine derivs_run()
    interface
      subroutine scl_derivs(x, y, dydx, n)
        use mconstant
        integer, intent(in) :: n
        real(dp), intent(in) :: x
        real(dp), dimension(n), intent(in) :: y
        real(dp), dimension(n), intent(out) :: dydx
      end subroutine scl_derivs
    end interface
    type(c_ptr) :: plan64, dns
    type(scl_dns_params) :: prm
    real(dp), allocatable, target :: y(:)
    complex(dp), pointer :: yc(:,:,:,:)
    integer :: ny
    plan64 = c_null_ptr
    dns = c_null_ptr
    call scl_check(scl_plan3d_create(nx, nx, nx, SCL_DP, plan64), 'plan3d_create')
    prm%nu = rnu; prm%dt = dt; prm%forced = 1; prm%fkmax = fkmax; prm%eps = eps
    call scl_check(scl_dns_create(plan64, prm, dns), 'dns_create')
    call scl_check(scl_derivs_bind(dns), 'derivs_bind')
    ny = 2 * lx1 * ly * lz * 3
    allocate(y(ny))
    call c_f_pointer(c_loc(y), yc, [lx1, ly, lz, 3])
    yc(:,:,:,1) = vx; yc(:,:,:,2) = vy; yc(:,:,:,3) = vz
    call rk4_classic(scl_derivs, y, ny, 0.0_dp, real(dt, dp), nsteps)
    open(11, file=trim(outfile), access='stream', form='unformatted', status='replace')
    write(11) y
    close(11)
    call scl_check(scl_dns_destroy(dns), 'dns_destroy')
    call scl_check(scl_plan3d_destroy(plan64), 'plan3d_destroy')
  end subroutine derivs_run

end program dns_deform_gpu

! A fixed-step classical RK4 written against the derivs(x, y, dydx, ny)
! interface of rk5.f90:3-10, as a user's driver would be; y is passed by
! address as ny real(dp).
subroutine rk4_classic(derivs, y, ny, x0, h, nsteps)
  use mconstant
  implicit none
  interface
    subroutine derivs(x, y, dydx, n)
      use mconstant
      integer, intent(in) :: n
      real(dp), intent(in) :: x
      real(dp), dimension(n), intent(in) :: y
      real(dp), dimension(n), intent(out) :: dydx
    end subroutine derivs
  end interface
  integer, intent(in) :: ny, nsteps
  real(dp), dimension(ny), intent(inout) :: y
  real(dp), intent(in) :: x0, h
  real(dp), allocatable :: k1(:), k2(:), k3(:), k4(:), yt(:)
  real(dp) :: x
  integer :: i
  allocate(k1(ny), k2(ny), k3(ny), k4(ny), yt(ny))
  x = x0
  do i = 1, nsteps
    call derivs(x, y, k1, ny)
    yt = y + 0.5_dp * h * k1
    call derivs(x + 0.5_dp * h, yt, k2, ny)
    yt = y + 0.5_dp * h * k2
    call derivs(x + 0.5_dp * h, yt, k3, ny)
    yt = y + h * k3
    call derivs(x + h, yt, k4, ny)
    y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    x = x + h
  end do
end subroutine rk4_classic
