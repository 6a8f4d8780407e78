/*
 * scl.h - C ABI of the MI355X pseudo-spectral RHS + IF-RK4 library.
 *
 * Every function returns SCL_OK (0) or a negative status; scl_last_error()
 * gives the message of the last failure on the calling thread.  Arrays are
 * plain pointers: a "field" is the reference's
 *     complex(sp|dp), dimension(lx1, ly, lz)      lx1 = nx/2 + 1
 * column-major, i.e. interleaved (re, im) pairs with x fastest.  Precision
 * is fixed per plan: SCL_SP -> float pairs, SCL_DP -> double pairs.  Sizes
 * are passed by value.  The reference's Fortran module set (fortran/) binds
 * these through iso_c_binding under the reference's own procedure names.
 *
 * Grids: nx, ny, nz powers of two in [16, 1024] (FP32 and FP64 plans).
 * The wavenumber arrays a reference routine takes (kx, ky, kz, k2) must be
 * the wavenumber() ones (wavenumber.f90:89-129); the library checks the 1D
 * arrays and builds k2 itself.
 *
 * Threading: a plan (and a solver built on it) is used from one host thread
 * at a time; distinct plans are independent.  Work runs on the plan's HIP
 * stream; host-array entry points return after their results are on the host.
 */
#ifndef SCL_H
#define SCL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SCL_OK = 0,
  SCL_EINVAL = -1,    /* bad argument (size, null pointer, k arrays) */
  SCL_ENOMEM = -2,    /* device allocation failed */
  SCL_EHIP = -3,      /* HIP runtime error */
  SCL_ENOPLAN = -4,   /* no plan: call scl_fftwplan3d first */
  SCL_ENONFINITE = -5,/* non-finite value found in the solution */
  SCL_ENODEV = -6     /* no usable GPU */
};

enum { SCL_SP = 0, SCL_DP = 1 };

typedef struct scl_plan3d scl_plan3d;
typedef struct scl_dns scl_dns;

const char *scl_last_error(void);
int32_t scl_version(void);

/* ---------------------------------------------------------------------
 * Plan layer.  Replaces module mfftwplan3d (fftwplan.f90:1-31): the
 * reference plans FFTW2 in-place r2c/c2r transforms into module globals
 * R2C3D/C2R3D; here a plan holds the twiddle table, the stream and scratch.
 * ------------------------------------------------------------------- */
/* explicit handle */
int32_t scl_plan3d_create(int32_t nx, int32_t ny, int32_t nz, int32_t prec, scl_plan3d **plan);
int32_t scl_plan3d_destroy(scl_plan3d *plan);
/* Use an existing HIP stream (hipStream_t passed as void*), NULL = own. */
int32_t scl_plan3d_set_stream(scl_plan3d *plan, void *stream);
/* fftwplan3d(ix,iy,iz) (fftwplan.f90:7-15) / fftwplan3de (:17-25): the
 * process-global plan that the reference-named entry points use when passed
 * plan = NULL.  scl_fftwplan3d makes an SP plan (constant_sp.f90). */
int32_t scl_fftwplan3d(int32_t ix, int32_t iy, int32_t iz);
int32_t scl_fftwplan3d_prec(int32_t ix, int32_t iy, int32_t iz, int32_t prec);
/* destroyplan3d (fftwplan.f90:27-30) */
int32_t scl_destroyplan3d(void);

/* rfftwnd_f77_one_complex_to_real(C2R3D, a, ignore) / _real_to_complex
 * (convec_dns.f90:22-27,45-47): unnormalised, in place, on a host array
 * whose real view is (nx+2, ny, nz) (FFTW's padded in-place layout).
 * c2r uses exp(+ik.x), r2c exp(-ik.x); c2r(r2c(a)) = nx*ny*nz*a. */
int32_t scl_c2r3d(scl_plan3d *plan, void *a);
int32_t scl_r2c3d(scl_plan3d *plan, void *a);
/* The same transforms under FFTW 2.x's Fortran-77 names, for callers left
 * unmodified (convec_dns.f90:22-27,45-47 call them by reference): the global
 * plan's grid, plan token and third argument unused, abort() on failure
 * (the reference has no error path there). */
void rfftwnd_f77_one_complex_to_real_(int64_t *plan, void *a, void *ignore);
void rfftwnd_f77_one_real_to_complex_(int64_t *plan, void *a, void *ignore);

/* ---------------------------------------------------------------------
 * RHS blocks on host arrays (reference signatures; plan = NULL uses the
 * global plan).
 * ------------------------------------------------------------------- */
/* wavenumber(kx,ky,kz,k2,lx1,ly,lz) (wavenumber.f90:89-129), host only.
 * kx[lx1], ky[ly], kz[lz], k2[lx1*ly*lz] (k2 may be NULL); real type of prec. */
int32_t scl_wavenumber3d(int32_t prec, void *kx, void *ky, void *kz, void *k2, int32_t lx1, int32_t ly, int32_t lz);

/* convec_dns(vx,vy,vz,wx,wy,wz,lx1,ly,lz,nx,ny,nz) (convec_dns.f90:1-55):
 * w <- FFT(u x w)/(nx ny nz), u = c2r(v), w = c2r(w). */
int32_t scl_convec_dns(scl_plan3d *plan, const void *vx, const void *vy, const void *vz, void *wx, void *wy,
                       void *wz, int32_t lx1, int32_t ly, int32_t lz, int32_t nx, int32_t ny, int32_t nz);

/* force(wx,wy,wz,vx,vy,vz,k2,lx1,ly,lz,fkmax,eps) (force.f90:3-31):
 * w += eps/(2 E_f) v on k2 < fkmax^2.  k2 must be wavenumber()'s. */
int32_t scl_force(scl_plan3d *plan, void *wx, void *wy, void *wz, const void *vx, const void *vy, const void *vz,
                  const void *k2, int32_t lx1, int32_t ly, int32_t lz, double fkmax, double eps);

/* force_heli(fx,fy,fz,vx,vy,vz,wx,wy,wz,k2,lx1,ly,lz,fkmax,eps,eta)
 * (force_heli.f90:1-57): f = a v + b w on k2 < fkmax^2, 0 elsewhere. */
int32_t scl_force_heli(scl_plan3d *plan, void *fx, void *fy, void *fz, const void *vx, const void *vy,
                       const void *vz, const void *wx, const void *wy, const void *wz, const void *k2, int32_t lx1,
                       int32_t ly, int32_t lz, double fkmax, double eps, double eta);

/* symmetrize(c,lx1,ly,lz) (symmetrize.f90:32-64) */
int32_t scl_symmetrize3d(scl_plan3d *plan, void *c, int32_t lx1, int32_t ly, int32_t lz);

/* dealiasing(c,k2,lx1,ly,lz) (dealiasing.f90:2-37) */
int32_t scl_dealiasing(scl_plan3d *plan, void *c, const void *k2, int32_t lx1, int32_t ly, int32_t lz);

/* projection(vx,vy,vz,kx,ky,kz,lx1,ly,lz) (projection.f90:2-31) */
int32_t scl_projection(scl_plan3d *plan, void *vx, void *vy, void *vz, const void *kx, const void *ky,
                       const void *kz, int32_t lx1, int32_t ly, int32_t lz);

/* ratesdudt (dns-deform.f90:259-288): the whole RHS N(v) = P S [u x w + f]
 * into n{x,y,z}.  forced: 0 none, 1 force(), 2 force_heli(). */
int32_t scl_ratesdudt(scl_plan3d *plan, const void *vx, const void *vy, const void *vz, void *nx_, void *ny_,
                      void *nz_, int32_t forced, double fkmax, double eps, double eta);

/* smag(vx,vy,vz,t11,t12,t13,t22,t23,t33,kx,ky,kz,lx1,ly,lz,nx,ny,nz,delta)
 * (smag.f90:1-78): tau_ij = FFT(-2 c1 delta^2 |S| S_ij)/N^3, c1 = 0.026. */
int32_t scl_smag(scl_plan3d *plan, const void *vx, const void *vy, const void *vz, void *t11, void *t12, void *t13,
                 void *t22, void *t23, void *t33, const void *kx, const void *ky, const void *kz, int32_t lx1,
                 int32_t ly, int32_t lz, int32_t nx, int32_t ny, int32_t nz, double delta);

/* dynsmag(ux,uy,uz,t11,...,t33,kx,ky,kz,g_test,lx1,ly,lz,nx,ny,nz,lx,delta,delta_test,mm,lm,cs2,update)
 * (dynsmag.f90:1-236).  kx, ky, kz, g_test are 3D real arrays (lx1,ly,lz);
 * kx/ky/kz must be wavenumber()'s, g_test is used as given.  update != 0:
 * Germano-Lilly cs2 = lm/mm (mm, lm, cs2 returned); update == 0: *cs2 is
 * read (the static branch, :196-233). */
int32_t scl_dynsmag(scl_plan3d *plan, const void *ux, const void *uy, const void *uz, void *t11, void *t12, void *t13,
                    void *t22, void *t23, void *t33, const void *kx, const void *ky, const void *kz,
                    const void *g_test, int32_t lx1, int32_t ly, int32_t lz, int32_t nx, int32_t ny, int32_t nz,
                    int32_t lx, double delta, double delta_test, double *mm, double *lm, double *cs2, int32_t update);

/* convecphi_dns(vx,vy,vz,phi,wphi,kx,ky,kz,lx1,ly,lz,nx,ny,nz)
 * (convecphi_dns.f90:1-49): wphi = i k . FFT(u phi)/N^3, the reference's
 * sign (+div(u phi)). */
int32_t scl_convecphi_dns(scl_plan3d *plan, const void *vx, const void *vy, const void *vz, const void *phi,
                          void *wphi, const void *kx, const void *ky, const void *kz, int32_t lx1, int32_t ly,
                          int32_t lz, int32_t nx, int32_t ny, int32_t nz);

/* 3/2-rule padded products.  padd(u,ub,lx1,ly,lz,lxb1,lyb,lzb)
 * (padd.f90:1-14) and unpadd(ub,u,...) (unpadd.f90:1-13): host-array index
 * copies between a (lx1,ly,lz) half spectrum and a zero-padded
 * (lxb1,lyb,lzb) one, any sizes, in the real type of prec. */
int32_t scl_padd(int32_t prec, const void *u, void *ub, int32_t lx1, int32_t ly, int32_t lz, int32_t lxb1,
                 int32_t lyb, int32_t lzb);
int32_t scl_unpadd(int32_t prec, const void *ub, void *u, int32_t lx1, int32_t ly, int32_t lz, int32_t lxb1,
                   int32_t lyb, int32_t lzb);
/* The padded convec(vx,vy,vz,wx,wy,wz,lx1,ly,lz,lxb1,lyb,lzb,nxb,nyb,nzb)
 * of spectral_heli_dynsmag.f90:229 (plans of fftwplan_les.f90:11-21):
 * w <- unpadd(FFT_b(u x w))/(nxb nyb nzb), u, w from the padded inverse
 * transforms.  The padded sizes must be at least 3/2 of the plan's; the
 * library pads to the next power of two per axis (2 n), on which the
 * product is alias-free on every mode unpadd keeps, as on the 3n/2 grid. */
int32_t scl_convec_padded(scl_plan3d *plan, const void *vx, const void *vy, const void *vz, void *wx, void *wy,
                          void *wz, int32_t lx1, int32_t ly, int32_t lz, int32_t lxb1, int32_t lyb, int32_t lzb,
                          int32_t nxb, int32_t nyb, int32_t nzb);
/* smag(vx,vy,vz,tau11,...,tau33,kx,ky,kz,lx1,ly,lz,lxb1,lyb,lzb,nxb,nyb,nzb,delta)
 * of smag_dealias.f90:1-78 (c1 = 0.026, kx/ky/kz 3D arrays), with that
 * file's strain (off-diagonal terms without the 1/2): tau_ij =
 * unpadd(FFT_b(-2 c1 delta^2 |S| S_ij))/(padded points), on the padded grid
 * of scl_convec_padded.  |S| S_ij is not a polynomial, so its aliasing on
 * that grid differs from the 3n/2 grid's at the level of the field's
 * spectrum near the padded cutoff. */
int32_t scl_smag_dealias(scl_plan3d *plan, const void *vx, const void *vy, const void *vz, void *t11, void *t12,
                         void *t13, void *t22, void *t23, void *t33, const void *kx, const void *ky, const void *kz,
                         int32_t lx1, int32_t ly, int32_t lz, int32_t lxb1, int32_t lyb, int32_t lzb, int32_t nxb,
                         int32_t nyb, int32_t nzb, double delta);

/* skewness(ux,kx,s,nx,ny,nz) (skewness.f90:1-48): s = <d^3>/<d^2>^(3/2),
 * d = du_x/dx in physical space, summed in FP64.  kx must be wavenumber()'s. */
int32_t scl_skewness(scl_plan3d *plan, const void *ux, const void *kx, double *s, int32_t nx, int32_t ny, int32_t nz);

/* Helical initial field (initialize_heli.f90:1-109): per mode two complex
 * Gaussians (am, then ap) from ran1 pairs in (z, y, x) order; am rescaled
 * per shell so that sum |am|^2 = (1 - hrel) sum |ap|^2 (the file's shell
 * masks); u = sqrt(3/2) (ap h+ + am h-) u0 sqrt(8 sqrt(2/pi)/(3 pi akp^5))
 * k exp(-k^2/akp^2) with the helical basis i k x h+ = |k| h+; symmetrize;
 * sum' |u|^2 = 1.5 u0^2.  Host arrays in the reference layout. */
int32_t scl_initialize_heli(int32_t prec, void *ux, void *uy, void *uz, int32_t nx, int32_t ny, int32_t nz,
                            int32_t iseed, double hrel, double akp, double u0);
/* initialize(vx,vy,vz,kx,ky,kz,k2,iseed,lx1,ly,lz,akp,u0,kcut)
 * (initialize.f90:373-463 draw order, :328-332 complex projection), host
 * only, in the real type of prec.  kcut <= 0 selects 2*lx/3. */
int32_t scl_initialize3d(int32_t prec, void *vx, void *vy, void *vz, int32_t nx, int32_t ny, int32_t nz,
                         int32_t iseed, double akp, double u0, double kcut);

/* One Lawson IF-RK4 step + 2/3 truncation on host arrays, the inline step
 * of dns-deform.f90:160-237 as a subroutine. */
int32_t scl_rk4step(scl_plan3d *plan, void *vx, void *vy, void *vz, double nu, double dt, int32_t forced,
                    double fkmax, double eps, double eta);

/* ---------------------------------------------------------------------
 * Device-resident solver (the dns-deform.f90 loop body, :113-241).
 * ------------------------------------------------------------------- */
typedef struct scl_dns_params {
  double nu;           /* rnu */
  double dt;
  int32_t forced;      /* 0 none, 1 constant power (force.f90), 2 helical (force_heli.f90) */
  double fkmax, eps, eta;
  int32_t dealias_rhs; /* 1: also zero k2 >= kcut^2 in every RHS (default 0: once per step) */
  int32_t use_graph;   /* 1: replay one captured step as a hipGraph (not with sgs = 2) */
  /* LES (dynsmag.f90, smag.f90; RHS as spectral_les.f90:205-238) */
  int32_t sgs;         /* 0 none, 1 static Smagorinsky, 2 dynamic (Germano-Lilly) */
  double cs2;          /* static C_s^2 (smag.f90's c1 = 0.026); dynamic: start value */
  double delta;        /* filter width; <= 0 selects 2 pi / nx */
  double delta_test;   /* dynamic: test-filter width; <= 0 selects 2 delta */
  double k_test;       /* dynamic: test filter passes k2 <= k_test^2; <= 0 selects (2 nx/6)/2 */
  int32_t iupdate;     /* dynamic: new cs2 at the first stage of every iupdate-th step */
  /* passive scalar (convecphi_dns.f90, spectral_dns_scalar.f90:236-272) */
  int32_t scalar;      /* 1: advance phi with -div(u phi) - grad * u_z, diffusivity nu/schmidt */
  double grad;         /* mean scalar gradient along z */
  double schmidt;      /* Sc (Prandtl number pr of the reference) */
  /* time stepper (DNS, LES and scalar alike: steppers 1 and 2 are also the
   * loops of spectral_les.f90:241-304 and spectral_dns_scalar.f90:257-338,
   * the scalar with the factors exp(-nu/Sc k^2 dt)) */
  int32_t stepper;     /* 0 Lawson IF-RK4 (dns-deform.f90:160-237);
                          1 IF-AB2 with CFL steps (spectral_dns_dealiased.f90:124-296):
                            an Euler step, a leapfrog step, then variable-step
                            Adams-Bashforth, dt computed on the device;
                          2 IF-AB2 with the fixed dt of .dt and a half-step bootstrap
                            (specdns-forced-omp-ft.f90:224-274) */
  double beta;         /* stepper 1: dt = min(beta (pi/lx) / umax, 3 dto) */
  double dtw;          /* stepper 1: steps end on the output times (twrite, advanced by
                          dtw each time it is reached); <= 0: no landing */
  int32_t ieout;       /* stepper 1: umax = 3.5 sqrt(2 E) at steps that are multiples of
                          ieout; <= 0: never */
  double umax;         /* stepper 1: umax until the first refresh (3 sqrt(3) u0 in the reference) */
  double dt_first;     /* stepper 1: the first step's dto (timemax in the reference); <= 0: none */
  double kcut;         /* truncation after each step, k^2 >= kcut^2 zeroed; <= 0: 2 lx/3
                          (spectral_dns_dealiased.f90 uses lx - 1) */
} scl_dns_params;

typedef struct scl_diag {
  double energy;       /* sum' |v|^2, weight 1/2 on kx=0 (dns-deform.f90:124-126) */
  double dissipation;  /* 2 nu sum' k^2 |v|^2 */
  double enstrophy;    /* sum' k^2 |v|^2 */
  double helicity;     /* sum' 2 Re(w . conj v) (helispec.f90:82) */
  double divmax;       /* max over the grid of |div u| (physical space) */
  double divrms_spec;  /* sqrt(sum' |k.v|^2) */
  int32_t nonfinite;   /* 1 if any NaN/Inf in v */
  double scalar_variance;    /* sum' |phi|^2 (= <phi^2>/2) */
  double scalar_dissipation; /* 2 (nu/Sc) sum' k^2 |phi|^2 */
  double cs2;                /* current Smagorinsky coefficient (sgs > 0) */
  double helicity_dissipation; /* 2 nu sum' k^2 2 Re(w . conj v) */
  double skewness;             /* <(du/dx)^3> / <(du/dx)^2>^(3/2) (skewness.f90:1-48) */
} scl_diag;

int32_t scl_dns_create(scl_plan3d *plan, const scl_dns_params *p, scl_dns **out);
int32_t scl_dns_destroy(scl_dns *d);
int32_t scl_dns_set_velocity(scl_dns *d, const void *vx, const void *vy, const void *vz);
/* passive scalar (scalar = 1), host array in the reference layout */
int32_t scl_dns_set_scalar(scl_dns *d, const void *phi);
int32_t scl_dns_get_scalar(scl_dns *d, void *phi);
int32_t scl_dns_get_velocity(scl_dns *d, void *vx, void *vy, void *vz);
int32_t scl_dns_initialize(scl_dns *d, int32_t iseed, double akp, double u0);
/* The solver's state from scl_initialize_heli (spectral_heli_dns.f90:92). */
int32_t scl_dns_initialize_heli(scl_dns *d, int32_t iseed, double hrel, double akp, double u0);
/* nsteps IF-RK4 steps, asynchronous on the plan's stream */
int32_t scl_dns_step(scl_dns *d, int32_t nsteps);
int32_t scl_dns_sync(scl_dns *d);
/* The solver's time (IF-AB2: read from the device after the queued steps). */
double scl_dns_time(const scl_dns *d);
/* Time, the last step's dt and the steps taken (IF-AB2: since the state was
 * set), after the queued steps. */
int32_t scl_dns_step_info(scl_dns *d, double *time, double *dt_last, int64_t *steps);
/* E(ll), ll = 1..nbins (nbins <= nx/2), FP64 (spec.f90:66-89).  The
 * diagnostics sum in an order fixed by the grid: repeated calls on one state
 * give the same bits. */
int32_t scl_dns_spectrum(scl_dns *d, double *ek, int32_t nbins);
/* H(ll) (helispec.f90:67-91) */
int32_t scl_dns_helicity_spectrum(scl_dns *d, double *hk, int32_t nbins);
int32_t scl_dns_diagnostics(scl_dns *d, scl_diag *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the spectra and totals of scl_dns_spectrum,
 * scl_dns_helicity_spectrum and scl_dns_diagnostics, from the same kernels
 * (bitwise the same values), with no device allocation.  out is a device
 * array of 2 nbins + 8 doubles the caller owns: E(1..nbins), H(1..nbins),
 * then sum'|v|^2, sum'k^2|v|^2, sum' 2 Re(w.conj v), sum'|k.v|^2, the count
 * of non-finite entries, sum'k^2 2 Re(w.conj v), 0, 0.  nbins in [1, nx/2],
 * at most 512 (the partial sums use a 4 MiB static device array per GPU).
 * Asynchronous on the plan's stream (no host synchronisation), capturable.
 * Calls that overlap on one GPU (two streams at once) share that array and
 * race; serialise them. */
int32_t scl_dns_spectra_noalloc(scl_dns *d, double *out, int32_t nbins);

/* Per-kernel timing of nsteps steps with HIP events on the plan's stream
 * around every launch (direct launches, also when use_graph is set: event
 * records captured into a graph cannot be timed).  kind_ms[k] /
 * kind_count[k] for k < SCL_NKIND, total_ms = event time of the whole run.
 * The DNS solver chains its stages: the z-forward of stages 1..3 also
 * z-inverses the next stage state (SCL_K_ZFWD_ZINV) and that of stage 4 the
 * new v for the next step (SCL_K_ZFWD4_ZINV), so only the first step after
 * the state is set runs a z-inverse pass (SCL_K_ZINV) of its own.
 * SCL_K_ZFWD: unchained z-forwards (host-array steps, LES / scalar). */
enum {
  SCL_K_FORCE = 0, SCL_K_ZINV = 1, SCL_K_YINV = 2, SCL_K_XPASS = 3, SCL_K_YFWD = 4, SCL_K_ZFWD = 5,
  SCL_K_HERM = 6, SCL_K_ZFWD_ZINV = 7, SCL_K_ZFWD4_ZINV = 8, SCL_NKIND = 9
};
int32_t scl_dns_step_timed(scl_dns *d, int32_t nsteps, double *kind_ms, int32_t *kind_count, double *total_ms);

/* ---------------------------------------------------------------------
 * RHS-callback form (rk5.f90:1-12: derivs(x, y, dydx, ny); the driver owns
 * the stepping, derivs the physics).
 * ------------------------------------------------------------------- */
/* dydt = N(y) - nu k^2 y (scalar: nu/Sc) for the solver's configuration:
 * y and dydt hold the solver's nc fields (3 velocity [+ scalar]) in the
 * reference layout, on the host (on_device = 0) or the device.  t is unused
 * (the system is autonomous). */
int32_t scl_dns_rhs(scl_dns *d, double t, const void *const *y, void *const *dydt, int32_t on_device);
/* The solver's state on the device: nc fields of the split layout, *n
 * complex entries in all (the layout of scl_dns_derivs_cb's arrays).  The
 * caller may change the state through the pointer before the next step. */
int32_t scl_dns_state(scl_dns *d, void **state, int64_t *n);
/* nu k^2 (scalar: nu/Sc k^2) per state entry into a device array of *n reals. */
int32_t scl_dns_rates(scl_dns *d, void *rate);
/* The callback type of scl_ifrk4_integrate: dydt = f(t, y) on device arrays;
 * returns after dydt is complete. */
typedef int32_t (*scl_derivs_fn)(void *ctx, double t, const void *y, void *dydt);
/* scl_dns_rhs's N(y) without the linear term, as a callback: ctx = scl_dns*,
 * arrays in the split layout of scl_dns_state. */
int32_t scl_dns_derivs_cb(void *ctx, double t, const void *y, void *dydt);
/* Lawson IF-RK4 (dns-deform.f90:160-237's scheme) of y' = -rate y + f(t, y)
 * over n complex device entries: y in place, rate a device array of n reals
 * in the precision prec (NULL: classical RK4), f called four times per step. */
int32_t scl_ifrk4_integrate(int32_t prec, int64_t n, void *y, const void *rate, double t0, double dt,
                            int32_t nsteps, scl_derivs_fn f, void *ctx);
/* derivs(x, y, dydx, ny) for Fortran callers (external scl_derivs): y and
 * dydx are the bound FP64 solver's nc fields as real(dp), ny reals;
 * failures abort, as the reference STOPs. */
int32_t scl_derivs_bind(scl_dns *d);
void scl_derivs_(const double *x, const double *y, double *dydx, const int32_t *ny);

/* ---------------------------------------------------------------------
 * 2D vorticity ensembles.  Replaces the time loop of spectral_dns_2d.f90
 * (:94 integrating factor, :161-188 RHS) for many independent members of
 * several grid sizes at once.  A member's field is the reference's
 * complex(sp|dp), dimension(lx1, ly) vorticity wz, lx1 = n/2 + 1, ly = n.
 *
 * RHS: N = -(i kx (u w)^ + i ky (v w)^) + band forcing eps/(conj(w) + tiny)
 * on fkmin^2 <= k^2 <= fkmax^2, then symmetrize2d and the 2/3 cut, with
 * u = i ky w/k^2, v = -i kx w/k^2.  The reference negates the divergence only
 * inside the forced band (:181-185); here the sign is that of -div(u w)
 * everywhere.  Step: Lawson IF-RK4 with E = exp(-rate dt/2),
 * rate = c_hyper k^n_hyper + c_hypo k^-n_hypo + damp (the reference's
 * (k sigma)^nhyper + (nu/k)^nhypo is c_hyper = sigma^nhyper,
 * c_hypo = nu^nhypo, damp = 0), in place of the reference's IF-AB2
 * (:192-219).  Sizes: powers of two in [32, 1024] (FP32 and FP64).
 *
 * per_member = 0: every size group advances as one batch (one set of
 * launches per group, the groups on parallel branches of one hipGraph per
 * step).  per_member = 1: every member is its own solver with its own
 * launches, for comparison.
 * ------------------------------------------------------------------- */
typedef struct scl_ens2d scl_ens2d;
typedef struct scl_ens2d_params {
  double dt;
  double c_hyper;   /* coefficient of k^n_hyper */
  double c_hypo;    /* coefficient of k^-n_hypo (0: none) */
  double damp;      /* linear damping rate */
  double fkmin;     /* forced band fkmin <= |k| <= fkmax */
  double fkmax;
  double eps;       /* forcing amplitude (spectral_dns_2d.f90:183) */
  int32_t n_hyper;
  int32_t n_hypo;
  int32_t forced;
  int32_t per_member;
  int32_t use_graph;
  /* scheme 0: spectral_dns_2d.f90's RHS (2/3 rule) stepped with IF-RK4,
   * the fields above.  scheme 1: turb_2d.f90 at one rank (turb_2d.f90:
   * 287-483): phase-shift dealiased nonlinear term with the 8/9 cut, rate
   * (k sigma)^16 + (k/nu)^-4, forcing etak/conj(w) on kf <= |k| <= kf + 3
   * (etak = eta/104), Euler and midpoint half steps then IF-AB2; members
   * are that file's wz(my, mx2) arrays.  Scheme 1 uses only dt, forced,
   * per_member, use_graph and the four fields below. */
  int32_t scheme;
  double sigma;
  double nu;
  double eta;
  double kf;
} scl_ens2d_params;

/* ngroups groups: sizes[g] x sizes[g] grids, members[g] members each. */
int32_t scl_ens2d_create(int32_t prec, int32_t ngroups, const int32_t *sizes, const int32_t *members,
                         const scl_ens2d_params *params, scl_ens2d **ens);
int32_t scl_ens2d_destroy(scl_ens2d *ens);
/* Host copy in/out of member j of group g: scheme 0 the reference's
 * w(lx1, ly); scheme 1 turb_2d.f90's wz(my, mx2) (column kx = 0 .. n/2-1
 * of my entries each), symmetrized as turb_2d.f90:220 on the way in. */
int32_t scl_ens2d_set_member(scl_ens2d *ens, int32_t group, int32_t member, const void *w);
int32_t scl_ens2d_get_member(scl_ens2d *ens, int32_t group, int32_t member, void *w);
/* Seeded random-phase vorticity (initialize.f90:465-500 draw order: one
 * gasdev per mode, y outer, x inner; symmetrize2d; 2/3 cut), shaped to
 * E(k) ~ k exp(-(k/k0)^2) and scaled to sum' |w|^2/k^2 = energy.  Member q
 * in group-major order draws from seed0 - q (seed0 < 0). */
int32_t scl_ens2d_initialize(scl_ens2d *ens, int32_t seed0, double k0, double energy);
/* Sets member j of group g to w, evaluates N for it, copies N to rhs
 * (scheme 1: vy of turb_2d.f90:428-429, in its layout and convention). */
int32_t scl_ens2d_rhs(scl_ens2d *ens, int32_t group, int32_t member, const void *w, void *rhs);
int32_t scl_ens2d_step(scl_ens2d *ens, int32_t nsteps);
int32_t scl_ens2d_sync(scl_ens2d *ens);
/* nsteps steps bracketed by HIP events; *ms = elapsed milliseconds. */
int32_t scl_ens2d_step_timed(scl_ens2d *ens, int32_t nsteps, double *ms);
int32_t scl_ens2d_time(scl_ens2d *ens, double *time);
/* FP64 per member, group-major: out[2q] = sum' |w|^2/k^2 (the reference's
 * ek, spectral_dns_2d.f90:125-128), out[2q+1] = sum' |w|^2 (kx = 0 halved). */
int32_t scl_ens2d_diagnostics(scl_ens2d *ens, double *out);
/* Column schedule of group g: points per thread e (8, 16), columns per
 * workgroup cw (8, 16), x-pass grid row_grid (0 persistent, 1 one
 * workgroup per row group). */
int32_t scl_ens2d_set_schedule(scl_ens2d *ens, int32_t group, int32_t e, int32_t cw, int32_t row_grid);
int32_t scl_ens2d_get_schedule(scl_ens2d *ens, int32_t group, int32_t *e, int32_t *cw, int32_t *row_grid);
/* Autotuner: times reps steps of every valid schedule of every group on the
 * device, keeps the fastest; best_ms[g] gets its time per step (may be
 * NULL).  The state is saved before and restored after each candidate, so
 * it is not changed. */
int32_t scl_ens2d_tune(scl_ens2d *ens, int32_t reps, double *best_ms);

/* ---------------------------------------------------------------------
 * Discrete adjoint of nsteps IF-RK4 steps (dns-deform.f90:160-237) of the
 * unforced momentum equation, DNS (sgs = 0) or static Smagorinsky LES
 * (sgs = 1: spectral_les.f90:205-238 with dynsmag.f90:196-233's static
 * branch), on the plan's grid and precision.  No counterpart in the
 * reference, which has no adjoint; the forward map is scl_dns_step's.
 * Gradients are in the real inner product of the stored half spectrum:
 * per stored entry dL/dRe + i dL/dIm.  Checkpoints every ckpt_every steps;
 * the backward sweep recomputes one segment at a time into a tape of stage
 * states.  on_device = 1: every array argument is a device pointer (field
 * in the reference layout), 0: host pointers.
 * ------------------------------------------------------------------- */
typedef struct scl_adjoint scl_adjoint;
int32_t scl_adjoint_create(scl_plan3d *plan, const scl_dns_params *params, int32_t nsteps, int32_t ckpt_every,
                           scl_adjoint **adj);
int32_t scl_adjoint_destroy(scl_adjoint *adj);
/* u(0) -> u(T) with cs2 = C_s^2; keeps the checkpoints for the backward sweep. */
int32_t scl_adjoint_forward(scl_adjoint *adj, double cs2, const void *u0x, const void *u0y, const void *u0z,
                            void *uTx, void *uTy, void *uTz, int32_t on_device);
/* dL/du(T) -> dL/du(0) and dL/dcs2 for the last forward's u(0) and cs2. */
int32_t scl_adjoint_backward(scl_adjoint *adj, const void *gx, const void *gy, const void *gz, void *g0x, void *g0y,
                             void *g0z, double *dcs2, int32_t on_device);
/* One vector-Jacobian product of the RHS: out = J(s)^T nbar, *dcs2 = <nbar, dN/dcs2>. */
int32_t scl_adjoint_vjp(scl_adjoint *adj, double cs2, const void *sx, const void *sy, const void *sz,
                        const void *nx, const void *ny, const void *nz, void *ox, void *oy, void *oz, double *dcs2,
                        int32_t on_device);
/* Event-timed milliseconds of the last forward and backward calls. */
int32_t scl_adjoint_timing(scl_adjoint *adj, double *forward_ms, double *backward_ms);

/* ---------------------------------------------------------------------
 * Files of the reference's runs (host side).
 * ------------------------------------------------------------------- */
/* One Fortran unformatted sequential record (4-byte length markers; records
 * over 2^31 - 9 bytes split into subrecords as gfortran writes them), as
 * output.f90:103-166 writes a field with write(10) and input.f90:106-170
 * reads it.  scl_record_read requires the file to be exactly one record of
 * nbytes: a short, truncated or corrupted file returns SCL_EINVAL. */
int32_t scl_record_write(const char *path, const void *data, int64_t nbytes);
int32_t scl_record_read(const char *path, void *data, int64_t nbytes);
/* output(vx,vy,vz[,phi],idump,...) (output.f90:103-166): the solver's
 * fields to <dir>/ux<idump>.dat, uy, uz (and phi with a scalar), one record
 * each in the reference layout; dir NULL = ".".  scl_dns_input
 * (input.f90:106-170) reads them back (every file checked before the state
 * is set); a dump written after step k and read into a solver of the same
 * configuration continues the IF-RK4 run bit-identically. */
int32_t scl_dns_output(scl_dns *d, const char *dir, int32_t idump);
int32_t scl_dns_input(scl_dns *d, const char *dir, int32_t idump);
/* Appends one line of ./post/ener_time.data: write(25,'(I6,20E12.4)')
 * istep, vals(1:nvals) (spectral_dns_dealiased.f90:136, nvals <= 20). */
int32_t scl_write_ener_time(const char *path, int32_t istep, const double *vals, int32_t nvals);
/* Appends one block of ./post/spectrum.data: the time, then "ll E(ll)" for
 * ll = 1..n, list-directed (spectral_dns_dealiased.f90:158-171). */
int32_t scl_write_spectrum(const char *path, double time, const double *ek, int32_t n);

/* ---------------------------------------------------------------------
 * Measurement.  No counterpart in the reference: the HBM rate a plain
 * streaming kernel reaches on the current device (16-byte vector accesses,
 * grid-stride, buffers of `bytes` each), the practical ceiling next to the
 * solver's per-pass rates.  kind 0: read `bytes`; 1: write `bytes`; 2: copy
 * (read and write `bytes`).  Allocates and frees its own buffers; *ms_per_rep
 * = event time of one pass, averaged over reps after one warm-up pass.
 * ------------------------------------------------------------------- */
int32_t scl_stream_probe(int32_t kind, int64_t bytes, int32_t reps, double *ms_per_rep);

#ifdef __cplusplus
}
#endif

#endif /* SCL_H */
