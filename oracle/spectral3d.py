"""Oracle: CPU restatement of the reference's pseudo-spectral RHS + IF-RK4 path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import this module, and only as the checker
(or, for the baseline, as the CPU path timed beside the GPU).  The product
package never imports it.

Array convention: a reference array ``complex(sp), dimension(lx1, ly, lz)``
(column-major) has exactly the bytes of a C-ordered numpy array of shape
``(lz, ly, lx1)``; every field here uses that shape, so index ``[z, y, x]``.

FFT convention (fftwplan.f90:7-15, FFTW 2.x rfftwnd, in place, unnormalised):
    c2r(X)[n] = sum_k X[k] exp(+i k.x_n)      (no 1/N)
    r2c(x)[k] = sum_n x[n] exp(-i k.x_n)
so c2r(r2c(x)) = nx*ny*nz * x.  The FFT engine is numpy.fft (pocketfft), an
independent library, or scipy.fft with worker threads for the CPU baseline.

Each function cites the reference lines it restates.  What pins each one
(closed forms, invariants, brute force) lives in tests/test_oracle_*.py; the
oracle header of every function says which test pins it.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------
# FFT backend
# --------------------------------------------------------------------------
_FFT = {"mod": np.fft, "kw": {}}


def set_fft_backend(name: str = "numpy", workers: int | None = None) -> None:
    """Select numpy.fft (default) or scipy.fft(workers=...) for c2r/r2c."""
    if name == "numpy":
        _FFT["mod"], _FFT["kw"] = np.fft, {}
    elif name == "scipy":
        import scipy.fft as sfft

        _FFT["mod"], _FFT["kw"] = sfft, ({"workers": workers} if workers else {})
    else:
        raise ValueError(name)


def c2r(c: np.ndarray, nx: int) -> np.ndarray:
    """Unnormalised complex-to-real 3D (rfftwnd_f77_one_complex_to_real,
    convec_dns.f90:22-27).  Input (nz, ny, nx/2+1) -> real (nz, ny, nx).
    Like FFTW's halfcomplex format, Im of the kx=0 and kx=nx/2 entries of each
    x-line is ignored.  Pinned by tests/test_oracle_fft.py (brute-force DFT)."""
    nz, ny, _ = c.shape
    return _FFT["mod"].irfftn(c, s=(nz, ny, nx), axes=(0, 1, 2), norm="forward", **_FFT["kw"])


def r2c(r: np.ndarray) -> np.ndarray:
    """Unnormalised real-to-complex 3D (rfftwnd_f77_one_real_to_complex,
    convec_dns.f90:45-47).  Pinned by tests/test_oracle_fft.py."""
    return _FFT["mod"].rfftn(r, axes=(0, 1, 2), **_FFT["kw"])


def c2r2(c: np.ndarray, nx: int) -> np.ndarray:
    """2D counterpart (fftwplan.f90:39-47): (ny, nx/2+1) -> (ny, nx)."""
    ny, _ = c.shape
    return _FFT["mod"].irfftn(c, s=(ny, nx), axes=(0, 1), norm="forward", **_FFT["kw"])


def r2c2(r: np.ndarray) -> np.ndarray:
    return _FFT["mod"].rfftn(r, axes=(0, 1), **_FFT["kw"])


# --------------------------------------------------------------------------
# Precision (constant_sp.f90 / constant_dp.f90)
# --------------------------------------------------------------------------
def real_dtype(prec: str):
    return np.float32 if prec == "sp" else np.float64


def cplx_dtype(prec: str):
    return np.complex64 if prec == "sp" else np.complex128


def tiny(prec: str) -> float:
    """mytiny = tiny(1.0_sp) (constant_sp.f90:14) or tiny(1.0_dp)."""
    return float(np.finfo(real_dtype(prec)).tiny)


# The reference's pi literal (constant_sp.f90:20, constant_dp.f90:17).
PI_REF = 3.1415926536


# --------------------------------------------------------------------------
# Wavenumbers (wavenumber.f90:89-129, wavenumberlowmem3d)
# --------------------------------------------------------------------------
def wavenumber(nx: int, ny: int, nz: int, prec: str = "dp"):
    """kx = 0..lx; ky, kz in FFTW order 0..n/2-1, -n/2..-1; k2 with
    k2(1,1,1) = tiny (wavenumber.f90:102-125).  Pinned by
    tests/test_oracle_basic.py::test_wavenumber_closed_form."""
    rd = real_dtype(prec)
    lx1 = nx // 2 + 1
    kx = np.arange(lx1, dtype=rd)
    ky = np.array([((j + ny // 2) % ny) - ny // 2 for j in range(ny)], dtype=rd)
    kz = np.array([((k + nz // 2) % nz) - nz // 2 for k in range(nz)], dtype=rd)
    k2 = (kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + kz[:, None, None] ** 2).astype(rd)
    k2[0, 0, 0] = tiny(prec)
    return kx, ky, kz, k2


# --------------------------------------------------------------------------
# Symmetrize / dealiasing (symmetrize.f90:32-64, dealiasing.f90:2-37)
# --------------------------------------------------------------------------
def symmetrize3d(c: np.ndarray) -> np.ndarray:
    """In place.  Zero ky=-ny/2, kz=-nz/2, kx=nx/2 planes and the mean
    (symmetrize.f90:41-48); on the kx=0 plane copy conj partners into
    kz<0 (symmetrize.f90:51-56) and, on kz=0, into ky<0 (:59-61).
    Pinned by tests/test_oracle_basic.py (real-consistency, idempotence)."""
    lz, ly, lx1 = c.shape
    c[:, ly // 2, :] = 0
    c[lz // 2, :, :] = 0
    c[:, :, lx1 - 1] = 0
    c[0, 0, 0] = 0
    z = np.arange(lz // 2 + 1, lz)
    y = np.arange(ly)
    c[z[:, None], y[None, :], 0] = np.conj(c[(lz - z)[:, None], ((ly - y) % ly)[None, :], 0])
    y = np.arange(ly // 2 + 1, ly)
    c[0, y, 0] = np.conj(c[0, ly - y, 0])
    return c


def symmetrize2d(c: np.ndarray) -> np.ndarray:
    """In place (symmetrize.f90:8-30): zero ky=-ny/2 row, kx=nx/2 column and
    the mean; kx=0 column gets conj partners for ky<0."""
    ly, lx1 = c.shape
    c[ly // 2, :] = 0
    c[:, lx1 - 1] = 0
    c[0, 0] = 0
    y = np.arange(ly // 2 + 1, ly)
    c[y, 0] = np.conj(c[ly - y, 0])
    return c


def dealiasing(c: np.ndarray, k2: np.ndarray) -> np.ndarray:
    """symmetrize3d then zero k2 >= ly^2/9 (dealiasing.f90:13-35)."""
    symmetrize3d(c)
    ly = c.shape[1]
    c[k2 >= np.asarray(ly * ly, k2.dtype) / 9] = 0
    return c


# --------------------------------------------------------------------------
# Projection (projection.f90:16-30)
# --------------------------------------------------------------------------
def projection(vx, vy, vz, kx, ky, kz, prec="dp"):
    """In place: v -= k (k.v)/(k^2 + mytiny).  Pinned by
    tests/test_oracle_basic.py (k.Pv = 0, P^2 = P, P(k phi) = 0)."""
    KX = kx[None, None, :]
    KY = ky[None, :, None]
    KZ = kz[:, None, None]
    t = (KX * vx + KY * vy + KZ * vz) / (KX * KX + KY * KY + KZ * KZ + np.asarray(tiny(prec), kx.dtype))
    vx -= KX * t
    vy -= KY * t
    vz -= KZ * t
    return vx, vy, vz


# --------------------------------------------------------------------------
# Vorticity (dns-deform.f90:265-273) and u x w (convec_dns.f90:15-53)
# --------------------------------------------------------------------------
def vorticity(vx, vy, vz, kx, ky, kz):
    """w = i k x v (dns-deform.f90:268-270)."""
    KX = kx[None, None, :]
    KY = ky[None, :, None]
    KZ = kz[:, None, None]
    wx = 1j * KY * vz - 1j * KZ * vy
    wy = 1j * KZ * vx - 1j * KX * vz
    wz = 1j * KX * vy - 1j * KY * vx
    dt = vx.dtype
    return wx.astype(dt), wy.astype(dt), wz.astype(dt)


def convec_dns(vx, vy, vz, wx, wy, wz, nx, ny, nz):
    """Returns FFT(u x w)/(nx ny nz): c2r of the six inputs (:22-27), the
    packed-real cross product (:30-39) — in the unpacked real domain it is the
    ordinary cross product — scaled by const = 1/(nx ny nz) (:15,40-42), then
    r2c (:45-47).  Pinned by tests/test_oracle_convec.py (Beltrami u x w = 0,
    Taylor-Green closed form, brute-force direct sums on 8^3)."""
    rd = np.float32 if vx.dtype == np.complex64 else np.float64
    const = rd(1.0) / rd(nx * ny * nz)
    ux, uy, uz = (c2r(a, nx) for a in (vx, vy, vz))
    ox, oy, oz = (c2r(a, nx) for a in (wx, wy, wz))
    px = (uy * oz - uz * oy) * const
    py = (uz * ox - ux * oz) * const
    pz = (ux * oy - uy * ox) * const
    return r2c(px).astype(vx.dtype), r2c(py).astype(vx.dtype), r2c(pz).astype(vx.dtype)


# --------------------------------------------------------------------------
# Forcing (force.f90:16-29, force_heli.f90:15-55)
# --------------------------------------------------------------------------
def _half_weight_sum(t: np.ndarray) -> float:
    """sum over the half spectrum with weight 1/2 on the kx=0 plane
    (force.f90:21-22)."""
    return float(t.sum() - 0.5 * t[:, :, 0].sum())


def force_coeff(vx, vy, vz, k2, fkmax, eps) -> float:
    """a = eps / (2 E_f), E_f = sum_{k2 < fkmax^2} |v|^2 (half weight kx=0)
    (force.f90:16-24).  Accumulated in FP64."""
    m = k2 < np.asarray(fkmax * fkmax, k2.dtype)
    t = np.where(m, (np.abs(vx) ** 2 + np.abs(vy) ** 2 + np.abs(vz) ** 2).astype(np.float64), 0.0)
    kinet = _half_weight_sum(t)
    return eps / (2.0 * kinet)


def force(wx, wy, wz, vx, vy, vz, k2, fkmax, eps):
    """In place: w += a v on k2 < fkmax^2 (force.f90:25-29).  Pinned by
    tests/test_oracle_force.py (injected power = eps)."""
    a = force_coeff(vx, vy, vz, k2, fkmax, eps)
    m = k2 < np.asarray(fkmax * fkmax, k2.dtype)
    a = np.asarray(a, dtype=k2.dtype)
    wx[m] += a * vx[m]
    wy[m] += a * vy[m]
    wz[m] += a * vz[m]
    return a


def force_heli_coeffs(vx, vy, vz, wx, wy, wz, k2, fkmax, eps, eta):
    """(a, b, kinet, enst, uw) of force_heli.f90:15-42, accumulated in FP64."""
    m = k2 < np.asarray(fkmax * fkmax, k2.dtype)

    def hs(t):
        return _half_weight_sum(np.where(m, t.astype(np.float64), 0.0))

    kinet = hs(np.abs(vx) ** 2 + np.abs(vy) ** 2 + np.abs(vz) ** 2)
    enst = hs(np.abs(wx) ** 2 + np.abs(wy) ** 2 + np.abs(wz) ** 2)
    uw = hs(2.0 * np.real(vx * np.conj(wx) + vy * np.conj(wy) + vz * np.conj(wz)))
    den = 4.0 * kinet * enst - uw * uw
    a = (4.0 * eps * enst - eta * uw) / (2.0 * den)
    b = (eta * kinet - eps * uw) / den
    return a, b, kinet, enst, uw


def force_heli(vx, vy, vz, wx, wy, wz, k2, fkmax, eps, eta):
    """f = a v + b w on k2 < fkmax^2, 0 elsewhere, f(0) = 0
    (force_heli.f90:44-55).  Pinned by tests/test_oracle_force.py (energy
    injection = eps, helicity injection = eta)."""
    a, b, *_ = force_heli_coeffs(vx, vy, vz, wx, wy, wz, k2, fkmax, eps, eta)
    m = k2 < np.asarray(fkmax * fkmax, k2.dtype)
    rd = k2.dtype.type
    out = []
    for v, w in ((vx, wx), (vy, wy), (vz, wz)):
        f = np.where(m, rd(a) * v + rd(b) * w, 0).astype(vx.dtype)
        f[0, 0, 0] = 0
        out.append(f)
    return tuple(out)


# --------------------------------------------------------------------------
# RHS and IF-RK4 (dns-deform.f90:160-237, 259-288)
# --------------------------------------------------------------------------
@dataclass
class DNSParams:
    """The dns-deform.f90 parameters (parameter_dns_deform.d, :28-46)."""

    nx: int
    ny: int | None = None
    nz: int | None = None
    nu: float = 0.01
    dt: float = 0.002
    forced: bool = True
    fkmax: float = 2.0
    eps: float = 0.1
    prec: str = "dp"
    # Truncate every RHS output at the 2/3 cut as well (off = dns-deform
    # semantics, where the cut is applied once per full step, :232-237).
    dealias_rhs: bool = False

    def __post_init__(self):
        self.ny = self.ny or self.nx
        self.nz = self.nz or self.nx

    @property
    def lx(self):
        return self.nx // 2

    @property
    def kcut2(self):
        rd = real_dtype(self.prec)
        kcut = rd(2.0) * rd(self.lx) / rd(3.0)  # dns-deform.f90:75-76
        return rd(kcut * kcut)


class DNS:
    """The dns-deform.f90 solver state: wavenumbers, k2e and the RHS."""

    def __init__(self, p: DNSParams):
        self.p = p
        rd = real_dtype(p.prec)
        self.kx, self.ky, self.kz, self.k2 = wavenumber(p.nx, p.ny, p.nz, p.prec)
        # k2e = exp(-rnu*k2*dt/2) (dns-deform.f90:90)
        self.k2e = np.exp(-rd(p.nu) * self.k2 * rd(p.dt) / rd(2)).astype(rd)

    def rhs(self, vx, vy, vz):
        """ratesdudt (dns-deform.f90:259-288).  Returns (Nx, Ny, Nz)."""
        p = self.p
        wx, wy, wz = vorticity(vx, vy, vz, self.kx, self.ky, self.kz)
        wx, wy, wz = convec_dns(vx, vy, vz, wx, wy, wz, p.nx, p.ny, p.nz)
        if p.forced:
            force(wx, wy, wz, vx, vy, vz, self.k2, p.fkmax, p.eps)
        for w in (wx, wy, wz):
            symmetrize3d(w)
        projection(wx, wy, wz, self.kx, self.ky, self.kz, p.prec)
        if p.dealias_rhs:
            for w in (wx, wy, wz):
                w[self.k2 >= p.kcut2] = 0
        return wx, wy, wz

    def rk4_step(self, vx, vy, vz):
        """One Lawson IF-RK4 step (dns-deform.f90:164-237).  Returns new v.
        Pinned by tests/test_oracle_rk4.py (exact viscous decay with N = 0,
        4th-order convergence)."""
        E = self.k2e
        dt = real_dtype(self.p.prec)(self.p.dt)
        # stage 1 (:164-172)
        n = self.rhs(vx, vy, vz)
        k = [dt * a for a in n]
        du = [a.copy() for a in k]
        # stage 2 (:177-189)
        s = [E * (v + a / 2) for v, a in zip((vx, vy, vz), k)]
        n = self.rhs(*s)
        k = [dt * a / E for a in n]
        du = [d + 2 * a for d, a in zip(du, k)]
        # stage 3 (:194-206)
        s = [E * (v + a / 2) for v, a in zip((vx, vy, vz), k)]
        n = self.rhs(*s)
        k = [dt * a / E for a in n]
        du = [d + 2 * a for d, a in zip(du, k)]
        # stage 4 (:211-223)
        s = [E * E * (v + a) for v, a in zip((vx, vy, vz), k)]
        n = self.rhs(*s)
        k = [dt * a / E / E for a in n]
        du = [d + a for d, a in zip(du, k)]
        # update (:226-228) and 2/3 truncation (:233-237)
        out = [E * E * (v + d / 6) for v, d in zip((vx, vy, vz), du)]
        m = self.k2 >= self.p.kcut2
        for a in out:
            a[m] = 0
        return tuple(a.astype(vx.dtype) for a in out)


# --------------------------------------------------------------------------
# Initial condition (initialize.f90:284-371 projection, :373-463 draw order)
# --------------------------------------------------------------------------
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle C library not built: {path} (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        lib.oracle_ran1_stream.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p]
        lib.oracle_gauss3_f64.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p]
        lib.oracle_gauss3_f32.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p]
        _LIB = lib
    return _LIB


def ran1_stream(seed: int, n: int) -> np.ndarray:
    """Integer iy values of n successive ran1 calls from a fresh state."""
    out = np.empty(n, np.int32)
    _lib().oracle_ran1_stream(seed, n, out.ctypes.data)
    return out


def gauss3(seed: int, nmodes: int, prec: str = "dp") -> np.ndarray:
    """(nmodes, 3) complex Gaussian draws in the reference's order."""
    if prec == "dp":
        g = np.empty((nmodes, 3), np.complex128)
        _lib().oracle_gauss3_f64(seed, nmodes, g.ctypes.data)
    else:
        g = np.empty((nmodes, 3), np.complex64)
        _lib().oracle_gauss3_f32(seed, nmodes, g.ctypes.data)
    return g


def initialize(nx, ny, nz, iseed, akp, u0, prec="dp", kcut=None):
    """Seeded random-phase solenoidal field, E(k) ~ k^4 exp(-2 k^2/akp^2),
    cut at kcut = 2 lx/3, normalised to sum' |v|^2 = 1.5 u0^2.
    Draw order of initialize3dvellowmemdealiased (initialize.f90:397-417),
    complex projection of initialize3dvelhighmemdealiased (:328-332),
    symmetrize (:429-431), truncation (:433-437), shaping (:441), rescale
    (:454-460).  Pinned by tests/test_oracle_ic.py."""
    rd = real_dtype(prec)
    kx, ky, kz, k2 = wavenumber(nx, ny, nz, prec)
    lx1 = nx // 2 + 1
    g = gauss3(iseed, nz * ny * lx1, prec).reshape(nz, ny, lx1, 3)
    vx, vy, vz = (np.ascontiguousarray(g[..., c]) for c in range(3))
    projection(vx, vy, vz, kx, ky, kz, prec)
    for v in (vx, vy, vz):
        symmetrize3d(v)
    if kcut is None:
        kcut = rd(2.0) * rd(nx // 2) / rd(3.0)
    kcut2 = rd(kcut) * rd(kcut)
    m = k2 >= kcut2
    for v in (vx, vy, vz):
        v[m] = 0
    pi = rd(PI_REF)
    akp = rd(akp)
    u0 = rd(u0)
    amp = u0 * np.sqrt(rd(8) * np.sqrt(rd(2) / pi) / (rd(3) * pi * akp**5)) * np.sqrt(k2) * np.exp(-k2 / (akp * akp))
    amp = amp.astype(rd)
    vx, vy, vz = vx * amp, vy * amp, vz * amp
    e = energy(vx, vy, vz)
    s = np.sqrt(rd(1.5) * u0 * u0 / rd(e)).astype(rd)
    return (vx * s).astype(cplx_dtype(prec)), (vy * s).astype(cplx_dtype(prec)), (vz * s).astype(cplx_dtype(prec))


# --------------------------------------------------------------------------
# Diagnostics (dns-deform.f90:124-156, spec.f90:66-89, helispec.f90:67-91)
# --------------------------------------------------------------------------
def _modal_energy(vx, vy, vz) -> np.ndarray:
    t = (np.abs(vx.astype(np.complex128)) ** 2 + np.abs(vy.astype(np.complex128)) ** 2 + np.abs(vz.astype(np.complex128)) ** 2)
    t[:, :, 0] *= 0.5
    return t


def energy(vx, vy, vz) -> float:
    """E = sum' |v|^2 with weight 1/2 at kx=0 (dns-deform.f90:124-126)."""
    return float(_modal_energy(vx, vy, vz).sum())


def shell_index(k2: np.ndarray) -> np.ndarray:
    """ll = floor(sqrt(k2) + 0.5) (spec.f90:188, dns-deform.f90:147)."""
    return np.floor(np.sqrt(k2.astype(np.float64)) + 0.5).astype(np.int64)


def spectrum(vx, vy, vz, k2) -> np.ndarray:
    """E(ll), ll = 1..lx, FP64 (dns-deform.f90:139-153; spec.f90:172-195).
    Returned array index i holds shell i+1."""
    lx = vx.shape[2] - 1
    ll = shell_index(k2)
    t = _modal_energy(vx, vy, vz)
    m = (ll >= 1) & (ll <= lx)
    return np.bincount(ll[m] - 1, weights=t[m], minlength=lx)[:lx]


def helicity_spectrum(vx, vy, vz, kx, ky, kz, k2) -> np.ndarray:
    """H(ll) = sum_shell 2 Re(w . conj v), half weight at kx=0
    (helispec.f90:76-89)."""
    lx = vx.shape[2] - 1
    c = [a.astype(np.complex128) for a in (vx, vy, vz)]
    w = vorticity(*c, kx.astype(np.float64), ky.astype(np.float64), kz.astype(np.float64))
    h = 2.0 * np.real(w[0] * np.conj(c[0]) + w[1] * np.conj(c[1]) + w[2] * np.conj(c[2]))
    h[:, :, 0] *= 0.5
    ll = shell_index(k2)
    m = (ll >= 1) & (ll <= lx)
    return np.bincount(ll[m] - 1, weights=h[m], minlength=lx)[:lx]


def skewness(ux, kx, nx) -> float:
    """Velocity-derivative skewness <d^3> / <d^2>^(3/2), d = du_x/dx in
    physical space (skewness.f90:16-45), FP64 moments.  Pinned by
    tests/test_oracle_pins.py."""
    d = c2r(1j * kx[None, None, :].astype(np.float64) * ux.astype(np.complex128), nx)
    return float(np.mean(d**3) / np.mean(d**2) ** 1.5)


def dissipation(vx, vy, vz, k2, nu) -> float:
    """eps = 2 nu sum' k^2 |v|^2 = <2 nu s_ij s_ij> (viscenerdiss.f90:16-47 by
    Parseval, for solenoidal v)."""
    k2d = k2.astype(np.float64).copy()
    k2d[0, 0, 0] = 0.0
    return float(2.0 * nu * (k2d * _modal_energy(vx, vy, vz)).sum())


def enstrophy_total(vx, vy, vz, k2) -> float:
    """Omega = 1/2 <|w|^2> = sum' k^2 |v|^2 (enstrophy.f90:16-32 averaged)."""
    k2d = k2.astype(np.float64).copy()
    k2d[0, 0, 0] = 0.0
    return float((k2d * _modal_energy(vx, vy, vz)).sum())


def divergence_max(vx, vy, vz, kx, ky, kz, nx) -> float:
    """max |div u| over the physical grid: c2r of i k.v in FP64."""
    KX = kx[None, None, :].astype(np.float64)
    KY = ky[None, :, None].astype(np.float64)
    KZ = kz[:, None, None].astype(np.float64)
    d = 1j * (KX * vx.astype(np.complex128) + KY * vy.astype(np.complex128) + KZ * vz.astype(np.complex128))
    return float(np.abs(c2r(d, nx)).max())


@dataclass
class Diag:
    energy: float
    dissipation: float
    enstrophy: float
    divmax: float
    spectrum: np.ndarray = field(repr=False)


def diagnostics(vx, vy, vz, dns: DNS) -> Diag:
    p = dns.p
    return Diag(
        energy=energy(vx, vy, vz),
        dissipation=dissipation(vx, vy, vz, dns.k2, p.nu),
        enstrophy=enstrophy_total(vx, vy, vz, dns.k2),
        divmax=divergence_max(vx, vy, vz, dns.kx, dns.ky, dns.kz, p.nx),
        spectrum=spectrum(vx, vy, vz, dns.k2),
    )


def initialize_heli(nx, ny, nz, iseed, hrel, akp, u0):
    """initialize_heli.f90:1-109 in FP64: per mode (x fastest) two complex
    Gaussians am then ap from ran1 pairs (:22-35); am rescaled on the shells
    |k - ii| < 1/2 by sqrt((1 - hrel) P/M), P and M the sums of |ap|^2 and
    |am|^2 over |k - ii - oneless/2| < 1/2 (:37-44, the two masks as
    written); u = sqrt(3/2) (ap h+ + am conj(h+)) A(k) with the basis of
    :46-88 and A(k) = u0 sqrt(8 sqrt(2/pi)/(3 pi akp^5)) k exp(-k^2/akp^2);
    the mean zero (the file's basis is 0/0 there); symmetrize; sum' |u|^2 =
    1.5 u0^2 (:90-101).  Pinned by tests/test_oracle_ic.py."""
    kx, ky, kz, k2 = wavenumber(nx, ny, nz)
    lx1 = nx // 2 + 1
    nm = nz * ny * lx1
    d = ran1_stream(iseed, 4 * nm).astype(np.float64).reshape(nm, 4)
    t = np.minimum(d / 2147483647.0, 1.0 - 1.2e-7)
    g = np.sqrt(-2.0 * np.log(t[:, [0, 2]])) * np.exp(1j * 2.0 * PI_REF * t[:, [1, 3]]) / np.sqrt(2.0)
    am = g[:, 0].reshape(nz, ny, lx1).copy()
    ap = g[:, 1].reshape(nz, ny, lx1).copy()
    kk = np.sqrt(np.round(k2).astype(np.float64))
    kk[0, 0, 0] = 0.0
    oneless = 1.0 - 2.0 * float(np.finfo(np.float32).eps)
    for ii in range(1, lx1):
        m = np.abs(kk - ii - 0.5 * oneless) < 0.5
        P = np.sum(np.abs(ap[m]) ** 2)
        M = np.sum(np.abs(am[m]) ** 2)
        am[np.abs(kk - ii) < 0.5] *= np.sqrt((1.0 - hrel) * P / M) if M > 0 else 0.0
    KX = np.broadcast_to(kx[None, None, :], kk.shape).astype(np.float64)
    KY = np.broadcast_to(ky[None, :, None], kk.shape).astype(np.float64)
    KZ = np.broadcast_to(kz[:, None, None], kk.shape).astype(np.float64)
    h = np.sqrt(2.0) / 2.0
    with np.errstate(divide="ignore", invalid="ignore"):
        axis = (KY == 0) & (KZ == 0)
        k2d = np.where(axis, np.maximum(np.sqrt(KX**2 + KY**2), 2.0 * float(np.finfo(np.float32).eps)),
                       np.sqrt(KZ**2 + KY**2))
        hpx = np.where(axis, h * (-KY / k2d - 1j * KX * KZ / kk / k2d), h * 1j * (KY**2 + KZ**2) / kk / k2d)
        hpy = np.where(axis, h * (KX / k2d - 1j * KY * KZ / kk / k2d), h * (-KZ / k2d - 1j * KX * KY / kk / k2d))
        hpz = np.where(axis, h * 1j * (KY**2 + KX**2) / kk / k2d, h * (KY / k2d - 1j * KZ * KX / kk / k2d))
    amp = u0 * np.sqrt(8.0 * np.sqrt(2.0 / PI_REF) / (3.0 * PI_REF * akp**5)) * kk * np.exp(-kk * kk / (akp * akp))
    u = [np.sqrt(1.5) * (ap * hp + am * np.conj(hp)) * amp for hp in (hpx, hpy, hpz)]
    for a in u:
        a[0, 0, 0] = 0
        symmetrize3d(a)
    s = np.sqrt(1.5 * u0 * u0 / energy(*u))
    return [a * s for a in u]
