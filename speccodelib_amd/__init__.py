"""speccodelib_amd: MI355X pseudo-spectral RHS + IF-RK4 path behind the
reference's routine names.

Host-side mirror of the reference interface over the C ABI (include/scl.h).
Arrays follow the reference layout: a Fortran ``complex(sp), dimension(lx1,
ly, lz)`` is a C-ordered numpy array of shape ``(lz, ly, lx1)``
(``complex64`` for SP plans, ``complex128`` for DP plans).

Nothing here computes on the CPU: every numerical routine runs through
``libscl.so`` on the GPU, and importing a routine without the built library
raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libscl.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "scl.h")

SP, DP = 0, 1
_STATUS = {0: "SCL_OK", -1: "SCL_EINVAL", -2: "SCL_ENOMEM", -3: "SCL_EHIP", -4: "SCL_ENOPLAN", -5: "SCL_ENONFINITE", -6: "SCL_ENODEV"}
KIND_NAMES = ["force_coef", "zinv_vort", "y_inverse", "x_pass", "y_forward", "zfwd_rk", "herm_fix", "zfwd_zinv",
              "zfwd4_zinv"]

_lib = None


class SclError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class NonFiniteError(SclError):
    """SCL_ENONFINITE: the solver's state holds a NaN or Inf."""


class ScDnsParams(ctypes.Structure):
    _fields_ = [
        ("nu", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("forced", ctypes.c_int32),
        ("fkmax", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("dealias_rhs", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("sgs", ctypes.c_int32),
        ("cs2", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("delta_test", ctypes.c_double),
        ("k_test", ctypes.c_double),
        ("iupdate", ctypes.c_int32),
        ("scalar", ctypes.c_int32),
        ("grad", ctypes.c_double),
        ("schmidt", ctypes.c_double),
        ("stepper", ctypes.c_int32),
        ("beta", ctypes.c_double),
        ("dtw", ctypes.c_double),
        ("ieout", ctypes.c_int32),
        ("umax", ctypes.c_double),
        ("dt_first", ctypes.c_double),
        ("kcut", ctypes.c_double),
    ]


class ScDiag(ctypes.Structure):
    _fields_ = [
        ("energy", ctypes.c_double),
        ("dissipation", ctypes.c_double),
        ("enstrophy", ctypes.c_double),
        ("helicity", ctypes.c_double),
        ("divmax", ctypes.c_double),
        ("divrms_spec", ctypes.c_double),
        ("nonfinite", ctypes.c_int32),
        ("scalar_variance", ctypes.c_double),
        ("scalar_dissipation", ctypes.c_double),
        ("cs2", ctypes.c_double),
        ("helicity_dissipation", ctypes.c_double),
        ("skewness", ctypes.c_double),
    ]


class ScEns2DParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("c_hyper", ctypes.c_double),
        ("c_hypo", ctypes.c_double),
        ("damp", ctypes.c_double),
        ("fkmin", ctypes.c_double),
        ("fkmax", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("n_hyper", ctypes.c_int32),
        ("n_hypo", ctypes.c_int32),
        ("forced", ctypes.c_int32),
        ("per_member", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("sigma", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("kf", ctypes.c_double),
    ]


def header_functions() -> list[str]:
    """Names of the functions include/scl.h declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(scl_\w+)\s*\(", txt, flags=re.M)))


def lib():
    """The loaded libscl.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        P, I32, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        sig = {
            "scl_last_error": ([], ctypes.c_char_p),
            "scl_version": ([], I32),
            "scl_plan3d_create": ([I32, I32, I32, I32, ctypes.POINTER(P)], I32),
            "scl_plan3d_destroy": ([P], I32),
            "scl_plan3d_set_stream": ([P, P], I32),
            "scl_fftwplan3d": ([I32, I32, I32], I32),
            "scl_fftwplan3d_prec": ([I32, I32, I32, I32], I32),
            "scl_destroyplan3d": ([], I32),
            "scl_c2r3d": ([P, P], I32),
            "scl_r2c3d": ([P, P], I32),
            "scl_wavenumber3d": ([I32, P, P, P, P, I32, I32, I32], I32),
            "scl_convec_dns": ([P] + [P] * 6 + [I32] * 6, I32),
            "scl_force": ([P] + [P] * 7 + [I32] * 3 + [D, D], I32),
            "scl_force_heli": ([P] + [P] * 10 + [I32] * 3 + [D, D, D], I32),
            "scl_symmetrize3d": ([P, P, I32, I32, I32], I32),
            "scl_dealiasing": ([P, P, P, I32, I32, I32], I32),
            "scl_projection": ([P] + [P] * 6 + [I32] * 3, I32),
            "scl_ratesdudt": ([P] + [P] * 6 + [I32, D, D, D], I32),
            "scl_initialize3d": ([I32, P, P, P, I32, I32, I32, I32, D, D, D], I32),
            "scl_initialize_heli": ([I32, P, P, P, I32, I32, I32, I32, D, D, D], I32),
            "scl_dns_initialize_heli": ([P, I32, D, D, D], I32),
            "scl_rk4step": ([P, P, P, P, D, D, I32, D, D, D], I32),
            "scl_dns_create": ([P, ctypes.POINTER(ScDnsParams), ctypes.POINTER(P)], I32),
            "scl_dns_destroy": ([P], I32),
            "scl_dns_set_velocity": ([P, P, P, P], I32),
            "scl_dns_get_velocity": ([P, P, P, P], I32),
            "scl_dns_initialize": ([P, I32, D, D], I32),
            "scl_dns_step": ([P, I32], I32),
            "scl_dns_sync": ([P], I32),
            "scl_dns_time": ([P], D),
            "scl_dns_step_info": ([P, P, P, P], I32),
            "scl_dns_spectrum": ([P, P, I32], I32),
            "scl_dns_helicity_spectrum": ([P, P, I32], I32),
            "scl_dns_diagnostics": ([P, ctypes.POINTER(ScDiag)], I32),
            "scl_dns_spectra_noalloc": ([P, P, I32], I32),
            "scl_dns_step_timed": ([P, I32, P, P, P], I32),
            "scl_dns_set_scalar": ([P, P], I32),
            "scl_dns_get_scalar": ([P, P], I32),
            "scl_smag": ([P] + [P] * 12 + [I32] * 6 + [D], I32),
            "scl_dynsmag": ([P] + [P] * 13 + [I32] * 7 + [D, D, P, P, P, I32], I32),
            "scl_convecphi_dns": ([P] + [P] * 8 + [I32] * 6, I32),
            "scl_skewness": ([P, P, P, P, I32, I32, I32], I32),
            "scl_padd": ([I32, P, P] + [I32] * 6, I32),
            "scl_unpadd": ([I32, P, P] + [I32] * 6, I32),
            "scl_convec_padded": ([P] + [P] * 6 + [I32] * 9, I32),
            "scl_smag_dealias": ([P] + [P] * 12 + [I32] * 9 + [D], I32),
            "scl_ens2d_create": ([I32, I32, P, P, ctypes.POINTER(ScEns2DParams), ctypes.POINTER(P)], I32),
            "scl_ens2d_destroy": ([P], I32),
            "scl_ens2d_set_member": ([P, I32, I32, P], I32),
            "scl_ens2d_get_member": ([P, I32, I32, P], I32),
            "scl_ens2d_initialize": ([P, I32, D, D], I32),
            "scl_ens2d_rhs": ([P, I32, I32, P, P], I32),
            "scl_ens2d_step": ([P, I32], I32),
            "scl_ens2d_sync": ([P], I32),
            "scl_ens2d_step_timed": ([P, I32, P], I32),
            "scl_ens2d_time": ([P, P], I32),
            "scl_ens2d_diagnostics": ([P, P], I32),
            "scl_ens2d_set_schedule": ([P, I32, I32, I32, I32], I32),
            "scl_ens2d_get_schedule": ([P, I32, P, P, P], I32),
            "scl_ens2d_tune": ([P, I32, P], I32),
            "scl_dns_rhs": ([P, D, P, P, I32], I32),
            "scl_dns_state": ([P, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int64)], I32),
            "scl_dns_rates": ([P, P], I32),
            "scl_dns_derivs_cb": ([P, D, P, P], I32),
            "scl_ifrk4_integrate": ([I32, ctypes.c_int64, P, P, D, D, I32, P, P], I32),
            "scl_derivs_bind": ([P], I32),
            "scl_derivs_": ([P, P, P, P], None),
            "scl_adjoint_create": ([P, ctypes.POINTER(ScDnsParams), I32, I32, ctypes.POINTER(P)], I32),
            "scl_adjoint_destroy": ([P], I32),
            "scl_adjoint_forward": ([P, D] + [P] * 6 + [I32], I32),
            "scl_adjoint_backward": ([P] + [P] * 6 + [P, I32], I32),
            "scl_adjoint_vjp": ([P, D] + [P] * 9 + [P, I32], I32),
            "scl_adjoint_timing": ([P, P, P], I32),
            "scl_stream_probe": ([I32, ctypes.c_int64, I32, P], I32),
            "scl_record_write": ([ctypes.c_char_p, P, ctypes.c_int64], I32),
            "scl_record_read": ([ctypes.c_char_p, P, ctypes.c_int64], I32),
            "scl_dns_output": ([P, ctypes.c_char_p, I32], I32),
            "scl_dns_input": ([P, ctypes.c_char_p, I32], I32),
            "scl_write_ener_time": ([ctypes.c_char_p, I32, P, I32], I32),
            "scl_write_spectrum": ([ctypes.c_char_p, D, P, I32], I32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        cls = NonFiniteError if status == -5 else SclError
        raise cls(status, lib().scl_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def _cdtype(prec):
    return np.complex64 if prec == SP else np.complex128


def _rdtype(prec):
    return np.float32 if prec == SP else np.float64


def _field(a, prec, shape):
    if a.dtype != _cdtype(prec) or a.shape != shape or not a.flags.c_contiguous:
        raise ValueError(f"expected C-contiguous {np.dtype(_cdtype(prec))} array of shape {shape}, got {a.dtype} {a.shape}")
    return a


class Plan3D:
    """A plan (fftwplan.f90:7-15): grid, precision, stream, twiddles."""

    def __init__(self, nx, ny=None, nz=None, prec=SP):
        self.h = None
        ny = ny or nx
        nz = nz or nx
        h = ctypes.c_void_p()
        _check(lib().scl_plan3d_create(nx, ny, nz, prec, ctypes.byref(h)))
        self.h, self.nx, self.ny, self.nz, self.prec = h, nx, ny, nz, prec
        self.shape = (nz, ny, nx // 2 + 1)

    def close(self):
        if getattr(self, "h", None):
            lib().scl_plan3d_destroy(self.h)
            self.h = None

    __del__ = close

    def zeros(self):
        return np.zeros(self.shape, _cdtype(self.prec))

    # -- FFTW-equivalent in-place transforms (real view (nx+2, ny, nz)) --
    def c2r(self, a):
        """rfftwnd_f77_one_complex_to_real: in place; returns the real view
        (nz, ny, nx) of the result."""
        _field(a, self.prec, self.shape)
        _check(lib().scl_c2r3d(self.h, _ptr(a)))
        return a.view(_rdtype(self.prec)).reshape(self.nz, self.ny, self.nx + 2)[:, :, : self.nx]

    def r2c(self, a):
        """rfftwnd_f77_one_real_to_complex: in place on the complex view."""
        _field(a, self.prec, self.shape)
        _check(lib().scl_r2c3d(self.h, _ptr(a)))
        return a

    def packed_real(self, r):
        """A complex array whose real view holds the real field r (nz, ny, nx)
        with FFTW's two padding reals per x-line."""
        a = self.zeros()
        a.view(_rdtype(self.prec)).reshape(self.nz, self.ny, self.nx + 2)[:, :, : self.nx] = r
        return a

    # -- RHS blocks (reference signatures, arrays updated in place) --
    def convec_dns(self, vx, vy, vz, wx, wy, wz):
        for a in (vx, vy, vz, wx, wy, wz):
            _field(a, self.prec, self.shape)
        lx1, ly, lz = self.nx // 2 + 1, self.ny, self.nz
        _check(lib().scl_convec_dns(self.h, *map(_ptr, (vx, vy, vz, wx, wy, wz)), lx1, ly, lz, self.nx, self.ny, self.nz))

    def force(self, wx, wy, wz, vx, vy, vz, k2, fkmax, eps):
        for a in (vx, vy, vz, wx, wy, wz):
            _field(a, self.prec, self.shape)
        _check(lib().scl_force(self.h, *map(_ptr, (wx, wy, wz, vx, vy, vz, k2)), *self._dims(), fkmax, eps))

    def force_heli(self, vx, vy, vz, wx, wy, wz, k2, fkmax, eps, eta):
        out = [self.zeros() for _ in range(3)]
        _check(lib().scl_force_heli(self.h, *map(_ptr, (*out, vx, vy, vz, wx, wy, wz, k2)), *self._dims(), fkmax, eps, eta))
        return tuple(out)

    def symmetrize(self, c):
        _field(c, self.prec, self.shape)
        _check(lib().scl_symmetrize3d(self.h, _ptr(c), *self._dims()))

    def dealiasing(self, c, k2):
        _field(c, self.prec, self.shape)
        _check(lib().scl_dealiasing(self.h, _ptr(c), _ptr(k2), *self._dims()))

    def projection(self, vx, vy, vz, kx, ky, kz):
        for a in (vx, vy, vz):
            _field(a, self.prec, self.shape)
        _check(lib().scl_projection(self.h, *map(_ptr, (vx, vy, vz, kx, ky, kz)), *self._dims()))

    def ratesdudt(self, vx, vy, vz, forced=1, fkmax=2.0, eps=0.1, eta=0.0):
        out = [self.zeros() for _ in range(3)]
        _check(lib().scl_ratesdudt(self.h, *map(_ptr, (vx, vy, vz, *out)), forced, fkmax, eps, eta))
        return tuple(out)

    def rk4step(self, vx, vy, vz, nu, dt, forced=1, fkmax=2.0, eps=0.1, eta=0.0):
        _check(lib().scl_rk4step(self.h, *map(_ptr, (vx, vy, vz)), nu, dt, forced, fkmax, eps, eta))

    def smag(self, vx, vy, vz, kx, ky, kz, delta):
        """smag.f90:1-78: returns (t11, t12, t13, t22, t23, t33)."""
        t = [self.zeros() for _ in range(6)]
        _check(lib().scl_smag(self.h, *map(_ptr, (vx, vy, vz, *t, kx, ky, kz)), *self._dims(), self.nx, self.ny, self.nz, delta))
        return tuple(t)

    def dynsmag(self, ux, uy, uz, kx3, ky3, kz3, g_test, delta, delta_test, cs2=0.0, update=True):
        """dynsmag.f90:1-236: returns (tau (6), mm, lm, cs2).  kx3/ky3/kz3 and
        g_test are real 3D arrays (lz, ly, lx1)."""
        t = [self.zeros() for _ in range(6)]
        mm, lm, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(cs2)
        _check(lib().scl_dynsmag(self.h, *map(_ptr, (ux, uy, uz, *t, kx3, ky3, kz3, g_test)), *self._dims(),
                                 self.nx, self.ny, self.nz, self.nx // 2, delta, delta_test,
                                 ctypes.byref(mm), ctypes.byref(lm), ctypes.byref(c), int(update)))
        return tuple(t), mm.value, lm.value, c.value

    def convecphi_dns(self, vx, vy, vz, phi, kx, ky, kz):
        """convecphi_dns.f90:1-49: wphi = i k . FFT(u phi)/N^3 (the reference's sign)."""
        w = self.zeros()
        _check(lib().scl_convecphi_dns(self.h, *map(_ptr, (vx, vy, vz, phi, w, kx, ky, kz)), *self._dims(),
                                       self.nx, self.ny, self.nz))
        return w

    def padded_shape(self):
        """The reference's 3/2-rule padded half-spectrum shape (lzb, lyb, lxb1)."""
        return (3 * self.nz // 2, 3 * self.ny // 2, 3 * self.nx // 4 + 1)

    def convec_padded(self, vx, vy, vz, wx, wy, wz):
        """The padded convec of spectral_heli_dynsmag.f90:229 (3/2 rule):
        w <- the unaliased Lamb vector, in place."""
        for a in (vx, vy, vz, wx, wy, wz):
            _field(a, self.prec, self.shape)
        nxb, nyb, nzb = 3 * self.nx // 2, 3 * self.ny // 2, 3 * self.nz // 2
        _check(lib().scl_convec_padded(self.h, *map(_ptr, (vx, vy, vz, wx, wy, wz)), *self._dims(),
                                       nxb // 2 + 1, nyb, nzb, nxb, nyb, nzb))

    def smag_dealias(self, vx, vy, vz, kx3, ky3, kz3, delta):
        """smag of smag_dealias.f90 (3/2-padded products): (t11 .. t33)."""
        t = [self.zeros() for _ in range(6)]
        nxb, nyb, nzb = 3 * self.nx // 2, 3 * self.ny // 2, 3 * self.nz // 2
        _check(lib().scl_smag_dealias(self.h, *map(_ptr, (vx, vy, vz, *t, kx3, ky3, kz3)), *self._dims(),
                                      nxb // 2 + 1, nyb, nzb, nxb, nyb, nzb, delta))
        return tuple(t)

    def skewness(self, ux, kx):
        """skewness.f90:1-48: <(du_x/dx)^3> / <(du_x/dx)^2>^(3/2), FP64."""
        _field(ux, self.prec, self.shape)
        s = ctypes.c_double()
        _check(lib().scl_skewness(self.h, _ptr(ux), _ptr(kx), ctypes.byref(s), self.nx, self.ny, self.nz))
        return s.value

    def wavenumber(self):
        return wavenumber(self.nx, self.ny, self.nz, self.prec)

    def _dims(self):
        return self.nx // 2 + 1, self.ny, self.nz


def wavenumber(nx, ny=None, nz=None, prec=SP):
    """wavenumber(kx,ky,kz,k2,lx1,ly,lz) (wavenumber.f90:89-129)."""
    ny = ny or nx
    nz = nz or nx
    rd = _rdtype(prec)
    kx = np.zeros(nx // 2 + 1, rd)
    ky = np.zeros(ny, rd)
    kz = np.zeros(nz, rd)
    k2 = np.zeros((nz, ny, nx // 2 + 1), rd)
    _check(lib().scl_wavenumber3d(prec, _ptr(kx), _ptr(ky), _ptr(kz), _ptr(k2), nx // 2 + 1, ny, nz))
    return kx, ky, kz, k2


def initialize(nx, ny=None, nz=None, iseed=1, akp=4.0, u0=1.0, prec=SP, kcut=-1.0):
    """initialize(vx,vy,vz,...,iseed,...,akp,u0,kcut) (initialize.f90:373-463,
    with the complex projection of :328-332)."""
    ny = ny or nx
    nz = nz or nx
    out = [np.zeros((nz, ny, nx // 2 + 1), _cdtype(prec)) for _ in range(3)]
    _check(lib().scl_initialize3d(prec, *map(_ptr, out), nx, ny, nz, iseed, akp, u0, kcut))
    return tuple(out)


def initialize_heli(nx, ny=None, nz=None, iseed=1, hrel=1.0, akp=4.0, u0=1.0, prec=SP):
    """initialize_heli(ux,uy,uz,kx,ky,kz,k2,lx1,ly,lz,iseed,hrel,akp,u0)
    (initialize_heli.f90:1-109): a helical field of relative helicity set by
    hrel (1: every mode of the rescaled shells positively helical)."""
    ny = ny or nx
    nz = nz or nx
    out = [np.zeros((nz, ny, nx // 2 + 1), _cdtype(prec)) for _ in range(3)]
    _check(lib().scl_initialize_heli(prec, *map(_ptr, out), nx, ny, nz, iseed, hrel, akp, u0))
    return tuple(out)


class DNS:
    """Device-resident IF-RK4 solver: the dns-deform.f90 loop body."""

    def __init__(self, plan: Plan3D, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, eta=0.0, dealias_rhs=False,
                 use_graph=True, sgs=0, cs2=0.0, delta=0.0, delta_test=0.0, k_test=0.0, iupdate=10, scalar=False,
                 grad=0.0, schmidt=1.0, stepper=0, beta=0.5, dtw=0.0, ieout=10, umax=3.0 * 3.0**0.5, dt_first=0.0,
                 kcut=0.0):
        """stepper: 0 IF-RK4 (dns-deform.f90), 1 IF-AB2 with CFL steps
        (spectral_dns_dealiased.f90, and with sgs / scalar the loops of
        spectral_les.f90 and spectral_dns_scalar.f90; beta, dtw, ieout, umax,
        dt_first), 2 IF-AB2 with fixed dt (specdns-forced-omp-ft.f90).
        kcut: truncation radius after each step (0: 2 lx/3)."""
        self.plan = plan
        self.params = ScDnsParams(nu, dt, int(forced), fkmax, eps, eta, int(dealias_rhs), int(use_graph), int(sgs), cs2,
                                  delta, delta_test, k_test, int(iupdate), int(scalar), grad, schmidt, int(stepper),
                                  beta, dtw, int(ieout), umax, dt_first, kcut)
        h = ctypes.c_void_p()
        _check(lib().scl_dns_create(plan.h, ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().scl_dns_destroy(self.h)
            self.h = None

    __del__ = close

    def set_velocity(self, vx, vy, vz):
        for a in (vx, vy, vz):
            _field(a, self.plan.prec, self.plan.shape)
        _check(lib().scl_dns_set_velocity(self.h, _ptr(vx), _ptr(vy), _ptr(vz)))

    def get_velocity(self):
        out = [self.plan.zeros() for _ in range(3)]
        _check(lib().scl_dns_get_velocity(self.h, *map(_ptr, out)))
        return tuple(out)

    def set_scalar(self, phi):
        _field(phi, self.plan.prec, self.plan.shape)
        _check(lib().scl_dns_set_scalar(self.h, _ptr(phi)))

    def get_scalar(self):
        out = self.plan.zeros()
        _check(lib().scl_dns_get_scalar(self.h, _ptr(out)))
        return out

    def initialize(self, iseed=1, akp=4.0, u0=1.0):
        _check(lib().scl_dns_initialize(self.h, iseed, akp, u0))

    def initialize_heli(self, iseed=1, hrel=1.0, akp=4.0, u0=1.0):
        """The state from initialize_heli (spectral_heli_dns.f90:92)."""
        _check(lib().scl_dns_initialize_heli(self.h, iseed, hrel, akp, u0))

    def step(self, n=1):
        _check(lib().scl_dns_step(self.h, n))

    def sync(self):
        _check(lib().scl_dns_sync(self.h))

    @property
    def time(self):
        return lib().scl_dns_time(self.h)

    def step_info(self):
        """(time, the last step's dt, steps since the state was set)."""
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _check(lib().scl_dns_step_info(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)))
        return t.value, dt.value, n.value

    def spectrum(self):
        ek = np.zeros(self.plan.nx // 2)
        _check(lib().scl_dns_spectrum(self.h, _ptr(ek), len(ek)))
        return ek

    def helicity_spectrum(self):
        hk = np.zeros(self.plan.nx // 2)
        _check(lib().scl_dns_helicity_spectrum(self.h, _ptr(hk), len(hk)))
        return hk

    def diagnostics(self, allow_nonfinite=False) -> dict:
        """FP64 totals of the state.  A non-finite state raises
        NonFiniteError unless allow_nonfinite (then ``nonfinite`` is 1)."""
        d = ScDiag()
        st = lib().scl_dns_diagnostics(self.h, ctypes.byref(d))
        if not (allow_nonfinite and st == -5):
            _check(st)
        return {k: getattr(d, k) for k, _ in ScDiag._fields_}

    def rhs(self, *y):
        """dydt = N(y) - nu k^2 y (the derivs form) for host fields y."""
        nc = len(y)
        for a in y:
            _field(a, self.plan.prec, self.plan.shape)
        out = [self.plan.zeros() for _ in range(nc)]
        yi = (ctypes.c_void_p * nc)(*[_ptr(a) for a in y])
        yo = (ctypes.c_void_p * nc)(*[_ptr(a) for a in out])
        _check(lib().scl_dns_rhs(self.h, 0.0, yi, yo, 0))
        return tuple(out)

    def state(self):
        """(device pointer, complex entries) of the state in the split layout."""
        ptr = ctypes.c_void_p()
        n = ctypes.c_int64()
        _check(lib().scl_dns_state(self.h, ctypes.byref(ptr), ctypes.byref(n)))
        return ptr.value, n.value

    def rates(self, rate_ptr):
        _check(lib().scl_dns_rates(self.h, rate_ptr))

    def output(self, directory=".", idump=0):
        """output.f90:103-166: <dir>/ux<idump>.dat, uy, uz [, phi] as Fortran
        unformatted sequential records."""
        _check(lib().scl_dns_output(self.h, os.fsencode(directory), idump))

    def input(self, directory=".", idump=0):
        """input.f90:106-170: sets the state from <dir>/ux<idump>.dat, ..."""
        _check(lib().scl_dns_input(self.h, os.fsencode(directory), idump))

    def step_timed(self, n):
        """Runs n steps with HIP events around every kernel launch; returns
        (total_ms, {kind: (ms, launches)})."""
        ms = np.zeros(len(KIND_NAMES))
        cnt = np.zeros(len(KIND_NAMES), np.int32)
        tot = ctypes.c_double()
        _check(lib().scl_dns_step_timed(self.h, n, _ptr(ms), _ptr(cnt), ctypes.byref(tot)))
        return tot.value, {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KIND_NAMES)}


class Ensemble2D:
    """Batched 2D vorticity ensembles.

    scheme 0 (default): spectral_dns_2d.f90's RHS stepped with IF-RK4; a
    member's field is the reference's ``complex(sp), dimension(lx1, ly)``,
    i.e. a numpy array of shape ``(n, n//2 + 1)``.
    scheme 1: turb_2d.f90 at one rank (phase-shift dealiasing, (k sigma)^16
    + (k/nu)^-4, forcing eta/104/conj(w) on kf <= |k| <= kf + 3, its
    half-step bootstrap then IF-AB2); a member is that file's
    ``wz(my, mx2)`` as an array of shape ``(n, n//2)`` indexed [iy, ix].
    ``groups`` is a list of ``(n, members)``.  ``per_member=True`` gives
    every member its own launches (the comparison case)."""

    def __init__(self, groups, prec=SP, dt=5e-4, c_hyper=1e-5, n_hyper=8, c_hypo=0.0, n_hypo=0, damp=0.1,
                 forced=True, fkmin=4.0, fkmax=5.0, eps=0.01, per_member=False, use_graph=True, scheme=0,
                 sigma=0.0, nu=1.0, eta=0.0, kf=4.0):
        self.groups = [(int(n), int(m)) for n, m in groups]
        self.prec = prec
        self.scheme = int(scheme)
        self.params = ScEns2DParams(dt, c_hyper, c_hypo, damp, fkmin, fkmax, eps, int(n_hyper), int(n_hypo),
                                    int(forced), int(per_member), int(use_graph), self.scheme, sigma, nu, eta, kf)
        sizes = np.array([n for n, _ in self.groups], np.int32)
        mem = np.array([m for _, m in self.groups], np.int32)
        h = ctypes.c_void_p()
        _check(lib().scl_ens2d_create(prec, len(self.groups), _ptr(sizes), _ptr(mem), ctypes.byref(self.params),
                                      ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().scl_ens2d_destroy(self.h)
            self.h = None

    __del__ = close

    def shape(self, g):
        n = self.groups[g][0]
        return (n, n // 2) if self.scheme == 1 else (n, n // 2 + 1)

    def _in(self, g, w):
        if self.scheme == 1:  # Fortran order of wz(my, mx2): iy fastest
            if w.dtype != _cdtype(self.prec) or w.shape != self.shape(g):
                raise ValueError(f"expected {np.dtype(_cdtype(self.prec))} array of shape {self.shape(g)}")
            return np.ascontiguousarray(w.T)
        return _field(w, self.prec, self.shape(g))

    def _out(self, g):
        s = self.shape(g)
        return np.zeros(s[::-1] if self.scheme == 1 else s, _cdtype(self.prec))

    def _view(self, out):
        return out.T if self.scheme == 1 else out

    def set_member(self, g, j, w):
        a = self._in(g, w)  # held until the call returns
        _check(lib().scl_ens2d_set_member(self.h, g, j, _ptr(a)))

    def get_member(self, g, j):
        out = self._out(g)
        _check(lib().scl_ens2d_get_member(self.h, g, j, _ptr(out)))
        return self._view(out)

    def initialize(self, seed0=-1, k0=8.0, energy=0.5):
        _check(lib().scl_ens2d_initialize(self.h, seed0, k0, energy))

    def rhs(self, g, j, w):
        """Sets member j of group g to w and returns N(w)."""
        out = self._out(g)
        a = self._in(g, w)
        _check(lib().scl_ens2d_rhs(self.h, g, j, _ptr(a), _ptr(out)))
        return self._view(out)

    def step(self, n=1):
        _check(lib().scl_ens2d_step(self.h, n))

    def sync(self):
        _check(lib().scl_ens2d_sync(self.h))

    def step_timed(self, n):
        ms = ctypes.c_double()
        _check(lib().scl_ens2d_step_timed(self.h, n, ctypes.byref(ms)))
        return ms.value

    @property
    def time(self):
        t = ctypes.c_double()
        _check(lib().scl_ens2d_time(self.h, ctypes.byref(t)))
        return t.value

    def diagnostics(self):
        """(members, 2) FP64: [sum' |w|^2/k^2, sum' |w|^2] per member, group-major."""
        tot = sum(m for _, m in self.groups)
        out = np.zeros((tot, 2))
        _check(lib().scl_ens2d_diagnostics(self.h, _ptr(out)))
        return out

    def set_schedule(self, g, e, cw, row_grid=0):
        _check(lib().scl_ens2d_set_schedule(self.h, g, e, cw, row_grid))

    def get_schedule(self, g):
        v = [ctypes.c_int32() for _ in range(3)]
        _check(lib().scl_ens2d_get_schedule(self.h, g, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def tune(self, reps=5):
        """Autotunes every group's schedule on whole steps (state restored);
        returns the best ms per step per group."""
        best = np.zeros(len(self.groups))
        _check(lib().scl_ens2d_tune(self.h, reps, _ptr(best)))
        return best


class Adjoint:
    """Discrete adjoint of nsteps IF-RK4 steps of the unforced momentum
    equation (DNS, or static Smagorinsky LES with sgs=1) on a plan.

    Fields are numpy arrays in the reference layout (host path) or device
    pointers (``on_device=True`` with integer addresses, e.g. torch
    ``tensor.data_ptr()``).  Gradients: dL/dRe + i dL/dIm per stored entry."""

    def __init__(self, plan: Plan3D, nsteps, ckpt_every=20, nu=0.005, dt=0.002, sgs=1, delta=0.0):
        self.plan = plan
        self.nsteps = nsteps
        self.params = ScDnsParams(nu, dt, 0, 0.0, 0.0, 0.0, 0, 0, int(sgs), 0.0, delta, 0.0, 0.0, 10, 0, 0.0, 1.0)
        h = ctypes.c_void_p()
        _check(lib().scl_adjoint_create(plan.h, ctypes.byref(self.params), nsteps, ckpt_every, ctypes.byref(h)))
        self.h = h
        self._fwd_token = None

    def close(self):
        if getattr(self, "h", None):
            lib().scl_adjoint_destroy(self.h)
            self.h = None

    __del__ = close

    def _in(self, arrs):
        for a in arrs:
            _field(a, self.plan.prec, self.plan.shape)
        return [_ptr(a) for a in arrs]

    def forward(self, u0, cs2):
        # the checkpoints now belong to this call, not to a torch forward
        self._fwd_token = None
        out = [self.plan.zeros() for _ in range(3)]
        _check(lib().scl_adjoint_forward(self.h, cs2, *self._in(u0), *map(_ptr, out), 0))
        return tuple(out)

    def backward(self, gT):
        out = [self.plan.zeros() for _ in range(3)]
        d = ctypes.c_double()
        _check(lib().scl_adjoint_backward(self.h, *self._in(gT), *map(_ptr, out), ctypes.byref(d), 0))
        return tuple(out), d.value

    def vjp(self, s, nbar, cs2):
        self._fwd_token = None
        out = [self.plan.zeros() for _ in range(3)]
        d = ctypes.c_double()
        _check(lib().scl_adjoint_vjp(self.h, cs2, *self._in(s), *self._in(nbar), *map(_ptr, out), ctypes.byref(d), 0))
        return tuple(out), d.value

    def forward_device(self, cs2, u0_ptrs, uT_ptrs):
        self._fwd_token = None
        _check(lib().scl_adjoint_forward(self.h, cs2, *u0_ptrs, *uT_ptrs, 1))

    def backward_device(self, gT_ptrs, g0_ptrs):
        d = ctypes.c_double()
        _check(lib().scl_adjoint_backward(self.h, *gT_ptrs, *g0_ptrs, ctypes.byref(d), 1))
        return d.value

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().scl_adjoint_timing(self.h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value


def padd(u, shape_b):
    """padd.f90: u (lz, ly, lx1) zero-padded into shape_b (lzb, lyb, lxb1)."""
    prec = SP if u.dtype == np.complex64 else DP
    ub = np.zeros(shape_b, u.dtype)
    lz, ly, lx1 = u.shape
    u = np.ascontiguousarray(u)  # held until the call returns
    _check(lib().scl_padd(prec, _ptr(u), _ptr(ub), lx1, ly, lz, shape_b[2], shape_b[1],
                          shape_b[0]))
    return ub


def unpadd(ub, shape):
    """unpadd.f90: the (lz, ly, lx1) part of a padded half spectrum."""
    prec = SP if ub.dtype == np.complex64 else DP
    u = np.zeros(shape, ub.dtype)
    lzb, lyb, lxb1 = ub.shape
    ub = np.ascontiguousarray(ub)
    _check(lib().scl_unpadd(prec, _ptr(ub), _ptr(u), shape[2], shape[1], shape[0], lxb1, lyb,
                            lzb))
    return u


def record_write(path, a):
    """One Fortran unformatted sequential record holding the array a."""
    a = np.ascontiguousarray(a)
    _check(lib().scl_record_write(os.fsencode(path), a.ctypes.data, a.nbytes))


def record_read(path, shape, dtype):
    """The array of one Fortran unformatted sequential record (exactly)."""
    out = np.zeros(shape, dtype)
    _check(lib().scl_record_read(os.fsencode(path), out.ctypes.data, out.nbytes))
    return out


def write_ener_time(path, istep, *vals):
    """Appends write(25,'(I6,20E12.4)') istep, vals (the ener_time.data line)."""
    v = np.array(vals, np.float64)
    _check(lib().scl_write_ener_time(os.fsencode(path), istep, v.ctypes.data, len(v)))


def write_spectrum(path, time, ek):
    """Appends the time and the "ll E(ll)" lines of spectrum.data."""
    e = np.ascontiguousarray(ek, np.float64)
    _check(lib().scl_write_spectrum(os.fsencode(path), time, e.ctypes.data, len(e)))


def stream_probe(kind, nbytes, reps=10):
    """HBM rate of a plain streaming kernel: kind 0 read, 1 write, 2 copy of
    nbytes; returns (ms per pass, GB/s of bytes moved: 2 nbytes for a copy)."""
    ms = ctypes.c_double()
    _check(lib().scl_stream_probe(kind, nbytes, reps, ctypes.byref(ms)))
    moved = nbytes * (2 if kind == 2 else 1)
    return ms.value, moved / (ms.value * 1e-3) / 1e9


def ifrk4_integrate(prec, n, y_ptr, rate_ptr, t0, dt, nsteps, fn_ptr, ctx):
    """Device Lawson IF-RK4 of y' = -rate y + f(t, y) with a C callback f
    (``fn_ptr``: address of an ``scl_derivs_fn``)."""
    _check(lib().scl_ifrk4_integrate(prec, n, y_ptr, rate_ptr, t0, dt, nsteps, fn_ptr, ctx))


def dns_derivs_cb_ptr():
    """Address of the built-in spectral RHS callback (ctx = a DNS handle)."""
    return ctypes.cast(lib().scl_dns_derivs_cb, ctypes.c_void_p).value
