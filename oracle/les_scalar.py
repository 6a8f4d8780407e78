"""Oracle (TEST INFRASTRUCTURE ONLY): Smagorinsky SGS stress (static and
dynamic), passive-scalar flux divergence, and the LES + scalar IF-RK4 step.

Only tests/ import this module, as the checker; the product never does.
Arrays follow oracle.spectral3d's convention, shape (nz, ny, nx/2+1).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import spectral3d as s3

PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]  # t11 t12 t13 t22 t23 t33


def _k(kx, ky, kz):
    return kx[None, None, :], ky[None, :, None], kz[:, None, None]


def strain(vx, vy, vz, kx, ky, kz):
    """S_ij = (i/2)(k_i v_j + k_j v_i) in (11,12,13,22,23,33) order
    (dynsmag.f90:198-203, smag.f90:24-29)."""
    K = _k(kx, ky, kz)
    v = (vx, vy, vz)
    out = []
    for i, j in PAIRS:
        if i == j:
            out.append(1j * K[i] * v[i])
        else:
            out.append(0.5j * (K[j] * v[i] + K[i] * v[j]))
    return [a.astype(vx.dtype) for a in out]


def _abs_strain(S):
    return np.sqrt(2.0 * (S[0] ** 2 + S[3] ** 2 + S[5] ** 2 + 2.0 * (S[1] ** 2 + S[2] ** 2 + S[4] ** 2)))


def smag_tau(vx, vy, vz, kx, ky, kz, cs2, delta, n):
    """Static Smagorinsky: tau_ij = FFT(-2 cs2 Delta^2 |S| S_ij)/N^3
    (dynsmag.f90:196-233 with update = .false.; smag.f90:19-75 is this with
    cs2 = c1 = 0.026).  Pinned by tests/test_oracle_les.py."""
    nx, ny, nz = n
    const = 1.0 / (nx * ny * nz)
    Sp = [s3.c2r(a, nx) for a in strain(vx, vy, vz, kx, ky, kz)]
    a = _abs_strain(Sp)
    return [s3.r2c(-cs2 * 2.0 * delta * delta * a * s * const).astype(vx.dtype) for s in Sp]


def g_test(k2, k_test):
    """Sharp spectral test filter: 1 where k2 <= k_test^2 (spectral_les.f90:105-109)."""
    return np.where(k2 <= k_test * k_test, 1.0, 0.0).astype(k2.dtype)


def dynsmag_update(vx, vy, vz, kx, ky, kz, g, delta, delta_test, n):
    """Germano-Lilly dynamic coefficient with a sharp test filter
    (dynsmag.f90:27-194): L_ij = filt(u_i u_j) - filt(u_i) filt(u_j),
    M_ij = filt(2 D^2 |S| S_ij) - 2 Dt^2 |filt S| filt(S_ij),
    cs2 = <L:M>/<M:M> over the grid, tau_ij = -cs2 FFT(2 D^2 |S| S_ij)/N^3.
    Returns (tau (6), mm, lm, cs2).  Pinned by tests/test_oracle_les.py."""
    nx, ny, nz = n
    const = 1.0 / (nx * ny * nz)
    v = (vx, vy, vz)
    u = [s3.c2r(a, nx) for a in v]
    ut = [s3.c2r(g * a, nx) for a in v]
    L = [s3.c2r(g * s3.r2c(u[i] * u[j]) * const, nx) - ut[i] * ut[j] for i, j in PAIRS]
    S = strain(vx, vy, vz, kx, ky, kz)
    Sp = [s3.c2r(a, nx) for a in S]
    a = _abs_strain(Sp)
    s2 = [s3.r2c(2.0 * delta * delta * a * s * const) for s in Sp]
    St = [s3.c2r(g * s, nx) for s in S]
    at = _abs_strain(St)
    M = [s3.c2r(g * s2[q], nx) - 2.0 * delta_test * delta_test * at * St[q] for q in range(6)]
    w = [1.0, 2.0, 2.0, 1.0, 2.0, 1.0]
    mm = float(sum(w[q] * (M[q] * M[q]).sum() for q in range(6)))
    lm = float(sum(w[q] * (L[q] * M[q]).sum() for q in range(6)))
    cs2 = lm / mm
    tau = [(-cs2 * s).astype(vx.dtype) for s in s2]
    return tau, mm, lm, cs2


def convecphi_dns(vx, vy, vz, phi, kx, ky, kz, n):
    """wphi = i k . FFT(u phi)/N^3, the reference's sign (convecphi_dns.f90:17-47:
    it returns +div(u phi), which the drivers add without negating).
    Pinned by tests/test_oracle_les.py."""
    nx, ny, nz = n
    const = 1.0 / (nx * ny * nz)
    K = _k(kx, ky, kz)
    p = s3.c2r(phi, nx)
    out = 0
    for q, a in enumerate((vx, vy, vz)):
        out = out + 1j * K[q] * s3.r2c(s3.c2r(a, nx) * p)
    return (out * const).astype(vx.dtype)


@dataclass
class LESParams:
    nx: int
    ny: int | None = None  # default nx
    nz: int | None = None  # default nx
    nu: float = 1e-4
    dt: float = 0.001
    forced: bool = True
    fkmax: float = 2.5
    eps: float = 0.1
    cs2: float = 0.17**2  # static coefficient; dynamic when dynamic=True
    delta: float | None = None  # default 2 pi / nx
    dynamic: bool = False
    iupdate: int = 10
    delta_ratio: float = 2.0
    scalar: bool = True
    G: float = 1.0  # mean scalar gradient along z
    Sc: float = 0.7

    def __post_init__(self):
        if self.delta is None:
            self.delta = 2 * np.pi / self.nx
        self.ny = self.ny or self.nx
        self.nz = self.nz or self.nx


class LESScalar:
    """IF-RK4 (dns-deform.f90:160-237 stepping) of the LES momentum RHS
    N = P S [u x w + f - i k_j tau_ij] (spectral_les.f90:205-238) and the
    scalar RHS R = -i k . FFT(u phi)/N^3 - G u_z (spectral_dns_scalar.f90:236,
    245; the physical sign of the advection), symmetrized, with integrating
    factors exp(-nu k^2 t) and exp(-nu k^2 t / Sc) (:272)."""

    def __init__(self, p: LESParams):
        self.p = p
        n = p.nx
        self.n = (n, p.ny, p.nz)
        self.kx, self.ky, self.kz, self.k2 = s3.wavenumber(*self.n)
        self.cs2 = p.cs2
        lx = n // 2
        kcut = 2.0 * lx / 3.0
        self.kcut2 = kcut * kcut
        k_test = kcut / p.delta_ratio
        self.g = g_test(self.k2, k_test)
        self.delta_test = p.delta_ratio * p.delta
        self.nrhs = 0
        self.rhs_per_step = 4  # the dynamic update's cadence: every iupdate steps (IF-AB2: 1 RHS a step)

    def rhs(self, vx, vy, vz, phi):
        p = self.p
        if p.dynamic and self.nrhs % (self.rhs_per_step * p.iupdate) == 0:
            _, _, _, self.cs2 = dynsmag_update(vx, vy, vz, self.kx, self.ky, self.kz, self.g, p.delta,
                                               self.delta_test, self.n)
        self.nrhs += 1
        K = _k(self.kx, self.ky, self.kz)
        wx, wy, wz = s3.vorticity(vx, vy, vz, self.kx, self.ky, self.kz)
        wx, wy, wz = s3.convec_dns(vx, vy, vz, wx, wy, wz, *self.n)
        if p.forced:
            s3.force(wx, wy, wz, vx, vy, vz, self.k2, p.fkmax, p.eps)
        t = smag_tau(vx, vy, vz, self.kx, self.ky, self.kz, self.cs2, p.delta, self.n)
        wx = wx - 1j * (K[0] * t[0] + K[1] * t[1] + K[2] * t[2])
        wy = wy - 1j * (K[0] * t[1] + K[1] * t[3] + K[2] * t[4])
        wz = wz - 1j * (K[0] * t[2] + K[1] * t[4] + K[2] * t[5])
        for a in (wx, wy, wz):
            s3.symmetrize3d(a)
        s3.projection(wx, wy, wz, self.kx, self.ky, self.kz)
        if p.scalar:
            r = -convecphi_dns(vx, vy, vz, phi, self.kx, self.ky, self.kz, self.n) - p.G * vz
            s3.symmetrize3d(r)
        else:
            r = np.zeros_like(phi)
        return wx, wy, wz, r

    def rk4_step(self, vx, vy, vz, phi):
        p = self.p
        E = np.exp(-p.nu * self.k2 * p.dt / 2)
        Ep = np.exp(-p.nu / p.Sc * self.k2 * p.dt / 2)
        Es = [E, E, E, Ep]
        v = (vx, vy, vz, phi)
        n = self.rhs(*v)
        k = [p.dt * a for a in n]
        du = [a.copy() for a in k]
        s = [e * (x + a / 2) for e, x, a in zip(Es, v, k)]
        n = self.rhs(*s)
        k = [p.dt * a / e for a, e in zip(n, Es)]
        du = [d + 2 * a for d, a in zip(du, k)]
        s = [e * (x + a / 2) for e, x, a in zip(Es, v, k)]
        n = self.rhs(*s)
        k = [p.dt * a / e for a, e in zip(n, Es)]
        du = [d + 2 * a for d, a in zip(du, k)]
        s = [e * e * (x + a) for e, x, a in zip(Es, v, k)]
        n = self.rhs(*s)
        k = [p.dt * a / e / e for a, e in zip(n, Es)]
        du = [d + a for d, a in zip(du, k)]
        out = [e * e * (x + d / 6) for e, x, d in zip(Es, v, du)]
        m = self.k2 >= self.kcut2
        for a in out:
            a[m] = 0
        return tuple(out)


def scalar_variance(phi):
    """<phi^2>/2 = sum' |phi|^2 (half weight on kx = 0)."""
    t = np.abs(phi.astype(np.complex128)) ** 2
    return float(t.sum() - 0.5 * t[:, :, 0].sum())
