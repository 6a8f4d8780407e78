"""Oracle (TEST INFRASTRUCTURE ONLY): the 2D vorticity RHS and IF-RK4 step
of the batched ensembles.  Only tests/ import it, as the checker.

Arrays are (ny, nx/2+1) complex (the reference's complex(sp)(lx1, ly)).

RHS (spectral_dns_2d.f90:161-188): u = i ky w/k^2, v = -i kx w/k^2; c2r of
u, v, w; F = (u w, v w)/N^2; A = i kx F_x + i ky F_y = div(u w).
The reference negates A only inside the forced band (:181-185), so outside
it the advection has the wrong sign; here N = -A + f everywhere, with the
band forcing f = eps/(conj(w) + tiny) on fkmin^2 <= k^2 <= fkmax^2
(:182-183), then symmetrize2d and the 2/3 cut (:188).
Integrating factor (:94) generalised with a linear damping term:
rate(k) = c_hyper k^n_hyper + c_hypo k^-n_hypo + damp
(the reference's (k sigma)^nhyper + (nu/k)^nhypo is c_hyper = sigma^nhyper,
c_hypo = nu^nhypo, damp = 0).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import spectral3d as s3


def wavenumber2d(n, prec="dp"):
    rd = s3.real_dtype(prec)
    kx = np.arange(n // 2 + 1, dtype=rd)
    ky = np.array([((j + n // 2) % n) - n // 2 for j in range(n)], dtype=rd)
    k2 = (kx[None, :] ** 2 + ky[:, None] ** 2).astype(rd)
    k2[0, 0] = s3.tiny(prec)
    return kx, ky, k2


@dataclass
class Vort2DParams:
    n: int
    dt: float = 5e-4
    c_hyper: float = 1e-5
    n_hyper: int = 8
    c_hypo: float = 0.0
    n_hypo: int = 0
    damp: float = 0.1
    forced: bool = True
    fkmin: float = 4.0
    fkmax: float = 5.0
    eps: float = 0.01


class Vort2D:
    def __init__(self, p: Vort2DParams):
        self.p = p
        self.kx, self.ky, self.k2 = wavenumber2d(p.n)
        k = np.sqrt(self.k2)
        rate = p.c_hyper * k ** p.n_hyper + p.damp
        if p.c_hypo:
            rate = rate + p.c_hypo * k ** (-p.n_hypo)
        self.rate = rate
        lx = p.n // 2
        kcut = 2.0 * lx / 3.0
        self.kcut2 = kcut * kcut

    def rhs(self, w):
        p, n = self.p, self.p.n
        KX, KY = self.kx[None, :], self.ky[:, None]
        u = 1j * KY * w / self.k2
        v = -1j * KX * w / self.k2
        uu, vv, ww = (s3.c2r2(a, n) for a in (u, v, w))
        const = 1.0 / (n * n)
        fx = s3.r2c2(uu * ww) * const
        fy = s3.r2c2(vv * ww) * const
        r = -(1j * KX * fx + 1j * KY * fy)
        if p.forced:
            band = (self.k2 >= p.fkmin**2) & (self.k2 <= p.fkmax**2)
            r[band] = r[band] + p.eps / (np.conj(w[band]) + s3.tiny("dp"))
        s3.symmetrize2d(r)
        r[self.k2 >= self.kcut2] = 0
        return r.astype(w.dtype)

    def rk4_step(self, w):
        """Lawson RK4 (dns-deform.f90:164-228) written without dividing by
        E: with k^8 hyperviscosity E = exp(-rate dt/2) underflows at high k
        and the stage increments dt N / E would overflow.  With a_i = dt N_i:
        s2 = E w + E a1/2, s3 = E w + a2/2, s4 = E^2 w + E a3,
        w' = E^2 w + (E^2 a1 + 2 E a2 + 2 E a3 + a4)/6 (the same scheme)."""
        dt = self.p.dt
        E = np.exp(-self.rate * dt / 2)
        E2 = E * E
        a1 = dt * self.rhs(w)
        a2 = dt * self.rhs(E * w + E * a1 / 2)
        a3 = dt * self.rhs(E * w + a2 / 2)
        a4 = dt * self.rhs(E2 * w + E * a3)
        out = E2 * w + (E2 * a1 + 2 * E * a2 + 2 * E * a3 + a4) / 6
        out[self.k2 >= self.kcut2] = 0
        return out


def initialize2d(n, iseed, k0=8.0, energy=0.5):
    """Seeded random-phase vorticity: one ran1/gasdev draw per mode in
    (jj, ii) order (initialize.f90:485-492), symmetrize2d, 2/3 cut, then
    |w| shaped so that E(k) ~ k exp(-(k/k0)^2) (|w| ~ k exp(-k^2/(2 k0^2)))
    and rescaled to sum' |w|^2/k^2 = energy."""
    lx1 = n // 2 + 1
    g = s3.gauss3(iseed, (n * lx1 + 2) // 3 + 1, "dp").reshape(-1)[: n * lx1].reshape(n, lx1)
    w = g.copy()
    s3.symmetrize2d(w)
    kx, ky, k2 = wavenumber2d(n)
    kcut = 2.0 * (n // 2) / 3.0
    w[k2 >= kcut * kcut] = 0
    k = np.sqrt(k2)
    w = w * k * np.exp(-k2 / (2 * k0 * k0))
    e = np.abs(w) ** 2 / k2
    e[:, 0] *= 0.5
    return w * np.sqrt(energy / e.sum())


def energy2d(w, k2):
    e = np.abs(w.astype(np.complex128)) ** 2 / k2
    e[:, 0] *= 0.5
    return float(e.sum())


def enstrophy2d(w):
    e = np.abs(w.astype(np.complex128)) ** 2
    e[:, 0] *= 0.5
    return float(e.sum())
