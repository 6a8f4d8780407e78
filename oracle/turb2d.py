"""Oracle (TEST INFRASTRUCTURE ONLY): the slab-decomposed 2D solver of
turb_2d.f90 at one rank (nproc = 1): its transform pair, the phase-shift
dealiased nonlinear term and its stepper.  Only tests/ import it, as the
checker.

Arrays are the reference's wz(my, mx2) indexed [iy, ix] (shape (n, n/2)):
kx = ix (0 .. n/2 - 1, no Nyquist column), ky = mod(iy + n/2, n) - n/2
(turb_2d.f90:524-545, k2(0, 0) = 0.5).

Transforms (newfft, turb_2d.f90:660-723, Ooura rdft/cdft):
  to k-space (isign = 1): rdft along x, / (n/2), a[1] (the x Nyquist)
    zeroed; cdft(+1) along y, / n.  Ooura's forward transforms use
    exp(+i k x), so C = (2/n^2) sum_x u(x) exp(+i k.x);
  to x-space (isign = -1): cdft(-1) along y, a[1] zeroed, rdft(-1) along x:
    u(x) = Re[ C0/2 + sum_{kx>=1} C exp(-i k.x) ] summed over ky.
  So C = 2 conj(D) for the usual normalised coefficients D of a real field.
symmetrize (:569-590): row ky = -n/2 and (0, 0) zeroed, the kx = 0 column
  averaged with its mirror for 1 <= iy <= n/2 - 2 (0-based; the loop stops
  short of ky = n/2 - 1).  dealiasing (:514-522): zero where
  k2 >= (2/9) n^2 (single precision constant).
Nonlinear term (:287-421): with W the vorticity, phi = pi (kx + ky)/n,
  u = inv(i ky W/k2), v = inv(-i kx W/k2), w = inv(W);
  vy  = -i/2 kx sd(fwd(u w)) - i/2 ky sd(fwd(v w))            (unshifted)
  vy += -i/2 kx s(e^{-i phi} fwd(u' w')) - i/2 ky s(e^{-i phi} fwd(v' w'))
  where u', v', w' come from the same spectra times e^{+i phi} (shifted),
  sd = symmetrize then dealiasing, s = symmetrize; then N = s(d(vy)).
Stepper (:440-483), L = (k sigma)^16 + (k/nu)^-4, f(a) = etak/conj(a) on
  kf <= |k| <= kf + 3, etak = eta/104 (:111):
  istep 0: w1/2 = w (1 - L dt/2) + dt/2 (N(w) + f(w)); keeps w0, N0.
  istep 1: w = w0 + dt (N(w1/2) + f(w1/2)) - L dt w1/2; owz = N0.
  istep >= 2: E = exp(-L dt) (1 at k = 0),
    w = [w + dt/2 (3 (N(w) + f(w)) - E (owz + f(owz)))] E,  owz = N(w).
  The forcing of the old term is evaluated at owz, the old nonlinear term
  (:470-471), not at the old vorticity; restated as written.
  time advances dt/2, dt/2, then dt per step.
Pinned by tests/test_oracle_turb2d.py."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def kgrid(n):
    """wavenumber (turb_2d.f90:524-545): kx, ky, k2 as [iy, ix] arrays."""
    kx = np.arange(n // 2, dtype=np.float64)[None, :] * np.ones((n, 1))
    ky = (np.mod(np.arange(n) + n // 2, n) - n // 2).astype(np.float64)[:, None] * np.ones((1, n // 2))
    k2 = kx * kx + ky * ky
    k2[0, 0] = 0.5
    return kx, ky, k2


def to_x(c):
    """newfft isign = -1: the real field u[y, x] of the spectrum c[iy, ix]."""
    n = c.shape[0]
    full = np.zeros((n, n), complex)
    full[:, : n // 2] = c
    full[:, 0] *= 0.5
    return np.real(np.fft.fft2(full))


def to_k(u):
    """newfft isign = 1: the spectrum c[iy, ix] of the real field u[y, x]."""
    n = u.shape[0]
    return 2.0 * np.fft.ifft2(u)[:, : n // 2]


def symmetrize(c):
    """turb_2d.f90:569-590."""
    c = c.copy()
    n = c.shape[0]
    c[n // 2, :] = 0
    c[0, 0] = 0
    for iy in range(1, n // 2 - 1):  # Fortran iy = 2 .. my/2 - 1
        c[iy, 0] = 0.5 * (c[iy, 0] + np.conj(c[n - iy, 0]))
    for iy in range(1, n // 2 - 1):
        c[n - iy, 0] = np.conj(c[iy, 0])
    return c


def dealias(c, k2):
    """turb_2d.f90:514-522: 8/9 rule, zero where k2 >= (2/9) my^2."""
    n = c.shape[0]
    ass = float(np.float32(np.float32(2.0) / np.float32(9.0)) * np.float32(n * n))
    return np.where(k2 >= ass, 0, c)


def nonlinear(w):
    """The dealiased nonlinear term vy of turb_2d.f90:287-421, then
    dealiasing and symmetrize (:428-429)."""
    n = w.shape[0]
    kx, ky, k2 = kgrid(n)
    ph = np.exp(1j * np.pi / n * (kx + ky))
    vx = symmetrize(1j * ky * w / k2)
    wt = to_x(w)
    a = dealias(symmetrize(to_k(to_x(vx) * wt)), k2)
    vy = a * (-0.5j) * kx
    vx = -1j * kx * w / k2
    b = dealias(symmetrize(to_k(to_x(vx) * wt)), k2)
    vy = vy + (-0.5j) * ky * b
    vxs = symmetrize(1j * ky * w / k2 * ph)
    wts = to_x(symmetrize(w * ph))
    c3 = symmetrize(to_k(to_x(vxs) * wts) * np.conj(ph))
    vy = vy + c3 * (-0.5j) * kx
    vxs = symmetrize(-1j * kx * w / k2 * ph)
    c4 = symmetrize(to_k(to_x(vxs) * wts) * np.conj(ph))
    vy = vy + c4 * (-0.5j) * ky
    return symmetrize(dealias(vy, k2))


@dataclass
class Turb2DParams:
    n: int
    dt: float
    sigma: float  # hyperviscous scale, (k sigma)^16
    nu: float     # hypoviscous scale, (k/nu)^-4
    eta: float    # enstrophy injection, etak = eta/104
    kf: float     # forced band kf <= |k| <= kf + 3


class Turb2D:
    """turb_2d.f90's main loop (:193-483) without its file output."""

    def __init__(self, p: Turb2DParams, w0, nonlinear_fn=None):
        self.p = p
        self.w = symmetrize(np.asarray(w0, complex))  # :220-221
        self.w0 = self.w.copy()
        self.owz = None
        self.istep = 0
        self.time = 0.0
        self.N = nonlinear_fn or nonlinear
        _, _, self.k2 = kgrid(p.n)
        k = np.sqrt(self.k2)
        self.L = (k * p.sigma) ** 16 + (k / p.nu) ** (-4)
        self.band = (k >= p.kf) & (k <= p.kf + 3.0)
        self.etak = p.eta / 104.0

    def f(self, a):
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(self.band, self.etak / np.conj(a), 0)

    def step(self):
        p, w, L = self.p, self.w, self.L
        dt, dth = p.dt, 0.5 * p.dt
        n = self.N(w)
        if self.istep == 0:
            self.w = w * (1.0 - L * dth) + dth * (n + self.f(w))
            self.n0 = n
            self.time += dth
        elif self.istep == 1:
            self.w = self.w0 + dt * (n + self.f(w)) - L * dt * w
            self.owz = self.n0
            self.time += dth
        else:
            E = np.where(self.k2 > 0.9, np.exp(-L * dt), 1.0)
            E[0, 0] = 1.0
            self.w = (w + dth * (3.0 * (n + self.f(w)) - E * (self.owz + self.f(self.owz)))) * E
            self.owz = n
            self.time += dt
        self.istep += 1
