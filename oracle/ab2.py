"""Oracle (TEST INFRASTRUCTURE ONLY): the integrating-factor Adams-Bashforth
stepper of the production drivers, with its bootstrap and CFL time step.

Only tests/ import this module, as the checker; the product never does.

spectral_dns_dealiased.f90:124-285 (variable dt):
  step 0 (:237-245): Euler, v1 = (v0 + dt N0) E(dt);
  step 1 (:247-266): v2 = [S E(dto) v0 + ek v1 + delta N1] E(dt) with
    dtt = dt/dto, S = (1 + dtt^2)/2, ek = (1 - dtt^2)/2,
    delta = dto (1 + dtt)^2 / 2; then dto <- dt + dto and the old RHS is N0;
  steps >= 2 (:268-285): v <- [v + (1 + dtt/2) dt N - (dtt/2) dt E(dto) O] E(dt),
    O <- N, dto <- dt;
  E(t) = exp(-nu k^2 t); after every step k^2 >= kcut^2 and the mean are
  zeroed (:288-296).  dt = min(beta (pi/lx) / umax, 3 dto) (:227), cut so
  that the step ends on the next output time twrite or halves the remaining
  interval (:228-232); umax = 3.5 sqrt(2 E) at steps that are multiples of
  ieout (:129-142, E with half weight on kx = 0); twrite advances by dtw at
  each output time (:146-149).  The first dto is timemax (:113), and
  umax starts at 3 sqrt(3) u0 (:111).
specdns-forced-omp-ft.f90:224-274 (fixed dt) is the same scheme with the
first two steps of dt/2 (a half Euler step and a midpoint step), then
v <- [v + (3 N - E(dt) O) dt/2] E(dt).

The RHS is oracle.spectral3d.DNS.rhs (ratesdudt: vorticity, convec_dns,
force, symmetrize, projection); the production drivers skip symmetrize,
which changes only the kx = 0 plane at rounding level and planes the cut
removes.  Pinned by tests/test_oracle_ab2.py (exact viscous decay with
N = 0 over a variable-dt sequence, second-order convergence to a fine
IF-RK4 run, a scalar linear ODE with variable steps)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import spectral3d as s3

MYEPS_SP = 2.0 * float(np.finfo(np.float32).eps)  # constant_sp.f90: myeps


@dataclass
class AB2Params:
    nx: int
    nu: float = 0.01
    forced: bool = True
    fkmax: float = 2.0
    eps: float = 0.1
    beta: float = 0.5        # CFL factor: dt = beta (pi/lx) / umax
    dtw: float = 1.0         # output interval (twrite advances by it)
    ieout: int = 10          # umax refresh interval in steps
    umax: float = 3.0 * np.sqrt(3.0)  # umax before the first refresh (3 sqrt(3) u0)
    dto0: float = 1.0e30     # the first step's dto (timemax in the reference)
    kcut: float | None = None  # truncation radius; default lx - 1 (spectral_dns_dealiased.f90:77)
    dt_fixed: float = 0.0    # > 0: specdns-forced-omp-ft.f90's fixed step
    time: float = 0.0
    myeps: float = MYEPS_SP

    @property
    def lx(self):
        return self.nx // 2


def ab2_update(istep, v, n, aux, dt, dto, E_dt, E_dto):
    """The update of one IF-AB2 step for every mode: v (current state), n
    (its RHS), aux (step 1: v0; steps >= 2: the old RHS).  Returns
    (new v, new aux, new dto) before the truncation."""
    if istep == 0:
        return (v + dt * n) * E_dt, v.copy(), dt
    dtt = dt / dto
    if istep == 1:
        S = 0.5 * (1.0 + dtt * dtt)
        ek = 0.5 * (1.0 - dtt * dtt)
        delta = 0.5 * dto * (1.0 + dtt) ** 2
        return (S * aux * E_dto + ek * v + delta * n) * E_dt, None, dt + dto
    S = (1.0 + 0.5 * dtt) * dt
    ek = -0.5 * dtt * dt
    return (v + S * n + ek * aux * E_dto) * E_dt, n, dt


class AB2:
    """The IF-AB2 loop of spectral_dns_dealiased.f90 on three velocity fields,
    or, with rhs (a function of the fields returning their RHS) and nus (one
    viscosity per field), the same loop of spectral_les.f90:241-304 and
    spectral_dns_scalar.f90:257-338 (velocity and scalar, the scalar with
    nu/Sc); umax comes from the first three fields."""

    def __init__(self, p: AB2Params, v0, rhs=None, nus=None):
        self.p = p
        n = p.nx
        self.dns = s3.DNS(s3.DNSParams(nx=n, nu=p.nu, dt=0.0, forced=p.forced, fkmax=p.fkmax, eps=p.eps))
        self.rhs = rhs or self.dns.rhs
        self.nus = list(nus) if nus is not None else [p.nu] * len(v0)
        self.k2 = self.dns.k2.astype(np.float64)
        kcut = p.kcut if p.kcut is not None else p.lx - 1.0
        self.kcut2 = kcut * kcut
        self.v = [a.astype(np.complex128).copy() for a in v0]
        self.aux = None
        self.n0 = None
        self.istep = -1
        self.time = p.time
        self.dto = p.dto0
        self.twrite = p.time
        self.umax = p.umax
        self.dts = []

    def _dt(self):
        p = self.p
        if p.dt_fixed > 0:
            return 0.5 * p.dt_fixed if self.istep < 2 else p.dt_fixed
        dt = min(p.beta * (np.pi / p.lx) / self.umax, 3.0 * self.dto)
        if p.dtw > 0:
            rem = self.twrite - self.time
            if dt >= rem:
                dt = rem
            elif rem - dt < 0.2 * dt:
                dt = 0.5 * rem
        return dt

    def step(self):
        p = self.p
        self.istep += 1
        if p.dt_fixed <= 0:
            if p.ieout > 0 and self.istep % p.ieout == 0:
                self.umax = 3.5 * np.sqrt(2.0 * s3.energy(*self.v[:3]))
            if p.dtw > 0 and abs(self.time - self.twrite) <= p.myeps:
                self.twrite += p.dtw
        n = self.rhs(*self.v)
        dt = self._dt()
        new = []
        aux = []
        for q in range(len(self.v)):
            E_dt = np.exp(-self.nus[q] * self.k2 * dt)
            E_dto = np.exp(-self.nus[q] * self.k2 * self.dto) if self.istep > 0 else None
            a = self.aux[q] if self.aux is not None else None
            vn, an, dto = ab2_update(self.istep, self.v[q], n[q], a, dt, self.dto, E_dt, E_dto)
            new.append(vn)
            aux.append(an)
        if self.istep == 0:
            self.aux = aux          # v0, for step 1
            self.n0 = [a.copy() for a in n]
        elif self.istep == 1:
            self.aux = self.n0      # the old RHS of the first AB2 step is N0
        else:
            self.aux = aux
        self.dto = dto
        m = self.k2 >= self.kcut2
        for a in new:
            a[m] = 0
            a[0, 0, 0] = 0
        self.v = new
        self.time += dt
        self.dts.append(dt)
        return self.v
