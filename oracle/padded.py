"""Oracle (TEST INFRASTRUCTURE ONLY): the 3/2-rule padded products of the
reference, at the reference's padded sizes.

Only tests/ import this module, as the checker; the product never does.

padd (padd.f90:9-13): copies the half spectrum (lx1, ly, lz) into a zeroed
(lxb1, lyb, lzb) one, the negative ky / kz halves at the top of their axes.
unpadd (unpadd.f90:9-12): the inverse selection.
convec_padded: the Lamb vector of spectral_heli_dynsmag.f90:229's padded
convec: u = c2r_b(padd(v)), w = c2r_b(padd(w)), u x w / (nxb nyb nzb),
r2c_b, unpadd (FFTW plans of fftwplan_les.f90:11-21 at 3 n/2).
smag_dealias: smag of smag_dealias.f90:19-75 (c1 = 0.026), with its strain
as written there: tau12 = i (kx vy + ky vx) without the 1/2 of the
symmetric strain, so S12b is twice S_12, |S| counts the off-diagonal
strain four times and the off-diagonal stresses come out doubled.
Pinned by tests/test_oracle_padded.py."""
from __future__ import annotations

import numpy as np

from . import spectral3d as s3


def padd(u, lxb1, lyb, lzb):
    """padd.f90:9-13 on a (lz, ly, lx1) array (C order of the Fortran (lx1, ly, lz))."""
    lz, ly, lx1 = u.shape
    ub = np.zeros((lzb, lyb, lxb1), u.dtype)
    hy, hz = ly // 2, lz // 2
    ub[:hz, :hy, :lx1] = u[:hz, :hy, :]
    ub[lzb - hz:, :hy, :lx1] = u[hz:, :hy, :]
    ub[:hz, lyb - hy:, :lx1] = u[:hz, hy:, :]
    ub[lzb - hz:, lyb - hy:, :lx1] = u[hz:, hy:, :]
    return ub


def unpadd(ub, lx1, ly, lz):
    """unpadd.f90:9-12."""
    lzb, lyb, _ = ub.shape
    hy, hz = ly // 2, lz // 2
    u = np.zeros((lz, ly, lx1), ub.dtype)
    u[:hz, :hy, :] = ub[:hz, :hy, :lx1]
    u[hz:, :hy, :] = ub[lzb - hz:, :hy, :lx1]
    u[:hz, hy:, :] = ub[:hz, lyb - hy:, :lx1]
    u[hz:, hy:, :] = ub[lzb - hz:, lyb - hy:, :lx1]
    return u


def _padded_sizes(n):
    nx, ny, nz = n
    return 3 * nx // 2, 3 * ny // 2, 3 * nz // 2


def convec_padded(vx, vy, vz, wx, wy, wz, n, nb=None):
    """The padded Lamb vector FFT(u x w) on the 3n/2 grid (or the padded grid
    nb), unpadded."""
    nx, ny, nz = n
    nxb, nyb, nzb = nb or _padded_sizes(n)
    lxb1 = nxb // 2 + 1
    lx1 = nx // 2 + 1
    u = [s3.c2r(padd(a.astype(np.complex128), lxb1, nyb, nzb), nxb) for a in (vx, vy, vz)]
    w = [s3.c2r(padd(a.astype(np.complex128), lxb1, nyb, nzb), nxb) for a in (wx, wy, wz)]
    c = 1.0 / (nxb * nyb * nzb)
    p = [(u[1] * w[2] - u[2] * w[1]) * c, (u[2] * w[0] - u[0] * w[2]) * c, (u[0] * w[1] - u[1] * w[0]) * c]
    return [unpadd(s3.r2c(a), lx1, ny, nz).astype(vx.dtype) for a in p]


def smag_dealias(vx, vy, vz, kx, ky, kz, n, delta, c1=float(np.float32(0.026)), nb=None):
    """smag of smag_dealias.f90:19-75; kx, ky, kz 1D wavenumbers; on the 3n/2
    grid or the padded grid nb."""
    nx, ny, nz = n
    nxb, nyb, nzb = nb or _padded_sizes(n)
    lxb1 = nxb // 2 + 1
    lx1 = nx // 2 + 1
    K = (kx[None, None, :], ky[None, :, None], kz[:, None, None])
    v = [a.astype(np.complex128) for a in (vx, vy, vz)]
    t = [1j * K[0] * v[0], 1j * (K[0] * v[1] + K[1] * v[0]), 1j * (K[0] * v[2] + K[2] * v[0]),
         1j * K[1] * v[1], 1j * (K[1] * v[2] + K[2] * v[1]), 1j * K[2] * v[2]]  # 11 12 13 22 23 33
    S = [s3.c2r(padd(a, lxb1, nyb, nzb), nxb) for a in t]
    a = np.sqrt(2.0 * (S[0] ** 2 + S[3] ** 2 + S[5] ** 2 + 2.0 * (S[1] ** 2 + S[2] ** 2 + S[4] ** 2)))
    const = 1.0 / (nxb * nyb * nzb)
    return [unpadd(s3.r2c(-2.0 * c1 * delta**2 * a * s * const), lx1, ny, nz).astype(vx.dtype) for s in S]
