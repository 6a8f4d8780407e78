"""GPU parity of the 3/2-rule padded products (the padded convec of
spectral_heli_dynsmag.f90:229 and smag of smag_dealias.f90) against
oracle/padded.py, through the C ABI.

The library pads each axis to 2n (the smallest power of two of at least
3n/2).  Bars:
* convec_padded of fields with empty Nyquist planes (what the solvers hold):
  the reference's 3n/2-grid product to 1e-12 (FP64) / 2e-6 (FP32); any
  fields: the oracle on the 2n grid to the same bars;
* smag_dealias: the oracle on the 2n grid to 1e-12 / 2e-6; against the
  reference's 3n/2 grid |S| is not a polynomial, so the padded grids differ
  by a few 1e-4 of the largest stress at 32^3 (bound 2e-3 here)."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import padded as pd
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return max(np.abs(np.asarray(x, np.complex128) - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


def fields(n, seed, nyquist):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(6):
        a = rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))
        if not nyquist:
            a[n // 2] = 0
            a[:, n // 2] = 0
            a[:, :, n // 2] = 0
        orc.symmetrize3d(a)
        out.append(a)
    return out


@pytest.mark.parametrize("prec,tol", [(scl.DP, 1e-12), (scl.SP, 2e-6)])
@pytest.mark.parametrize("n", [16, 32])
def test_convec_padded(prec, tol, n):
    print(f"case convec_padded n={n} prec={prec}", flush=True)
    dt = np.complex128 if prec == scl.DP else np.complex64
    p = scl.Plan3D(n, n, n, prec)
    for nyq in (False, True):
        f = fields(n, 10 + n, nyq)
        g = [a.astype(dt) for a in f]
        p.convec_padded(*g)
        want = pd.convec_padded(*f, (n, n, n), nb=(2 * n,) * 3)
        e2 = relerr(g[3:], want)
        print(f"nyquist={nyq}: vs oracle on the 2n grid {e2:.2e}", flush=True)
        assert e2 < tol
        if not nyq:
            e3 = relerr(g[3:], pd.convec_padded(*f, (n, n, n)))
            print(f"vs oracle on the 3n/2 grid {e3:.2e}", flush=True)
            assert e3 < tol
    p.close()


@pytest.mark.parametrize("prec,tol", [(scl.DP, 1e-12), (scl.SP, 2e-6)])
def test_smag_dealias(prec, tol):
    n = 32
    print(f"case smag_dealias n={n} prec={prec}", flush=True)
    p = scl.Plan3D(n, n, n, prec)
    v = scl.initialize(n, n, n, 1, 4.0, 1.0, prec)
    k1 = orc.wavenumber(n, n, n)
    K = [np.ascontiguousarray(np.broadcast_to(a, (n, n, n // 2 + 1)), dtype=v[0].real.dtype)
         for a in (k1[0][None, None, :], k1[1][None, :, None], k1[2][:, None, None])]
    delta = 2 * np.pi / n
    got = p.smag_dealias(*v, *K, delta)
    vd = [a.astype(np.complex128) for a in v]
    e2 = relerr(got, pd.smag_dealias(*vd, *k1[:3], (n, n, n), delta, nb=(2 * n,) * 3))
    e3 = relerr(got, pd.smag_dealias(*vd, *k1[:3], (n, n, n), delta))
    print(f"vs oracle on the 2n grid {e2:.2e}, on the 3n/2 grid {e3:.2e}", flush=True)
    assert e2 < tol and e3 < 2e-3
    p.close()


def test_padded_size_checks():
    """Padded sizes below 3n/2, or inconsistent ones, are SCL_EINVAL."""
    print("case padded size checks", flush=True)
    p = scl.Plan3D(16, 16, 16, scl.SP)
    z = [p.zeros() for _ in range(6)]
    for dims in ((9, 16, 16, 16, 16, 16), (13, 24, 23, 24, 24, 24), (13, 24, 24, 24, 24, 22)):
        assert scl.lib().scl_convec_padded(p.h, *[scl._ptr(a) for a in z], 9, 16, 16, *dims) == -1
    p.close()
