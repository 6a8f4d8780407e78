"""scl_dns_spectra_noalloc through the C ABI: the same bits as
scl_dns_spectrum, scl_dns_helicity_spectrum and scl_dns_diagnostics' totals,
on two calls in a row, the second on a smaller grid with fewer shells (the
static partial-sum array keeps the first call's larger contents); sizes
beyond the array are rejected."""
import ctypes

import numpy as np
import pytest

import speccodelib_amd as scl

pytestmark = pytest.mark.gpu


def _noalloc(s, nbins, torch):
    out = torch.full((2 * nbins + 8,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = scl.lib().scl_dns_spectra_noalloc(s.h, ctypes.c_void_p(out.data_ptr()), nbins)
    assert st == 0, scl.lib().scl_last_error().decode()
    s.sync()
    return out.cpu().numpy()


def test_noalloc_matches_allocating_entry_points():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    for n, nbins, prec in ((64, 32, scl.SP), (32, 12, scl.DP)):
        print(f"case spectra_noalloc n={n} nbins={nbins} prec={prec}", flush=True)
        p = scl.Plan3D(n, n, n, prec)
        s = scl.DNS(p, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1)
        s.initialize(-2, 4.0, 1.0)
        s.step(2)
        s.sync()
        got = _noalloc(s, nbins, torch)
        ek = np.zeros(nbins)
        hk = np.zeros(nbins)
        assert scl.lib().scl_dns_spectrum(s.h, ek.ctypes.data, nbins) == 0
        assert scl.lib().scl_dns_helicity_spectrum(s.h, hk.ctypes.data, nbins) == 0
        dg = s.diagnostics()
        assert np.array_equal(got[:nbins], ek)
        assert np.array_equal(got[nbins:2 * nbins], hk)
        t = got[2 * nbins:]
        assert t[0] == dg["energy"] and t[1] == dg["enstrophy"] and t[2] == dg["helicity"]
        assert t[4] == 0 and t[6] == 0 and t[7] == 0
        assert 2 * 0.01 * t[5] == dg["helicity_dissipation"]
        # beyond nx/2 (and beyond the array) is rejected
        assert scl.lib().scl_dns_spectra_noalloc(s.h, ctypes.c_void_p(0), nbins) == -1
        assert scl.lib().scl_dns_spectra_noalloc(s.h, ctypes.c_void_p(1 << 20), n // 2 + 1) == -1
        s.close()
        p.close()
