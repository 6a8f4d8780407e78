"""The solver's sticky non-finite status through the C ABI: a state holding a
NaN, stepped once, makes scl_dns_sync and scl_dns_diagnostics return
SCL_ENONFINITE (diagnostics still filled, divmax = inf); a finite state set
afterwards clears it.  DNS and LES + scalar paths, FP32 and FP64."""
import ctypes

import numpy as np
import pytest

import speccodelib_amd as scl

pytestmark = pytest.mark.gpu

ENONFINITE = -5


@pytest.mark.parametrize("prec,sgs,scalar", [(scl.SP, 0, False), (scl.DP, 0, False), (scl.SP, 1, True)])
def test_nan_state_reports_enonfinite(prec, sgs, scalar):
    n = 32
    print(f"case nonfinite status prec={prec} sgs={sgs} scalar={scalar}", flush=True)
    p = scl.Plan3D(n, n, n, prec)
    d = scl.DNS(p, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, sgs=sgs, cs2=0.0289 if sgs else 0.0,
                scalar=scalar, grad=1.0, schmidt=0.7, use_graph=True)
    v = list(scl.initialize(n, n, n, 1, 4.0, 1.0, prec))
    if scalar:
        d.set_scalar(p.zeros())
    d.set_velocity(*v)
    d.step(2)
    d.sync()  # finite: no error
    assert d.diagnostics()["nonfinite"] == 0
    bad = [a.copy() for a in v]
    bad[1][3, 2, 1] = np.nan
    d.set_velocity(*bad)
    d.step(1)
    assert scl.lib().scl_dns_sync(d.h) == ENONFINITE
    with pytest.raises(scl.NonFiniteError):
        d.sync()
    dg = scl.ScDiag()
    assert scl.lib().scl_dns_diagnostics(d.h, ctypes.byref(dg)) == ENONFINITE
    assert dg.nonfinite == 1 and not np.isfinite(dg.divmax)
    with pytest.raises(scl.NonFiniteError):
        d.diagnostics()
    assert d.diagnostics(allow_nonfinite=True)["nonfinite"] == 1
    # a finite state clears the flag (the NaN has reached the scalar too)
    if scalar:
        d.set_scalar(p.zeros())
    d.set_velocity(*v)
    d.step(1)
    d.sync()
    assert d.diagnostics()["nonfinite"] == 0
    d.close()
    p.close()
