"""The Fortran drop-in: a dns-deform-style program (tests/fortran/dns_deform_gpu.f90)
linked against the library's Fortran module set (reference routine names,
iso_c_binding over the C ABI) and run on the GPU, checked against the FP64
oracle.  Modes: 1 = RHS from convec_dns/force/symmetrize/projection with the
RK4 stages in Fortran; 2 = rk4step; 3 = as 1 with u x w built from the
rfftwnd_f77_one_* entry points.  Bar: 2e-5 of max|v| after 5 steps at 32^3
(FP32 field from the library's FP32 initial condition)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "speccodelib_amd", "fortran", "_build", "dns_deform_gpu")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (amdflang)")
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_fortran_driver_matches_oracle(mode, tmp_path):
    n, nsteps = 32, 5
    out = tmp_path / f"v{mode}.bin"
    print(f"case fortran driver mode={mode}", flush=True)
    r = subprocess.run([EXE, str(n), str(nsteps), str(mode), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lx1 = n // 2 + 1
    v = np.fromfile(out, np.complex64).reshape(3, n, n, lx1)
    d = orc.DNS(orc.DNSParams(nx=n, nu=0.01, dt=0.002, forced=True, fkmax=2.0, eps=0.1))
    w = orc.initialize(n, n, n, 1, 4.0, 1.0)
    for _ in range(nsteps):
        w = d.rk4_step(*w)
    for a, b in zip(v, w):
        assert np.abs(a - b).max() < 2e-5 * np.abs(b).max()
