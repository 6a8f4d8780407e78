"""Kill-and-restart through the reference's dump files (output.f90:103-166,
input.f90:106-170), on the GPU: an IF-RK4 run dumped after step k and read
into a new solver continues bit-identically to the run that went straight
through (DNS with the cross-step stage chain, graph and eager; static LES +
scalar); a truncated or corrupted dump is rejected with SCL_EINVAL and the
solver's state is left as it was."""
import numpy as np
import pytest

import speccodelib_amd as scl

pytestmark = pytest.mark.gpu


def _dns(p, graph, les=False):
    if les:
        return scl.DNS(p, nu=1e-3, dt=0.002, forced=1, fkmax=2.5, eps=0.1, sgs=1, cs2=0.0289, scalar=True, grad=1.0,
                       schmidt=0.7, use_graph=graph)
    return scl.DNS(p, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, use_graph=graph)


@pytest.mark.parametrize("graph,les", [(True, False), (False, False), (True, True)])
def test_restart_is_bit_identical(tmp_path, graph, les):
    n, k, K = 32, 3, 7
    print(f"case restart graph={graph} les={les}", flush=True)
    p = scl.Plan3D(n, n, n, scl.SP)
    v0 = scl.initialize(n, n, n, 1, 4.0, 1.0, scl.SP)
    phi0 = scl.initialize(n, n, n, -9, 3.0, 0.3, scl.SP)[0]
    a = _dns(p, graph, les)
    a.set_velocity(*v0)
    if les:
        a.set_scalar(phi0)
    a.step(K)
    a.sync()
    straight = a.get_velocity() + ((a.get_scalar(),) if les else ())
    a.close()
    b = _dns(p, graph, les)
    b.set_velocity(*v0)
    if les:
        b.set_scalar(phi0)
    b.step(k)
    b.sync()
    b.output(str(tmp_path), k)
    b.close()
    names = ["ux", "uy", "uz"] + (["phi"] if les else [])
    for nm in names:
        assert (tmp_path / f"{nm}{k}.dat").stat().st_size == 8 * (n // 2 + 1) * n * n + 8
    c = _dns(p, graph, les)
    c.input(str(tmp_path), k)
    c.step(K - k)
    c.sync()
    restarted = c.get_velocity() + ((c.get_scalar(),) if les else ())
    for x, y in zip(straight, restarted):
        np.testing.assert_array_equal(x, y)
    c.close()
    p.close()


def test_bad_dump_rejected_state_kept(tmp_path):
    n = 32
    p = scl.Plan3D(n, n, n, scl.SP)
    s = _dns(p, True)
    s.initialize(1, 4.0, 1.0)
    s.step(2)
    s.output(str(tmp_path), 5)
    before = s.get_velocity()
    uz = tmp_path / "uz5.dat"
    raw = uz.read_bytes()
    for bad in (raw[: len(raw) // 2], raw[:-4] + b"\x01\x02\x03\x04", b""):
        uz.write_bytes(bad)
        with pytest.raises(scl.SclError) as e:
            s.input(str(tmp_path), 5)
        assert e.value.status == -1, str(e.value)
        for x, y in zip(s.get_velocity(), before):
            np.testing.assert_array_equal(x, y)
    with pytest.raises(scl.SclError):
        s.input(str(tmp_path), 6)  # no such dump
    s.close()
    p.close()
