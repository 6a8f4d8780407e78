"""Pins of the oracle's diagnostics and dynamic Smagorinsky coefficient that
do not restate the oracle's own formulas:

* shell binning of E(k) and H(k) (spec.f90:82, helispec.f90:82): exact
  single-mode placements (k^2 = 3 -> shell 2, 12 -> 3, 13 -> 4, ...), the
  half weight on kx = 0, and an ABC (Beltrami) field whose k = 1 shell holds
  E(1) = <|u|^2>/2 = 3/2 and H(1) = 2 E(1);
* dissipation and enstrophy totals against <2 nu s_ij s_ij> and <|w|^2>/2
  computed in physical space (viscenerdiss.f90:16-47, enstrophy.f90:15-32);
* dynsmag_update (dynsmag.f90:27-194) against a brute-force physical-space
  evaluation on 8^3: direct DFT sums (no FFT), full 3x3 tensor contractions
  for L:M and M:M (so the 1,2,2,1,2,1 weights of :170-184 are not retyped).

Each check is also run on copies of the oracle function with a plausible
mistake planted (floor or ceil binning, a dropped half weight, a flipped
vorticity sign, the sign of M's second term, unit off-diagonal weights, a
dropped 1/N^3) and must fail on every one."""
import inspect

import numpy as np
import pytest

from oracle import les_scalar as les
from oracle import spectral3d as orc


# --------------------------------------------------------------------------
# shell binning
def _single_mode(n, kx, ky, kz, amp=1.0 + 0.0j):
    v = [np.zeros((n, n, n // 2 + 1), complex) for _ in range(3)]
    v[0][kz % n, ky % n, kx] = amp
    return v


def _check_binning(spectrum):
    n = 16
    k2 = orc.wavenumber(n, n, n)[3]
    # (kx, ky, kz) -> expected shell ll = floor(|k| + 1/2)
    cases = [((1, 1, 0), 1), ((1, 1, 1), 2), ((2, 1, 1), 2), ((2, 1, -1), 2), ((2, 2, 1), 3),
             ((2, -2, 2), 3), ((3, 2, 0), 4), ((0, 3, -2), 4), ((3, 3, 3), 5), ((4, 4, 2), 6),
             ((5, 0, 0), 5), ((0, 0, -7), 7)]
    for (kx, ky, kz), ll in cases:
        e = spectrum(*_single_mode(n, kx, ky, kz), k2)
        w = 0.5 if kx == 0 else 1.0
        want = np.zeros(n // 2)
        want[ll - 1] = w
        assert np.array_equal(e, want), ((kx, ky, kz), ll, e)
    # the mean and modes beyond lx = n/2 are in no shell
    assert not spectrum(*_single_mode(n, 0, 0, 0), k2).any()
    assert not spectrum(*_single_mode(n, 8, 7, 7), k2).any()


def test_shell_binning_exact():
    _check_binning(orc.spectrum)


def _abc_field(n):
    """ABC flow (A = B = C = 1) on n^3: curl u = u, <|u|^2> = 3."""
    x = 2 * np.pi * np.arange(n) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = [np.sin(Z) + np.cos(Y), np.sin(X) + np.cos(Z), np.sin(Y) + np.cos(X)]
    return [orc.r2c(a) / n**3 for a in u]


def _check_abc(spectrum, helicity_spectrum):
    n = 16
    kx, ky, kz, k2 = orc.wavenumber(n, n, n)
    v = _abc_field(n)
    e = spectrum(*v, k2)
    h = helicity_spectrum(*v, kx, ky, kz, k2)
    assert abs(e[0] - 1.5) < 1e-13 and np.abs(e[1:]).max() < 1e-13
    assert abs(h[0] - 3.0) < 1e-13 and np.abs(h[1:]).max() < 1e-13


def test_abc_spectrum_and_helicity():
    _check_abc(orc.spectrum, orc.helicity_spectrum)


def _random_solenoidal(n, seed, kmax):
    rng = np.random.default_rng(seed)
    kx, ky, kz, k2 = orc.wavenumber(n, n, n)
    v = [rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape) for _ in range(3)]
    for a in v:
        a[k2 > kmax * kmax] = 0
        orc.symmetrize3d(a)
    orc.projection(*v, kx, ky, kz)
    return v, (kx, ky, kz, k2)


def _phys_grad(v, n):
    """du_j/dx_i in physical space by numpy's full complex FFT of the
    Hermitian-extended field (independent of the oracle's c2r)."""
    k = np.fft.fftfreq(n, 1.0 / n)
    K = [k[None, None, :], k[None, :, None], k[:, None, None]]
    full = []
    for a in v:
        f = np.zeros((n, n, n), complex)
        f[:, :, : n // 2 + 1] = a
        for x in range(n // 2 + 1, n):
            f[:, :, x] = np.conj(a[(-np.arange(n)) % n][:, (-np.arange(n)) % n][:, :, n - x])
        full.append(f)
    return [[np.real(np.fft.ifftn(1j * K[i] * full[j]) * n**3) for j in range(3)] for i in range(3)]


def _check_dissipation(dissipation, enstrophy_total):
    n, nu = 16, 0.013
    v, (kx, ky, kz, k2) = _random_solenoidal(n, 5, 5)
    G = _phys_grad(v, n)
    s = [[0.5 * (G[i][j] + G[j][i]) for j in range(3)] for i in range(3)]
    eps = 2 * nu * np.mean(sum(s[i][j] ** 2 for i in range(3) for j in range(3)))
    w = [G[1][2] - G[2][1], G[2][0] - G[0][2], G[0][1] - G[1][0]]
    omega = 0.5 * np.mean(sum(a**2 for a in w))
    assert abs(dissipation(*v, k2, nu) - eps) < 1e-12 * eps
    assert abs(enstrophy_total(*v, k2) - omega) < 1e-12 * omega


def test_dissipation_and_enstrophy_parseval():
    _check_dissipation(orc.dissipation, orc.enstrophy_total)


def test_shells_partition_the_energy():
    n = 16
    v, (kx, ky, kz, k2) = _random_solenoidal(n, 6, 7)
    e = orc.spectrum(*v, k2)
    assert abs(e.sum() - orc.energy(*v)) < 1e-12 * orc.energy(*v)


def _check_skewness(skewness):
    """u_x = a sin x + b sin 2x: d = a cos x + 2b cos 2x, <d^2> = a^2/2 + 2 b^2,
    <d^3> = 6 a^2 b <cos^2 x cos 2x> = 3 a^2 b / 2 (the other terms average
    to zero); u_y, u_z do not enter."""
    n = 16
    kx = orc.wavenumber(n, n, n)[0]
    x = 2 * np.pi * np.arange(n) / n
    for a, b in ((1.0, 0.5), (0.7, -1.3), (2.0, 0.0)):
        ux = np.broadcast_to(a * np.sin(x) + b * np.sin(2 * x), (n, n, n))
        want = 1.5 * a * a * b / (a * a / 2 + 2 * b * b) ** 1.5
        got = skewness(orc.r2c(np.ascontiguousarray(ux)) / n**3, kx, n)
        assert abs(got - want) < 1e-13, (a, b, got, want)


def test_skewness_closed_form():
    _check_skewness(orc.skewness)


@pytest.mark.parametrize("old,new", [("1j * kx[None", "-1j * kx[None"), ("** 1.5", "** 2.0"),
                                     ("np.mean(d**3)", "np.mean(np.abs(d) ** 3)")])
def test_skewness_check_catches_planted_mistakes(old, new):
    with pytest.raises(AssertionError):
        _check_skewness(_mutant(orc.skewness, old, new))


# --------------------------------------------------------------------------
# dynamic Smagorinsky update
def _dft_mats(n):
    x = 2 * np.pi * np.arange(n) / n
    k = np.fft.fftfreq(n, 1.0 / n)
    return np.exp(1j * np.outer(k, x)), k  # E[k, x] = exp(i k x)


def _inv(fh, E):
    """f(x) = sum_k fh_k exp(i k.x), full (nz, ny, nx) spectrum, by direct sums."""
    return np.einsum("cz,by,ax,cba->zyx", E, E, E, fh, optimize=True)


def _fwd(f, E, n):
    """fh_k = N^-3 sum_x f(x) exp(-i k.x), by direct sums."""
    return np.einsum("cz,by,ax,zyx->cba", E.conj(), E.conj(), E.conj(), f, optimize=True) / n**3


def _full_from_half(a, n):
    f = np.zeros((n, n, n), complex)
    f[:, :, : n // 2 + 1] = a
    for x in range(n // 2 + 1, n):
        f[:, :, x] = np.conj(a[(-np.arange(n)) % n][:, (-np.arange(n)) % n][:, :, n - x])
    return f


def _dynsmag_brute(v, n, kt, delta, dtest):
    E, k = _dft_mats(n)
    K = [k[None, None, :], k[None, :, None], k[:, None, None]]
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    g = (k2 <= kt * kt).astype(float)
    vh = [_full_from_half(a, n) for a in v]
    u = [np.real(_inv(a, E)) for a in vh]
    filt = lambda f: np.real(_inv(g * _fwd(f, E, n), E))  # noqa: E731
    grad = [[np.real(_inv(1j * K[i] * vh[j], E)) for j in range(3)] for i in range(3)]
    S = np.array([[0.5 * (grad[i][j] + grad[j][i]) for j in range(3)] for i in range(3)])
    aS = np.sqrt(2 * np.einsum("ij...,ij...->...", S, S))
    uf = [filt(a) for a in u]
    L = np.array([[filt(u[i] * u[j]) - uf[i] * uf[j] for j in range(3)] for i in range(3)])
    Sf = np.array([[filt(S[i, j]) for j in range(3)] for i in range(3)])
    aSf = np.sqrt(2 * np.einsum("ij...,ij...->...", Sf, Sf))
    M = np.array([[filt(2 * delta**2 * aS * S[i, j]) - 2 * dtest**2 * aSf * Sf[i, j] for j in range(3)]
                  for i in range(3)])
    mm = float((M * M).sum())  # full 3x3 contraction over i, j and the grid
    lm = float((L * M).sum())
    cs2 = lm / mm
    # tau_ij spectral (half spectrum, oracle's normalisation): -cs2 FFT(2 D^2 |S| S_ij)/N^3
    tau = [-cs2 * _fwd(2 * delta**2 * aS * S[i, j], E, n)[:, :, : n // 2 + 1] for i, j in les.PAIRS]
    return tau, mm, lm, cs2


def _check_dynsmag(fn):
    n, kt = 8, 1.8
    v, (kx, ky, kz, k2) = _random_solenoidal(n, 11, 2.5)
    delta, dtest = 2 * np.pi / n, 4 * np.pi / n
    g = les.g_test(k2, kt)
    tau, mm, lm, cs2 = fn(*v, kx, ky, kz, g, delta, dtest, (n, n, n))
    btau, bmm, blm, bcs2 = _dynsmag_brute(v, n, kt, delta, dtest)
    assert abs(mm - bmm) < 1e-11 * abs(bmm), (mm, bmm)
    assert abs(lm - blm) < 1e-11 * abs(blm), (lm, blm)
    assert abs(cs2 - bcs2) < 1e-11 * abs(bcs2), (cs2, bcs2)
    scale = max(np.abs(b).max() for b in btau)
    for a, b in zip(tau, btau):
        assert np.abs(a - b).max() < 1e-11 * scale


def test_dynsmag_update_brute_force():
    _check_dynsmag(les.dynsmag_update)


# --------------------------------------------------------------------------
# planted mistakes: every check above must fail on each
def _mutant(fn, old, new):
    src = inspect.getsource(fn)
    assert old in src, old
    ns = dict(inspect.getmodule(fn).__dict__)
    exec(compile(src.replace(old, new), "<mutant>", "exec"), ns)
    return ns[fn.__name__]


@pytest.mark.parametrize("new", ["np.floor(np.sqrt(k2.astype(np.float64)))",
                                 "np.ceil(np.sqrt(k2.astype(np.float64)))",
                                 "np.rint(np.sqrt(k2.astype(np.float64)) + 0.5)"])
def test_binning_check_catches_wrong_binning(monkeypatch, new):
    bad = _mutant(orc.shell_index, "np.floor(np.sqrt(k2.astype(np.float64)) + 0.5)", new)
    monkeypatch.setattr(orc, "shell_index", bad)
    with pytest.raises(AssertionError):
        _check_binning(orc.spectrum)


def test_binning_check_catches_dropped_half_weight():
    bad = _mutant(orc._modal_energy, "t[:, :, 0] *= 0.5", "pass")
    ns_spec = _mutant(orc.spectrum, "t = _modal_energy(vx, vy, vz)", "t = _me(vx, vy, vz)")
    ns = inspect.getmodule(orc.spectrum).__dict__
    g = dict(ns)
    g["_me"] = bad
    spec = type(ns_spec)(ns_spec.__code__, g)
    with pytest.raises(AssertionError):
        _check_binning(spec)


@pytest.mark.parametrize("old,new", [("h[:, :, 0] *= 0.5", "pass"),
                                     ("h = 2.0 * np.real(", "h = -2.0 * np.real("),
                                     ("h = 2.0 * np.real(", "h = 1.0 * np.real(")])
def test_abc_check_catches_helicity_mistakes(old, new):
    bad = _mutant(orc.helicity_spectrum, old, new)
    with pytest.raises(AssertionError):
        _check_abc(orc.spectrum, bad)


def test_dissipation_check_catches_missing_factor():
    bad = _mutant(orc.dissipation, "2.0 * nu *", "nu *")
    with pytest.raises(AssertionError):
        _check_dissipation(bad, orc.enstrophy_total)


@pytest.mark.parametrize("old,new", [
    ("- 2.0 * delta_test * delta_test * at * St[q]", "+ 2.0 * delta_test * delta_test * at * St[q]"),
    ("w = [1.0, 2.0, 2.0, 1.0, 2.0, 1.0]", "w = [1.0, 1.0, 1.0, 1.0, 1.0, 1.0]"),
    ("s3.r2c(u[i] * u[j]) * const", "s3.r2c(u[i] * u[j])"),
    ("2.0 * delta_test * delta_test * at", "2.0 * delta * delta * at"),
    ("ut[i] * ut[j]", "u[i] * u[j]"),
])
def test_dynsmag_check_catches_planted_mistakes(old, new):
    bad = _mutant(les.dynsmag_update, old, new)
    with pytest.raises(AssertionError):
        _check_dynsmag(bad)
