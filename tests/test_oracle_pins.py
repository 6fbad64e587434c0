"""CPU checks that pin the oracle's closed forms against facts their own
formulas do not restate (no GPU).

  * fwi_wavelet (fwi_wavelet.m:24-28): for t0 = 0 the amplitude f^2
    exp(-(f/f0)^2) peaks at f = f0 with value f0^2 / e (a dropped square
    peaks at f0 / sqrt(2), a dropped square in the Gaussian never peaks);
    the phase is -2 pi f t0 (a sign error flips it); without f0 the spectrum
    is flat.
  * ls_misfit (LS_misfit.m:21-26): a value computed by hand (a missing 1/2
    doubles it) and g = x - y (not y - x).
  * param2wavenum (param2wavenum.m:20-49): the wavenumber it returns must
    make a plane wave of the physical wavenumber omega / v an approximate
    null vector of the oracle stencil (the discrete Helmholtz equation),
    which fixes both the scale (2 pi f)^2 and the exponent of v; the three
    units must agree on the same medium (s2/m2 = v^-2, s2/km2 = 1e6 v^-2);
    dwn and ddwn must be the first and second derivatives of wn (central
    differences).
"""
import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------- fwi_wavelet

def test_wavelet_peaks_at_f0_with_value_f0_squared_over_e():
    f0 = 10.0
    f = np.linspace(0.01, 40.0, 400001)
    w = O.fwi_wavelet(f, 0.0, f0)
    a = np.abs(w)
    k = int(np.argmax(a))
    assert abs(f[k] - f0) < 1e-3
    assert abs(a[k] - f0 ** 2 / np.e) < 1e-9 * f0 ** 2
    # t0 = 0: real and non-negative
    assert np.max(np.abs(w.imag)) == 0.0 and np.min(w.real) >= 0.0


@pytest.mark.parametrize("t0", [0.05, 0.3, -0.12])
def test_wavelet_phase_is_minus_two_pi_f_t0(t0):
    f = np.array([0.5, 1.0, 2.5, 4.0, 6.0, 7.25])
    w = O.fwi_wavelet(f, t0, 10.0)
    # removing the delay exp(-2 pi i f t0) leaves the real t0 = 0 spectrum
    undone = w * np.exp(2j * np.pi * f * t0)
    assert np.allclose(undone, O.fwi_wavelet(f, 0.0, 10.0), rtol=1e-14, atol=0)
    # a time shift only changes the phase, by -2 pi f t0 (mod 2 pi)
    dphi = np.angle(w / O.fwi_wavelet(f, 0.0, 10.0))
    expect = np.angle(np.exp(-2j * np.pi * f * t0))
    assert np.allclose(dphi, expect, atol=1e-12)


def test_wavelet_without_f0_is_flat():
    f = np.linspace(1.0, 30.0, 50)
    assert np.allclose(np.abs(O.fwi_wavelet(f, 0.2, 0.0)), 1.0, atol=1e-15)
    assert np.allclose(O.fwi_wavelet(f, 0.0, None), 1.0, atol=0)


# ------------------------------------------------------------------ ls_misfit

def test_ls_misfit_hand_value():
    x = np.array([1 + 2j, 3 + 0j, -1j])
    y = np.array([0 + 0j, 1 - 1j, -1j])
    # r = x - y = [1+2i, 2+i, 0]; ||r||^2 = 1 + 4 + 4 + 1 = 10; f = 5
    f, g = O.ls_misfit(x, y)
    assert f == 5.0
    assert np.array_equal(g, np.array([1 + 2j, 2 + 1j, 0j]))


def test_ls_misfit_matrix_is_frobenius():
    # LS_misfit.m:24 takes norm(r,'fro') of an nrec x nsrc block
    x = np.array([[1.0, 2.0], [0.0, -2.0]]) + 0j
    y = np.zeros((2, 2), dtype=complex)
    f, _ = O.ls_misfit(x, y)
    assert f == 0.5 * (1 + 4 + 0 + 4)


# -------------------------------------------------------------- param2wavenum

def _plane_wave_residual(wn_value, kphys, h, n=(24, 24, 24)):
    """|(H u)[p] / u[p]| relative to the stiffness part alone, for the plane
    wave u = exp(i k.x) with |k| = kphys along a fixed oblique direction,
    read in the interior (the eigenvalue of the constant-coefficient row)."""
    d = np.array([0.48, 0.6, 0.64])
    d = d / np.linalg.norm(d)
    kv = kphys * d * np.asarray(h)  # phase per grid step
    I, J, K = np.meshgrid(*(np.arange(a) for a in n), indexing="ij")
    u = np.exp(1j * (kv[0] * I + kv[1] * J + kv[2] * K)).reshape(-1, order="F")
    npml = np.array([[2, 2, 2], [2, 2, 2]])
    core = (slice(5, n[0] - 5),) * 3
    ph = u.reshape(n, order="F")[core]
    y = O.hmvp("forw", np.full(u.size, wn_value), h, n, npml, u).reshape(n, order="F")[core]
    s = O.hmvp("forw", np.zeros(u.size), h, n, npml, u).reshape(n, order="F")[core]
    lam = np.median((y / ph).real)
    stiff = np.median((s / ph).real)
    return abs(lam) / abs(stiff)


@pytest.mark.parametrize("v,freq,h", [(2000.0, 4.0, 10.0), (1500.0, 6.0, 10.0), (3000.0, 5.0, 12.5)])
def test_wavenumber_annihilates_physical_plane_wave(v, freq, h):
    """H(wn) exp(i k.x) ~ 0 for |k| = 2 pi f / v: the discrete operator is
    the Helmholtz operator of velocity v exactly when wn = (2 pi f / v)^2.
    Well resolved (>= 20 points per wavelength) the eigenvalue is a small
    fraction of the stiffness term; the plausible mistakes miss by O(1)."""
    hh = (h, h, h)
    kphys = 2 * np.pi * freq / v
    wn, _, _ = O.param2wavenum(np.array([v]), freq, "m/s")
    good = _plane_wave_residual(float(wn[0]), kphys, hh)
    assert good < 2e-3
    c = (2 * np.pi * freq) ** 2
    for wrong in (c / v, (2 * np.pi * freq / v), c * v ** -2 / (2 * np.pi), freq ** 2 / v ** 2):
        assert _plane_wave_residual(wrong, kphys, hh) > 0.3
    # even a 10 % scale error is 25x above the bar
    for s in (0.9, 1.1):
        assert _plane_wave_residual(s * float(wn[0]), kphys, hh) > 0.05


def test_units_agree_on_one_medium():
    v = np.array([1500.0, 2000.0, 3333.0, 4500.0])
    for f in (3.0, 6.0):
        w_ms, _, _ = O.param2wavenum(v, f, "m/s")
        w_s2, _, _ = O.param2wavenum(v ** -2, f, "s2/m2")
        w_km, _, _ = O.param2wavenum(1e6 * v ** -2, f, "s2/km2")
        assert np.allclose(w_s2, w_ms, rtol=1e-14, atol=0)
        assert np.allclose(w_km, w_ms, rtol=1e-14, atol=0)


@pytest.mark.parametrize("unit,p", [("m/s", np.array([1500.0, 2500.0, 4000.0])),
                                    ("s2/m2", np.array([4e-7, 1.6e-7, 6.25e-8])),
                                    ("s2/km2", np.array([0.4, 0.16, 0.0625]))])
def test_derivatives_match_central_differences(unit, p):
    f = 5.0
    wn, dwn, ddwn = O.param2wavenum(p, f, unit)
    eps = 1e-4 * p
    wp, dp, _ = O.param2wavenum(p + eps, f, unit)
    wm, dm, _ = O.param2wavenum(p - eps, f, unit)
    assert np.allclose(dwn, (wp - wm) / (2 * eps), rtol=1e-7, atol=0)
    scale = np.max(np.abs(dwn / p))
    assert np.allclose(ddwn, (dp - dm) / (2 * eps), rtol=1e-6, atol=1e-9 * scale)


def test_unknown_unit_raises():
    with pytest.raises(ValueError):
        O.param2wavenum(np.ones(2), 5.0, "km/s")
