"""Pins for the oracle's scalar numerics: the smooth profile family
(utility.cpp:47-106), the spectral multipliers (utility.cpp:80-93) and the
FFT that replaces MKL DFTI (PSTD_QScat.cpp:502-505)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_ksi_endpoints_and_symmetry():
    assert O.ksi(0.0) == 0.0 and O.ksi(1.0) == 1.0
    assert O.ksi(-0.3) == 0.0 and O.ksi(1.7) == 1.0
    assert O.ksi(0.5) == pytest.approx(0.5, abs=1e-15)
    for x in np.linspace(0.01, 0.99, 37):
        assert O.ksi(x) + O.ksi(1 - x) == pytest.approx(1.0, abs=1e-14)
        assert O.dksi(x) == pytest.approx(O.dksi(1 - x), abs=1e-13)
        assert O.d2ksi(x) == pytest.approx(-O.d2ksi(1 - x), abs=1e-12)
    # 4th-order smoothness at the ends: xi' and xi'' vanish (clamped to 0)
    assert O.dksi(0.0) == 0.0 and O.d2ksi(1.0) == 0.0
    assert abs(O.dksi(1e-6)) < 1e-9 and abs(O.d2ksi(1e-6)) < 1e-4


def test_ksi_known_values():
    # xi(1/4) = 1/4 - (4/3)/(2 pi) * sin(pi/2) + (1/3)/(4 pi) * sin(pi)
    assert O.ksi(0.25) == pytest.approx(0.25 - 2.0 / (3.0 * math.pi), abs=1e-15)
    # xi'(1/2) = 1 + 4/3 + 1/3 = 8/3 ; xi''(1/4) = -2 pi (-4/3) = 8 pi / 3
    assert O.dksi(0.5) == pytest.approx(8.0 / 3.0, abs=1e-14)
    assert O.d2ksi(0.25) == pytest.approx(8.0 * math.pi / 3.0, rel=1e-14)


@pytest.mark.parametrize("x", np.linspace(0.05, 0.95, 19))
def test_ksi_derivatives_numerically(x):
    h = 1e-4
    # 4th-order central differences of the oracle's own xi and xi'
    num1 = (-O.ksi(x + 2 * h) + 8 * O.ksi(x + h) - 8 * O.ksi(x - h) + O.ksi(x - 2 * h)) / (12 * h)
    num2 = (-O.dksi(x + 2 * h) + 8 * O.dksi(x + h) - 8 * O.dksi(x - h) + O.dksi(x - 2 * h)) / (12 * h)
    assert O.dksi(x) == pytest.approx(num1, abs=1e-9)
    assert O.d2ksi(x) == pytest.approx(num2, abs=1e-8)


def test_window_weights_long_double():
    """m_weight[i] = weight((i+1)/9) (PSTD_QScat.cpp:104-105), against a
    long-double evaluation of the closed form."""
    for i in range(8):
        x = np.longdouble(i + 1) / np.longdouble(9)
        pi = np.longdouble("3.14159265358979323846264338327950288")
        ref = x - np.longdouble(4) / 3 / (2 * pi) * np.sin(2 * pi * x) + np.longdouble(1) / 3 / (4 * pi) * np.sin(4 * pi * x)
        assert O.weight(float(x)) == pytest.approx(float(ref), abs=2e-16)


def test_ibh_profile():
    O.lib().orc_set_profile(1)
    try:
        # integrated Blackman-Harris: xi'(1/2) = 1 + 1.3611 + 0.3938 + 0.0326
        assert O.dksi(0.5) == pytest.approx(2.7875, abs=1e-14)
        assert O.ksi(0.5) == pytest.approx(0.5, abs=1e-15)
        h = 1e-4
        for x in (0.2, 0.6):
            num = (O.ksi(x + h) - O.ksi(x - h)) / (2 * h)
            assert O.dksi(x) == pytest.approx(num, abs=1e-7)
    finally:
        O.lib().orc_set_profile(0)


@pytest.mark.parametrize("n,d", [(64, 0.5), (80, 0.6), (96, 0.5), (7, 0.3), (160, 0.31)])
def test_kcoeff_closed_form(n, d):
    k1, k2 = O.kcoeff(d, n)
    m = np.arange(n)
    kap = np.where(m <= n // 2, 2 * np.pi / d * m / n, 2 * np.pi / d * (m / n - 1))
    assert np.allclose(k2, kap ** 2, rtol=1e-14, atol=0)
    z = 1j * kap * np.exp(1j * kap * d / 2)
    assert np.allclose(k1[0::2], z.real, rtol=1e-13, atol=1e-13)
    assert np.allclose(k1[1::2], z.imag, rtol=1e-13, atol=1e-13)
    if n % 2 == 0:
        assert kap[n // 2] > 0  # Nyquist bin taken positive
        assert np.allclose(k2[1:], k2[1:][::-1])  # even -> D2 real symmetric


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 12, 16, 22, 30, 64, 80, 96, 97, 160, 288])
def test_fft_against_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = O.fft(x, -1)
    assert np.allclose(f, np.fft.fft(x), rtol=1e-13, atol=1e-12 * max(1, n))
    b = O.fft(x, +1)
    assert np.allclose(b, np.fft.ifft(x), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [64, 80, 96])
def test_spectral_second_derivative_exact_on_modes(n):
    """IFFT(k2 FFT(e^{i kappa_m x})) = kappa_m^2 e^{i kappa_m x} for every
    lattice mode m (Trefethen, Spectral Methods in MATLAB, ch. 3)."""
    d = 0.5
    _, k2 = O.kcoeff(d, n)
    j = np.arange(n)
    for m in (0, 1, 5, n // 2 - 1, n // 2, n // 2 + 1, n - 1):
        e = np.exp(2j * np.pi * m * j / n)
        u = O.fft(O.fft(e, -1) * k2, +1)
        assert np.allclose(u, k2[m] * e, rtol=1e-12, atol=1e-10)
