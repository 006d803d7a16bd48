"""CPU tests of the host side of the warm start (warmstart.py): the Born
feature against the outgoing spherical wave of a compact source, and the
layout of the network input."""
import numpy as np

import pstd_amd  # noqa: F401
from pstd_amd import warmstart as WS
from pstd_amd import workloads as W


class _Grid:
    d = (0.5, 0.5, 0.5)


def test_born_field_is_the_outgoing_wave_of_a_compact_source():
    # (lap + 1) A = S with S a narrow Gaussian: far from it A ~ -(int S) F e^{i k r} / (4 pi r),
    # F = exp(-sigma^2 / 2) the source's form factor at |q| = 1, k = sqrt(1 + i eps)
    N, h, sig, eps = 64, 0.5, 0.6, 0.05
    x = (np.arange(N) - N // 2) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    src = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * sig ** 2))
    A = WS.born_field(_Grid(), src, eps=eps)
    k = np.sqrt(1 + 1j * eps)
    for r in (5.0, 8.0):
        i = N // 2 + int(r / h)
        want = -src.sum() * h ** 3 * np.exp(-sig ** 2 / 2) * np.exp(1j * k * r) / (4 * np.pi * r)
        got = A[i, N // 2, N // 2]
        # zero padding to twice the grid leaves periodic images at ~10-20 %
        assert abs(got - want) < 0.2 * abs(want), (r, got, want)
        # outgoing: the phase advances with r as e^{+i r}
    a1, a2 = A[N // 2 + 12, N // 2, N // 2], A[N // 2 + 13, N // 2, N // 2]
    assert abs(np.angle(a2 / a1) - h) < 0.1


def test_features_layout():
    p, V = W.batch_problem(3, 64, 2.7, seed_base=200000)
    a0 = np.zeros((64, 64, 64), complex)
    a0[20:44, 20:44, 20:44] = np.exp(1j * 0.3 * np.arange(24))[None, None, :]
    x = WS.features(p, V, a0)
    assert x.shape == (64, 64, 64, WS.IN_CHANNELS) and x.dtype == np.float32
    Vg = WS.potential_grid(p, V) / p.E0
    np.testing.assert_allclose(x[..., 0], Vg, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(x[..., 1] + 1j * x[..., 2], a0, atol=1e-6)
    np.testing.assert_allclose(x[..., 3], Vg * a0.real, rtol=1e-5, atol=1e-7)
    assert np.allclose(x[..., 5], np.sqrt(2 * p.E0) / 4)
    assert abs(np.linalg.norm(x[0, 0, 0, 6:9]) - 1) < 1e-6
    assert np.all(x[..., 11:] == 0)
    # nonzero potential only inside the potential box
    assert np.count_nonzero(Vg) > 0 and np.all(Vg[:2] == 0)
