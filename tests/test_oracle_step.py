"""Pins for the oracle's time step (PSTD_QScat.cpp:1233-1446), initial field
(:959-1118) and TF/SF source (:617-955) against independent formulations."""
import math

import numpy as np
import pytest

from pstd_amd import workloads as W
from oracle import oracle as O
import bruteforce as BF


PROFILES = [0, 1]   # SMOOTH4 and the integrated Blackman-Harris variant (__IBH__)


@pytest.mark.parametrize("prof", PROFILES)
@pytest.mark.parametrize("cfg", [
    dict(nprocs=(2, 2, 2), L=(32, 32, 32), h=10),
    dict(nprocs=(3, 1, 2), L=(40, 24, 30), h=9),
    dict(nprocs=(1, 1, 1), L=(24, 20, 22), h=10),
])
def test_step_against_dense_global(cfg, prof):
    """One Update from a random state equals the dense global restatement
    (windowed D2 matrices from numpy's FFT), to 1e-12."""
    p, V = W.config_tiny(**cfg)
    p.profile = prof
    o = O.Oracle(p, V)
    rng = np.random.default_rng(11)
    a = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    b = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    o.set_psi(0, a)
    o.set_psi(1, b)
    for ntau in (5, 1400):  # inside and after the potential ramp
        o.set_psi(0, a)
        o.set_psi(1, b)
        o.set_time_index(ntau)
        ref = BF.step(a, b, p, ntau, o.get_vtilde(), o.get_source(), [o.get_g(x) for x in range(3)])
        o.update(1)
        got = o.get_psi(0)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-12, err
        assert np.array_equal(o.get_psi(1), a)  # swap: old psi^n is now psi^{n-1}
        assert o.time_index() == ntau + 1


@pytest.mark.parametrize("profile", PROFILES)
def test_initial_field_closed_form(profile):
    """psi(0) = Xi e^{i k.(r - r0)}, psi(dtau) = psi(0) e^{-i dtau} on the TF
    box, zero outside; m_nTau = 1."""
    p, _ = W.config_tiny(theta=0.7, phi=2.2, well=None)
    p.profile = profile
    o = O.Oracle(p, None)
    _xi_np = lambda s: BF.xi(s, profile)
    assert o.time_index() == 1
    N = o.N
    n0t1 = p.n_ksi - 1
    khat = (math.sin(p.theta) * math.cos(p.phi), math.sin(p.theta) * math.sin(p.phi), math.cos(p.theta))
    prof, ph = [], []
    for a in range(3):
        i = np.arange(N[a])
        s1 = (i - o.info["ksi1a"][a]) / n0t1
        s2 = (o.info["ksi2b"][a] - i) / n0t1
        m = (i >= o.info["ksi1a"][a]) & (i <= o.info["ksi2b"][a])
        prof.append(np.where(m, _xi_np(s1) * _xi_np(s2), 0.0))
        ph.append((i - o.info["inj"][a]) * p.d[a] * khat[a])
    Xi = prof[0][:, None, None] * prof[1][None, :, None] * prof[2][None, None, :]
    r = ph[0][:, None, None] + ph[1][None, :, None] + ph[2][None, None, :]
    psi0 = Xi * np.exp(1j * r)
    assert np.allclose(o.get_psi(1), psi0, rtol=0, atol=1e-14)
    assert np.allclose(o.get_psi(0), psi0 * np.exp(-1j * p.dtau), rtol=0, atol=1e-14)


@pytest.mark.parametrize("profile", PROFILES)
@pytest.mark.parametrize("theta,phi", [(0.4, 0.9), (2.5, -2.0), (0.0, 0.0)])
def test_tfsf_source_is_commutator(theta, phi, profile):
    """Sum of the six Inc slabs = 2 dtau (lap Xi + 2 i khat.grad Xi) e^{i r}
    with Xi the two-sided closed-form profile; derivatives of Xi taken by
    5-point finite differences of xi at continuous arguments."""
    p, _ = W.config_tiny(theta=theta, phi=phi, well=None, nprocs=(2, 1, 2), L=(36, 40, 36), h=10)
    p.profile = profile
    o = O.Oracle(p, None)
    _xi_np = lambda s: BF.xi(s, profile)
    S = o.get_source()
    N = o.N
    n0t1 = p.n_ksi - 1
    khat = (math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta))
    eps = 1e-3

    def prof(a, i):
        s1 = (i - o.info["ksi1a"][a]) / n0t1
        s2 = (o.info["ksi2b"][a] - i) / n0t1
        return _xi_np(s1) * _xi_np(s2)

    P, dP, d2P, ph = [], [], [], []
    for a in range(3):
        i = np.arange(N[a]).astype(float)
        inside = (i >= o.info["ksi1a"][a]) & (i <= o.info["ksi2b"][a])
        f = [prof(a, i + k * eps) for k in (-2, -1, 0, 1, 2)]
        d1 = (f[0] - 8 * f[1] + 8 * f[3] - f[4]) / (12 * eps) / p.d[a]
        d2 = (-f[0] + 16 * f[1] - 30 * f[2] + 16 * f[3] - f[4]) / (12 * eps * eps) / p.d[a] ** 2
        P.append(np.where(inside, f[2], 0.0))
        dP.append(np.where(inside, d1, 0.0))
        d2P.append(np.where(inside, d2, 0.0))
        ph.append((i - o.info["inj"][a]) * p.d[a] * khat[a])
    e = lambda v, a: v.reshape([-1 if b == a else 1 for b in range(3)])
    lap = (e(d2P[0], 0) * e(P[1], 1) * e(P[2], 2) + e(P[0], 0) * e(d2P[1], 1) * e(P[2], 2)
           + e(P[0], 0) * e(P[1], 1) * e(d2P[2], 2))
    grad = (khat[0] * e(dP[0], 0) * e(P[1], 1) * e(P[2], 2) + khat[1] * e(P[0], 0) * e(dP[1], 1) * e(P[2], 2)
            + khat[2] * e(P[0], 0) * e(P[1], 1) * e(dP[2], 2))
    r = e(ph[0], 0) + e(ph[1], 1) + e(ph[2], 2)
    ref = 2 * p.dtau * (lap + 2j * grad) * np.exp(1j * r)
    # the inner region (strictly between transition layers) has no source
    scale = np.abs(ref).max()
    assert scale > 0
    keep = np.ones(S.shape, bool)
    if profile == 1:
        # the integrated Blackman-Harris xi has xi'(0) = xi'(1) = A0+A1+A2+A3 = 1e-4,
        # which the clamp at the layer ends sets to 0: a kink that a finite-difference
        # stencil across the ends cannot see, so those planes are left out
        for a in range(3):
            ends = np.isin(np.arange(N[a]), [o.info[k][a] for k in ("ksi1a", "ksi1b", "ksi2a", "ksi2b")])
            keep &= ~e(ends, a)
    assert np.abs(S - ref)[keep].max() / scale < 2e-6
