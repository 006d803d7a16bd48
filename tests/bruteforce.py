"""Dense numpy restatement of one PSTD step in the *global* formulation.

Used to pin the oracle's per-rank composition (local boxes + exchange +
FFT lines, PSTD_QScat.cpp:1233-1731) against an independent formulation:
every axis operator is a dense windowed second-derivative matrix built from
numpy's FFT of unit vectors, applied to segments of one global array.  The
decomposition index maths follows PSTD_QScat.cpp:277-331; the window is
plane 0 -> 0, planes 1..8 -> xi((m)/9) (PSTD_QScat.cpp:104-105, :1554-1603).
"""
import math

import numpy as np


# A1, A2, A3 of SMOOTH4 (profile 0) and of the integrated Blackman-Harris set
# (profile 1), utility.cpp:32-45
PROFILES = [(-4.0 / 3.0, 1.0 / 3.0, 0.0), (-1.3611, 0.3938, -0.0326)]


def xi(x, prof=0):
    """Transition profile, closed form (utility.cpp:47-58)."""
    a1, a2, a3 = PROFILES[prof]
    x = np.asarray(x, dtype=np.float64)
    t = 2 * math.pi * x
    v = x + a1 / (2 * math.pi) * np.sin(t) + a2 / (4 * math.pi) * np.sin(2 * t) + a3 / (6 * math.pi) * np.sin(3 * t)
    return np.where(x < 1e-10, 0.0, np.where(x > 1 - 1e-10, 1.0, v))


def segments(P, L, h):
    """(fft1, core1, core2) per subdomain along one axis."""
    out = []
    for r in range(P):
        if P == 1:
            out.append((0, 0, L - 1))
        elif r == 0:
            out.append((0, 0, L - 1 - h))
        else:
            f1 = r * (L - 2 * h)
            out.append((f1, f1 + h, f1 + L - 1 if r == P - 1 else f1 + L - 1 - h))
    return out


def window(P, L, r, prof=0):
    W = np.ones(L)
    w = xi(np.arange(1, 9) / 9.0, prof)
    if r > 0:
        W[0] = 0.0
        W[1:9] = w
    if r < P - 1:
        W[L - 1] = 0.0
        W[L - 9:L - 1] = w[::-1]
    return W


def kappa(L, d):
    m = np.arange(L)
    return np.where(m <= L // 2, 2 * math.pi / d * m / L, 2 * math.pi / d * (m / L - 1.0))


def d2_matrix(L, d):
    """Dense IFFT(kappa^2 FFT(.)) matrix from numpy's FFT of unit vectors."""
    k2 = kappa(L, d) ** 2
    E = np.eye(L)
    return np.real(np.fft.ifft(k2[:, None] * np.fft.fft(E, axis=0), axis=0))


def d1_matrix(L, d):
    """Dense IFFT(i kappa e^{i kappa d/2} FFT(.)) matrix: derivative at x+d/2."""
    k = kappa(L, d)
    k1 = 1j * k * np.exp(1j * k * d / 2)
    E = np.eye(L)
    return np.real(np.fft.ifft(k1[:, None] * np.fft.fft(E, axis=0), axis=0))


def laplacian_term(psi, p):
    """sum over axes of D2_local (W . psi) on core points (u = -lap psi)."""
    out = np.zeros_like(psi)
    for a in range(3):
        P, L, h, d = p.nprocs[a], p.L[a], p.n_halo, p.d[a]
        D = d2_matrix(L, d)
        x = np.moveaxis(psi, a, 0)
        o = np.moveaxis(out, a, 0)
        for r, (f1, c1, c2) in enumerate(segments(P, L, h)):
            seg = x[f1:f1 + L] * window(P, L, r, p.profile)[:, None, None]
            u = np.tensordot(D, seg, axes=(1, 0))
            o[c1:c2 + 1] += u[c1 - f1:c2 - f1 + 1]
    return out


def step(psi_n, psi_nm1, p, ntau, vtilde, source, g):
    """One Update: psi^{n+1} from psi^n, psi^{n-1} (global arrays)."""
    alpha = 2.0 * math.sin(p.dtau)
    u = laplacian_term(psi_n, p)
    out = psi_nm1 - 1j * alpha * u
    t = ntau * p.dtau
    f = t / (2 * math.pi)
    f = f if f < 0.997 else 1.0
    out = out - 1j * f * vtilde * psi_n
    out = out - source * 1j * np.exp(-1j * t)
    out = out * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    return out
