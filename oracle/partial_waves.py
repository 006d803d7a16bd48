"""Partial-wave scattering amplitudes (test infrastructure: closed forms and
a radial integrator against which the far-field post-processor and the
solver behind it are checked; never part of the product path).

Units are the solver's scaled ones: lengths in lambdabar = 1/k, so the
steady state A of psi = A e^{-i tau} solves (lap + 1 - U) A = 0 with
U = V / E0 (PSTD_QScat.cpp:1136 divides the hook's V by E0; the time
scaling of :134 makes the free dispersion omega = |k|^2 = 1).  For a
central U(r) and incidence along khat,

    f(theta) = sum_l (2l + 1) e^{i delta_l} sin(delta_l) P_l(cos theta)

(k = 1), theta measured from khat; the optical theorem reads
sigma_tot = 4 pi Im f(0).

phase_shifts integrates u'' = (l(l+1)/r^2 + U(r) - 1) u by Numerov from
u ~ r^{l+1} and matches to r [j_l cos delta - y_l sin delta] at two radii
past the potential's range; square_well_phase_shifts is the textbook closed
form (continuity of the log-derivative at R) that pins it.
"""
import numpy as np
from scipy.special import eval_legendre, spherical_jn, spherical_yn


def square_well_phase_shifts(U0, R, lmax):
    """U = U0 for r < R, 0 outside: tan delta_l =
    (j_l'(R) j_l(qR) - q j_l(R) j_l'(qR)) / (y_l'(R) j_l(qR) - q y_l(R) j_l'(qR)),
    q = sqrt(1 - U0) (U0 < 1)."""
    q = np.sqrt(1.0 - U0)
    out = []
    for l in range(lmax + 1):
        jq, djq = spherical_jn(l, q * R), spherical_jn(l, q * R, derivative=True)
        j, dj = spherical_jn(l, R), spherical_jn(l, R, derivative=True)
        y, dy = spherical_yn(l, R), spherical_yn(l, R, derivative=True)
        num = dj * jq - q * j * djq
        den = dy * jq - q * y * djq
        out.append(np.arctan(num / den))
    # delta is fixed modulo pi, in (-pi/2, pi/2); e^{i delta} sin(delta) is invariant under delta -> delta + pi
    return np.array(out)


def phase_shifts(U, lmax, r_match, h=1e-3):
    """Numerov integration of the radial equation for a central U(r)
    (vectorised callable), matched at r_match and r_match - 0.5."""
    r = np.arange(1, int(np.ceil(r_match / h)) + 1) * h
    Ur = U(r)
    out = []
    for l in range(lmax + 1):
        k2 = 1.0 - Ur - l * (l + 1) / r ** 2          # u'' = -k2 u
        c = 1.0 + h * h * k2 / 12.0
        u = np.zeros_like(r)
        u[0], u[1] = r[0] ** (l + 1), r[1] ** (l + 1)
        for n in range(1, len(r) - 1):
            u[n + 1] = ((12.0 - 10.0 * c[n]) * u[n] - c[n - 1] * u[n - 1]) / c[n + 1]
            if abs(u[n + 1]) > 1e200:                 # keep the scale in range
                u[: n + 2] *= 1e-200
        i2 = len(r) - 1
        i1 = i2 - int(round(0.5 / h))
        r1, r2 = r[i1], r[i2]
        K = u[i1] / u[i2]
        num = K * r2 * spherical_jn(l, r2) - r1 * spherical_jn(l, r1)
        den = K * r2 * spherical_yn(l, r2) - r1 * spherical_yn(l, r1)
        out.append(np.arctan(num / den))
    return np.array(out)


def amplitude(deltas, cos_theta):
    """f(theta) = sum_l (2l+1) e^{i delta_l} sin delta_l P_l(cos theta)."""
    ct = np.asarray(cos_theta, dtype=np.float64)
    f = np.zeros(ct.shape, dtype=np.complex128)
    for l, d in enumerate(deltas):
        f += (2 * l + 1) * np.exp(1j * d) * np.sin(d) * eval_legendre(l, ct)
    return f


def total_cross_section(deltas):
    """sigma = 4 pi sum_l (2l+1) sin^2 delta_l (k = 1)."""
    l = np.arange(len(deltas))
    return float(4 * np.pi * np.sum((2 * l + 1) * np.sin(deltas) ** 2))
