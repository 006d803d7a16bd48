"""CPU tests that pin the far-field oracle (oracle/farfield_oracle.py) and
the partial-wave reference (oracle/partial_waves.py) to closed forms.

Far field: an outgoing spherical wave from r0 inside the box has
f(rhat) = e^{-i rhat.r0}; the z-dipole (d/dz0 of it) has
f = -i rhat_z e^{-i rhat.r0}; a plane wave, regular inside the box, has
f = 0.  The quadrature is 4th order, so the error must fall about 16x per
halving of the spacing.  A flipped normal, a lost i, a transposed face or
a wrong origin each break one of these by O(1).

Partial waves: Numerov phase shifts against the square well's closed form
and, for a weak Gaussian, the Born limit delta_l = -int U r^2 j_l^2 dr."""
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import spherical_jn

from oracle import farfield_oracle as F
from oracle import partial_waves as PW


def monopole(r0):
    def fn(x, y, z):
        X, Y, Z = x - r0[0], y - r0[1], z - r0[2]
        R = np.sqrt(X * X + Y * Y + Z * Z)
        A = np.exp(1j * R) / R
        dp = (1j - 1 / R) * A
        return A, dp * X / R, dp * Y / R, dp * Z / R
    return fn


def dipole_z(r0):
    def fn(x, y, z):
        X, Y, Z = x - r0[0], y - r0[1], z - r0[2]
        R = np.sqrt(X * X + Y * Y + Z * Z)
        ph = np.exp(1j * R) / R
        p1 = (1j - 1 / R) * ph
        p2 = ph / R ** 2 + (1j - 1 / R) ** 2 * ph
        c = -(p2 - p1 / R) * Z / R ** 2
        return -p1 * Z / R, c * X, c * Y, c * Z - p1 / R
    return fn


def plane_wave(kh):
    def fn(x, y, z):
        A = np.exp(1j * (kh[0] * x + kh[1] * y + kh[2] * z))
        return A, 1j * kh[0] * A, 1j * kh[1] * A, 1j * kh[2] * A
    return fn


def box(ppw, half_wavelengths=3):
    """A grid with ppw points per wavelength and a virtual box about
    half_wavelengths wavelengths wide, off-centre in the grid."""
    d = 2 * np.pi / ppw
    nb = int(round(half_wavelengths * 2 * np.pi / d))
    N = (nb + 12, nb + 8, nb + 16)
    orig = [(n - 1) / 2.0 for n in N]
    vr = []
    for a in range(3):
        lo = int(orig[a] - nb / 2)
        vr += [lo, lo + nb]
    return dict(N=N, vrtl=vr, d=(d, d, d), orig=orig)


def _dirs(n=24, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def _errors(ppw):
    h = box(ppw)
    dirs = _dirs()
    r0 = np.array([0.3, -0.2, 0.5])
    body = lambda fn: F.surface_body(h["N"], h["vrtl"], h["d"], h["orig"], fn)
    em = np.abs(F.far_field(h, body(monopole(r0)), dirs) - np.exp(-1j * dirs @ r0)).max()
    ed = np.abs(F.far_field(h, body(dipole_z(r0)), dirs) + 1j * dirs[:, 2] * np.exp(-1j * dirs @ r0)).max()
    ep = np.abs(F.far_field(h, body(plane_wave(np.array([0.6, 0.0, 0.8]))), dirs)).max()
    return em, ed, ep


def test_far_field_closed_forms_and_order():
    e12, e24 = _errors(12), _errors(24)
    assert e12[0] < 1e-3 and e12[1] < 1e-3 and e12[2] < 3e-2
    for a, b in zip(e12, e24):
        assert a / b > 12, (a, b)           # 4th order: 16x per halving


def test_quad_weights():
    for n in (8, 9, 20):
        w = F.quad_weights(n)
        x = np.arange(n, dtype=float)
        for p in range(4):                  # exact for cubics
            assert abs(np.dot(w, x ** p) - (n - 1) ** (p + 1) / (p + 1)) < 1e-9 * (n - 1) ** (p + 1)
    assert np.allclose(F.quad_weights(3), [0.5, 1, 0.5])


def test_square_well_phase_shifts_numerov():
    for U0, R in ((-1.0, 3.0), (0.5, 2.0), (-3.0, 1.5)):
        ds = PW.square_well_phase_shifts(U0, R, 12)
        U = lambda r: np.where(np.abs(r - R) < 1e-9, U0 / 2, np.where(r < R, U0, 0.0))
        dn = PW.phase_shifts(U, 12, 8.0, h=1e-3)
        diff = np.angle(np.exp(2j * (ds - dn))) / 2          # modulo pi
        assert np.abs(diff).max() < 1e-6


def test_gaussian_phase_shifts_born_limit():
    a, U0 = 1.5, 1e-4
    U = lambda r: U0 * np.exp(-r ** 2 / (2 * a * a))
    dn = PW.phase_shifts(U, 6, 12.0, h=2e-3)
    born = np.array([-quad(lambda r: U(r) * r * r * spherical_jn(l, r) ** 2, 0, 15, limit=200)[0]
                     for l in range(7)])
    np.testing.assert_allclose(dn, born, rtol=3e-4)


def test_partial_wave_optical_theorem():
    d = PW.square_well_phase_shifts(-1.0, 3.0, 20)
    sig = PW.total_cross_section(d)
    assert abs(4 * np.pi * PW.amplitude(d, 1.0).imag - sig) < 1e-10 * sig
    # and the angular integral of |f|^2
    x, w = np.polynomial.legendre.leggauss(64)
    assert abs(2 * np.pi * np.sum(w * np.abs(PW.amplitude(d, x)) ** 2) - sig) < 1e-9 * sig
