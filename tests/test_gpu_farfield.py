"""GPU tests of the far-field post-processor (include/pstd_farfield.h).

Parity: the HIP transform against the oracle (oracle/farfield_oracle.py,
itself pinned to closed forms in tests/test_farfield_oracle.py) on the same
surface data -- synthetic files in both header precisions, short and long
box edges, and a solver's live accumulators and its saved file -- to
1e-12 relative (both sides FP64; they differ in summation order and in
Horner vs direct phases).

Physics: a Gaussian well solved to steady state on the GPU, transformed on
the GPU, against the partial-wave amplitude from the Numerov phase shifts
(oracle/partial_waves.py), and the optical theorem on the computed f."""
import numpy as np
import pytest

import pstd_amd
from pstd_amd import farfield as FF
from pstd_amd import workloads as W
from pstd_amd.surface_io import header_bytes
from oracle import farfield_oracle as F
from oracle import partial_waves as PW

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _field(r0, kh, c):
    """Outgoing wave from r0 plus c times a plane wave along kh."""
    def fn(x, y, z):
        X, Y, Z = x - r0[0], y - r0[1], z - r0[2]
        R = np.sqrt(X * X + Y * Y + Z * Z)
        A = np.exp(1j * R) / R
        dp = (1j - 1 / R) * A
        Pw = c * np.exp(1j * (kh[0] * x + kh[1] * y + kh[2] * z))
        return A + Pw, dp * X / R + 1j * kh[0] * Pw, dp * Y / R + 1j * kh[1] * Pw, dp * Z / R + 1j * kh[2] * Pw
    return fn


def _write(path, N, vr, d, orig, body, precision):
    dt = np.float64 if precision == 8 else np.float32
    with open(path, "wb") as fp:
        fp.write(header_bytes(N, vr, 8, d, orig, 0.5, precision))
        fp.write(np.asarray(body, dtype=dt).tobytes())


def _dirs(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


@pytest.mark.parametrize("case", [
    dict(N=(40, 36, 44), vr=(5, 33, 4, 30, 6, 37), d=(0.45, 0.5, 0.4), prec=8, ndir=200),
    dict(N=(40, 36, 44), vr=(5, 33, 4, 30, 6, 37), d=(0.45, 0.5, 0.4), prec=4, ndir=67),
    dict(N=(12, 10, 14), vr=(2, 8, 3, 6, 1, 12), d=(0.6, 0.6, 0.6), prec=8, ndir=1),   # edges of 4..12 points
])
def test_file_parity(case, tmp_path):
    N, vr, d = case["N"], list(case["vr"]), case["d"]
    orig = [(n - 1) / 2.0 for n in N]
    print(f"case far field file N={N} vr={vr} precision={case['prec']} ndir={case['ndir']}", flush=True)
    body = F.surface_body(N, vr, d, orig, _field((0.3, -0.4, 0.2), (0.0, 0.6, 0.8), 0.3))
    path = tmp_path / "surf.dat"
    _write(path, N, vr, d, orig, body, case["prec"])
    if case["prec"] == 4:   # the header's reals and the body are float in such a file
        body = body.astype(np.float32).astype(np.float64)
        d = tuple(float(np.float32(v)) for v in d)
        orig = [float(np.float32(v)) for v in orig]
    dirs = _dirs(case["ndir"], 1)
    got = FF.far_field(path, dirs)
    want = F.far_field(dict(N=N, vrtl=vr, d=d, orig=orig), body, dirs)
    assert rel(got, want) < 1e-12


def test_rejects_bad_input(tmp_path):
    N, vr, d = (12, 10, 14), [2, 8, 3, 6, 1, 12], (0.6,) * 3
    orig = [(n - 1) / 2.0 for n in N]
    path = tmp_path / "s.dat"
    _write(path, N, vr, d, orig, F.surface_body(N, vr, d, orig, _field((0, 0, 0), (0, 0, 1), 0)), 8)
    with pytest.raises(pstd_amd.PSTDError) as e:
        FF.far_field(path, [[1.0, 1.0, 0.0]])                   # not a unit vector
    assert e.value.code == pstd_amd.PSTD_EINVAL
    raw = path.read_bytes()
    (tmp_path / "t.dat").write_bytes(raw[:-8])
    with pytest.raises(pstd_amd.PSTDError) as e:
        FF.far_field(tmp_path / "t.dat", [[0.0, 0.0, 1.0]])     # truncated body
    assert e.value.code == pstd_amd.PSTD_EIO
    with pytest.raises(pstd_amd.PSTDError) as e:
        FF.far_field(tmp_path / "missing.dat", [[0.0, 0.0, 1.0]])
    assert e.value.code == pstd_amd.PSTD_EIO


def test_solver_accumulators_and_file(tmp_path):
    p, V = W.config_tiny(nprocs=(2, 1, 2), L=(32, 40, 32), h=10)
    print("case far field solver tiny 2x1x2", flush=True)
    s = pstd_amd.PSTDQScat(p, V)
    with pytest.raises(pstd_amd.PSTDError):
        FF.far_field(s, [[0.0, 0.0, 1.0]])                      # nothing accumulated yet
    s.Update(20)
    s.update_accumulate(30)
    g = s.geometry
    hdr = dict(N=s.N, vrtl=[g.vrtl1[0], g.vrtl2[0], g.vrtl1[1], g.vrtl2[1], g.vrtl1[2], g.vrtl2[2]],
               d=tuple(p.d), orig=tuple(g.orig))
    dirs = _dirs(150, 2)
    inc = FF.incident_amplitude(p)
    for deriv in (pstd_amd.PSTD_DERIV_PS, pstd_amd.PSTD_DERIV_FD):
        got = FF.far_field(s, dirs, deriv)        # normalised to the solver's incident wave
        want = F.far_field(hdr, s.get_surfaces(deriv), dirs) / inc
        assert rel(got, want) < 1e-12
    f_ps = FF.far_field(s, dirs)
    s.SaveVirtualSurfaces(str(tmp_path / "w"))
    n0 = s.time_index() - 30
    assert rel(FF.far_field(tmp_path / f"w_PS_{n0}.dat", dirs, params=p), f_ps) < 1e-13
    assert rel(FF.far_field(tmp_path / f"w_PS_{n0}.dat", dirs) / inc, f_ps) < 1e-13
    s.close()


def _gaussian_problem(U0, a, precision=pstd_amd.PSTD_FP64):
    """96^3, dx_ref = 0.5, 2x2x2 subdomains (L = 64, h = 16), 12-cell
    absorber; U = U0 exp(-r^2 / 2 a^2) in E0 units; oblique incidence."""
    dxr = 0.5
    p = pstd_amd.make_params((2, 2, 2), W.stable_dtau(dxr), (dxr,) * 3, (64, 64, 64), 12, 7, 12, 16,
                             0.5, 1.1, 0.5, 5.0, 4.0 / 12, 1000, "gauss", precision)
    V = W.sample_hook(p, lambda X, Y, Z: p.E0 * U0 * np.exp(-(X * X + Y * Y + Z * Z) / (2 * a * a)))
    return p, V


def test_gaussian_well_against_partial_waves():
    U0, a = -0.5, 1.2
    p, V = _gaussian_problem(U0, a)
    print(f"case far field Gaussian well U0={U0} a={a} 96^3", flush=True)
    s = pstd_amd.PSTDQScat(p, V)
    conv, periods, _, hist = s.run_to_steady_state(tol=1e-4, max_periods=40)
    print("periods", periods, "converged", conv, "changes", np.array2string(np.asarray(hist[:periods]), precision=2),
          flush=True)
    assert conv
    # the detector leaves its last one-period window in the accumulators
    kh = FF.incident_direction(p)
    dirs, w, ct, ph = FF.sphere_quadrature(48, 24, axis=kh)
    f = FF.far_field(s, dirs)
    deltas = PW.phase_shifts(lambda r: U0 * np.exp(-r * r / (2 * a * a)), 14, 12.0, h=2e-3)
    f_pw = PW.amplitude(deltas, ct)
    e_rel = rel(f, f_pw)
    # azimuthal symmetry about khat
    spread = float(np.abs(f.reshape(48, 24) - f.reshape(48, 24).mean(axis=1, keepdims=True)).max())
    sig = FF.cross_section(f, w)
    sig_ot = FF.optical_theorem(FF.far_field(s, kh[None, :])[0])
    sig_pw = PW.total_cross_section(deltas)
    print(f"f vs partial waves rel L2 {e_rel:.3e}; azimuthal spread {spread:.3e}; "
          f"sigma {sig:.6f} optical theorem {sig_ot:.6f} partial waves {sig_pw:.6f}", flush=True)
    assert e_rel < 2e-2
    assert abs(sig - sig_pw) < 3e-2 * sig_pw
    assert abs(sig_ot - sig) < 3e-2 * sig
    s.close()
