"""GPU tests of the steady-state detector (pstd_run_to_steady_state): the
per-window change history of the HIP path against the same back-to-back
one-period windows run on the CPU oracle (reset, Update +
AccumulateSurfaceTerms for a period, compare surface psi with the previous
window), and the last window's surfaces against the oracle's."""
import numpy as np
import pytest

import pstd_amd
from pstd_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def psi_parts(body, N):
    """Surface psi of the six faces from a surface-file body (each face holds
    its psi plane then its dpsi plane, complex interleaved)."""
    Nx, Ny, Nz = N
    out, o = [], 0
    for n in (Ny * Nz, Ny * Nz, Nx * Nz, Nx * Nz, Nx * Ny, Nx * Ny):
        out.append(body[o:o + 2 * n])
        o += 4 * n
    return np.concatenate(out)


def oracle_history(o, N, tol, max_periods):
    nper = o.steps_per_period()
    prev, hist = None, []
    for k in range(max_periods):
        o.reset_surface_terms()
        o.update_accumulate(nper)
        cur = psi_parts(o.get_surfaces(0), N)
        hist.append(1.0 if prev is None else float(np.linalg.norm(cur - prev) / np.linalg.norm(cur)))
        prev = cur
        if k > 0 and hist[-1] <= tol:
            return True, hist
    return False, hist


def _case(precision=pstd_amd.PSTD_FP64):
    # dx_ref = 1: 372 steps per period on a 44 x 24 x 24 grid
    p, V = W.config_tiny(nprocs=(2, 1, 1), L=(32, 24, 24), h=10, k=4.0, dx=0.25, precision=precision)
    print(f"case steady-state nprocs={tuple(p.nprocs)} L={tuple(p.L)} prec={precision}", flush=True)
    return p, V


def test_steady_state_matches_oracle_windows():
    p, V = _case()
    s = pstd_amd.PSTDQScat(p, V)
    conv, periods, steps, hist = s.run_to_steady_state(tol=0.05, max_periods=10)
    o = O.Oracle(p, V)
    oconv, ohist = oracle_history(o, s.N, 0.05, 10)
    assert conv and oconv
    assert periods == len(ohist)
    assert steps == periods * s.IterationsToAccumulateSurfaceTerms()
    np.testing.assert_allclose(hist, ohist, rtol=1e-7, atol=1e-10)
    assert hist[-1] <= 0.05 < hist[-2]
    # the accumulators hold the last window, as the oracle's do
    for deriv in (pstd_amd.PSTD_DERIV_PS, pstd_amd.PSTD_DERIV_FD):
        a, b = s.get_surfaces(deriv), o.get_surfaces(deriv)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-10
    assert s.surface_count() == s.IterationsToAccumulateSurfaceTerms()


def test_steady_state_reports_no_convergence():
    p, V = _case()
    s = pstd_amd.PSTDQScat(p, V)
    conv, periods, steps, hist = s.run_to_steady_state(tol=1e-12, max_periods=3)
    assert not conv and periods == 3 and len(hist) == 3
    assert hist[0] == 1.0 and all(h > 1e-12 for h in hist)
    with pytest.raises(pstd_amd.PSTDError) as e:
        s.run_to_steady_state(tol=0.0, max_periods=3)
    assert e.value.code == pstd_amd.PSTD_EINVAL


def test_steady_state_fp32_history_close_to_fp64_oracle():
    p, V = _case(pstd_amd.PSTD_FP32)
    s = pstd_amd.PSTDQScat(p, V)
    conv, periods, _, hist = s.run_to_steady_state(tol=0.05, max_periods=10)
    p64, V64 = _case()
    oconv, ohist = oracle_history(O.Oracle(p64, V64), s.N, 0.05, 10)
    assert conv == oconv and periods == len(ohist)
    np.testing.assert_allclose(hist, ohist, atol=1e-4)


def test_rre_steady_state_closer_than_plain_windows():
    """Reduced-rank extrapolation of the window sequence on the 96^3
    Gaussian well (no bound state): after 40 periods the estimate is at
    least 4x closer to the 150th plain window than the 40th plain window is,
    the accumulators hold that estimate (all derivative kinds, one-window
    normalisation), and the plain change history equals the plain run's."""
    from test_gpu_farfield import _gaussian_problem
    p, V = _gaussian_problem(-0.5, 1.2)
    print("case RRE steady state, 96^3 Gaussian well, 40 vs 150 periods", flush=True)
    s = pstd_amd.PSTDQScat(p, V)
    nper = s.IterationsToAccumulateSurfaceTerms()
    plain = {}
    for k in range(1, 151):
        s.reset_surface_terms()
        s.update_accumulate(nper)
        if k in (39, 40, 150):
            plain[k] = (s.get_surfaces(0), s.get_surfaces(1))
    s.close()
    ref = plain[150][0]
    g = pstd_amd.PSTDQScat(p, V)
    conv, periods, steps, hist, raw = g.run_to_steady_state_rre(tol=1e-12, max_periods=40, rank=10)
    assert not conv and periods == 40 and steps == 40 * nper
    est_ps, est_fd = g.get_surfaces(0), g.get_surfaces(1)
    e_rre, e_plain = rel(est_ps, ref), rel(plain[40][0], ref)
    print(f"error vs window 150: RRE {e_rre:.2e}, plain window 40 {e_plain:.2e}; "
          f"estimate change {hist[-1]:.2e}, plain change {raw[-1]:.2e}", flush=True)
    assert e_rre < e_plain / 4
    # the same weights act on the derivative planes: FD moves with PS
    assert rel(est_fd, plain[150][1]) < rel(plain[40][1], plain[150][1]) / 4
    # the plain change of the last window is the change between windows 39 and 40
    assert abs(raw[-1] - _psi_change(plain[40][0], plain[39][0], g.N)) < 1e-9
    g.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _psi_change(a, b, N):
    """Relative L2 change of the surface psi planes between two bodies (the
    detector's measure: every psi plane of the six faces)."""
    Nx, Ny, Nz = N
    dims = [Ny * Nz, Nx * Nz, Nx * Ny]
    ia, off = [], 0
    for ax in range(3):
        for _ in range(2):
            ia.append(np.arange(off, off + 2 * dims[ax]))
            off += 4 * dims[ax]
    idx = np.concatenate(ia)
    return float(np.linalg.norm(a[idx] - b[idx]) / np.linalg.norm(a[idx]))


def test_mixed_precision_steady_state_matches_fp64():
    """pstd_run_to_steady_state_mixed on the 96^3 Gaussian well: FP32
    extrapolated windows to a 1e-4 estimate change, then FP64 windows from
    the handed-over state to 1e-6.  The estimate agrees with the FP64-only
    run's to 2e-6 (both stop at a 1e-6 change), although the FP32 phase
    alone is about 1e-4 off; the FP64 solver's clock continues the FP32 one;
    mismatched solver pairs are rejected."""
    from test_gpu_farfield import _gaussian_problem
    p64, V = _gaussian_problem(-0.5, 1.2)
    p32, _ = _gaussian_problem(-0.5, 1.2, precision=pstd_amd.PSTD_FP32)
    print("case mixed-precision steady state, 96^3 Gaussian well", flush=True)
    a = pstd_amd.PSTDQScat(p64, V)
    conv, n64, _, _, _ = a.run_to_steady_state_rre(tol=1e-6, max_periods=150, rank=10)
    assert conv
    ref = (a.get_surfaces(0), a.get_surfaces(1))
    a.close()
    s32 = pstd_amd.PSTDQScat(p32, V)
    s64 = pstd_amd.PSTDQScat(p64, V)
    nper = s64.IterationsToAccumulateSurfaceTerms()
    conv, n1, n2, hist, raw = s64.run_to_steady_state_mixed(s32, switch_tol=1e-4, switch_max=80, tol=1e-6,
                                                             max_periods=150, rank=10)
    e_ps, e_fd = rel(s64.get_surfaces(0), ref[0]), rel(s64.get_surfaces(1), ref[1])
    print(f"FP64 alone {n64} periods; mixed {n1} FP32 + {n2} FP64 periods; estimate vs FP64-only: "
          f"PS {e_ps:.2e} FD {e_fd:.2e}", flush=True)
    assert conv and n1 > 0 and n2 > 0 and len(hist) == n2 and hist[-1] <= 1e-6
    assert e_ps < 2e-6 and e_fd < 2e-6
    t32, t64 = s32.time_index(), s64.time_index()
    assert t32 > n1 * nper and t64 - t32 in (n2 * nper, (n2 + 1) * nper)
    with pytest.raises(pstd_amd.PSTDError) as e:      # roles swapped: the FP32 solver as the FP64 one
        s32.run_to_steady_state_mixed(s64, max_periods=2)
    assert e.value.code == pstd_amd.PSTD_EINVAL
    other, Vo = _gaussian_problem(-0.5, 1.2, precision=pstd_amd.PSTD_FP32)
    other.theta = 0.7                                   # a different incidence
    o32 = pstd_amd.PSTDQScat(other, Vo)
    with pytest.raises(pstd_amd.PSTDError) as e:
        s64.run_to_steady_state_mixed(o32, max_periods=2)
    assert e.value.code == pstd_amd.PSTD_EINVAL
    for g in (s32, s64, o32):
        g.close()
