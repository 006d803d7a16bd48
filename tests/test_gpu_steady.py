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
