"""GPU tests of batched independent solves (pstd_batch_*): every pass is one
launch over all problems and steps replay a captured HIP graph.  Each
problem must match its own solo solve, graph replays included."""
import numpy as np
import pytest

import pstd_amd
from pstd_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _problems(precision):
    specs = [(0, 64, 1.3), (1, 96, 2.7), (2, 64, 3.4), (3, 128, 1.9), (4, 96, 2.2)]
    return [W.batch_problem(i, N, k, precision=precision, steps=100) for i, N, k in specs]


@pytest.mark.parametrize("precision", [pstd_amd.PSTD_FP64, pstd_amd.PSTD_FP32])
def test_batch_matches_solo_runs(precision):
    probs = _problems(precision)
    print(f"case batch of {len(probs)} prec={precision}", flush=True)
    batch = pstd_amd.PSTDBatch([p for p, _ in probs], [v for _, v in probs])
    solo = [pstd_amd.PSTDQScat(p, v) for p, v in probs]
    assert all(s.pass_fusion() == 1 for s in solo)   # the batch fuses the y and z passes as they do
    tol = 1e-13 if precision == pstd_amd.PSTD_FP64 else 1e-6
    # graph replays (2 steps each) with eager host work between them, then an
    # odd remainder stepped eagerly
    for chunk in (4, 2, 6, 3):
        batch.Update(chunk)
        for s in solo:
            s.Update(chunk)
        batch.synchronize()
        for i, s in enumerate(solo):
            assert batch[i].time_index() == s.time_index()
            assert rel(batch[i].get_psi(0), s.get_psi(0)) < tol, (chunk, i)
    # accumulation steps (eager) and surfaces
    batch.update_accumulate(3)
    for s in solo:
        s.update_accumulate(3)
    for i, s in enumerate(solo):
        assert batch[i].surface_count() == 3
        for deriv in (0, 1):
            assert rel(batch[i].get_surfaces(deriv), s.get_surfaces(deriv)) < tol


def test_batch_rejects_long_lines():
    p, V = W.config_tiny(nprocs=(1, 2, 1), L=(48, 160, 40), h=15, n_abc=8, n_sca=9, n_ksi=8)
    with pytest.raises(pstd_amd.PSTDError) as e:
        pstd_amd.PSTDBatch([p], [V])
    assert e.value.code == pstd_amd.PSTD_EINVAL


def test_streamed_batch_matches_solo_runs():
    """The 48-problem benchmark batch (grids 64..160^3, FP32, 113 M points):
    its psi and psi^(n-1) together exceed the last-level cache, so the batch
    runs the streamed-psi^(n-1) kernels; members of every grid size must
    match their solo runs, which at their own sizes take the plain kernels."""
    probs = W.config_batch48()
    batch = pstd_amd.PSTDBatch([q for q, _ in probs], [v for _, v in probs])
    npts = sum(int(np.prod(m.N)) for m in batch.members)
    print(f"case streamed batch48, {npts} points", flush=True)
    assert 2 * npts * 2 * 4 > 256 * 2 ** 20
    pick = {}
    for i, m in enumerate(batch.members):
        pick.setdefault(m.N[0], i)                  # one member per grid size
    solo = {i: pstd_amd.PSTDQScat(*probs[i]) for i in pick.values()}
    batch.Update(5)
    batch.update_accumulate(2)
    for i, s in solo.items():
        s.Update(5)
        s.update_accumulate(2)
        assert rel(batch[i].get_psi(0), s.get_psi(0)) < 1e-6, i
        assert rel(batch[i].get_surfaces(0), s.get_surfaces(0)) < 1e-6, i
