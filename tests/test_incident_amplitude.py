"""CPU test of pstd_incident_amplitude (host code of libpstd_amd.so): the
solver's incident wave at the grid centre, pinned to the oracle's initial
field (PSTD_QScat.cpp:959-1118, phase origin at the injection corner),
which away from the ksi taper is e^{i khat.(r - r_inj)} at every grid
point: there psi^0 = inc * e^{i khat.(i - orig) d}."""
import numpy as np
import pytest

import pstd_amd
from pstd_amd import farfield as FF
from pstd_amd import workloads as W
from oracle import oracle as O


@pytest.mark.parametrize("cfg", [dict(nprocs=(2, 2, 2), L=(32, 32, 32), h=10, theta=0.3, phi=0.7),
                                 dict(nprocs=(1, 3, 1), L=(36, 44, 28), h=12, theta=2.6, phi=-2.1)])
def test_incident_amplitude_matches_initial_field(cfg):
    p, V = W.config_tiny(**cfg)
    inc = FF.incident_amplitude(p)
    assert abs(abs(inc) - 1) < 1e-14
    o = O.Oracle(p, V)
    psi0 = o.get_psi(1)                  # m_Psi0: the field at tau = 0
    g = pstd_amd.compute_geometry(p)
    kh = FF.incident_direction(p)
    c = [int(round(g.orig[a])) for a in range(3)]
    for off in [(0, 0, 0), (2, -1, 1), (-2, 2, -1)]:
        ijk = [c[a] + off[a] for a in range(3)]
        r = np.array([(ijk[a] - g.orig[a]) * p.d[a] for a in range(3)])
        want = inc * np.exp(1j * kh @ r)
        assert abs(psi0[tuple(ijk)] - want) < 1e-12, (ijk, psi0[tuple(ijk)], want)
