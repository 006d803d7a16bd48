"""GPU test of the qscat_amd driver (qscat.cpp:48-206) on a small input file
in the reference's format: the files it writes at the reference's cadence
(Save_g and the origin planes at start-up, the planes after every tenth of
the run, the surface files of the last period) and the surface file against
the same run through the library API with potential.cpp's barrier sampled
on the potential box."""
import math
import subprocess

import numpy as np
import pytest

import pstd_amd
from pstd_amd import workloads as W
from pstd_amd.surface_io import read_surface_file
from conftest import PKG_DIR

pytestmark = pytest.mark.gpu

DRIVER = PKG_DIR / "qscat_amd"
ITER = 40

INPUT = f"""# numbers of processes in x, y, z directions, respectively
1 2 1
# time step and grid sizes
0.0066 {2 * math.pi / 10!r} {2 * math.pi / 10!r} {2 * math.pi / 10!r}
# length of 1-D fft in x, y, z directions, respectively
48 40 48
# thickness of absorbing boundary, scattering field zone, ksi transition zone, domain overlapping halos
4 3 4 13
# incident angle in unit of radiance
1.2 0.4
# Energy in unit of micro-electron-volts
0.327216805
# absorbing boundary ComplexPotential_U0
5.0
# absorbing boundary ComplexPotential_alpha
1.0
# number of iterations
{ITER}
# data-save-to file name for final result
drv
"""


def test_driver_files_and_surfaces(tmp_path):
    if not DRIVER.exists():
        pytest.fail(f"{DRIVER} is not built")
    (tmp_path / "in.txt").write_text(INPUT)
    print("case qscat_amd 1x2x1 L=48,40,48 h=13 Iter=40", flush=True)
    r = subprocess.run([str(DRIVER), "-i", "in.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p = pstd_amd.load_input(tmp_path / "in.txt")
    s = pstd_amd.PSTDQScat(p, W.example_barrier(p))
    g = s.geometry
    # Save_g: three core-length files per subdomain
    for rank in range(2):
        for a in range(3):
            f = tmp_path / f"Rank{rank}_g{a + 1}.dat"
            assert len(f.read_text().splitlines()) == g.core2[a][rank if a == 1 else 0] - \
                g.core1[a][rank if a == 1 else 0] + 1
    # planes through the origin, at nTau = 1 and after steps 4, 8, ..., 40
    origin = [int(0.0 / p.d[a] + g.orig[a]) for a in range(3)]
    tenth = ITER // 10
    stamps = [("Init", 1)] + [("Iter", i + 1) for i in range(tenth, ITER + 1, tenth)]
    per_stamp = 0
    for prefix, ntau in stamps:
        for a, tag in enumerate("xyz"):
            files = sorted(tmp_path.glob(f"{prefix}{ntau}_Rank*_r{tag}{origin[a]}.dat"))
            owners = [rank for rank in range(2)
                      if g.core1[a][rank if a == 1 else 0] <= origin[a] <= g.core2[a][rank if a == 1 else 0]]
            assert [f.name for f in files] == [f"{prefix}{ntau}_Rank{rk}_r{tag}{origin[a]}.dat" for rk in owners]
            per_stamp += 2 * len(owners) if prefix == "Init" else 0
    # x and z planes cross both subdomains' cores, the y plane one; real and imaginary files
    assert per_stamp == 2 * (2 + 1 + 2)
    assert len(list(tmp_path.glob("Iter*_Rank*"))) == 10 * per_stamp
    # the last period's surface files against the library run
    s.update_accumulate(ITER)
    s.SaveVirtualSurfaces(str(tmp_path / "lib"))
    n0 = ITER + 1 - ITER
    for kind in ("PS", "FD"):
        hd, dd = read_surface_file(tmp_path / f"drv_{kind}_{n0}.dat")
        hl, dl = read_surface_file(tmp_path / f"lib_{kind}_{n0}.dat")
        assert hd == hl
        assert np.linalg.norm(dd - dl) <= 1e-12 * np.linalg.norm(dl)
    s.close()
