"""CPU tests of the qscat_amd driver's argument handling and input-file
errors (qscat.cpp:82-104, :214-297).  Every case ends before the driver
touches a device: the help text, a missing -i, -i without a file, a file
that cannot be opened (Error 3) and a file with fewer than ten parameter
lines (Error 4), with the reference's messages and exit codes."""
import subprocess

import pytest

from conftest import PKG_DIR, ROOT

DRIVER = PKG_DIR / "qscat_amd"


def run(*args, cwd=None):
    if not DRIVER.exists():
        pytest.fail(f"{DRIVER} is not built")
    return subprocess.run([str(DRIVER), *args], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_help():
    r = run("-h")
    assert r.returncode == 0
    assert "Usage:" in r.stdout and "-i FILE" in r.stdout


def test_missing_input_flag():
    r = run()
    assert r.returncode == 1
    assert "Error: mandatory -i FILE for model paramters" in r.stderr   # the reference's spelling
    assert "Usage:" in r.stdout


def test_input_flag_without_file():
    for args in (("-i",), ("-i", "--fp32")):
        r = run(*args)
        assert r.returncode == 1
        assert "Error: -i requires an argument" in r.stderr


def test_unreadable_input_file(tmp_path):
    r = run("-i", str(tmp_path / "absent.txt"))
    assert r.returncode == 1
    assert f"Error 3: Cannot open input file {tmp_path / 'absent.txt'}" in r.stderr


def test_incomplete_input_file(tmp_path):
    lines = (ROOT / "tests" / "golden" / "InputExample.txt").read_text().splitlines()
    data = [ln for ln in lines if ln.strip() and not ln.lstrip().startswith("#")]
    assert len(data) == 10
    f = tmp_path / "short.txt"
    f.write_text("\n".join(data[:9]) + "\n")      # no prefix line
    r = run("-i", str(f))
    assert r.returncode == 1
    assert f"Error 4: incomplete parameter file {f}" in r.stderr
