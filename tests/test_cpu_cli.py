"""krans_nullspace_hip argument handling (apps/krans_nullspace.cpp's flags and
exit codes), without touching the GPU: every case here exits before the
device is opened."""
import os
import subprocess

import pytest

import krans_amd as K

EXE = os.path.join(os.path.dirname(K.LIB_PATH), "krans_nullspace_hip")


def run(*args):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(EXE), reason="CLI not built")
def test_cli_arguments(tmp_path):
    r = run("--help")
    assert r.returncode == 0 and "--solver" in r.stderr and "--lobpcg-block" in r.stderr
    assert run("--solver", "cyclic").returncode == 2                       # no --input
    r = run("--solver", "svd", "--input", "x.ija")
    assert r.returncode == 2 and "Unknown solver" in r.stderr               # not on this backend
    assert run("--bogus").returncode == 2
    assert run("--solver", "cyclic", "--input", "x.txt").returncode == 2    # extension
    assert run("--solver", "cyclic", "--input", "x.ija", "--k", "abc").returncode == 2
    assert run("--solver", "cyclic", "--input", "x.ija", "--backend", "cpu").returncode == 2
    assert run("--solver", "cyclic", "--input", "x.ija", "--precision", "fp16").returncode == 2
    r = run("--solver", "cyclic", "--input", str(tmp_path / "missing.ija"))
    assert r.returncode == 2 and "Matrix load error" in r.stderr
