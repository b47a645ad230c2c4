"""The library's host-only C++ (the reference-compatible random stream and
its threads, the projected-problem kernels with the two-thread split SVD,
the matrix loaders) and the oracle's solvers, built with g++ under
AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer and
run on the CPU (tests/cxx/host_sanitize.cpp). A sanitizer finding aborts the
driver; a wrong value makes it exit non-zero."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = ["krans_amd/csrc/refstream.cpp", "krans_amd/csrc/dense.cpp", "krans_amd/csrc/ingest.cpp"]
ORACLE = ["oracle/dense.cpp", "oracle/cyclic.cpp", "oracle/lobpcg.cpp", "oracle/loader.cpp",
          "oracle/capi.cpp"]
BASE = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-Iinclude",
        "-Ikrans_amd/csrc", "-Ioracle", "tests/cxx/host_sanitize.cpp"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


def build_and_run(tmp_path, name, flags, extra):
    exe = str(tmp_path / name)
    cmd = BASE + flags + HOST + extra + ["-pthread", "-o", exe]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, TMPDIR=str(tmp_path))
    r = subprocess.run([exe, "tests/golden/mat1.ija"], cwd=ROOT, capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "host_sanitize: 0 failure(s)" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr


def test_address_and_undefined_behaviour(tmp_path):
    build_and_run(tmp_path, "host_asan",
                  ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-DWITH_ORACLE"],
                  ORACLE)


def test_threads(tmp_path):
    build_and_run(tmp_path, "host_tsan", ["-fsanitize=thread"], [])
