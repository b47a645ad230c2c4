"""krans_nullspace_hip end to end on the reference's data matrices: the
outputs of apps/krans_nullspace.cpp (N.csv as ExportIO.hpp's write_csv_basis
writes it, report.json as write_json_report writes it) and the null spaces."""
import json
import os
import subprocess

import numpy as np
import pytest

import krans_amd as K
from conftest import principal_sine

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(K.LIB_PATH), "krans_nullspace_hip")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def dense_null(path, k):
    A = K.load_ija(path)
    _, s, Vt = np.linalg.svd(A.to_scipy().toarray())
    return A, Vt[-k:].T


@pytest.mark.parametrize("args,name,k", [
    (["--solver", "cyclic", "--k", "5", "--ncv", "60", "--max-restarts", "200",
      "--cg-maxit", "60", "--cg-rel", "1e-8", "--left"], "mat1", 5),
    (["--solver", "lobpcg", "--k", "10", "--lobpcg-block", "16", "--tol-rel", "1e-8",
      "--precision", "mixed"], "mat3", 10)])
def test_cli_solves_and_writes_outputs(tmp_path, args, name, k):
    path = os.path.join(GOLDEN, name + ".ija")
    print(f"[case] cli {name} {args[1]}", flush=True)
    r = subprocess.run([EXE, "--input", path, "--seed", "42", *args], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith(f"OK [{args[1]}]")
    A, N = dense_null(path, k)
    rows = (tmp_path / "N.csv").read_text().splitlines()
    assert len(rows) == A.n
    # fixed notation, 17 digits after the point
    assert all(len(f.split(".")[1]) == 17 for f in rows[0].split(","))
    U = np.array([[float(f) for f in line.split(",")] for line in rows])
    assert U.shape == (A.n, k)
    rep = json.loads((tmp_path / "report.json").read_text())
    assert rep["solver"] == args[1] and rep["rows"] == A.m and rep["cols"] == A.n
    assert rep["nullity"] == k and rep["rank"] == A.n - k
    assert set(rep["timings_ms"]) == {"load", "solve"}
    res = np.linalg.norm(A.to_scipy() @ U, axis=0)
    assert np.allclose(rep["residuals"], res, rtol=1e-6, atol=1e-15)
    assert rep["max_offdiag_abs_dot"] <= 1e-8
    assert principal_sine(U, N) <= 1e-6
    if "--left" in args:
        assert (tmp_path / "N_left.csv").exists()
