"""The reference's own solver scenarios (tests/test_cyclic_lanczos.cpp tests
1-6, tests/test_lobpcg.cpp tests 1-5) compiled against the unchanged KraNS
C++ API with the HIP adapters swapped in (tests/cxx/reference_scenarios_hip.cpp):
the reference's options, counting decorators, initial vectors,
preconditioner callbacks (identity and steepest-descent precond_AAt,
steepest-descent M_inv) and CHECK thresholds, on mat1 and mat2 from the
reference's data directory (fixtures in tests/golden/)."""
import os
import subprocess

import pytest

import krans_amd as K

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(K.LIB_PATH), "reference_scenarios_hip")
SCENARIOS = ["cyclic_small", "cyclic_mat1", "cyclic_identity", "cyclic_mat2", "cyclic_rank1",
             "cyclic_scale", "lobpcg_small", "lobpcg_mat1", "lobpcg_identity", "lobpcg_mat2",
             "lobpcg_scale"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_reference_scenario(scenario, precision):
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    print(f"[case] c++ reference scenario {scenario} {precision}", flush=True)
    r = subprocess.run([EXE, os.path.join(HERE, "golden"), precision, scenario],
                       capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "ALL OK" in r.stdout
