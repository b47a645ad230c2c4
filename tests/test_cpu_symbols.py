"""libkrans_hip.so defines the HIP solvers under krans::hip only, so it links
beside the reference's library, whose krans::solvers::find_nullspace_*
(CyclicLanczosSolver.cpp:663-668, LOBPCGSolver.cpp:837-842) it must not
redefine; libkrans_solvers_hip.so provides those names for programs without
the reference's library."""
import os
import subprocess

import krans_amd as K

BUILD = os.path.dirname(K.LIB_PATH)
ENTRY = ["find_nullspace_cyclic_lanczos_ex", "find_nullspace_cyclic_lanczos",
         "find_nullspace_lobpcg_ex", "find_nullspace_lobpcg"]


def defined(lib):
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(BUILD, lib)],
                         capture_output=True, text=True, check=True).stdout
    return [ln.split(" ", 2)[2] for ln in out.splitlines() if " T " in ln]


def test_hip_library_leaves_the_reference_names_free():
    syms = defined("libkrans_hip.so")
    for name in ENTRY:
        assert any(s.startswith(f"krans::hip::{name}(") for s in syms), name
        assert not any(s.startswith(f"krans::solvers::{name}(") for s in syms), name


def test_entry_library_defines_the_reference_names():
    syms = defined("libkrans_solvers_hip.so")
    for name in ENTRY:
        assert any(s.startswith(f"krans::solvers::{name}(") for s in syms), name
