"""Build libkrans_hip.so (all HIP/C++ sources of krans_amd/csrc) for gfx950.

Usage: python -m krans_amd.build  (or krans_amd.build.build()).
Objects and the library go to krans_amd/_build/ (in-tree, git-ignored).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libkrans_hip.so")
ARCH = os.environ.get("KRANS_HIP_ARCH", "gfx950")

CXXFLAGS = [
    "-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(os.path.dirname(HERE), "include"),
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (hipcc only; the host pass warns)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(f for f in os.listdir(SRC) if f.endswith((".hip", ".cpp")))


def _compile(f):
    src = os.path.join(SRC, f)
    obj = os.path.join(OUT, f + ".o")
    cmd = ["hipcc", *CXXFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, sources()))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_solvers_entry()
    build_examples()
    if verbose:
        print("built", LIB)
    return LIB


ENTRY_SRC = os.path.join(SRC, "solvers_entry", "krans_solvers_hip.cpp")
ENTRY_LIB = os.path.join(OUT, "libkrans_solvers_hip.so")


def build_solvers_entry():
    """libkrans_solvers_hip.so: the reference's krans::solvers entry points
    forwarding to libkrans_hip.so (plain host C++, g++)."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    tmp = ENTRY_LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I", inc, ENTRY_SRC, "-o",
           tmp, "-L", OUT, "-lkrans_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"entry library build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, ENTRY_LIB)


EXAMPLES = {"quickstart_hip": os.path.join(os.path.dirname(HERE), "tests", "cxx",
                                           "quickstart_hip.cpp"),
            "reference_scenarios_hip": os.path.join(os.path.dirname(HERE), "tests", "cxx",
                                                    "reference_scenarios_hip.cpp"),
            "krans_nullspace_hip": os.path.join(HERE, "apps", "krans_nullspace_hip.cpp")}


def build_examples():
    """C++ programs written against the KraNS API + HIP adapters (plain g++,
    linked to libkrans_hip.so): the quick-start drop-in check of the C++
    boundary and the krans_nullspace command line on the HIP backend."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    for name, src in EXAMPLES.items():
        if not os.path.exists(src):
            continue
        out = os.path.join(OUT, name)
        cmd = ["g++", "-std=c++17", "-O2", "-I", inc, src, "-o", out, "-L", OUT,
               "-lkrans_solvers_hip", "-lkrans_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
