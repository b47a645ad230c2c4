"""The ctypes mirrors of the C ABI's option and result structs
(krans_amd/__init__.py) match include/krans_hip.h field for field: a small C
program compiled against the header prints sizeof and every offsetof, and
the ctypes layout must give the same numbers (a field added on one side
only would shift every later field)."""
import ctypes as C
import os
import subprocess

import pytest

import krans_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {"kh_cyclic_options": K._CyclicOpts, "kh_cyclic_result": K._CyclicRes,
           "kh_lobpcg_options": K._LobpcgOpts, "kh_lobpcg_result": K._LobpcgRes}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "krans_hip.h"',
             "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append(f'  printf("{name} sizeof %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{name} {fname} %zu\\n", offsetof({name}, {fname}));')
    lines += ["  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    table = {}
    for ln in out.splitlines():
        struct, field, value = ln.split()
        table[struct, field] = int(value)
    return table


@pytest.mark.parametrize("name", list(STRUCTS))
def test_ctypes_mirror_matches_header(c_layout, name):
    cls = STRUCTS[name]
    assert C.sizeof(cls) == c_layout[name, "sizeof"], name
    for fname, _ in cls._fields_:
        assert getattr(cls, fname).offset == c_layout[name, fname], (name, fname)
