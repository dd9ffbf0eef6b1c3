"""The ctypes mirrors of the C-ABI structs must have the layout the header
gives them: a C program compiled against include/mgsb.h prints sizeof and
every member's offsetof, which are compared with the ctypes fields (a field
added, dropped or reordered on one side only would otherwise pass every
call silently with shifted members)."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

import multigridsoftbody_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = ["mgsb_mesh", "mgsb_params", "mgsb_step_stats", "mgsb_run_stats_t", "mgsb_tune_result", "mgsb_energy_t"]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    d = tmp_path_factory.mktemp("layout")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mgsb.h"', "int main(void) {"]
    for name in STRUCTS:
        lines.append(f'    printf("{name} sizeof %zu\\n", sizeof({name}));')
        for field, _ in getattr(M, name)._fields_:
            lines.append(f'    printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["    return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {(a, b): int(c) for a, b, c in (ln.split() for ln in out.splitlines())}


@pytest.mark.parametrize("name", STRUCTS)
def test_ctypes_struct_matches_header(c_layout, name):
    cls = getattr(M, name)
    assert C.sizeof(cls) == c_layout[(name, "sizeof")]
    for field, _ in cls._fields_:
        assert getattr(cls, field).offset == c_layout[(name, field)], field
