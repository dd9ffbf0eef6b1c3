"""The C-ABI library loads and exports every entry point include/mgsb.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "mgsb.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mgsb_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "mgsb_create" in names and "mgsb_step" in names and "mgsb_read_positions" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    import multigridsoftbody_amd as M

    lib = ctypes.CDLL(M.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    import multigridsoftbody_amd as M

    assert set(_declared()) == set(M.SIGNATURES)


def test_errors_are_codes_not_exits(tmp_path):
    import multigridsoftbody_amd as M

    m = M.mgsb_mesh()
    rc = M.lib.mgsb_load_msh(str(tmp_path / "missing.msh").encode(), ctypes.byref(m))
    assert rc == -2
    assert b"cannot open" in M.lib.mgsb_last_error()
    bad = tmp_path / "bad.msh"
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n3\n1 0 0 0\n")
    assert M.lib.mgsb_load_msh(str(bad).encode(), ctypes.byref(m)) == -3


def test_diagnostics_reject_bad_arguments_without_a_device():
    """The run aggregates and the per-step check history validate their
    arguments before touching the device: a null handle or output, or a
    negative length, is MGSB_ERR_ARG with a message."""
    import multigridsoftbody_amd as M

    n = ctypes.c_int32(-7)
    buf = (ctypes.c_double * 4)()
    assert M.lib.mgsb_step_history(None, buf, 4, ctypes.byref(n)) == -1
    assert b"mgsb_step_history" in M.lib.mgsb_last_error()
    assert M.lib.mgsb_step_history(None, None, 0, None) == -1
    assert M.lib.mgsb_run_stats(None, None, 0) == -1
    assert b"mgsb_run_stats" in M.lib.mgsb_last_error()
