"""The C ABI library loads and exports every entry point that include/mclgpu.h declares (no GPU needed)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mclgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mcl_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mcl_create" in names and "mcl_sweeps" in names and len(names) >= 20


def test_library_exports_every_declared_symbol(mcl):
    lib = ctypes.CDLL(mcl.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header(mcl):
    assert sorted(mcl.ABI) == declared_functions()


def test_abi_version_and_error_path_without_gpu(mcl):
    lib = mcl.load()
    assert lib.mcl_abi_version() == 1
    # argument errors are reported before any device call
    h = ctypes.c_void_p()
    assert lib.mcl_create(None, 0, ctypes.byref(h)) == 1
    assert b"no lattices" in lib.mcl_last_error()
