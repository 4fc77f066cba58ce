"""The C ABI library loads and exports every function include/prism_hip.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

from prism_executor import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "prism_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(prism_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "prism_join_probe_i64" in names and "prism_groupby_update" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    bound = {s[0] for s in _lib.signatures()}
    assert set(declared_functions()) <= bound


def test_version_and_error_without_device():
    lib = _lib.load()
    assert lib.prism_version().startswith(b"prism-hip")
    assert lib.prism_last_error() == b""
