"""The C-ABI library loads and exports every function that
include/pandemic/*.h declares.  Nothing is called, so no GPU is needed."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pandemicthrust_amd", "lib", "libpt_hip.so")


def declared_functions():
    names = set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "pandemic", "*.h"))):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"\b(pt_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_engine_entry_points():
    names = declared_functions()
    for n in ("pt_sim_create", "pt_sim_run", "pt_batch_create", "pt_batch_run", "pt_last_error"):
        assert n in names


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = sorted(n for n in declared_functions() if not hasattr(lib, n))
    assert not missing, missing


def test_python_binding_covers_the_headers():
    from pandemicthrust_amd.api import EXPORTS
    assert set(EXPORTS) == declared_functions()
