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


def enum_members(header, name):
    """(member, value) of `enum name { ... }` in an include/pandemic header,
    comments stripped, implicit values counted up from the last explicit."""
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "pandemic", header)).read(), flags=re.S)
    body = re.search(r"enum\s+%s\s*\{(.*?)\}" % name, text, flags=re.S).group(1)
    out, v = [], -1
    for item in (x.strip() for x in body.split(",")):
        if not item:
            continue
        m = re.match(r"([A-Z0-9_]+)\s*(?:=\s*(\d+))?$", item)
        assert m, item
        v = int(m.group(2)) if m.group(2) else v + 1
        out.append((m.group(1), v))
    return out


def test_python_constants_follow_the_header_enums():
    import pandemicthrust_amd.api as api
    arrays = enum_members("pt_abi.h", "pt_array")
    assert len(arrays) >= 18
    for member, value in arrays:
        assert getattr(api, member) == value, member
    # launch settings: the binding's names in the header's order
    knobs = [m for m, _ in enum_members("pt_abi.h", "pt_launch_knob")]
    assert knobs[-1] == "PT_LAUNCH_KNOBS" and len(knobs) - 1 == len(api.LAUNCH_KNOBS)
    short = {"PT_LAUNCH_CONTACT_BLOCKS_PER_CU": "contact_blocks_per_cu",
             "PT_LAUNCH_GEN_BLOCKS_PER_CU": "gen_blocks_per_cu", "PT_LAUNCH_SORT_WIDE_BITS": "sort_wide_bits",
             "PT_LAUNCH_MARK_LIMIT": "mark_limit_pct", "PT_LAUNCH_SUBSET_SHARE": "subset_share_pct",
             "PT_LAUNCH_SORT_MAX_BITS": "sort_max_bits", "PT_LAUNCH_SEGMENTED_TABLES": "segmented_tables"}
    assert [short[k] for k in knobs[:-1]] == api.LAUNCH_KNOBS
