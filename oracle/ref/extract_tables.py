"""Build step for the reference harness (test infrastructure only).

Reads the reference's setup_loadParameters() body from PandemicSim.cu and
emits a C++ header into oracle/_ref/ holding the literal table assignments
found there, wrapped in ref_load_tables().  The harness compiles against the
reference's own host_functions.cpp, so the golden vectors it writes follow
the reference's own table values, not this repository's copy of them.

Usage: python3 extract_tables.py <path to PandemicSim.cu> > ref_tables.h
"""
import re
import sys

ALLOWED_INDEX = re.compile(r"^(\d+|PROFILE_[A-Z0-9]+|BUSINESS_TYPE_[A-Z]+)$")
ASSIGN = re.compile(r"^\s*([A-Z_]+_HOST)((?:\[[A-Za-z0-9_]+\])+)\s*=\s*([^;]+);")


def main(path):
    text = open(path, encoding="latin-1").read()
    start = text.index("void PandemicSim::setup_loadParameters()")
    end = text.index("void PandemicSim::setup_pushDeviceData()")
    body = text[start:end]
    out = ["// generated from the reference setup_loadParameters(); do not commit",
           "static void ref_load_tables(void) {"]
    for line in body.splitlines():
        m = ASSIGN.match(line)
        if not m:
            continue
        idx = re.findall(r"\[([A-Za-z0-9_]+)\]", m.group(2))
        if not all(ALLOWED_INDEX.match(i) for i in idx):
            continue
        out.append("    %s%s = %s;" % (m.group(1), m.group(2), m.group(3).strip()))
    out.append("}")
    print("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1])
