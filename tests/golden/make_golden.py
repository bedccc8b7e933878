"""Regenerates the golden fixtures in this directory.

Builds oracle/ref/Makefile (a harness around the reference's own
host_functions.cpp and vendored Random123 threefry.h, compiled from
/root/reference) and runs it.  The fixtures it writes:

  threefry_ref.csv   Threefry2x64-20 / 2x32-20 outputs of the vendored header,
                     including the published known-answer inputs and the
                     reference's {off,off} counter / SEED key convention.
  ref_tables.csv     every literal table the reference's
                     setup_loadParameters() assigns.
  ref_scaling.csv    scaled per-type workplace counts and offsets.
  host_workplace.csv host_recalcWorkplace outputs.
  host_weekday.csv   host_recalc_weekdayErrandDests_assignProfile outputs.
  host_weekend.csv   host_recalc_weekendErrandDests outputs.
  host_fish.csv      the four fishing functions over boundary-probing draws.

Only this container has /root/reference; the fixtures are committed so that
tests elsewhere never need it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    refdir = os.path.join(ROOT, "oracle", "ref")
    subprocess.check_call(["make", "-s"], cwd=refdir)
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    subprocess.check_call([harness, HERE])
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
