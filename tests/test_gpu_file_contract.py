"""The reference's process and file contract on the GPU.

pandemicthrust_amd/bin/pandemicthrust (the C++ host with the PandemicSim flow,
main.cu:12-31) reads constants.csv and seed.txt and must write
output_infected_stats.csv, output_rn.csv and infected_max.txt byte-identical
to the CPU oracle's driver run on the same inputs.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pandemicthrust_amd", "bin", "pandemicthrust")
ORACLE = os.path.join(ROOT, "oracle", "build", "pt_oracle")
HEADER = ("constants,max_days,reproduction_number_pandemic,reproduction_number_seasonal,initial_infected_pandemic,"
          "initial_infected_seasonal,people_scaling_factor,location_scaling_factor,percent_symptomatic,asymp_factor")
FILES = ["output_infected_stats.csv", "output_rn.csv", "infected_max.txt"]


def _inputs(d, row, seed):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "constants.csv"), "w") as f:
        f.write(HEADER + "\n" + row + "\n")
    with open(os.path.join(d, "seed.txt"), "w") as f:
        f.write("%d\n" % seed)


@pytest.mark.gpu
@pytest.mark.parametrize("row,seed", [
    ("constants,30,1.8,1.3,10,10,0.0425,0.0425,1,0", 1),
    ("constants,60,2.2,1.5,15,5,0.1,0.1,0.7,0.3", 4),
    ("constants,25,0,0,5,5,0.0425,0.0425,1,0", 2),  # dies out: terminal I == 0 row, nan/inf in R
])
def test_output_files_match_oracle(tmp_path, row, seed):
    g, o = str(tmp_path / "gpu"), str(tmp_path / "oracle")
    _inputs(g, row, seed)
    _inputs(o, row, seed)
    rg = subprocess.run([EXE], cwd=g, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert rg.returncode == 0, rg.stderr
    ro = subprocess.run([ORACLE], cwd=o, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert ro.returncode == 0, ro.stderr
    for name in FILES:
        with open(os.path.join(g, name)) as a, open(os.path.join(o, name)) as b:
            assert a.read() == b.read(), name
    day_lines = [l for l in rg.stdout.splitlines() if l.startswith("Day ")]
    assert day_lines == [l for l in ro.stdout.splitlines() if l.startswith("Day ")]
