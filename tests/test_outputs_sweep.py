"""Output-file writers and the batcher-compatible sweep driver on the CPU.

The Python writers (pandemicthrust_amd/outputs.py) must produce the same
bytes as the oracle's C writer (printf-based, like the reference's
PandemicSim.cu:2771-2818, :855-866, :2276-2297) from the same run.  The
sweep's constants files must match the reference batcher's text
(PythonBatcher/write_gpu_constants.py:13-14, :39, :46-47).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
from pandemicthrust_amd import outputs, sweep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pandemicthrust_amd", "bin", "pandemicthrust")


@pytest.mark.parametrize("row,seed", [
    (["30", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"], 1),
    (["25", "0", "0", "5", "5", "0.0425", "0.0425", "1", "0"], 2),   # dies out: -nan and inf in output_rn.csv
    (["12", "2.2", "1.5", "15", "5", "0.02", "0.05", "0.7", "0.3"], 4),
])
def test_python_writers_match_the_oracle_files(tmp_path, row, seed):
    sim = O.OracleSim(O.params_from_row(row), seed)
    sim.setup()
    counts, inf = sim.run()
    odir, pdir = tmp_path / "oracle", tmp_path / "python"
    odir.mkdir()
    pdir.mkdir()
    sim.write_outputs(str(odir), counts, inf)
    cp, cs = sim.generation_offsets()
    md = int(row[0])
    outputs.write_run(str(pdir), counts, inf, sim.n_people, cp, cs, md)
    for name in ("output_infected_stats.csv", "infected_max.txt", "output_rn.csv"):
        a = (odir / name).read_text()
        b = (pdir / name).read_text()
        assert a == b, name
    if row[1] == "0":
        assert "-nan" in (pdir / "output_rn.csv").read_text()


def test_c_float_formatting():
    assert outputs.c_float(np.float32(1.8)) == "1.800000"
    assert outputs.c_float(np.float32(0.1)) == "0.100000"
    assert outputs.c_float(np.float32(2.0 / 3.0)) == "0.666667"
    assert outputs._rn_text(0, 0) == "-nan"
    assert outputs._rn_text(3, 0) == "inf"
    assert outputs._rn_text(1, 3) == "0.333333"


def test_constants_file_text(tmp_path):
    out = str(tmp_path / "g")
    p = sweep.write_constants_file("p0.1_l1_30_60", initial_pandemic=30, initial_seasonal=60, people_scale=0.1,
                                   location_scale=1, output_path=out)
    text = open(p).read()
    # write_gpu_constants.py:13-14 with the defaults of :6
    assert text == ("constants,max_days,reproduction_number_pandemic,reproduction_number_seasonal,"
                    "initial_infected_pandemic,initial_infected_seasonal,people_scaling_factor,"
                    "location_scaling_factor,percent_symptomatic,asymp_factor\n"
                    "constants,100,1.8,1.3,30,60,0.1,1,1,0\n")
    params = sweep.read_constants(p)
    assert params.max_days == 100 and params.init_seasonal == 60
    assert params.people_scaling == np.float32(0.1) and params.r_pandemic == np.float32(1.8)


def test_experiment_lists(tmp_path):
    out = str(tmp_path / "gpu_experiment_files")
    names = sweep.write_experiments(out)
    assert len(names) == 32
    assert names[:4] == ["p0.1_l0.1_30_30", "p0.1_l0.1_30_60", "p0.1_l0.1_60_30", "p0.1_l0.1_60_60"]
    assert names[-1] == "p20_l200_60_60"
    files = open(os.path.join(out, "experiment_files.txt")).read().splitlines()
    assert files[0] == os.path.join(out, "p0.1_l0.1_30_30", "constants.csv")
    assert open(os.path.join(out, "experiment_names.txt")).read().splitlines() == names
    paths = sweep.read_experiment_files(os.path.join(out, "experiment_files.txt"))
    assert all(os.path.exists(p) for p in paths) and len(paths) == 32


def test_relative_experiment_lists(tmp_path):
    # The batcher writes paths relative to its working directory
    # ("./gpu_experiment_files/<name>/constants.csv").
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        sweep.write_experiments(scales=[[0.1, 0.1]], init_p=[30], init_s=[30])
    finally:
        os.chdir(cwd)
    paths = sweep.read_experiment_files(str(tmp_path / "gpu_experiment_files" / "experiment_files.txt"))
    assert paths == [str(tmp_path / "gpu_experiment_files" / "p0.1_l0.1_30_30" / "constants.csv")]


def test_chunks_respect_the_budget():
    from pandemicthrust_amd import Params
    runs = [{"params": Params.from_row(["10", "1.8", "1.3", "1", "1", s, "1", "1", "0"])} for s in
            ("0.1", "0.1", "1", "0.1")]
    per = [sweep.estimated_people(r["params"]) * sweep.BYTES_PER_AGENT for r in runs]
    chunks = list(sweep._chunks(runs, per[0] * 2 + 1))
    assert [len(c) for c in chunks] == [2, 1, 1]
    assert sum(len(c) for c in sweep._chunks(runs, 10 ** 15)) == 4


def test_host_program_rejects_bad_arguments():
    if not os.path.exists(EXE):
        pytest.skip("host program not built")
    for args in (["--bogus"], ["--checkpoint", "x"], ["--checkpoint-day", "3"], ["a", "b"]):
        r = subprocess.run([EXE] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=30)
        assert r.returncode == 2 and "usage" in r.stderr, args


def _header_define(name):
    import re
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pandemic",
                             "pt_batch.h")).read()
    return int(re.search(r"#define\s+%s\s+(\d+)" % name, text).group(1))


def test_chunks_fit_the_batch_slot_limit():
    """The reference batcher's 32 experiments x 3 seeds on a 230 GB budget:
    every chunk stays within the batch engine's agent-slot limit."""
    from pandemicthrust_amd import Params
    assert sweep.MAX_SLOTS == _header_define("PT_BATCH_MAX_SLOTS")
    assert sweep.SLOT_TILE == _header_define("PT_BATCH_SLOT_TILE")
    runs = []
    for ps, ls in sweep.SCALES:
        for ip in sweep.INITIAL_PANDEMIC:
            for is_ in sweep.INITIAL_SEASONAL:
                p = Params.from_row(["100", "1.8", "1.3", str(ip), str(is_), str(ps), str(ls), "1", "0"])
                runs += [{"params": p, "seed": s} for s in (1, 2, 3)]
    assert len(runs) == 96
    chunks = list(sweep._chunks(runs, 230 * 10 ** 9))
    assert sum(len(c) for c in chunks) == 96
    for c in chunks:
        assert sum(sweep.slots(r["params"]) for r in c) <= sweep.MAX_SLOTS or len(c) == 1
        assert sum(sweep.estimated_people(r["params"]) * sweep.BYTES_PER_AGENT for r in c) <= 230 * 10 ** 9 \
            or len(c) == 1


def test_population_bound_holds():
    """max_people bounds the population setup actually builds."""
    import oracle_ctypes as O
    from pandemicthrust_amd import Params
    for ps, seed in (("0.0001", 1), ("0.001", 2), ("0.0425", 3), ("0.2", 4), ("1", 5)):
        row = ["1", "1.8", "1.3", "1", "1", ps, ps, "1", "0"]
        sim = O.OracleSim(O.params_from_row(row), seed)
        sim.setup()
        assert sim.n_people <= sweep.max_people(Params.from_row(row)), ps
        assert sim.scalar(1) == sweep.households(Params.from_row(row))
