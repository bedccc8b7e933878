"""Checkpoint/resume, the profiling and resource logs, and the sweep driver
on the GPU.

Resume is exact: a run saved at a day boundary and continued in a fresh
handle (or a fresh process) ends with the same daily rows, agent arrays and
output files as the uninterrupted run.  Log formats follow the reference's
fprintf lines: profile_log.csv (profiler.cu:13, :72-76),
output_resource_log.csv (resource_logging.cu:105, :112-113) and
output_mem.csv (resource_logging.cu:48, :53, :78-79).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from pandemicthrust_amd import Interventions, Params, PandemicSim, PTError, sweep

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pandemicthrust_amd", "bin", "pandemicthrust")
ROW = ["40", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"]
FILES = ["output_infected_stats.csv", "output_rn.csv", "infected_max.txt"]


def _oracle_full(row, seed, iv=None):
    """The same run on the CPU oracle."""
    import oracle_ctypes as O
    r = O.OracleSim(O.params_from_row(row), seed)
    if iv is not None:
        r.set_interventions(iv)
    r.setup()
    counts, inf = r.run()
    return counts, inf, [r.array(k) for k in range(6)], r.generation_offsets()


def _run_oracle_driver(run_dir, tmp):
    """The oracle's driver on a copy of run_dir's constants.csv and seed.txt;
    returns the directory holding its output files."""
    import shutil
    os.makedirs(tmp, exist_ok=True)
    for name in ("constants.csv", "seed.txt"):
        shutil.copy(os.path.join(run_dir, name), os.path.join(tmp, name))
    r = subprocess.run([os.path.join(ROOT, "oracle", "build", "pt_oracle")], cwd=tmp, stdout=subprocess.DEVNULL,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return tmp


def _full(row, seed, iv=None):
    s = PandemicSim(Params.from_row(row), seed)
    if iv is not None:
        s.set_interventions(iv)
    s.setup_sim()
    counts, inf = s.run()
    arrays = [s.array(k) for k in range(6)]
    gens = s.generation_offsets()
    s.close()
    return counts, inf, arrays, gens


@pytest.mark.parametrize("stop", [0, 1, 6, 12])
def test_resume_matches_uninterrupted_run(stop):
    counts, inf, arrays, gens = _full(ROW, 3)
    oc, oi, oa, og = _oracle_full(ROW, 3)
    # the uninterrupted run is the oracle's; the resumed run must equal both
    assert np.array_equal(counts, oc) and np.array_equal(inf, oi)
    assert all(np.array_equal(a, b) for a, b in zip(arrays, oa))
    assert np.array_equal(gens[0], og[0]) and np.array_equal(gens[1], og[1])
    a = PandemicSim(Params.from_row(ROW), 3)
    a.setup_sim()
    for _ in range(stop):
        a.step()
    blob = a.save_checkpoint()
    a.close()
    b = PandemicSim(Params.from_row(ROW), 3)
    b.setup_only()
    b.load_checkpoint(blob)
    assert int(b.scalars()[6]) == stop
    tail_counts, tail_inf = b.run()
    rows_c, rows_i = b.daily_rows()
    assert np.array_equal(rows_c, counts) and np.array_equal(rows_i, inf)
    assert np.array_equal(tail_counts, counts[stop:]) and np.array_equal(tail_inf, inf[stop:])
    for k in range(6):
        assert np.array_equal(b.array(k), arrays[k]), k
    cp, cs = b.generation_offsets()
    assert np.array_equal(cp, gens[0]) and np.array_equal(cs, gens[1])
    b.close()


def test_resume_with_interventions_and_after_the_end():
    iv = Interventions(school_close_start=3, school_close_days=10, workplace_reduction=0.3, vaccination=(0.1, 0.0))
    counts, inf, arrays, _ = _full(ROW, 5, iv)
    oc, oi, oa, _ = _oracle_full(ROW, 5, iv)
    assert np.array_equal(counts, oc) and np.array_equal(inf, oi)
    assert all(np.array_equal(a, b) for a, b in zip(arrays, oa))
    a = PandemicSim(Params.from_row(ROW), 5)
    a.set_interventions(iv)
    a.setup_sim()
    for _ in range(9):
        a.step()
    mid = a.save_checkpoint()
    a.run()
    end = a.save_checkpoint()
    a.close()
    b = PandemicSim(Params.from_row(ROW), 5)
    b.set_interventions(iv)
    b.setup_only()
    b.load_checkpoint(mid)
    b.run()
    assert np.array_equal(b.daily_rows()[0], counts) and np.array_equal(b.daily_rows()[1], inf)
    b.load_checkpoint(end)                       # a finished run: nothing left to do
    c2, i2 = b.run()
    assert len(c2) == 0 and np.array_equal(b.daily_rows()[1], inf)
    for k in range(6):
        assert np.array_equal(b.array(k), arrays[k]), k
    # the same inputs without the interventions are another model
    c = PandemicSim(Params.from_row(ROW), 5)
    c.setup_only()
    with pytest.raises(PTError, match="other constants"):
        c.load_checkpoint(mid)
    c.close()
    b.close()


def test_checkpoint_rejects_other_inputs_and_corruption():
    a = PandemicSim(Params.from_row(ROW), 3)
    a.setup_sim()
    a.step()
    blob = bytearray(a.save_checkpoint())
    a.close()
    # the seed also draws the population: another seed is another population size
    other_seed = PandemicSim(Params.from_row(ROW), 4)
    other_seed.setup_only()
    with pytest.raises(PTError, match="agents"):
        other_seed.load_checkpoint(bytes(blob))
    other_seed.close()
    other_r = PandemicSim(Params.from_row([ROW[0], "1.9"] + ROW[2:]), 3)   # same population, other R
    other_r.setup_only()
    with pytest.raises(PTError, match="other constants"):
        other_r.load_checkpoint(bytes(blob))
    other_r.close()
    b = PandemicSim(Params.from_row(ROW), 3)
    b.setup_only()
    bad = bytearray(blob)
    bad[-5] ^= 1
    with pytest.raises(PTError, match="checksum"):
        b.load_checkpoint(bytes(bad))
    with pytest.raises(PTError, match="truncated"):
        b.load_checkpoint(bytes(blob[:-1]))
    bad = bytearray(blob)
    bad[0] = ord("X")
    with pytest.raises(PTError, match="magic"):
        b.load_checkpoint(bytes(bad))
    bigger = PandemicSim(Params.from_row(["40", "1.8", "1.3", "10", "10", "0.05", "0.0425", "1", "0"]), 3)
    bigger.setup_only()
    with pytest.raises(PTError, match="agents"):
        bigger.load_checkpoint(bytes(blob))
    bigger.close()
    b.load_checkpoint(bytes(blob))               # the intact blob still loads
    b.close()


def _inputs(d, row, seed):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "constants.csv"), "w") as f:
        f.write(sweep.CONSTANTS_HEADER + "\nconstants," + ",".join(row) + "\n")
    with open(os.path.join(d, "seed.txt"), "w") as f:
        f.write("%d\n" % seed)


def _exe(d, *args):
    r = subprocess.run([EXE] + list(args), cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_host_program_checkpoint_resume(tmp_path):
    full, part = str(tmp_path / "full"), str(tmp_path / "part")
    _inputs(full, ROW, 2)
    _inputs(part, ROW, 2)
    _exe(full)
    out = _exe(part, "--checkpoint", "state.bin", "--checkpoint-day", "8")
    assert "checkpoint written" in out and not os.path.exists(os.path.join(part, "output_rn.csv"))
    out = _exe(part, "--resume", "state.bin")
    assert "resumed from state.bin at day 8" in out
    o = _run_oracle_driver(full, str(tmp_path / "oracle"))
    for name in FILES:
        text = open(os.path.join(o, name)).read()
        assert open(os.path.join(full, name)).read() == text, name
        assert open(os.path.join(part, name)).read() == text, name


ROW_RE = re.compile(r"^(-?\d+), (\d+), ([A-Za-z_]+), (-?\d+), (-?\d+\.\d{6}), (-?\d+\.\d{6})$")


def _infected_per_day(d):
    """I of each day: the problem size of daily_buildInfectedArray_global."""
    out = {}
    for l in open(os.path.join(d, "profile_log.csv")).read().splitlines()[1:]:
        f = [x.strip() for x in l.split(",")]
        if f[2] == "daily_buildInfectedArray_global":
            out[int(f[0])] = int(f[3])
    return [out[k] for k in sorted(out)]


def test_profile_log(tmp_path):
    d = str(tmp_path / "run")
    _inputs(d, ROW, 1)
    _exe(d)
    lines = open(os.path.join(d, "profile_log.csv")).read().splitlines()
    assert lines[0] == ("day,stack_depth,function_name,problem_size,inclusive_time_milliseconds,"
                        "exclusive_time_milliseconds")
    rows = []
    for l in lines[1:]:
        m = ROW_RE.match(l)
        assert m, l
        rows.append((int(m[1]), int(m[2]), m[3], int(m[4]), float(m[5]), float(m[6])))
    names = [r[2] for r in rows]
    assert names[:3] == ["setup_loadParameters", "setup_scaleSimulation", "setup_pushDeviceData"]
    assert names[3:9] == ["setup_calcPopulationSize_thrust", "setup_sizeGlobalArrays", "setup_generateHouseholds",
                          "setup_assignWorkplaces", "setup_initialInfected", "setupSim"]
    assert names[-4:] == ["final_countReproduction_genSearch", "final_countReproduction_hostCode",
                          "final_countReproduction", "runToCompletion"]
    assert rows[-1][1] == 0 and rows[8][1] == 0
    inf = _infected_per_day(d)
    n_days = len(open(os.path.join(d, "output_infected_stats.csv")).read().splitlines()) - 1
    assert len(inf) == n_days
    day_rows = [r for r in rows if 0 <= r[0] < 255 and r[2] != "setup_initialInfected"]
    assert sorted({r[0] for r in day_rows}) == list(range(n_days))
    for day in range(n_days):
        mine = [r[2] for r in day_rows if r[0] == day]
        head = ["daily_writeInfectedStats", "daily_countAndRecover", "daily_buildInfectedArray_global"]
        if day == n_days - 1 and inf[day] == 0:
            assert mine == head
        elif day % 7 >= 5:
            assert mine == head + ["weekend_assignErrands", "doWeekend_wholeDay_setup_sort",
                                   "doWeekend_wholeDay_setup_locationSearch", "doWeekend_wholeDay_kernel",
                                   "doWeekend_wholeDay"], day
        else:
            assert mine == head + ["weekday_generateAfterschoolAndErrandDestinations",
                                   "doWeekday_wholeDay_generateSchedulesKernel", "doWeekday_wholeDay_setup_sort",
                                   "doWeekday_wholeDay_setup_locationSearch", "doWeekday_wholeDay_kernel",
                                   "doWeekday_wholeDay"], day
    for r in rows:
        assert r[4] >= 0 and r[5] <= r[4] + 1e-5, r
    # a parent's inclusive time covers its children (profiler.cu:65-69)
    by = {(r[0], r[2]): r for r in rows}
    for day in range(n_days - 1):
        kind = "Weekend" if day % 7 >= 5 else "Weekday"
        whole = by[(day, "do%s_wholeDay" % kind)]
        kids = [r for r in day_rows if r[0] == day and r[1] == 2 and r[2] != "daily_writeInfectedStats"]
        assert abs(whole[4] - sum(k[4] for k in kids) - whole[5]) < 1e-3, day
    setup = rows[8]
    assert abs(setup[4] - sum(r[4] for r in rows[3:8]) - setup[5]) < 1e-3


def test_resource_and_memory_logs(tmp_path):
    d = str(tmp_path / "run")
    _inputs(d, ROW, 1)
    _exe(d, "--poll-memory")
    _exe(d, "--no-profile")
    res = open(os.path.join(d, "output_resource_log.csv")).read().splitlines()
    assert res[0] == ("sim_type,sim_device,people_sim_scale,location_sim_scale,seed,runtime_milliseconds,"
                      "runtime_seconds,bytes_used,megabytes_used")
    assert len(res) == 3                       # appended, one header
    pat = re.compile(r"^gfx950_hip,[^,]+,0\.042500,0\.042500,1,(\d+\.\d{6}),(\d+\.\d{6}),(\d+),(\d+)$")
    for l in res[1:]:
        m = pat.match(l)
        assert m, l
        assert abs(float(m[1]) / 1000 - float(m[2])) < 1e-3
        assert int(m[3]) > 0 and int(m[4]) == int(m[3]) >> 20
    mem = open(os.path.join(d, "output_mem.csv")).read().splitlines()
    assert mem[0] == "day,freeBytes,bytesUsed,totalBytes,megabytesUsed"
    assert re.match(r"^INITIAL,\d+,0,\d+,0$", mem[1])
    inf = _infected_per_day(d)
    days = [int(l.split(",")[0]) for l in mem[2:]]
    # setupSim's sample (day 0), one per day with infected agents, then final_releaseMemory's (day_t)-1
    assert days == [0] + [dd for dd, i in enumerate(inf) if i > 0] + [255]
    for l in mem[2:]:
        _, fr, used, tot, mb = (int(x) for x in l.split(","))
        assert fr + used <= tot and mb == used >> 20


def test_sweep_batch_matches_single_runs(tmp_path):
    out = str(tmp_path / "gpu_experiment_files")
    sweep.write_experiments(out, scales=[[0.03, 0.05], [0.0425, 0.0425]], init_p=[10], init_s=[5, 20], max_days=30)
    paths = sweep.read_experiment_files(os.path.join(out, "experiment_files.txt"))
    assert len(paths) == 4
    batch = sweep.run_sweep(paths, [1, 7], out_root=str(tmp_path / "batch"))
    single = sweep.run_sweep(paths, [1, 7], out_root=str(tmp_path / "single"), mode="process")
    assert len(batch) == len(single) == 8
    for k, (a, b) in enumerate(zip(batch, single)):
        o = _run_oracle_driver(a, str(tmp_path / "oracle" / str(k)))
        for name in FILES:
            text = open(os.path.join(o, name)).read()
            assert open(os.path.join(a, name)).read() == text, (a, name)
            assert open(os.path.join(b, name)).read() == text, (b, name)
        for name in ("output_resource_log.csv", "profile_log.csv"):
            assert os.path.exists(os.path.join(a, name)), name
    # a tiny budget splits the sweep into several batches; results are the same
    again = sweep.run_sweep(paths, [1, 7], out_root=str(tmp_path / "chunked"), memory_budget=1)
    for a, b in zip(again, single):
        for name in FILES:
            assert open(os.path.join(a, name)).read() == open(os.path.join(b, name)).read(), (a, name)
