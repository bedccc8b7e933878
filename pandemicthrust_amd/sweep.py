"""Experiment sweeps in the layout of the reference's batcher
(PythonBatcher/write_gpu_constants.py:6-49): one directory per experiment
under gpu_experiment_files/ holding constants.csv, listed in
experiment_files.txt and experiment_names.txt.  The reference runs its
executable once per experiment and seed, and the client collects the files
of PythonBatcher/client_output_files.txt from each run.

Here every (experiment, seed) run of a sweep goes through one batched
handle on one GPU (include/pandemic/pt_batch.h), in chunks that fit the
device, and each run's directory receives the files the executable writes:
output_infected_stats.csv, output_rn.csv and infected_max.txt (byte-identical
to a single run of bin/pandemicthrust on the same constants.csv and
seed.txt), output_resource_log.csv and profile_log.csv (a batch has no
per-run kernel times: the rows carry the batch's times).  mode="process"
runs bin/pandemicthrust once per run instead, as the reference's batcher
does.

    python -m pandemicthrust_amd.sweep write [--output-path DIR]
    python -m pandemicthrust_amd.sweep run DIR/experiment_files.txt --seeds 1,2,3
"""
import argparse
import os
import subprocess
import time

import numpy as np

from . import outputs

OUTPUT_PATH = "./gpu_experiment_files"
CONSTANTS_HEADER = ("constants,max_days,reproduction_number_pandemic,reproduction_number_seasonal,"
                    "initial_infected_pandemic,initial_infected_seasonal,people_scaling_factor,"
                    "location_scaling_factor,percent_symptomatic,asymp_factor")
# The reference batcher's sweep (write_gpu_constants.py:30-32).
SCALES = [[0.1, 0.1], [0.1, 1], [1, 1], [1, 10], [10, 10], [10, 100], [20, 20], [20, 200]]
INITIAL_PANDEMIC = [30, 60]
INITIAL_SEASONAL = [30, 60]
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "pandemicthrust")
# Device bytes per agent a batch needs, with headroom (agent arrays, the
# 3N-pair errand tables and their sort scratch).
BYTES_PER_AGENT = 200
# include/pandemic/pt_batch.h: each run takes its population rounded up to
# SLOT_TILE agent slots, and one batch holds at most MAX_SLOTS slots.
SLOT_TILE = 8192
MAX_SLOTS = 715827882


def write_constants_file(experiment_name, max_days=100, rn_pandemic=1.8, rn_seasonal=1.3, initial_pandemic=30,
                         initial_seasonal=30, people_scale=1, location_scale=1, percent_symptomatic=1,
                         asymp_factor=0, output_path=OUTPUT_PATH):
    """constants.csv of one experiment (write_gpu_constants.py:6-17); the
    values are written with str(), as the batcher does."""
    d = os.path.join(output_path, experiment_name)
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, "constants.csv")
    vals = [max_days, rn_pandemic, rn_seasonal, initial_pandemic, initial_seasonal, people_scale, location_scale,
            percent_symptomatic, asymp_factor]
    with open(path, "w") as f:
        f.write(CONSTANTS_HEADER + "\n")
        f.write("constants," + ",".join(str(v) for v in vals) + "\n")
    return path


def experiment_name(p_scale, l_scale, init_p, init_s):
    return "p" + str(p_scale) + "_l" + str(l_scale) + "_" + str(init_p) + "_" + str(init_s)


def write_experiments(output_path=OUTPUT_PATH, scales=SCALES, init_p=INITIAL_PANDEMIC, init_s=INITIAL_SEASONAL,
                      **constants):
    """The batcher's main (write_gpu_constants.py:21-50): one experiment per
    (scales, initial pandemic, initial seasonal) and the two list files.
    Returns the experiment names."""
    os.makedirs(output_path, exist_ok=True)
    names = []
    with open(os.path.join(output_path, "experiment_files.txt"), "w") as ff, \
            open(os.path.join(output_path, "experiment_names.txt"), "w") as fn:
        for p_scale, l_scale in scales:
            for ip in init_p:
                for is_ in init_s:
                    name = experiment_name(p_scale, l_scale, ip, is_)
                    write_constants_file(name, initial_pandemic=ip, initial_seasonal=is_, people_scale=p_scale,
                                         location_scale=l_scale, output_path=output_path, **constants)
                    ff.write(os.path.join(output_path, name, "constants.csv") + "\n")
                    fn.write(name + "\n")
                    names.append(name)
    return names


def read_experiment_files(path):
    """constants.csv paths listed in experiment_files.txt (relative paths
    are taken from the list's directory, then the working directory)."""
    base = os.path.dirname(os.path.abspath(path))
    out = []
    with open(path) as f:
        for line in f:
            p = line.strip()
            if not p:
                continue
            if not os.path.isabs(p) and not os.path.exists(p):
                cand = os.path.join(base, p)
                if not os.path.exists(cand):
                    cand = os.path.join(base, os.path.relpath(p, os.path.basename(base)))
                p = cand
            out.append(os.path.abspath(p))
    return out


def read_constants(path):
    """The data row of constants.csv after its label column, as the
    reference parses it (PandemicSim.cu:268-323)."""
    from .api import Params
    with open(path) as f:
        f.readline()
        fields = f.readline().strip().split(",")
    if len(fields) < 10:
        raise ValueError("%s: malformed data row" % path)
    return Params.from_row(fields[1:10])


def households(params):
    """setup_scaleSimulation's household count (PandemicSim.cu:1133)."""
    return int(np.floor(np.float64(np.float32(params.people_scaling)) * 1e6 + 0.5))


def estimated_people(params):
    """About 2.4 agents per household (the reference's household tables)."""
    return int(2.4 * households(params)) + 16


def max_people(params):
    """An upper bound on the population: the household tables' mean of 2.354
    agents plus 8 standard deviations of a sum of H household sizes (each
    1..5, standard deviation below 1.2), and never above 5 per household."""
    H = households(params)
    return int(min(5 * H, 2.354 * H + 8 * 1.2 * np.sqrt(H) + 16))


def slots(params):
    """Agent slots a run takes in a batch (pt_batch.h PT_BATCH_SLOT_TILE)."""
    return -(-max_people(params) // SLOT_TILE) * SLOT_TILE


def _chunks(runs, budget):
    """Consecutive runs grouped so that each group fits the memory budget
    and the batch engine's slot limit (a run larger than either alone forms
    a group of its own)."""
    chunk, used, used_slots = [], 0, 0
    for r in runs:
        need = estimated_people(r["params"]) * BYTES_PER_AGENT
        sl = slots(r["params"])
        if chunk and (used + need > budget or used_slots + sl > MAX_SLOTS):
            yield chunk
            chunk, used, used_slots = [], 0, 0
        chunk.append(r)
        used += need
        used_slots += sl
    if chunk:
        yield chunk


def _run_dir(constants_path, seed, out_root):
    exp_dir = os.path.dirname(constants_path)
    name = os.path.basename(exp_dir)
    base = os.path.join(out_root, name) if out_root else exp_dir
    return os.path.join(base, "seed_%d" % seed)


def _prepare(constants_path, seed, out_root):
    d = _run_dir(constants_path, seed, out_root)
    os.makedirs(d, exist_ok=True)
    with open(constants_path) as f:
        text = f.read()
    with open(os.path.join(d, "constants.csv"), "w") as f:
        f.write(text)
    with open(os.path.join(d, "seed.txt"), "w") as f:
        f.write("%d\n" % seed)
    return d


def _append_resource_row(d, sim_type, device, params, seed, ms, used):
    path = os.path.join(d, "output_resource_log.csv")
    exists = os.path.exists(path)
    with open(path, "a") as f:
        if not exists:
            f.write("sim_type,sim_device,people_sim_scale,location_sim_scale,seed,runtime_milliseconds,"
                    "runtime_seconds,bytes_used,megabytes_used\n")
        ms32 = np.float32(ms)
        f.write("%s,%s,%s,%s,%d,%s,%s,%d,%d\n" % (
            sim_type, device.replace(",", " "), outputs.c_float(params.people_scaling),
            outputs.c_float(params.location_scaling), seed, outputs.c_float(ms32),
            outputs.c_float(np.float32(ms32 / np.float32(1000))), used, used >> 20))


def _profile(d, n_people, setup_ms, run_ms, final_ms):
    with open(os.path.join(d, "profile_log.csv"), "w") as f:
        f.write("day,stack_depth,function_name,problem_size,inclusive_time_milliseconds,"
                "exclusive_time_milliseconds\n")
        f.write("%d, %d, %s, %d, %f, %f\n" % (-1, 0, "setupSim", n_people, setup_ms, setup_ms))
        f.write("%d, %d, %s, %d, %f, %f\n" % (-1, 1, "final_countReproduction", n_people, final_ms, final_ms))
        f.write("%d, %d, %s, %d, %f, %f\n" % (-1, 0, "runToCompletion", n_people, run_ms + final_ms, run_ms))


def _run_batch(chunk, device):
    from .api import PandemicBatch, device_memory, device_name
    free0, _ = device_memory(device)
    t0 = time.perf_counter()
    b = PandemicBatch([r["params"] for r in chunk], [r["seed"] for r in chunk], device=device)
    try:
        b.setup()
        t1 = time.perf_counter()
        free1, _ = device_memory(device)
        results = b.run()
        t2 = time.perf_counter()
        setup_ms, run_ms = (t1 - t0) * 1e3, (t2 - t1) * 1e3
        dev = device_name(device)
        for i, (r, (counts, inf), n) in enumerate(zip(chunk, results, b.n_people)):
            t3 = time.perf_counter()
            cp, cs = b.generation_offsets(i)
            md = r["params"].max_days
            outputs.write_run(r["dir"], counts, inf, n, cp, cs, md)
            final_ms = (time.perf_counter() - t3) * 1e3
            _profile(r["dir"], n, setup_ms, run_ms, final_ms)
            _append_resource_row(r["dir"], "gfx950_batch", dev, r["params"], r["seed"], setup_ms + run_ms,
                                 max(0, free0 - free1))
    finally:
        b.close()


def run_sweep(constants_paths, seeds, out_root=None, device=0, mode="batch", memory_budget=None):
    """Runs every (experiment, seed).  Each run's directory is
    <experiment dir>/seed_<seed> (or <out_root>/<experiment>/seed_<seed>)
    and gets a copy of constants.csv, seed.txt and the output files.
    Returns the run directories in (experiment, seed) order."""
    runs = []
    for path in constants_paths:
        params = read_constants(path)
        for s in seeds:
            runs.append({"params": params, "seed": int(s), "dir": _prepare(path, int(s), out_root)})
    if mode == "process":
        for r in runs:
            subprocess.run([EXE], cwd=r["dir"], check=True, stdout=subprocess.DEVNULL)
        return [r["dir"] for r in runs]
    if mode != "batch":
        raise ValueError("mode must be 'batch' or 'process'")
    if memory_budget is None:
        from .api import device_memory
        memory_budget = int(device_memory(device)[0] * 0.8)
    for chunk in _chunks(runs, memory_budget):
        _run_batch(chunk, device)
    return [r["dir"] for r in runs]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pandemicthrust_amd.sweep")
    sub = ap.add_subparsers(dest="cmd", required=True)
    w = sub.add_parser("write", help="write the batcher's experiment directories")
    w.add_argument("--output-path", default=OUTPUT_PATH)
    r = sub.add_parser("run", help="run every experiment of an experiment_files.txt for each seed")
    r.add_argument("experiment_files")
    r.add_argument("--seeds", default="1", help="comma-separated core seeds")
    r.add_argument("--out-root", default=None)
    r.add_argument("--device", type=int, default=0)
    r.add_argument("--mode", choices=["batch", "process"], default="batch")
    a = ap.parse_args(argv)
    if a.cmd == "write":
        for name in write_experiments(a.output_path):
            print(name)
        return 0
    seeds = [int(s) for s in a.seeds.split(",") if s]
    t0 = time.perf_counter()
    dirs = run_sweep(read_experiment_files(a.experiment_files), seeds, a.out_root, a.device, a.mode)
    print("%d runs in %.2f s" % (len(dirs), time.perf_counter() - t0))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
