#!/usr/bin/env python3
"""Benchmark of the daily simulation loop on one MI355X.

Workload: 50,000,000-agent population x 100 days, the reference's
constants.csv row `constants,100,1.8,1.3,10,10,21.24,21.24,1,0` with seed 1
(50,002,479 agents).  One step = one whole 100-day loop (runToCompletion) on
the GPU, started from the post-setup state kept in HBM (restored on the device
before each step; setup itself is input preparation and not timed).

value = person-days per second over the timed steps
      = steps * N * days_simulated / wall seconds of those steps.

roofline: the dominant kernel class by HIP-event time (the radix downsweep
of the full errand tables).  Its algorithmic bytes follow the loop model's
table term: an ideal single-pass table build reads the E*N (key, id) pairs
once (8 bytes each) and writes the sorted ids once (4 bytes), 12 bytes per
pair, E = 2 on weekdays and 3 on weekends, counted on the days that build
the full table (from the engine's per-day table log) and attributed to the
downsweep launches; `achieved` = those bytes / the class's launch time over
the timed steps.  `traffic` is the class's measured HBM bytes per launch
from the committed rocprofv3 summary (profiles/daily_loop_50M_100d.md:
FETCH_SIZE + WRITE_SIZE runs of this same command, calibrated on the copy
kernel), and `traffic_over_algorithmic` their ratio.  `per_pass` keeps the
engine's own byte model of the radix passes (every pass's reads and writes).
loop_roofline: the loop model over the whole loop,
B_wd = 43N + 89I + 3X + 8W + 4 and B_we = 62N + 62I + 3X + 40W + 4 per day.

setup_ms: the population setup (households, workplaces, workplace table,
status arrays) on the GPU; whole_program: bin/pandemicthrust run on the same
constants.csv / seed.txt (process start to exit, setup, 100 days, output
files), beside the reference's published whole-program time for 50M x 100
days on a K20 (README.md:3, about 24 s).

replicates: the 256-scenario set of pandemicthrust_amd/data/
config2_scenarios.csv run to the end, setup included, once through the batch
engine and once one simulation at a time; replicates per second both ways.

ensemble: scenario vectors with interventions, each run 8 times on the batch
engine at 1M agents x 120 days (surrogate training data generation).

surrogate: a 5-member MLP ensemble trained with the BF16 MFMA kernels on
simulated scenario ensembles (100k agents); training TFLOP/s against the
MFMA peak, held-out accuracy, and a screening sweep (surrogate inference,
re-simulation of flagged vectors) against simulating every vector, both
plainly (one run at a time, timed on a sample) and on the batch engine.

cpu_baseline: the CPU oracle (a sequential port of the reference semantics)
on one host core, timed on a bounded sample of the same workload: the first
CPU_DAYS days after setup, plus the epidemic's peak day (day PEAK_DAY,
12.4M infected) continued from the GPU's state at that day boundary; the
oracle's result of the peak day is compared with the GPU's (all six agent
arrays).  The rate is person-days over those days / their seconds (setup
excluded); the host's CPU model and core count are recorded.
"""
import argparse
import json
import re
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "person-days/sec at 50M×100d; batched vs plain replicates/sec; HBM GB/s"
ROW = ["100", "1.8", "1.3", "10", "10", "21.24", "21.24", "1", "0"]
SEED = 1
HBM_PEAK_GBS = 8000.0
CPU_DAYS = 4
PEAK_DAY = 95
K20_WHOLE_PROGRAM_S = 24.0   # README.md:3, 50M persons x 100 days on an NVIDIA K20
CONSTANTS_HEADER = ("constants,max_days,reproduction_number_pandemic,reproduction_number_seasonal,"
                    "initial_infected_pandemic,initial_infected_seasonal,people_scaling_factor,"
                    "location_scaling_factor,percent_symptomatic,asymp_factor")


def algorithmic_bytes(N, W, lay, inf, counts, full, pairs):
    """Algorithmic bytes per kernel class for one loop, and the loop total of
    the per-day model B_wd = 43N + 89I + 3X + 8W + 4, B_we = 62N + 62I + 3X +
    40W + 4 (terminal I = 0 day: 2N).

    full[d], pairs[d]: the day's errand table (every bucket: n = 2N or 3N
    keys, pairs[d] of them kept; or the queried buckets only, pairs[d]
    selected pairs) as the engine logged it.  Full-table days account to
    gen/sort/offsets, subset days to select/subset_sort.  Per radix pass:
    the first upsweep reads the keys (4 B each), later ones the (key, id)
    pairs (8 B); the first downsweep reads n keys and writes key and id of
    the kept pairs (4n + 8k), later ones read and write key and id (16k)."""
    per = {}

    def add(k, b):
        per[k] = per.get(k, 0.0) + float(b)

    # the loop model's table term: 12 bytes per pair on full-table days
    add("table_model", 0.0)

    loop = 0.0
    days = len(inf)
    for d in range(days):
        I = int(inf[d])
        add("sweep", 2 * N + 2 * I + 4 * I)          # statuses, days of the infected, staged ids
        add("plan", 0)
        if I == 0:
            loop += 2 * N
            continue
        add("compact", 8 * I)
        weekend = (d % 7) >= 5
        if d + 1 < days:
            X = (int(counts[d][1]) - int(counts[d + 1][1])) + (int(counts[d][9]) - int(counts[d + 1][9]))
        else:
            X = 0
        passes = lay["we_passes"] if weekend else lay["wd_passes"]
        buckets = lay["we_buckets"] if weekend else lay["wd_buckets"]
        if weekend:
            add("contacts_weekend", 56 * I)
            loop += 62 * N + 62 * I + 3 * X + 40 * W + 4
        else:
            add("contacts_weekday", 83 * I)
            loop += 43 * N + 89 * I + 3 * X + 8 * W + 4
        if d < len(full) and full[d]:
            n = 3 * N if weekend else 2 * N
            add("table_model", 12 * n)
            # pairs kept: the weekday table leaves out the never-queried
            # hour-1 bucket (the first pass reads every key, writes the rest)
            k = int(pairs[d]) if d < len(pairs) and int(pairs[d]) > 0 else n
            add("gen_weekend" if weekend else "gen_weekday", 4 * n + (0 if weekend else N))
            for p in range(passes):
                add("sort_upsweep", 8 * k if p else 4 * n)    # later passes read (key, id) pairs
                add("sort_downsweep", (4 * n + 8 * k) if p == 0 else 16 * k)
            add("offsets", 4 * k + 4 * (buckets + 1))
        elif d < len(full):
            m = int(pairs[d])
            add("select", (0 if weekend else N) + 8 * m + 16 * m)   # ages; staged pairs; compaction
            for p in range(passes):
                add("subset_sort", (4 if p == 0 else 8) * m + 16 * m)
            add("subset_sort", 4 * m + 4 * (buckets + 1))
        add("apply", 3 * X + 16 * X)
    return per, loop


PROFILE = os.path.join(ROOT, "profiles", "daily_loop_50M_100d.md")


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    summary (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    try:
        text = open(PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_per_launch %s: (\d+) bytes" % re.escape(kernel), text)
    return int(m.group(1)) if m else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(peak):
    """The oracle on one core: days 0..CPU_DAYS-1 after setup, then the peak
    day from the GPU's state.  peak = (day, rand_offset, arrays before the
    day, arrays after the GPU ran it)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    if not os.path.exists(O.LIB_PATH):
        return None
    t0 = time.perf_counter()
    ref = O.OracleSim(O.params_from_row(ROW), SEED)
    ref.setup()
    setup_s = time.perf_counter() - t0
    N = ref.n_people
    t0 = time.perf_counter()
    for _ in range(CPU_DAYS):
        ref.step()
    early_s = time.perf_counter() - t0
    day, rand_offset, before, after = peak
    ref.set_state(day, rand_offset, before)
    t0 = time.perf_counter()
    rc, _, inf = ref.step()
    peak_s = time.perf_counter() - t0
    identical = rc == 0 and all(np.array_equal(ref.array(k), after[k]) for k in range(6))
    ref.close()
    days = CPU_DAYS + 1
    try:
        cores_visible = len(os.sched_getaffinity(0))
    except AttributeError:
        cores_visible = os.cpu_count()
    return {"value": round(N * days / (early_s + peak_s), 1), "unit": "person-days/s", "cores": 1, "kind": "port",
            "sample": "oracle on one thread: days 0-%d of the 50M x 100d workload (%d agents; setup %.1f s untimed) "
                      "in %.2f s, and day %d (the peak, %d infected) continued from the GPU's state in %.2f s"
                      % (CPU_DAYS - 1, N, setup_s, early_s, day, inf, peak_s),
            "early_days_per_s": round(N * CPU_DAYS / early_s, 1), "peak_day_s": round(peak_s, 3),
            "peak_day_identical_to_gpu": bool(identical),
            "cpu_model": cpu_model(), "host_cores": cores_visible}


def peak_state(sim, day):
    """The GPU run's state at the start of `day` and after it."""
    sim.restore()
    for _ in range(day):
        sim.step()
    sc = sim.scalars()
    before = [sim.array(k) for k in range(6)]
    sim.step()
    after = [sim.array(k) for k in range(6)]
    return int(sc[6]), int(sc[3]), before, after


def whole_program():
    """bin/pandemicthrust on the workload's constants.csv and seed.txt."""
    exe = os.path.join(ROOT, "pandemicthrust_amd", "bin", "pandemicthrust")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "constants.csv"), "w") as f:
            f.write(CONSTANTS_HEADER + "\nconstants," + ",".join(ROW) + "\n")
        with open(os.path.join(d, "seed.txt"), "w") as f:
            f.write("%d\n" % SEED)
        t0 = time.perf_counter()
        subprocess.run([exe], cwd=d, stdout=subprocess.DEVNULL, timeout=600, check=True)
        t = time.perf_counter() - t0
    return {"seconds": round(t, 3), "command": "pandemicthrust_amd/bin/pandemicthrust (constants.csv + seed.txt, "
                                               "profile_log.csv on, as the reference's default build)",
            "reference_k20_s": K20_WHOLE_PROGRAM_S, "reference_source": "README.md:3 (NVIDIA K20)",
            "speedup_vs_reference": round(K20_WHOLE_PROGRAM_S / t, 1)}


def replicate_bench(n_reps, plain=True, tune=True, solo_launch=None):
    """256-scenario ensemble (pandemicthrust_amd/data/config2_scenarios.csv),
    each replicate run to its end (180 days or extinction), setup included:
    plain = one single-simulation handle after another; batched = one batch.
    Every replicate's daily counts must agree between the two.  The batch's
    launch shapes come from the autotuner run on the first 64 rows (untimed);
    plain runs use solo_launch (the daily loop's tuned shapes)."""
    from pandemicthrust_amd import PandemicBatch, PandemicSim, Params, autotune
    from pandemicthrust_amd.scenarios import load_config2
    rows = load_config2()[:n_reps]
    params = [Params.from_row(f) for _, f in rows]
    seeds = [s for s, _ in rows]
    batch_launch, tune_log = None, None
    if tune:
        tb = PandemicBatch(params[:64], seeds[:64])
        tb.setup()
        batch_launch, tune_log = autotune.tune_batch(tb)
        tb.close()
        print("replicates: tuned launch %s" % batch_launch, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = PandemicBatch(params, seeds)
    for k, v in (batch_launch or {}).items():
        b.set_launch(k, v)
    n_people = b.setup()
    res = b.run()
    batched_s = time.perf_counter() - t0
    batched_loop_ms = b.last_run_ms()
    info = b.info()
    b.close()
    print("replicates: batched %d in %.3f s (loop %.1f ms)" % (len(rows), batched_s, batched_loop_ms),
          file=sys.stderr, flush=True)
    person_days = int(sum(int(n_people[r]) * len(res[r][1]) for r in range(len(rows))))
    out = {"workload": "%d scenarios of pandemicthrust_amd/data/config2_scenarios.csv (N 50k-2M log-uniform, "
                       "max_days 180), setup + run to the end" % len(rows),
           "agents": info["agents"], "person_days": person_days,
           "batched_per_s": round(len(rows) / batched_s, 2), "batched_s": round(batched_s, 3),
           "batched_loop_ms": round(batched_loop_ms, 1),
           "batched_person_days_per_s": round(person_days / (batched_loop_ms * 1e-3), 1),
           "launch": {"batch": batch_launch, "plain": solo_launch, "tuning_ms": tune_log}}
    if not plain:
        return out
    identical = True
    t0 = time.perf_counter()
    for r, (p, s) in enumerate(zip(params, seeds)):
        sim = PandemicSim(p, s)
        for k, v in (solo_launch or {}).items():
            sim.set_launch(k, v)
        sim.setup_sim()
        counts, inf = sim.run()
        sim.close()
        identical &= bool(np.array_equal(counts, res[r][0]) and np.array_equal(inf, res[r][1]))
    plain_s = time.perf_counter() - t0
    print("replicates: plain %d in %.3f s" % (len(rows), plain_s), file=sys.stderr, flush=True)
    out.update({"plain_per_s": round(len(rows) / plain_s, 2), "plain_s": round(plain_s, 3),
                "speedup": round(plain_s / batched_s, 2), "identical_results": identical})
    return out


def ensemble_bench(scenarios, replicates, agents, days):
    """Scenario-ensemble generation (surrogate training data): scenario
    vectors with interventions, each run `replicates` times on the batch
    engine, summarised as log-incidence quantiles."""
    from pandemicthrust_amd import ensemble as E
    X = E.sample_scenarios(scenarios, 20261203)
    Y, info = E.run_ensemble(X, replicates, agents, days)
    full = 4096 * 64 * agents * days / max(info["person_days_per_s"], 1.0)
    print("ensemble: %d runs in %.2f s" % (info["runs"], info["seconds"]), file=sys.stderr, flush=True)
    return {"workload": "%d scenarios x %d replicates, %d agents, %d days (16-dim scenario vectors with "
                        "interventions)" % (scenarios, replicates, agents, days),
            "runs_per_s": round(info["runs_per_s"], 2), "person_days_per_s": round(info["person_days_per_s"], 1),
            "seconds": round(info["seconds"], 3),
            "projected_s_4096x64": round(full, 1)}


MFMA_BF16_PEAK_TFLOPS = 2500.0


def curve_stats(logmean):
    """Peak day, peak height and attack (sum) of incidence curves from
    log(1 + incidence) means: ... x days -> three arrays."""
    inc = np.expm1(np.clip(logmean, 0.0, None))
    return inc.argmax(axis=-1), inc.max(axis=-1), inc.sum(axis=-1)


def surrogate_bench(n_train, n_screen):
    """Surrogate training on simulated scenario ensembles, held-out accuracy,
    and surrogate screening of a scenario sweep with re-simulation of the
    flagged vectors."""
    from pandemicthrust_amd import ensemble as E
    from pandemicthrust_amd import surrogate as S
    agents, days, reps, test_reps = 100000, 120, 8, 32
    Xtr, Xte = E.sample_scenarios(n_train, 101), E.sample_scenarios(128, 202)
    Ytr, itr = E.run_ensemble(Xtr, reps, agents, days, seed0=1)
    # Held-out targets: 32-replicate means, so that their own sampling noise
    # (8 replicates: peak day +-2.3 days, attack 1%) does not dominate.
    Yte, _ = E.run_ensemble(Xte, test_reps, agents, days, seed0=10 ** 6)
    sur = S.Surrogate(members=5, batch=1024, seed=1)
    sur.set_data(S.normalize(Xtr), S.targets(Ytr))
    steps, train_ms = 0, 0.0
    for n, lr in ((2000, 1e-3), (2000, 3e-4), (2000, 1e-4)):
        sur.train(n, lr=lr, seed=n)
        steps += n
        train_ms += sur.last_train_ms()
    tflops = S.training_flops(1024, 5) * steps / (train_ms * 1e-3) / 1e12
    P = sur.predict(S.normalize(Xte)).mean(axis=0).reshape(len(Xte), 2, days, 3)
    rmse = float(np.sqrt(((P - Yte) ** 2).mean()))
    pd, ph, pa = curve_stats(P[:, 0, :, 0])
    sd, sh, sa = curve_stats(Yte[:, 0, :, 0])
    acc = {"heldout_scenarios": len(Xte), "heldout_replicates": test_reps,
           "heldout_rmse_log_incidence": round(rmse, 4), "target_std": round(float(Yte.std()), 3),
           "peak_day_mae": round(float(np.abs(pd - sd).mean()), 2),
           "peak_day_rel_err": round(float(np.median(np.abs(pd - sd) / np.maximum(sd, 1))), 4),
           "peak_day_rel_err_mean": round(float(np.mean(np.abs(pd - sd) / np.maximum(sd, 1))), 4),
           "peak_height_rel_err": round(float(np.median(np.abs(ph - sh) / np.maximum(sh, 1.0))), 4),
           "attack_rel_err": round(float(np.median(np.abs(pa - sa) / np.maximum(sa, 1.0))), 4),
           "attack_rel_err_mean": round(float(np.mean(np.abs(pa - sa) / np.maximum(sa, 1.0))), 4)}
    # Screening: surrogate inference on a sweep, re-simulation of the flagged rows.
    Xs = E.sample_scenarios(n_screen, 303)
    t0 = time.perf_counter()
    sc = S.screen(sur, Xs, S.normalize(Xtr))
    infer_s = time.perf_counter() - t0
    flagged = np.nonzero(sc["resimulate"])[0]
    chk = flagged[:64]
    Yf, info = E.run_ensemble(Xs[chk], reps, agents, days, seed0=2 * 10 ** 6)
    resim_s = info["seconds"] * len(flagged) / max(len(chk), 1)
    pf = sc["mean"][chk].reshape(len(chk), 2, days, 3)
    fd, fh, fa = curve_stats(pf[:, 0, :, 0])
    rd, rh, ra = curve_stats(Yf[:, 0, :, 0])
    batched_s = n_screen * reps / itr["runs_per_s"]
    # Plain simulation, one run at a time as in the replicate leg: a sample of
    # the screened scenarios, every replicate its own handle (create, set up,
    # run to the end).
    from pandemicthrust_amd import PandemicSim
    n_plain = min(8, n_screen)
    t0 = time.perf_counter()
    for i in range(n_plain):
        p, iv = E.scenario_inputs(Xs[i], agents, days)
        for r in range(reps):
            sim = PandemicSim(p, 3 * 10 ** 6 + i * reps + r)
            sim.set_interventions(iv)
            sim.setup_sim()
            sim.run()
            sim.close()
    plain_s = n_screen * (time.perf_counter() - t0) / n_plain
    sur.close()
    print("surrogate: %d steps in %.2f s (%.1f TFLOP/s), held-out rmse %.3f" % (steps, train_ms / 1e3, tflops, rmse),
          file=sys.stderr, flush=True)
    return {"workload": "5-member MLP 16-512x4-720 (GELU), BF16 MFMA / FP32 accumulate, batch 1024; trained on "
                        "%d scenarios x %d replicates (%d agents, %d days, log-incidence mean/q10/q90)" % (
                            n_train, reps, agents, days),
            "train_steps": steps, "train_ms_per_step": round(train_ms / steps, 4),
            "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5)},
            "accuracy": acc,
            "screening": {"vectors": n_screen, "inference_s": round(infer_s, 3),
                          "inference_per_s": round(n_screen / infer_s, 1),
                          "flagged": int(len(flagged)), "flagged_disagreement": int((sc["spread"] > 0.5).sum()),
                          "flagged_outside_hull": int(sc["outside"].sum()), "flagged_top": int(sc["top"].sum()),
                          "resimulated_checked": int(len(chk)),
                          "flagged_peak_day_mae": round(float(np.abs(fd - rd).mean()), 2) if len(chk) else None,
                          "flagged_attack_rel_err": round(float(np.median(np.abs(fa - ra) / np.maximum(ra, 1.0))), 4)
                          if len(chk) else None,
                          "end_to_end_per_s": round(n_screen / (infer_s + resim_s), 1),
                          "plain_simulation_per_s": round(n_screen / plain_s, 1),
                          "plain_sample": "%d scenarios x %d replicates, one run at a time" % (n_plain, reps),
                          "batched_simulation_per_s": round(n_screen / batched_s, 1),
                          "speedup_vs_plain": round(plain_s / (infer_s + resim_s), 1),
                          "speedup_vs_batched": round(batched_s / (infer_s + resim_s), 1)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--replicates", type=int, default=256, help="scenarios in the replicate leg (0 skips it)")
    ap.add_argument("--ensemble", type=int, default=32, help="scenarios in the ensemble leg (0 skips it)")
    ap.add_argument("--surrogate", type=int, default=8192,
                    help="training scenarios of the surrogate leg (0 skips it)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched replicate leg (for profiling the batch engine)")
    ap.add_argument("--kernels", action="store_true", help="print the per-kernel table to stderr")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the daily loop of one simulation runs on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build); use --gpus 1", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0
    if args.batched_only:
        print(json.dumps({"replicates": replicate_bench(args.replicates, plain=False, tune=not args.no_autotune)}),
              flush=True)
        return 0

    from pandemicthrust_amd import Params, PandemicSim
    sim = PandemicSim(Params.from_row(ROW), SEED)
    t0 = time.perf_counter()
    N = sim.setup_sim()
    setup_s = time.perf_counter() - t0
    setup_ms = sim.setup_times()
    sim.snapshot()
    lay = sim.layout()
    W = int(sim.scalars()[2])
    print("setup: %d agents in %.2f s" % (N, setup_s), file=sys.stderr, flush=True)
    launch, tune_log = sim.launch(), None
    if not args.no_autotune:
        from pandemicthrust_amd import autotune
        launch, tune_log = autotune.tune_sim(sim)
        print("tuned launch %s" % launch, file=sys.stderr, flush=True)

    # The first warmup step times every kernel class (the per-kernel table
    # and the choice of the dominant class); the timed steps carry events
    # only on the dominant class.
    profile = None
    for i in range(args.warmup):
        sim.enable_timing(i == 0)
        sim.restore()
        sim.run()
        if i == 0:
            profile = sim.kernel_times()
        print("warmup %d done" % i, file=sys.stderr, flush=True)
    if profile is not None:
        cand = {k: v for k, v in profile.items() if v[1] and k != "setup"}
        sim.enable_timing(True, only=max(cand, key=lambda k: cand[k][0]))
    else:
        sim.enable_timing(True)
    sim.synchronize()
    days_total = 0
    loop_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        sim.restore()
        counts, inf = sim.run()
        days_total += len(inf)
        loop_ms.append(sim.last_run_ms())
    sim.synchronize()
    wall = time.perf_counter() - t0
    kt = sim.kernel_times()
    sim.enable_timing(False)
    print("steps: loop ms %s" % ["%.1f" % x for x in loop_ms], file=sys.stderr, flush=True)
    print("infected per day: %s" % " ".join(str(int(x)) for x in inf[::5]), file=sys.stderr, flush=True)

    days_per_step = days_total / max(args.steps, 1)
    value = N * days_total / wall
    full, pairs = sim.table_log()
    per_loop, loop_bytes = algorithmic_bytes(N, W, lay, inf, counts, full, pairs)

    # Per-kernel table from the profiled warmup loop (one loop's worth).
    table = {}
    for name, (ms, launches) in (profile or kt).items():
        if launches == 0 or name == "setup":
            continue
        b = per_loop.get(name, 0.0) * (1 if profile else args.steps)
        table[name] = {"ms": round(ms, 3), "launches": launches, "avg_us": round(1000 * ms / launches, 2),
                       "GBps": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    timed = {k: v for k, v in kt.items() if v[1] and k != "setup"}
    dom = max(timed, key=lambda k: timed[k][0])
    dom_ms, dom_launches = kt[dom]
    # Algorithmic bytes of the dominant class: the loop model's table term
    # when it is a table-build class, else the class's own byte model.
    table_classes = ("sort_downsweep", "sort_upsweep")
    model_bytes = per_loop["table_model"] if dom in table_classes else per_loop.get(dom, 0.0)
    dom_bytes = model_bytes * args.steps
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    pass_bytes = per_loop.get(dom, 0.0) * args.steps
    traffic = measured_traffic(dom)
    if args.kernels:
        for k, v in sorted(table.items(), key=lambda kv: -kv[1]["ms"]):
            print("%-18s %s" % (k, v), file=sys.stderr)

    loop_s = sum(loop_ms) / 1000.0
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "person-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * wall / max(args.steps, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/i32 state, f32 draws",
        "data": "synthetic (the reference's own generator: households, workplaces, errands from Threefry draws)",
        "config": {"workload": "50M agents x 100 days: constants,100,1.8,1.3,10,10,21.24,21.24,1,0 seed 1",
                   "n_people": N, "days_per_step": days_per_step, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": os.path.relpath(PROFILE, ROOT),
                     "avg_launch_us": round(1000 * dom_ms / dom_launches, 2),
                     "algorithmic_bytes_per_launch": round(dom_bytes / dom_launches, 1),
                     "algorithmic_model": "12 B per pair (read pairs 8 + write ids 4) x E*N pairs on full-table days"
                     if dom in table_classes else "class byte model",
                     "traffic_over_algorithmic": round(traffic / (dom_bytes / dom_launches), 2)
                     if traffic and dom_bytes else None,
                     "per_pass": {"bytes_per_launch": round(pass_bytes / dom_launches, 1),
                                  "achieved": round(pass_bytes / (dom_ms * 1e-3) / 1e9, 1),
                                  "frac": round(pass_bytes / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "loop_roofline": {"algorithmic_GB_per_loop": round(loop_bytes / 1e9, 3),
                          "achieved": round(loop_bytes * args.steps / loop_s / 1e9, 1),
                          "frac": round(loop_bytes * args.steps / loop_s / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": table,
        "launch": {"settings": launch, "tuning_ms": tune_log},
        "setup_ms": {"total_wall": round(1000 * setup_s, 2),
                     **{k: round(v, 3) for k, v in setup_ms.items()}},
    }
    peak = None if args.no_cpu_baseline else peak_state(sim, PEAK_DAY)
    sim.close()
    line["whole_program"] = whole_program()
    if args.replicates > 0:
        line["replicates"] = replicate_bench(args.replicates, tune=not args.no_autotune, solo_launch=launch)
    if args.ensemble > 0:
        line["ensemble"] = ensemble_bench(args.ensemble, 8, 1000000, 120)
    if args.surrogate > 0:
        line["surrogate"] = surrogate_bench(args.surrogate, 10000)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(peak)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
