"""Launch-shape and table-strategy autotuner for the daily loop and the
replicate batch.

The engines expose their launch shapes as settings (pt_launch_knob in
include/pandemic/pt_abi.h): blocks per compute unit of the contact kernels
and of the errand-key kernels, the widest digit of a radix pass, the digit
width from which radix passes run 1024-thread blocks, and the two thresholds of the errand-table strategy (up
to which infected count a day marks its queried buckets, and below which
share of marked buckets it sorts only those).  They change only speed —
every kernel strides over its work, the table sort is stable at any tile
size and digit width, and a table of the queried buckets answers every query of the day —
so the tuner may pick them by measurement alone.  It runs the workload once per candidate
from the same snapshot (coordinate descent: one setting at a time, the
others at their best so far) and keeps the fastest by the engine's own
loop timer (HIP events around the enqueued days).

    python -m pandemicthrust_amd.autotune --rows 64 --out tuning.json
"""
import argparse
import json

CANDIDATES = {
    "contact_blocks_per_cu": [2, 4, 8, 16, 32],
    "gen_blocks_per_cu": [4, 8, 16, 32],
    "sort_wide_bits": [6, 7, 8, 9],
    "subset_share_pct": [60, 75, 90],
    "mark_limit_pct": [20, 50, 100, 200],
    "sort_max_bits": [8, 9, 10],
}
SIM_KNOBS = ["contact_blocks_per_cu", "gen_blocks_per_cu", "sort_max_bits", "sort_wide_bits", "subset_share_pct",
             "mark_limit_pct"]
BATCH_KNOBS = ["contact_blocks_per_cu", "sort_max_bits", "sort_wide_bits", "subset_share_pct", "mark_limit_pct"]


# A candidate replaces the current value only when it is faster by more
# than this share (run-to-run noise of a loop time is a few tenths of a %).
MIN_GAIN = 0.005


def tune(handle, run, knobs, repeats=2):
    """handle: a set-up PandemicSim or PandemicBatch with a snapshot taken;
    run(): restores the snapshot, runs the loop and returns its time in ms.
    Returns (best settings, [(setting, value, ms), ...])."""
    best = dict(handle.launch())
    log = []
    for name in knobs:
        current = best[name]
        times = {}
        for v in CANDIDATES[name]:
            handle.set_launch(name, v)
            times[v] = min(run() for _ in range(repeats))
            log.append((name, v, round(times[v], 3)))
        if current not in times:
            handle.set_launch(name, current)
            times[current] = min(run() for _ in range(repeats))
        fastest = min(times, key=times.get)
        best[name] = fastest if times[fastest] < times[current] * (1.0 - MIN_GAIN) else current
        handle.set_launch(name, best[name])
    return best, log


def tune_sim(sim, repeats=1):
    """The daily loop of a set-up, seeded PandemicSim (snapshot taken here)."""
    sim.snapshot()

    def run():
        sim.restore()
        sim.run()
        return sim.last_run_ms()

    run()  # first-run costs (code objects, caches) out of the measurement
    return tune(sim, run, SIM_KNOBS, repeats)


def tune_batch(batch, repeats=1):
    """A set-up, seeded PandemicBatch (snapshot taken here)."""
    batch.snapshot()

    def run():
        batch.restore()
        batch.run()
        return batch.last_run_ms()

    run()
    return tune(batch, run, BATCH_KNOBS, repeats)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pandemicthrust_amd.autotune")
    ap.add_argument("--rows", type=int, default=64, help="replicate rows of the 256-scenario set to tune the batch on")
    ap.add_argument("--out", default=None, help="write the chosen settings as JSON here")
    a = ap.parse_args(argv)
    from .api import PandemicBatch, Params
    from .scenarios import load_replicate_set
    rows = load_replicate_set()[:a.rows]
    b = PandemicBatch([Params.from_row(f) for _, f in rows], [s for s, _ in rows])
    b.setup()
    best, log = tune_batch(b)
    b.close()
    for name, v, ms in log:
        print("%-22s %3d  %9.3f ms" % (name, v, ms))
    print("best", best)
    if a.out:
        with open(a.out, "w") as f:
            json.dump({"batch": best, "log": log}, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
