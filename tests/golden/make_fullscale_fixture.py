"""Writes whole-run fixtures of the CPU oracle for workloads too large to run
the oracle inside a test: the 50M-agent x 100-day workload
(constants,100,1.8,1.3,10,10,21.24,21.24,1,0, seed 1) and the largest rows of
the 256-scenario replicate set (pandemicthrust_amd/data/replicate_scenarios_256.csv,
about 2M agents x 180 days).

Each fixture directory holds what the reference's executable writes
(output_infected_stats.csv, output_rn.csv, infected_max.txt; formats of
PandemicSim.cu:2771-2818, :2282-2295, :863) and state.json: the population
size, the days run, the generation offsets of both strains and a SHA-256 of
each of the six final per-agent arrays (status, day, generation per strain,
uint8).  The GPU tests compare a run of the HIP engine with them byte for
byte.

  python tests/golden/make_fullscale_fixture.py loop
  python tests/golden/make_fullscale_fixture.py replicate ROW [ROW ...]
  python tests/golden/make_fullscale_fixture.py replicate-largest K

Runs single-threaded; the 50M x 100-day run takes several minutes and about
6 GB of memory.
"""
import csv
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_ctypes as O  # noqa: E402

LOOP_ROW = ["100", "1.8", "1.3", "10", "10", "21.24", "21.24", "1", "0"]
SCENARIOS = os.path.join(ROOT, "pandemicthrust_amd", "data", "replicate_scenarios_256.csv")
ARRAY_NAMES = ["status_p", "status_s", "day_p", "day_s", "gen_p", "gen_s"]


def scenario_rows():
    """(row index, constants fields, seed) of the replicate set."""
    out = []
    with open(SCENARIOS) as f:
        for i, r in enumerate(csv.DictReader(f)):
            fields = [r[k] for k in ("max_days", "reproduction_number_pandemic", "reproduction_number_seasonal",
                                     "initial_infected_pandemic", "initial_infected_seasonal",
                                     "people_scaling_factor", "location_scaling_factor", "percent_symptomatic",
                                     "asymp_factor")]
            out.append((i, fields, int(r["seed"])))
    return out


def state_digest(sim, days):
    cp, cs = sim.generation_offsets()
    return {
        "n_people": sim.n_people,
        "days": int(days),
        "gen_offsets_p": [int(x) for x in cp],
        "gen_offsets_s": [int(x) for x in cs],
        "sha256": {name: hashlib.sha256(sim.array(i).tobytes()).hexdigest() for i, name in enumerate(ARRAY_NAMES)},
    }


def write_fixture(name, row, seed):
    out = os.path.join(HERE, name)
    os.makedirs(out, exist_ok=True)
    t0 = time.time()
    sim = O.OracleSim(O.params_from_row(row), seed)
    sim.setup()
    counts, inf = sim.run()
    sim.write_outputs(out, counts, inf)
    st = state_digest(sim, len(inf))
    st["constants"] = ",".join(["constants"] + list(row))
    st["seed"] = seed
    with open(os.path.join(out, "state.json"), "w") as f:
        json.dump(st, f, indent=1)
        f.write("\n")
    # the oracle's own resource and profile logs are not part of the fixture
    for extra in ("output_resource_log.csv", "profile_log.csv"):
        p = os.path.join(out, extra)
        if os.path.exists(p):
            os.remove(p)
    print("%s: %d agents, %d days, %.0f s" % (name, sim.n_people, len(inf), time.time() - t0), flush=True)


def main(argv):
    if not argv:
        print(__doc__)
        return 2
    if argv[0] == "loop":
        write_fixture("loop_50M_100d", LOOP_ROW, 1)
        return 0
    rows = scenario_rows()
    if argv[0] == "replicate-largest":
        k = int(argv[1])
        pick = sorted(rows, key=lambda t: -float(t[1][5]))[:k]
    elif argv[0] == "replicate":
        pick = [rows[int(x)] for x in argv[1:]]
    else:
        print(__doc__)
        return 2
    for i, fields, seed in pick:
        write_fixture("replicate_row%03d" % i, fields, seed)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
