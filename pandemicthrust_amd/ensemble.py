"""Scenario ensembles on the replicate-batch engine (surrogate training data).

A scenario is a 16-dimensional vector (SCENARIO_DIMS); each runs as several
replicates (distinct seeds) of one population size in one PandemicBatch.  The
result per scenario is the daily incidence of each strain, incidence_d =
S_d - S_{d+1} (the susceptible column of output_infected_stats.csv,
PandemicSim.cu:2806-2818), summarised over replicates as the mean and the
10% / 90% quantiles of log(1 + incidence).

Scenario dimensions map onto the reference's constants row (transmission,
initial infected, symptomatic share, asymptomatic factor) and onto the
engine's interventions (pt_interventions).  The reference calibrates with R
rather than a per-contact p, so p maps to R = 100 p (p in [0.010, 0.025] ->
R in [1.0, 2.5]).  Scenario vectors are drawn from one std::mt19937_64 stream
(scenarios.MT19937_64), u = (x >> 11) 2^-53 per dimension in order.

    python -m pandemicthrust_amd.ensemble --scenarios 64 --replicates 8 --agents 100000 --days 120 --out ens.npz

The surrogate's training set (4,096 scenarios x 64 replicates at 1,000,000
agents x 120 days, vectors from seed TRAINING_SET_SEED) is generated in parts and
merged into pandemicthrust_amd/data/surrogate_ensemble.npz:

    python -m pandemicthrust_amd.ensemble --scenarios 4096 --replicates 64 --agents 1000000 --days 120 \
        --seed 20261204 --range 0:2048 --out part0.npz
    python -m pandemicthrust_amd.ensemble --merge part0.npz part1.npz --out surrogate_ensemble.npz
"""
import argparse
import math
import time

import numpy as np

from .api import Interventions, PandemicBatch, Params
from .scenarios import MT19937_64

# (name, low, high, integer)
SCENARIO_DIMS = [
    ("p_pandemic", 0.010, 0.025, False),
    ("p_seasonal", 0.010, 0.025, False),
    ("initial_infected_pandemic", 5, 50, True),
    ("initial_infected_seasonal", 5, 50, True),
    ("latent_days_pandemic", 0, 3, True),
    ("latent_days_seasonal", 0, 3, True),
    ("infectious_days_pandemic", 5, 10, True),
    ("infectious_days_seasonal", 5, 10, True),
    ("school_close_start", 0, 60, True),
    ("school_close_days", 0, 42, True),
    ("workplace_reduction", 0.0, 0.5, False),
    ("errand_reduction", 0.0, 0.5, False),
    ("vaccination_pandemic", 0.0, 0.4, False),
    ("vaccination_seasonal", 0.0, 0.4, False),
    ("percent_symptomatic", 0.5, 1.0, False),
    ("asymp_factor", 0.0, 0.5, False),
]
NDIM = len(SCENARIO_DIMS)
DEFAULT_AGENTS_PER_SCALE = 2.354e6   # agents at people_scaling_factor 1


def sample_scenarios(n, seed):
    """n x 16 scenario vectors (float64; integer dimensions hold integers)."""
    g = MT19937_64(seed)
    out = np.zeros((n, NDIM))
    for i in range(n):
        for j, (_, lo, hi, integer) in enumerate(SCENARIO_DIMS):
            u = g.unit()
            out[i, j] = lo + math.floor(u * (hi - lo + 1)) if integer else lo + u * (hi - lo)
    return out


def neutral_vector(p_pandemic=0.018, p_seasonal=0.013, init=(10, 10)):
    """The scenario with no interventions (the reference model)."""
    return np.array([p_pandemic, p_seasonal, init[0], init[1], 0, 0, 10, 10, 0, 0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0])


# Dimension pairs (pandemic, seasonal) of the per-strain parameters.  The
# model treats the two strains alike: each has its own transmission factor,
# initial infected, latent and infectious days and vaccination share, and
# both use the same contacts, the same infection-profile function and
# identically distributed random words (PandemicSim.cu:3106-3233, :708-758),
# so a scenario with these pairs swapped has the same distribution of
# outcomes with the strains swapped.
STRAIN_PAIRS = [(0, 1), (2, 3), (4, 5), (6, 7), (12, 13)]


def swap_strains(X):
    """Scenario vectors with every per-strain dimension pair swapped."""
    X = np.array(X, copy=True)
    for a, b in STRAIN_PAIRS:
        X[..., [a, b]] = X[..., [b, a]]
    return X


def scenario_row(vec, n_agents, days, seeds_per_million=False):
    """(constants row, Interventions) of one scenario vector at a target
    population.  With seeds_per_million the initial-infected dimensions
    count per million agents (the vector's value times n_agents / 1e6,
    rounded), which keeps the per-capita start of a scenario learnt at 1M
    agents at any size."""
    v = [float(x) for x in vec]
    scale = "%.6f" % (n_agents / DEFAULT_AGENTS_PER_SCALE)
    seeds = [int(v[2]), int(v[3])]
    if seeds_per_million:
        seeds = [max(1, int(round(x * n_agents / 1e6))) for x in seeds]
    row = [str(int(days)), "%.4f" % (100.0 * v[0]), "%.4f" % (100.0 * v[1]), str(seeds[0]), str(seeds[1]), scale,
           scale, "%.4f" % v[14], "%.4f" % v[15]]
    iv = Interventions(latent=(int(v[4]), int(v[5])), infectious=(int(v[6]), int(v[7])),
                       school_close_start=int(v[8]), school_close_days=int(v[9]), workplace_reduction=v[10],
                       errand_reduction=v[11], vaccination=(v[12], v[13]))
    return row, iv


def scenario_inputs(vec, n_agents, days, seeds_per_million=False):
    """(Params, Interventions) of one scenario vector (scenario_row)."""
    row, iv = scenario_row(vec, n_agents, days, seeds_per_million)
    return Params.from_row(row), iv


def incidence(counts, days):
    """incidence[strain, d] = S_d - S_{d+1} for d < days (0 once the run has ended)."""
    out = np.zeros((2, days), np.float64)
    n = len(counts)
    for k, col in enumerate((1, 9)):
        s = counts[:, col].astype(np.int64)
        m = min(n - 1, days)
        if m > 0:
            out[k, :m] = s[:m] - s[1:m + 1]
    return out


def summarise(inc):
    """inc: replicates x 2 x days -> 2 x days x 3 (mean, q10, q90 of log1p)."""
    li = np.log1p(inc)
    return np.stack([li.mean(axis=0), np.quantile(li, 0.1, axis=0), np.quantile(li, 0.9, axis=0)], axis=-1)


def run_ensemble(vectors, replicates, n_agents, days, seed0=1, max_agents_per_batch=200_000_000, device=0,
                 return_raw=False, seeds_per_million=False):
    """Runs every scenario with `replicates` seeds (seed0 + scenario * replicates
    + r) in as few batches as fit max_agents_per_batch.  Returns Y
    (scenarios x 2 x days x 3) and timing; with return_raw also the raw
    (counts, infected) of every run in scenario-major order."""
    vectors = np.asarray(vectors)
    n = len(vectors)
    per_batch = max(1, int(max_agents_per_batch // (n_agents * replicates)))
    # replicates of one scenario may be split over batches when they do not fit one
    rep_split = max(1, -(-n_agents * replicates // max_agents_per_batch))
    Y = np.zeros((n, 2, days, 3))
    raw = []
    person_days = 0
    t0 = time.perf_counter()
    for s0 in range(0, n, per_batch):
        chunk = range(s0, min(n, s0 + per_batch))
        params, ivs, seeds = [], [], []
        for i in chunk:
            p, iv = scenario_inputs(vectors[i], n_agents, days, seeds_per_million)
            for r in range(replicates):
                params.append(p)
                ivs.append(iv)
                seeds.append(seed0 + i * replicates + r)
        res, n_people = [], []
        step = -(-len(params) // rep_split)
        for b0 in range(0, len(params), step):
            b = PandemicBatch(params[b0:b0 + step], seeds[b0:b0 + step], device)
            b.set_interventions(ivs[b0:b0 + step])
            n_people.extend(int(x) for x in b.setup())
            res.extend(b.run())
            b.close()
        for j, i in enumerate(chunk):
            runs = res[j * replicates:(j + 1) * replicates]
            inc = np.stack([incidence(c, days) for c, _ in runs])
            Y[i] = summarise(inc)
            for r, (c, inf) in enumerate(runs):
                person_days += int(n_people[j * replicates + r]) * len(inf)
            if return_raw:
                raw.extend(runs)
    wall = time.perf_counter() - t0
    info = {"scenarios": n, "replicates": replicates, "runs": n * replicates, "seconds": wall,
            "runs_per_s": n * replicates / wall, "person_days": person_days, "person_days_per_s": person_days / wall}
    return (Y, info, raw) if return_raw else (Y, info)


TRAINING_SET_SEED = 20261204          # scenario vectors of the surrogate's training set
TRAINING_SET_SEED0 = 1                # replicate seeds: TRAINING_SET_SEED0 + scenario * replicates + r


def save_part(path, X, Y, first, info, replicates, agents, days, seed):
    """A generated part: scenario rows [first, first + len(X)) of the vector
    stream `seed`, targets in float16 (log-incidence values up to ~15)."""
    np.savez_compressed(path, X=X, Y=Y.astype(np.float16), first=np.int64(first), replicates=np.int64(replicates),
                        agents=np.int64(agents), days=np.int64(days), seed=np.int64(seed),
                        seconds=np.float64(info["seconds"]), person_days=np.int64(info["person_days"]))


def merge_parts(paths, out):
    """Concatenates parts in row order (they must tile [0, n) of one stream)."""
    parts = sorted((np.load(p) for p in paths), key=lambda z: int(z["first"]))
    nxt = 0
    for z in parts:
        assert int(z["first"]) == nxt, "parts do not tile the scenario range"
        nxt += len(z["X"])
    meta = {k: int(parts[0][k]) for k in ("replicates", "agents", "days", "seed")}
    np.savez_compressed(out, X=np.concatenate([z["X"] for z in parts]), Y=np.concatenate([z["Y"] for z in parts]),
                        seconds=np.float64(sum(float(z["seconds"]) for z in parts)),
                        person_days=np.int64(sum(int(z["person_days"]) for z in parts)), **meta)


def load_training_set(path):
    """(X float64, Y float32 n x 2 x days x 3, metadata) of a merged set."""
    z = np.load(path)
    meta = {k: (float(z[k]) if k == "seconds" else int(z[k])) for k in
            ("replicates", "agents", "days", "seed", "seconds", "person_days")}
    return z["X"], z["Y"].astype(np.float32), meta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenarios", type=int, default=64)
    ap.add_argument("--replicates", type=int, default=8)
    ap.add_argument("--agents", type=int, default=100000)
    ap.add_argument("--days", type=int, default=120)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--seed0", type=int, default=1, help="first replicate seed")
    ap.add_argument("--range", default=None, help="first:stop scenario rows to generate (default: all)")
    ap.add_argument("--chunk", type=int, default=64, help="scenarios per progress line")
    ap.add_argument("--merge", nargs="+", default=None, help="merge generated parts instead")
    ap.add_argument("--out", default="ensemble.npz")
    a = ap.parse_args()
    if a.merge:
        merge_parts(a.merge, a.out)
        print("merged %d parts into %s" % (len(a.merge), a.out))
        return
    X = sample_scenarios(a.scenarios, a.seed)
    first, stop = (0, a.scenarios) if a.range is None else (int(x) for x in a.range.split(":"))
    Ys, secs, pdays = [], 0.0, 0
    for c0 in range(first, stop, a.chunk):
        c1 = min(stop, c0 + a.chunk)
        # replicate seeds follow the scenario's row in the whole stream
        Y, info = run_ensemble(X[c0:c1], a.replicates, a.agents, a.days, seed0=a.seed0 + c0 * a.replicates)
        Ys.append(Y)
        secs += info["seconds"]
        pdays += info["person_days"]
        print("rows %d-%d: %d runs in %.1f s (%.1f runs/s)" % (c0, c1, info["runs"], info["seconds"],
                                                                 info["runs_per_s"]), flush=True)
    save_part(a.out, X[first:stop], np.concatenate(Ys), first, {"seconds": secs, "person_days": pdays}, a.replicates,
              a.agents, a.days, a.seed)
    print("%d scenarios in %.1f s -> %s" % (stop - first, secs, a.out), flush=True)


if __name__ == "__main__":
    main()
