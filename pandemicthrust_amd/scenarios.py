"""Replicate scenario sets for the batch engine.

The replicate set: 256 constants rows for a replicate ensemble on one GPU.  Replicate r
(r = 0..255) has core seed r + 1, max_days 180 and four draws u1..u4 in that
order from one std::mt19937_64 stream seeded with 20261202, each draw taken
as u = (x >> 11) * 2^-53:

    N_r   = exp(ln 5e4 + u1 (ln 2e6 - ln 5e4))      target population, log-uniform
    pscale = lscale = N_r / 2.354e6                  (the default scale has 2.354M agents)
    R_p   = 100 (0.010 + 0.015 u2)                   in [1.0, 2.5]; R_s = 1.3
    init_p = 5 + floor(46 u3), init_s = 5 + floor(46 u4)   in 5..50
    percent_symptomatic = 1, asymp_factor = 0       (the reference's constants.csv)

The rows are written to data/replicate_scenarios_256.csv (python -m
pandemicthrust_amd.scenarios writes it); the realised population of a row is
what setup draws for it (within about 0.3% of N_r).
"""
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REPLICATE_SET_CSV = os.path.join(HERE, "data", "replicate_scenarios_256.csv")
REPLICATE_SET_SEED = 20261202
HEADER = ("replicate,seed,max_days,reproduction_number_pandemic,reproduction_number_seasonal,"
          "initial_infected_pandemic,initial_infected_seasonal,people_scaling_factor,location_scaling_factor,"
          "percent_symptomatic,asymp_factor")

_MASK = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto and Nishimura, 2000), with the
    parameters and seeding of C++'s std::mt19937_64."""
    n, m = 312, 156
    a = 0xB5026F5AA96619E9
    upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed=5489):
        mt = [0] * self.n
        mt[0] = seed & _MASK
        for i in range(1, self.n):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & _MASK
        self.mt = mt
        self.i = self.n

    def _twist(self):
        mt, n, m = self.mt, self.n, self.m
        for k in range(n):
            y = (mt[k] & self.upper) | (mt[(k + 1) % n] & self.lower)
            v = mt[(k + m) % n] ^ (y >> 1)
            if y & 1:
                v ^= self.a
            mt[k] = v
        self.i = 0

    def __call__(self):
        if self.i >= self.n:
            self._twist()
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & _MASK

    def unit(self):
        return (self() >> 11) * (1.0 / 9007199254740992.0)


def replicate_set_rows(n=256):
    """[(core_seed, [9 constants fields as text]), ...]."""
    g = MT19937_64(REPLICATE_SET_SEED)
    lo, hi = math.log(5e4), math.log(2e6)
    rows = []
    for r in range(n):
        u1, u2, u3, u4 = g.unit(), g.unit(), g.unit(), g.unit()
        n_target = math.exp(lo + u1 * (hi - lo))
        scale = "%.6f" % (n_target / 2.354e6)
        rp = "%.4f" % (100.0 * (0.010 + 0.015 * u2))
        ip, is_ = 5 + int(46 * u3), 5 + int(46 * u4)
        rows.append((r + 1, ["180", rp, "1.3", str(ip), str(is_), scale, scale, "1", "0"]))
    return rows


def replicate_set_text():
    lines = [HEADER]
    for r, (seed, f) in enumerate(replicate_set_rows()):
        lines.append("%d,%d,%s" % (r, seed, ",".join(f)))
    return "\n".join(lines) + "\n"


def load_replicate_set(path=REPLICATE_SET_CSV):
    """Reads the committed rows: [(core_seed, fields), ...]."""
    rows = []
    with open(path) as f:
        next(f)
        for line in f:
            p = line.strip().split(",")
            if len(p) == 11:
                rows.append((int(p[1]), p[2:]))
    return rows


if __name__ == "__main__":
    os.makedirs(os.path.dirname(REPLICATE_SET_CSV), exist_ok=True)
    with open(REPLICATE_SET_CSV, "w") as f:
        f.write(replicate_set_text())
    print("wrote", REPLICATE_SET_CSV)
