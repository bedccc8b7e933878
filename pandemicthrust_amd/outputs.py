"""The reference's per-run output files, written from a run's daily rows and
generation offsets (the same text the host program and the reference's
PandemicSim write):

  output_infected_stats.csv   daily_writeInfectedStats (PandemicSim.cu:263, :2771-2818)
  infected_max.txt            logging_closeOutputStreams (:855-866) with the
                              maximum tracked in runToCompletion (:923-931)
  output_rn.csv               final_countReproduction (:2276-2297)

Numbers follow C's printf: %d, and %f of a float promoted to double.  The
reproduction number is a float division of two ints; 0/0 is the x86
default NaN (sign bit set), which glibc prints as "-nan", and x/0 prints
"inf".
"""
import numpy as np

STATS_HEADER = ("day,pandemic_susceptible,pandemic_infectious,pandemic_symptomatic,pandemic_asymptomatic,"
                "pandemic_recovered,seasonal_susceptible,seasonal_infectious,seasonal_symptomatic,"
                "seasonal_asymptomatic,seasonal_recovered")
RN_HEADER = "gen,gen_size_p,rn_p,gen_size_s,rn_s"
MAX_HEADER = "max_proportion,day"


def c_float(x):
    """printf("%f", (double)x) for a float32 x."""
    x = np.float32(x)
    if np.isnan(x):
        return "-nan" if np.signbit(x) else "nan"
    return "%f" % float(x)


def rn_value(num, den):
    """(float)num / den with the reference's int operands."""
    if den == 0:
        return float("nan") if num == 0 else float("inf")
    return float(np.float32(num) / np.float32(den))


def _rn_text(num, den):
    if den == 0 and num == 0:
        return "-nan"
    return c_float(rn_value(num, den))


def infected_stats_text(counts, infected):
    """counts: rows x 16 status counts (pandemic codes 0..7, seasonal 0..7);
    infected: rows infected counts.  Rows stop after the first I == 0 row."""
    counts = np.asarray(counts)
    lines = [STATS_HEADER]
    for d in range(len(counts)):
        c = [int(v) for v in counts[d]]
        psym, pasym = c[2] + c[3] + c[4], c[5] + c[6] + c[7]
        ssym, sasym = c[10] + c[11] + c[12], c[13] + c[14] + c[15]
        lines.append("%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d" % (d, c[1], psym + pasym, psym, pasym, c[0], c[9],
                                                           ssym + sasym, ssym, sasym, c[8]))
        if int(infected[d]) == 0:
            break
    return "\n".join(lines) + "\n"


def infected_max(infected, n_people):
    """(float32 proportion, day) of the first maximum of infected / N."""
    best, day = np.float32(0), 0
    for d, i in enumerate(infected):
        if int(i) == 0:
            break
        p = np.float32(np.float32(int(i)) / np.float32(int(n_people)))
        if p > best:
            best, day = p, d
    return best, day


def infected_max_text(infected, n_people):
    p, d = infected_max(infected, n_people)
    return "%s\n%s,%d\n" % (MAX_HEADER, c_float(p), d)


def rn_text(c_pandemic, c_seasonal, max_days):
    """c_*: generation offsets (max_days + 1 entries, last = N)."""
    cp = [int(v) for v in c_pandemic]
    cs = [int(v) for v in c_seasonal]
    lines = [RN_HEADER]
    for g in range(max_days - 1):
        sp, np_ = cp[g + 1] - cp[g], cp[g + 2] - cp[g + 1]
        ss, ns = cs[g + 1] - cs[g], cs[g + 2] - cs[g + 1]
        lines.append("%d,%d,%s,%d,%s" % (g, sp, _rn_text(np_, sp), ss, _rn_text(ns, ss)))
    return "\n".join(lines) + "\n"


def write_run(directory, counts, infected, n_people, c_pandemic, c_seasonal, max_days):
    """Writes the three files of one run into directory."""
    import os
    with open(os.path.join(directory, "output_infected_stats.csv"), "w") as f:
        f.write(infected_stats_text(counts, infected))
    with open(os.path.join(directory, "infected_max.txt"), "w") as f:
        f.write(infected_max_text(infected, n_people))
    with open(os.path.join(directory, "output_rn.csv"), "w") as f:
        f.write(rn_text(c_pandemic, c_seasonal, max_days))
