"""Host logic behind the bench line: the byte model that turns the engine's
per-day table log into the roofline's algorithmic bytes, the traffic figure
read from the committed rocprofv3 summary, and the dispatch-order labelling
of a kernel trace into the engine's timing classes."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


LAY = {"wd_passes": 2, "we_passes": 3, "wd_buckets": 61_597, "we_buckets": 500_000}


def _week(days, inf=1000):
    return [inf] * days, np.zeros((days + 1, 16), np.int64)


def test_table_term_counts_each_days_table():
    """12 bytes per pair of each day's table: E*N pairs on full-table days,
    the logged selected pairs on subset days."""
    b = _bench()
    N, W = 1_000_000, 5_000
    full = [1, 0, 1, 0, 1, 1, 0]          # days 5 and 6 are the weekend
    pairs = [1_690_000, 12_345, 1_690_000, 777, 1_690_000, 3_000_000, 4_321]
    inf, counts = _week(len(full))
    per, loop = b.algorithmic_bytes(N, W, LAY, inf, counts, full, pairs)
    want = 12 * (2 * N + 12_345 + 2 * N + 777 + 2 * N + 3 * N + 4_321)
    assert per["table_model"] == want
    # The loop model: 43N + 89I + 3X + 8W + 4 per weekday, 62N + 62I + 3X +
    # 40W + 4 per weekend day (no new infections here: X = 0).
    I = inf[0]
    assert loop == 5 * (43 * N + 89 * I + 8 * W + 4) + 2 * (62 * N + 62 * I + 40 * W + 4)


def test_radix_bytes_follow_the_pass_structure():
    """A full weekday of 61,597 buckets: two 8-bit passes, the last writing
    ids plus a 1-byte low key; a full weekend day of 500,000 buckets: three
    passes (7, 6, 6 bits), the last writing ids plus a 2-byte low key."""
    b = _bench()
    N, W = 1_000_000, 5_000
    k = 1_690_000
    per, _ = b.algorithmic_bytes(N, W, LAY, [10], np.zeros((2, 16)), [1], [k])
    n = 2 * N
    assert per["sort_upsweep"] == 4 * n + 8 * k
    assert per["sort_downsweep"] == (4 * n + 8 * k) + (8 * k + 4 * k + 1 * k)
    full = [0, 0, 0, 0, 0, 1]
    pairs = [0, 0, 0, 0, 0, 3 * N]
    inf = [10] * 6
    per, _ = b.algorithmic_bytes(N, W, LAY, inf, np.zeros((7, 16)), full, pairs)
    n = k = 3 * N
    assert per["sort_downsweep"] == (4 * n + 8 * k) + 16 * k + (8 * k + 4 * k + 2 * k)
    assert per["sort_upsweep"] == 4 * n + 8 * k + 8 * k


def test_one_pass_table_writes_ids_only():
    b = _bench()
    lay = {"wd_passes": 1, "we_passes": 1, "wd_buckets": 200, "we_buckets": 250}
    per, _ = b.algorithmic_bytes(100, 20, lay, [5], np.zeros((2, 16)), [1], [150])
    assert per["sort_downsweep"] == 4 * 200 + 4 * 150   # reads 200 keys, writes 150 ids


def test_traffic_comes_from_the_committed_profile():
    b = _bench()
    t = b.measured_traffic("sort_downsweep")
    text = open(os.path.join(ROOT, "profiles", "daily_loop_50M_100d.md")).read()
    assert t is not None and ("traffic_per_launch sort_downsweep: %d bytes" % t) in text


def test_dispatch_labels_put_both_table_paths_in_the_sort_classes():
    from pandemicthrust_amd.profile_summary import label
    seq = ["k_household_types", "k_copy16",                          # setup, restore
           "k_sweep", "k_plan", "k_compact", "k_mark_weekday", "k_decide",
           "k_gen_weekday", "k_select_weekday", "k_scan_reduce", "k_scan_down<true>", "k_compact_runs",
           "k_upsweep<8, false, false, 1024>", "k_scan_reduce", "k_scan_down<true>",
           "k_downsweep<8, 3, 0, 1024>", "k_upsweep<8, false, true, 1024>", "k_scan_reduce", "k_scan_down<true>",
           "k_downsweep<8, 0, 2, 1024>", "k_offsets<1>", "k_contacts<false>", "k_apply", "k_end_day"]
    got = label(seq)
    want = ["setup", "restore",
            "sweep", "plan", "compact", "mark", "mark",
            "gen_weekday", "select", "select", "select", "select",
            "sort_upsweep", "sort_scan", "sort_scan",
            "sort_downsweep", "sort_upsweep", "sort_scan", "sort_scan",
            "sort_downsweep", "offsets", "contacts_weekday", "apply", "end_day"]
    assert got == want
