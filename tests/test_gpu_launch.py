"""Launch shapes and table-strategy thresholds (pt_launch_knob) change speed
only: every setting the autotuner may pick gives bit-identical days, agent
state and tables."""
import numpy as np
import pytest

from pandemicthrust_amd import PandemicBatch, PandemicSim, Params, PTError, autotune

pytestmark = pytest.mark.gpu

ROW = ["40", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"]
DEFAULTS = {"contact_blocks_per_cu": 8, "gen_blocks_per_cu": 16, "sort_wide_bits": 8, "mark_limit_pct": 100,
            "subset_share_pct": 75, "sort_max_bits": 8, "segmented_tables": 1}
ROWS = [(["30", "2.0", "1.4", "10", "10", "0.03", "0.03", "1", "0"], 1),
        (["35", "1.7", "1.3", "20", "5", "0.05", "0.04", "0.9", "0.2"], 2),
        (["25", "2.4", "1.2", "8", "12", "0.02", "0.025", "1", "0"], 3)]


def _sim_result(settings, mode=0):
    s = PandemicSim(Params.from_row(ROW), 3)
    for k, v in settings.items():
        s.set_launch(k, v)
    s.set_table_mode(mode)
    s.setup_sim()
    counts, inf = s.run()
    state = [s.array(k) for k in range(6)]
    s.close()
    return counts, inf, state


@pytest.mark.parametrize("mode", [0, 1])
def test_sim_launch_settings_do_not_change_results(mode):
    ref = _sim_result({}, mode)
    for settings in ({"contact_blocks_per_cu": 1, "gen_blocks_per_cu": 1, "sort_wide_bits": 9},
                     {"contact_blocks_per_cu": 32, "gen_blocks_per_cu": 64, "sort_wide_bits": 1},
                     {"contact_blocks_per_cu": 3, "gen_blocks_per_cu": 7, "sort_wide_bits": 6},
                     {"mark_limit_pct": 0, "subset_share_pct": 0},          # never a subset table
                     {"mark_limit_pct": 1000, "subset_share_pct": 100},     # a subset table on nearly every day
                     {"mark_limit_pct": 50, "subset_share_pct": 90},
                     {"sort_max_bits": 4}, {"sort_max_bits": 11, "sort_wide_bits": 1}):
        got = _sim_result(settings, mode)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), settings
        for a, b in zip(got[2], ref[2]):
            assert np.array_equal(a, b), settings


def test_batch_launch_settings_do_not_change_results():
    b = PandemicBatch([Params.from_row(r) for r, _ in ROWS], [s for _, s in ROWS])
    b.setup()
    b.snapshot()
    ref = b.run()
    for settings in ({"contact_blocks_per_cu": 1, "sort_wide_bits": 9}, {"contact_blocks_per_cu": 32, "sort_wide_bits": 2},
                     {"mark_limit_pct": 1000, "subset_share_pct": 100}, {"mark_limit_pct": 0, "subset_share_pct": 0},
                     {"sort_max_bits": 4}, {"sort_max_bits": 11}, {"segmented_tables": 0}, {"sort_max_bits": 9}):
        for k, v in settings.items():
            b.set_launch(k, v)
        b.restore()
        got = b.run()       # graphs recaptured with the new shapes
        for r in range(len(ROWS)):
            assert np.array_equal(got[r][0], ref[r][0]) and np.array_equal(got[r][1], ref[r][1]), (settings, r)
    b.close()


def test_launch_settings_are_checked():
    s = PandemicSim(Params.from_row(ROW), 3)
    assert s.launch() == DEFAULTS
    with pytest.raises(PTError, match="outside"):
        s.set_launch("sort_wide_bits", 10)
    with pytest.raises(PTError, match="outside"):
        s.set_launch("contact_blocks_per_cu", 0)
    with pytest.raises(PTError, match="outside"):
        s.set_launch("mark_limit_pct", 1001)
    with pytest.raises(PTError, match="outside"):
        s.set_launch("subset_share_pct", 101)
    with pytest.raises(PTError, match="outside"):
        s.set_launch("sort_max_bits", 12)
    with pytest.raises(PTError, match="outside"):
        s.set_launch("segmented_tables", 2)
    s.close()


def test_autotuner_keeps_results():
    s = PandemicSim(Params.from_row(ROW), 3)
    s.setup_sim()
    ref_c, ref_i = None, None
    best, log = autotune.tune_sim(s)
    # every setting comes back; all but segmented_tables (batch only) are tuned
    assert set(best) == set(DEFAULTS)
    assert set(autotune.SIM_KNOBS) == set(DEFAULTS) - {"segmented_tables"}
    assert len(log) == sum(len(autotune.CANDIDATES[k]) for k in autotune.SIM_KNOBS)
    for k, v in best.items():         # the default unless a candidate was clearly faster
        ms = {val: t for name, val, t in log if name == k}
        assert v == DEFAULTS[k] or \
            ms[v] == min(ms.values()), (k, v, ms)
    assert s.launch() == best
    s.restore()
    c, i = s.run()
    s.close()
    ref = _sim_result({})
    assert np.array_equal(c, ref[0]) and np.array_equal(i, ref[1])
