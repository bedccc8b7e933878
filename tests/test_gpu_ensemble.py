"""Ensemble generator on the GPU: every run equals its own single run, and
the neutral scenario equals the oracle (the reference model)."""
import numpy as np
import pytest

import oracle_ctypes as O

pytestmark = pytest.mark.gpu


def test_ensemble_runs_equal_single_runs():
    from pandemicthrust_amd import PandemicSim
    from pandemicthrust_amd import ensemble as E
    X = np.vstack([E.neutral_vector(), E.sample_scenarios(2, 99)])
    reps, agents, days = 2, 30000, 30
    Y, info, raw = E.run_ensemble(X, reps, agents, days, seed0=5, max_agents_per_batch=150000, return_raw=True)
    assert Y.shape == (3, 2, days, 3) and info["runs"] == 6 and len(raw) == 6
    for i in range(3):
        p, iv = E.scenario_inputs(X[i], agents, days)
        for r in range(reps):
            s = PandemicSim(p, 5 + i * reps + r)
            s.set_interventions(iv)
            s.setup_sim()
            c, inf = s.run()
            s.close()
            assert np.array_equal(c, raw[i * reps + r][0]) and np.array_equal(inf, raw[i * reps + r][1]), (i, r)
    # Neutral scenario, first replicate: the reference model itself.
    p, _ = E.scenario_inputs(X[0], agents, days)
    ref = O.OracleSim(O.params_from_row([str(p.max_days), "%.4f" % p.r_pandemic, "%.4f" % p.r_seasonal,
                                         str(p.init_pandemic), str(p.init_seasonal), "%.6f" % p.people_scaling,
                                         "%.6f" % p.location_scaling, "1", "0"]), 5)
    ref.setup()
    c, inf = ref.run()
    assert np.array_equal(c, raw[0][0]) and np.array_equal(inf, raw[0][1])
    # Summary of scenario 0 from its two runs.
    inc = np.stack([E.incidence(raw[k][0], days) for k in (0, 1)])
    assert np.allclose(Y[0], E.summarise(inc))
