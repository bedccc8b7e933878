"""Scenario ensembles: sampling, the scenario -> inputs map, incidence."""
import numpy as np

from pandemicthrust_amd import ensemble as E


def test_sampling_is_deterministic_and_in_range():
    a = E.sample_scenarios(200, 7)
    b = E.sample_scenarios(200, 7)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, E.sample_scenarios(200, 8))
    for j, (name, lo, hi, integer) in enumerate(E.SCENARIO_DIMS):
        col = a[:, j]
        assert col.min() >= lo and col.max() <= hi, name
        if integer:
            assert np.array_equal(col, np.round(col)), name
            assert len(np.unique(col)) > 1, name


def test_neutral_vector_maps_to_the_reference_inputs():
    p, iv = E.scenario_inputs(E.neutral_vector(), 100000, 30)
    assert abs(p.r_pandemic - 1.8) < 1e-6 and abs(p.r_seasonal - 1.3) < 1e-6
    assert p.init_pandemic == 10 and p.init_seasonal == 10 and p.max_days == 30
    assert p.percent_symptomatic == 1.0 and p.asymp_factor == 0.0
    assert abs(p.people_scaling - 100000 / 2.354e6) < 1e-6
    assert list(iv.latent_days) == [0, 0] and list(iv.infectious_days) == [10, 10]
    assert iv.school_close_days == 0 and iv.workplace_reduction == 0 and iv.errand_reduction == 0
    assert list(iv.vaccination) == [0.0, 0.0]


def test_incidence_is_the_drop_in_susceptibles():
    counts = np.zeros((5, 16), np.int32)
    counts[:, 1] = [100, 90, 70, 65, 65]
    counts[:, 9] = [100, 99, 97, 97, 90]
    inc = E.incidence(counts, 6)
    assert np.array_equal(inc[0], [10, 20, 5, 0, 0, 0])
    assert np.array_equal(inc[1], [1, 2, 0, 7, 0, 0])
    inc3 = E.incidence(counts, 3)
    assert np.array_equal(inc3[0], [10, 20, 5])


def test_summary_of_identical_replicates():
    inc = np.tile(np.arange(12, dtype=float).reshape(1, 2, 6), (4, 1, 1))
    y = E.summarise(inc)
    assert y.shape == (2, 6, 3)
    assert np.allclose(y[..., 0], np.log1p(inc[0])) and np.allclose(y[..., 1], y[..., 0]) and np.allclose(y[..., 2], y[..., 0])
