"""The replicate scenario set (pandemicthrust_amd/scenarios.py) and its
committed CSV."""
import math

from pandemicthrust_amd import scenarios as S


def test_mt19937_64_standard_check_value():
    # C++ [rand.predef]: the 10000th output of a default-constructed
    # std::mt19937_64 is 9981545732273789042.
    g = S.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_mt19937_64_seeded_stream():
    # std::mt19937_64(20261202) from libstdc++ (g++ 11): outputs 1, 2, 3 and 1001.
    g = S.MT19937_64(20261202)
    assert [g(), g(), g()] == [7380320939607606016, 12064863923457509506, 1707063066545808684]
    for _ in range(3, 1000):
        g()
    assert g() == 16491777865684072910


def test_committed_csv_is_the_generator_output():
    assert open(S.REPLICATE_SET_CSV).read() == S.replicate_set_text()


def test_replicate_set_ranges():
    rows = S.load_replicate_set()
    assert len(rows) == 256
    assert [s for s, _ in rows] == list(range(1, 257))
    for seed, f in rows:
        assert f[0] == "180" and f[2] == "1.3" and f[7] == "1" and f[8] == "0"
        assert 1.0 <= float(f[1]) <= 2.5
        assert 5 <= int(f[3]) <= 50 and 5 <= int(f[4]) <= 50
        assert f[5] == f[6]
        n = float(f[5]) * 2.354e6
        assert 5e4 * 0.999 <= n <= 2e6 * 1.001
    # log-uniform: the mean of ln N sits near the middle of [ln 5e4, ln 2e6]
    mid = 0.5 * (math.log(5e4) + math.log(2e6))
    mean = sum(math.log(float(f[5]) * 2.354e6) for _, f in rows) / len(rows)
    assert abs(mean - mid) < 0.25
