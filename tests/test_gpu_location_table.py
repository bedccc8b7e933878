"""pt_location_table: the reference's cub::DeviceRadixSort::SortPairs +
thrust::lower_bound (PandemicSim.cu:1212-1248, :3835-3862) on caller-owned
device memory, bit-exact against the closed form of a stable bucketing: ids
grouped by key in ascending order (numpy's stable argsort) and offsets[k] =
number of keys below k."""
import ctypes as C

import numpy as np
import pytest

from hip_mem import DeviceArray

pytestmark = pytest.mark.gpu


def closed_form(keys, K):
    people = np.argsort(keys, kind="stable").astype(np.int32)
    offsets = np.zeros(K + 1, np.int64)
    offsets[1:] = np.cumsum(np.bincount(keys, minlength=K))
    return people, offsets.astype(np.int32)


def build(keys, K, entry="pt_location_table"):
    from pandemicthrust_amd.api import PTError, lib
    n = len(keys)
    dk = DeviceArray.from_host(keys.astype(np.uint32))
    dp = DeviceArray(n, np.int32)
    do = DeviceArray(K + 1, np.int32)
    rc = getattr(lib(), entry)(dk.ptr, n, K, dp.ptr, do.ptr, 0, None)
    if rc != 0:
        raise PTError(lib().pt_last_error().decode())
    return dp.host(), do.host()


# Bucket counts cover every form of the last radix pass's key output: one
# pass (the digit is the key: K <= 256), and the key bits below the last
# digit kept in 1 byte (K = 545, 61,597), 2 bytes (276,120 up to 2,718,720)
# and 4 bytes (2^25 + 3: 26 bits in 4 passes, 20 bits below the last digit).
CASES = [
    (0, 1), (1, 1), (5, 3), (4096, 1), (1_000_000, 256), (100_000, 545), (1_000_000, 61_597),
    (3_000_000, 276_120), (2_000_000, 1 << 20), (777_777, 2_718_720), (500_000, (1 << 25) + 3),
]


@pytest.mark.parametrize("n,K", CASES)
def test_location_table_equals_stable_bucketing(n, K):
    rng = np.random.default_rng(n + K)
    keys = rng.integers(0, K, n, dtype=np.uint32)
    if n > 1000:
        keys[: n // 10] = K - 1            # a crowded last bucket
        keys[n // 10: n // 5] = 0          # and a crowded first one
        rng.shuffle(keys)
    print("case: location table n=%d K=%d" % (n, K), flush=True)
    people, offsets = build(keys, K)
    want_p, want_o = closed_form(keys, K)
    assert np.array_equal(offsets, want_o)
    assert np.array_equal(people, want_p)


def test_location_table_rejects_bad_arguments():
    from pandemicthrust_amd.api import lib
    assert lib().pt_location_table(None, 10, 5, None, None, 0, None) != 0
    assert lib().pt_location_table(None, -1, 5, None, None, 0, None) != 0
