"""pt_location_table_noalloc, the allocation-free location-table build (its
radix scratch is a file-scope device array; no counterpart in the
reference), against pt_location_table on the same inputs: two calls in a
row, the second smaller, so the second reuses scratch the first left behind;
and inputs beyond its fixed limit are refused."""
import numpy as np
import pytest

from hip_mem import DeviceArray
from test_gpu_location_table import build, closed_form

pytestmark = pytest.mark.gpu


def test_noalloc_equals_the_allocating_build_twice():
    from pandemicthrust_amd.api import PT_LOCATION_TABLE_NOALLOC_MAX_PAIRS as MAXP
    rng = np.random.default_rng(7)
    for n, K in [(MAXP, 276_120), (1_000_003, 61_597), (5, 3)]:
        keys = rng.integers(0, K, n, dtype=np.uint32)
        print("case: noalloc n=%d K=%d" % (n, K), flush=True)
        p_alloc, o_alloc = build(keys, K, "pt_location_table")
        p_free, o_free = build(keys, K, "pt_location_table_noalloc")
        assert np.array_equal(o_free, o_alloc) and np.array_equal(p_free, p_alloc), (n, K)
        want_p, want_o = closed_form(keys, K)
        assert np.array_equal(o_free, want_o) and np.array_equal(p_free, want_p), (n, K)


def test_noalloc_refuses_more_pairs_than_its_scratch_holds():
    from pandemicthrust_amd.api import PT_LOCATION_TABLE_NOALLOC_MAX_PAIRS as MAXP, lib
    n = MAXP + 1
    dk = DeviceArray(n, np.uint32)
    dp = DeviceArray(n, np.int32)
    do = DeviceArray(11, np.int32)
    rc = lib().pt_location_table_noalloc(dk.ptr, n, 10, dp.ptr, do.ptr, 0, None)
    assert rc != 0 and b"allocation-free limit" in lib().pt_last_error()
