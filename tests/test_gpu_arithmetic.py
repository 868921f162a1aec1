"""Device arithmetic the tracer relies on, checked exhaustively on the MI355X."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_fast_reciprocal_equals_ieee_division_for_every_float(gpu):
    """The tracer's 1/x (hardware estimate + one fused Newton step) must round exactly as 1.0f / x does
    wherever it is used (2^-126 <= |x| < 2^126): all 2^32 bit patterns are tried."""
    from neraulgraphics_amd import _lib
    lib = _lib.load()
    a, b, ex = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
    print("reciprocal check over 2^32 patterns", flush=True)
    _lib.check(lib.ngx_debug_check_reciprocal(0, ctypes.byref(a), ctypes.byref(b), ctypes.byref(ex)))
    print(f"mismatches in range {a.value}, anywhere {b.value}, example 0x{ex.value:08x}", flush=True)
    assert a.value == 0, f"{a.value} mismatches, e.g. bits 0x{ex.value:08x}"


def test_wave_prefix_sum_is_an_inclusive_scan_of_each_wave(gpu):
    """The walks place each leaf's (ray, triangle) pairs at the wave's inclusive prefix sum of the leaves' triangle
    counts: the scan (DPP row shifts, then row broadcasts) must equal a running sum over each run of 64 lanes, whatever
    the values' pattern across the four rows of 16."""
    import numpy as np
    from neraulgraphics_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(5)
    runs = [np.zeros(64), np.ones(64), np.arange(64), np.eye(64)[0], np.eye(64)[15], np.eye(64)[16], np.eye(64)[31],
            np.eye(64)[32], np.eye(64)[47], np.eye(64)[63], np.repeat([1, 2, 3, 4], 16), (np.arange(64) % 16 == 15) * 1000]
    runs += list(rng.integers(0, 256, size=(500, 64))) + list(rng.integers(-(1 << 20), 1 << 20, size=(500, 64)))
    values = np.ascontiguousarray(np.stack(runs).astype(np.int32))
    sums = np.zeros_like(values)
    print(f"wave scan of {len(values)} waves", flush=True)
    _lib.check(lib.ngx_debug_wave_prefix_sum(0, values.ctypes.data, len(values), sums.ctypes.data))
    want = np.cumsum(values.astype(np.int64), axis=1).astype(np.int32)
    bad = np.nonzero(np.any(sums != want, axis=1))[0]
    assert bad.size == 0, f"{bad.size} waves differ, first {bad[0]}: {sums[bad[0]].tolist()} vs {want[bad[0]].tolist()}"

