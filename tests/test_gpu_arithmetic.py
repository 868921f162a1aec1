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



def _direction_cells(d):
    """The 12-bit cell the ray order sorts by, in float32: the octahedral map of the direction (the lower hemisphere
    folded over the upper), 64 x 64 cells, the two 6-bit coordinates' bits interleaved (u's above v's); 4095 for a
    zero or non-finite direction."""
    import numpy as np
    f = np.float32
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    with np.errstate(all="ignore"):
        s = (np.abs(x) + np.abs(y)) + np.abs(z)
        bad = ~(s > 0) | ~np.isfinite(s)
        u, v = x / s, y / s
        uu = (f(1) - np.abs(v)) * np.where(u >= 0, f(1), f(-1))
        vv = (f(1) - np.abs(u)) * np.where(v >= 0, f(1), f(-1))
        u, v = np.where(z < 0, uu, u), np.where(z < 0, vv, v)
        q = [np.fmin(f(63), np.fmax(f(0), (w * f(0.5) + f(0.5)) * f(64))) for w in (u, v)]
    q = [np.where(bad, 0, w).astype(np.uint32) for w in q]
    spread = [sum(((w >> b) & 1) << (2 * b) for b in range(6)) for w in q]
    return np.where(bad, 4095, (spread[0] << 1) | spread[1]).astype(np.uint32)


def test_ray_order_is_the_stable_order_of_direction_cells(gpu):
    """The order a ray query traces each world's rays in: a permutation of the world's indices, non-decreasing in
    direction cell, equal cells in index order. Up to 65,536 rays a world it comes from one counting pass (16
    wavefronts, each over a sixteenth of the world); above, from a radix sort in three passes. Worlds of random
    directions, of one direction repeated (the order must be the identity), and of zero, infinite, NaN, axis and
    lower-hemisphere directions."""
    import numpy as np
    from neraulgraphics_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(21)
    special = np.array([[0, 0, 0], [np.inf, 0, 1], [np.nan, 1, 0], [1, 0, 0], [0, -1, 0], [0, 0, -1], [-1e-30, 0, 1e-30],
                        [1, 1, -1], [-1, -1, -1], [3, -4, 0]], np.float32)
    for worlds, rays in ((3, 65536), (2, 5000), (4, 1024), (2, 70000)):
        d = rng.normal(size=(worlds, rays, 3)).astype(np.float32)
        d[1] = np.float32([0.3, -0.2, 0.9])                    # one direction: the identity order
        if rays >= 4096:
            d[0, ::7] = special[rng.integers(0, len(special), size=d[0, ::7].shape[0])]
        order = np.zeros((worlds, rays), np.int32)
        print(f"ray order: {worlds} worlds x {rays} rays", flush=True)
        _lib.check(lib.ngx_debug_ray_order(0, np.ascontiguousarray(d).ctypes.data, rays, worlds, order.ctypes.data))
        for w in range(worlds):
            cells = _direction_cells(d[w])
            assert np.array_equal(np.sort(order[w]), np.arange(rays)), (worlds, rays, w)
            assert np.array_equal(order[w], np.argsort(cells, kind="stable")), (worlds, rays, w)
        assert np.array_equal(order[1], np.arange(rays))
