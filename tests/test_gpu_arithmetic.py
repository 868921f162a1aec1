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
