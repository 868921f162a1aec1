"""NumPy views of the 32-byte node layouts (the reference's src/bvh.h:30-49) and of the 64-byte SkipNode of the
per-lane bottom-level walk."""
import numpy as np

TLAS_NODE = np.dtype([("lo", "<f4", (3,)), ("left", "<i4"), ("hi", "<f4", (3,)), ("right", "<i4")])
BLAS_NODE = np.dtype([("lo", "<f4", (3,)), ("a", "<i4"), ("hi", "<f4", (3,)), ("count", "<i4")])
assert TLAS_NODE.itemsize == 32 and BLAS_NODE.itemsize == 32

SKIP_NODE = np.dtype([("lo", "<f4", (3,)), ("a", "<i4"), ("hi", "<f4", (3,)), ("meta", "<i4"), ("escape", "<i4", (8,))])
assert SKIP_NODE.itemsize == 64
