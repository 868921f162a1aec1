"""Zero-copy views of the backend's device buffers as DLPack capsules, for torch.from_dlpack.

The capsule's manager keeps its owner (the renderer) alive until the consumer releases the tensor, which
is how the reference's Python edge hands out its buffers (python/bindings.cpp:192-202)."""
import ctypes

import numpy as np

kDLROCM = 10
_CODES = {np.dtype(np.float32): (2, 32), np.dtype(np.int32): (0, 32), np.dtype(np.uint8): (1, 8)}


class DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", DLDevice), ("ndim", ctypes.c_int32), ("dtype", DLDataType),
                ("shape", ctypes.POINTER(ctypes.c_int64)), ("strides", ctypes.POINTER(ctypes.c_int64)),
                ("byte_offset", ctypes.c_uint64)]


_DELETER = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


class DLManagedTensor(ctypes.Structure):
    _fields_ = [("dl_tensor", DLTensor), ("manager_ctx", ctypes.c_void_p), ("deleter", _DELETER)]


_live = {}  # address of a DLManagedTensor -> (struct, shape array, owner)


@_DELETER
def _release(address):
    _live.pop(address, None)


_PyCapsule_New = ctypes.pythonapi.PyCapsule_New
_PyCapsule_New.restype = ctypes.py_object
_PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]


def capsule(pointer, shape, dtype, device_id, owner):
    """A 'dltensor' capsule viewing `pointer` (device memory) as a C-contiguous array."""
    code, bits = _CODES[np.dtype(dtype)]
    shp = (ctypes.c_int64 * len(shape))(*shape)
    mt = DLManagedTensor()
    mt.dl_tensor.data = pointer
    mt.dl_tensor.device = DLDevice(kDLROCM, device_id)
    mt.dl_tensor.ndim = len(shape)
    mt.dl_tensor.dtype = DLDataType(code, bits, 1)
    mt.dl_tensor.shape = shp
    mt.dl_tensor.strides = None
    mt.dl_tensor.byte_offset = 0
    mt.manager_ctx = None
    mt.deleter = _release
    address = ctypes.addressof(mt)
    _live[address] = (mt, shp, owner)
    return _PyCapsule_New(address, b"dltensor", None)


def to_torch(pointer, shape, dtype, device_id, owner):
    from torch.utils.dlpack import from_dlpack
    return from_dlpack(capsule(pointer, shape, dtype, device_id, owner))
