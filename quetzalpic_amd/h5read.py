"""Read the HDF5 output of the host layer without h5py (ctypes over the
read-back helpers in libqpic_hdf5.so).  Mirrors utils/reader.py's use:
read_field(path, name) -> numpy array, plus attributes."""
import ctypes
import os

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libqpic_hdf5.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(_LIB)
        lib.qpic_h5_dataset_shape.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_ulonglong)]
        lib.qpic_h5_read_dataset.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                             ctypes.c_ulonglong]
        lib.qpic_h5_read_attr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_ulonglong]
        lib.qpic_h5_read_attr_int.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                              ctypes.POINTER(ctypes.c_int)]
        _lib = lib
    return _lib


def read_field(path, name):
    lib = _load()
    rank = ctypes.c_int()
    dims = (ctypes.c_ulonglong * 4)()
    if lib.qpic_h5_dataset_shape(path.encode(), name.encode(), ctypes.byref(rank), dims) != 0:
        raise KeyError(f"{name} not in {path}")
    shape = tuple(int(dims[i]) for i in range(rank.value))
    out = np.zeros(shape)
    if lib.qpic_h5_read_dataset(path.encode(), name.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                out.size) != 0:
        raise IOError(f"cannot read {name} from {path}")
    return out


def read_attr(path, name, attr, count):
    out = np.zeros(count)
    if _load().qpic_h5_read_attr(path.encode(), name.encode(), attr.encode(),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), count) != 0:
        raise KeyError(f"{attr} of {name} in {path}")
    return out
