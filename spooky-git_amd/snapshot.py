"""Snapshot files of a SPOOKY run (data/snapNNNN.h5) from Python, through
libspkio.so (include/spk_io.h): the reference's layout (src/io/hdf5_io.cu),
written and read without h5py."""
import ctypes
import os

import numpy as np

_LIB = None
LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libspkio.so")


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("%s is not built (python spooky-git_amd/build.py)" % LIBPATH)
        L = ctypes.CDLL(LIBPATH)
        L.spkio_last_error.restype = ctypes.c_char_p
        L.spkio_write_snapshot.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_void_p),
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int]
        L.spkio_read_header.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_double)] * 3 + [
            ctypes.POINTER(ctypes.c_int)]
        L.spkio_read_field.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t]
        L.spkio_last_snapshot.argtypes = [ctypes.c_char_p]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().spkio_last_error().decode())


def snapshot_path(data_dir, num):
    return os.path.join(data_dir, "snap%04d.h5" % num)


def write_snapshot(path, fields, t_start=0.0, t_save=0.0, t_lastsnap=0.0, t_lastvar=0.0, step=0):
    """fields: {name: real array [NX, NY, NZ]} (all the same shape)."""
    names = list(fields)
    arrs = [np.ascontiguousarray(fields[n], dtype=np.float64) for n in names]
    shape = arrs[0].shape
    if len(shape) != 3 or any(a.shape != shape for a in arrs):
        raise ValueError("fields must share one 3D shape")
    cn = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    cp = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    _check(lib().spkio_write_snapshot(path.encode(), shape[0], shape[1], shape[2], len(names), cn, cp, t_start,
                                      t_save, t_lastsnap, t_lastvar, step))


def read_snapshot(path, names, shape):
    """(header dict, {name: array [NX, NY, NZ]})."""
    t_save, t_lastsnap, t_lastvar = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    step = ctypes.c_int()
    _check(lib().spkio_read_header(path.encode(), ctypes.byref(t_save), ctypes.byref(t_lastsnap),
                                   ctypes.byref(t_lastvar), ctypes.byref(step)))
    out = {}
    for n in names:
        a = np.empty(shape, dtype=np.float64)
        _check(lib().spkio_read_field(path.encode(), n.encode(), a.ctypes.data, a.size))
        out[n] = a
    hdr = dict(t_save=t_save.value, t_lastsnap=t_lastsnap.value, t_lastvar=t_lastvar.value, step=step.value)
    return hdr, out


def last_snapshot(data_dir):
    """Index of the last of the consecutive snapNNNN.h5 files, -1 if none."""
    return lib().spkio_last_snapshot(data_dir.encode())
