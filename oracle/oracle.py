"""ctypes wrapper around oracle/liboracle.so.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the checker or as the
timed CPU restatement; the product path never does.  See oracle.h for what each function restates and which
reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

METADATA_DTYPE = np.dtype(
    [
        ("baseline_offset", "<i4"),
        ("time_offset", "<i4"),
        ("nr_timesteps", "<i4"),
        ("aterm_index", "<i4"),
        ("station1", "<u4"),
        ("station2", "<u4"),
        ("x", "<i4"),
        ("y", "<i4"),
        ("z", "<i4"),
    ]
)
CHANNELS_DTYPE = np.dtype([("channel_begin", "<i4"), ("nr_channels", "<i4")])

ATERM_ADJOINT = 0
ATERM_REFERENCE = 1


class _Params(ctypes.Structure):
    _fields_ = [
        ("nr_subgrids", ctypes.c_int32),
        ("subgrid_size", ctypes.c_int32),
        ("grid_size", ctypes.c_int32),
        ("nr_channels", ctypes.c_int32),
        ("nr_stations", ctypes.c_int32),
        ("image_size", ctypes.c_float),
        ("w_step_in_lambda", ctypes.c_float),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    assert a.dtype == np.dtype(dtype), (a.dtype, dtype)
    assert a.flags.c_contiguous
    return a


def _params(nr_subgrids, subgrid_size, grid_size, nr_channels, nr_stations, image_size, w_step):
    return _Params(nr_subgrids, subgrid_size, grid_size, nr_channels, nr_stations, image_size, w_step)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with {rc}")


def grid_fp32(metadata, uvw, wavenumbers, visibilities, spheroidal, aterms, *, subgrid_size, grid_size,
              nr_stations, image_size, w_step=0.0, channels=None, nthreads=0):
    """FP32 restatement of the reference CPU gridder; returns complex64[n][4][S][S]."""
    n = metadata.shape[0]
    C = wavenumbers.shape[0]
    S = subgrid_size
    out = np.zeros((n, 4, S, S), np.complex64)
    p = _params(n, S, grid_size, C, nr_stations, image_size, w_step)
    rc = lib().oracle_grid_fp32(ctypes.byref(p), _ptr(_c(metadata, METADATA_DTYPE)),
                                _ptr(None if channels is None else _c(channels, CHANNELS_DTYPE)),
                                _ptr(_c(uvw, np.float32)), _ptr(_c(wavenumbers, np.float32)),
                                _ptr(_c(visibilities, np.complex64)), _ptr(_c(spheroidal, np.float32)),
                                _ptr(_c(aterms, np.complex64)), _ptr(out), ctypes.c_int(nthreads))
    _check(rc, "grid_fp32")
    return out


def degrid_fp32(metadata, uvw, wavenumbers, subgrids, spheroidal, aterms, *, grid_size, nr_stations, image_size,
                nr_rows, w_step=0.0, channels=None, convention=ATERM_ADJOINT, nthreads=0, out=None):
    """FP32 restatement of the reference CPU degridder; returns complex64[nr_rows][C][4]."""
    n = metadata.shape[0]
    C = wavenumbers.shape[0]
    S = subgrids.shape[-1]
    if out is None:
        out = np.zeros((nr_rows, C, 4), np.complex64)
    p = _params(n, S, grid_size, C, nr_stations, image_size, w_step)
    rc = lib().oracle_degrid_fp32(ctypes.byref(p), _ptr(_c(metadata, METADATA_DTYPE)),
                                  _ptr(None if channels is None else _c(channels, CHANNELS_DTYPE)),
                                  _ptr(_c(uvw, np.float32)), _ptr(_c(wavenumbers, np.float32)),
                                  _ptr(_c(subgrids, np.complex64)), _ptr(_c(spheroidal, np.float32)),
                                  _ptr(_c(aterms, np.complex64)), _ptr(out), ctypes.c_int(convention),
                                  ctypes.c_int(nthreads))
    _check(rc, "degrid_fp32")
    return out


def grid_fp64(metadata, uvw, wavenumbers, visibilities, spheroidal, aterms, *, subgrid_size, grid_size,
              nr_stations, image_size, w_step=0.0, channels=None, nthreads=0):
    """FP64 oracle gridder; returns complex128[n][4][S][S]."""
    n = metadata.shape[0]
    C = wavenumbers.shape[0]
    S = subgrid_size
    out = np.zeros((n, 4, S, S), np.complex128)
    p = _params(n, S, grid_size, C, nr_stations, image_size, w_step)
    rc = lib().oracle_grid_fp64(ctypes.byref(p), _ptr(_c(metadata, METADATA_DTYPE)),
                                _ptr(None if channels is None else _c(channels, CHANNELS_DTYPE)),
                                _ptr(_c(uvw, np.float32)), _ptr(_c(wavenumbers, np.float32)),
                                _ptr(_c(visibilities, np.complex64)), _ptr(_c(spheroidal, np.float32)),
                                _ptr(_c(aterms, np.complex64)), _ptr(out), ctypes.c_int(nthreads))
    _check(rc, "grid_fp64")
    return out


def degrid_fp64(metadata, uvw, wavenumbers, subgrids, spheroidal, aterms, *, grid_size, nr_stations, image_size,
                nr_rows, w_step=0.0, channels=None, nthreads=0, out=None):
    """FP64 oracle degridder (adjoint A-term); returns complex128[nr_rows][C][4]."""
    n = metadata.shape[0]
    C = wavenumbers.shape[0]
    S = subgrids.shape[-1]
    if out is None:
        out = np.zeros((nr_rows, C, 4), np.complex128)
    p = _params(n, S, grid_size, C, nr_stations, image_size, w_step)
    rc = lib().oracle_degrid_fp64(ctypes.byref(p), _ptr(_c(metadata, METADATA_DTYPE)),
                                  _ptr(None if channels is None else _c(channels, CHANNELS_DTYPE)),
                                  _ptr(_c(uvw, np.float32)), _ptr(_c(wavenumbers, np.float32)),
                                  _ptr(_c(subgrids, np.complex64)), _ptr(_c(spheroidal, np.float32)),
                                  _ptr(_c(aterms, np.complex64)), _ptr(out), ctypes.c_int(nthreads))
    _check(rc, "degrid_fp64")
    return out


def dft2d_fp64(data, sign, nthreads=0):
    """Brute-force batched 2-D DFT over the last two axes (sign -1 forward, +1 inverse unscaled)."""
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    S = a.shape[-1]
    assert a.shape[-2] == S
    batch = a.size // (S * S)
    _check(lib().oracle_dft2d_fp64(_ptr(a), ctypes.c_int(batch), ctypes.c_int(S), ctypes.c_int(sign),
                                   ctypes.c_int(nthreads)), "dft2d")
    return a


def adder_fp64(metadata, subgrids, grid_size):
    n, _, S, _ = subgrids.shape
    grid = np.zeros((4, grid_size, grid_size), np.complex128)
    _check(lib().oracle_adder_fp64(_ptr(_c(metadata, METADATA_DTYPE)), ctypes.c_int(n), ctypes.c_int(S),
                                   ctypes.c_int(grid_size), _ptr(_c(subgrids, np.complex64)), _ptr(grid)), "adder")
    return grid


def splitter_fp64(metadata, grid, subgrid_size):
    n = metadata.shape[0]
    G = grid.shape[-1]
    out = np.zeros((n, 4, subgrid_size, subgrid_size), np.complex128)
    _check(lib().oracle_splitter_fp64(_ptr(_c(metadata, METADATA_DTYPE)), ctypes.c_int(n),
                                      ctypes.c_int(subgrid_size), ctypes.c_int(G),
                                      _ptr(_c(np.ascontiguousarray(grid, np.complex128), np.complex128)), _ptr(out)),
           "splitter")
    return out
