"""Host-side synthetic workloads for the gridding hot path.

reference_workload(): the reference benchmark's own inputs (srand(0) and the initialize_* routines of
gridding/src/gridding_initialization.cpp), computed by libidgx_host (C++, so glibc's rand/cos/cexpf and the
reference's float/double promotions are reproduced exactly).
"""
import ctypes
import os

import numpy as np

from . import METADATA_DTYPE, PKG

HOST_LIB_PATH = os.path.join(PKG, "lib", "libidgx_host.so")
_hlib = None


class ReferenceShape(ctypes.Structure):
    _fields_ = [
        ("nstations", ctypes.c_int32),
        ("ntimeslots", ctypes.c_int32),
        ("ntimesteps_per_subgrid", ctypes.c_int32),
        ("nchannels", ctypes.c_int32),
        ("subgrid_size", ctypes.c_int32),
        ("grid_size", ctypes.c_int32),
    ]


def host_lib():
    global _hlib
    if _hlib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"libidgx_host is not built ({HOST_LIB_PATH} missing): run __graft_entry__.build()")
        _hlib = ctypes.CDLL(HOST_LIB_PATH)
    return _hlib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def reference_workload(nstations, ntimeslots, ntimesteps_per_subgrid, nchannels, subgrid_size, grid_size):
    """The reference generator's arrays, shaped as the C ABI consumes them (rows = subgrid * T + t)."""
    nbl = nstations * (nstations - 1) // 2
    nsub = nbl * ntimeslots
    T, C, S = ntimesteps_per_subgrid, nchannels, subgrid_size
    w = {
        "uvw": np.zeros((nsub * T, 3), np.float32),
        "frequencies": np.zeros(C, np.float32),
        "wavenumbers": np.zeros(C, np.float32),
        "visibilities": np.zeros((nsub * T, C, 4), np.complex64),
        "baselines": np.zeros((nbl, 2), np.uint32),
        "spheroidal": np.zeros((S, S), np.float32),
        "aterms": np.zeros((ntimeslots, nstations, S, S, 4), np.complex64),
        "metadata": np.zeros(nsub, METADATA_DTYPE),
    }
    shape = ReferenceShape(nstations, ntimeslots, T, C, S, grid_size)
    rc = host_lib().idgx_reference_workload(ctypes.byref(shape), *(_p(w[k]) for k in (
        "uvw", "frequencies", "wavenumbers", "visibilities", "baselines", "spheroidal", "aterms", "metadata")))
    if rc != 0:
        raise ValueError(f"invalid reference shape {nstations, ntimeslots, T, C, S, grid_size}")
    w.update(nr_subgrids=nsub, subgrid_size=S, grid_size=grid_size, nr_stations=nstations,
             nr_aterm_slots=ntimeslots, image_size=np.float32(0.1), w_step=0.0, max_nr_timesteps=T, channels=None)
    return w


def reference_subgrids(nr_subgrids, subgrid_size):
    out = np.zeros((nr_subgrids, 4, subgrid_size, subgrid_size), np.complex64)
    host_lib().idgx_reference_subgrids(ctypes.c_int(nr_subgrids), ctypes.c_int(subgrid_size), _p(out))
    return out
