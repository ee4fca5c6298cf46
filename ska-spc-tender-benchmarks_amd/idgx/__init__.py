"""Python binding of libidgx (the MI355X IDG gridding/degridding library) over its C ABI (include/idgx.h).

Device buffers are passed as raw device pointers; torch tensors on a ROCm device are accepted anywhere a device
pointer is expected (their data_ptr() is used), torch being only the allocator and stream provider here.  The
library itself is HIP code: nothing in this module computes gridding results, and loading fails loudly when the
built library is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libidgx.so")
HEADER = os.path.join(ROOT, "include", "idgx.h")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_INVALID_METADATA = 2
ERR_UNSUPPORTED = 3
ERR_DEVICE = 4
ERR_OUT_OF_MEMORY = 5

PLAIN_FP32 = 0
MIXED_BF16X3 = 1
MIXED_FP16X3 = 2
MIXED_BF16 = 3
MIXED_FP16 = 4
PRECISIONS = {"fp32": PLAIN_FP32, "bf16x3": MIXED_BF16X3, "fp16x3": MIXED_FP16X3, "bf16": MIXED_BF16,
              "fp16": MIXED_FP16}

ATERM_ADJOINT = 0
ATERM_REFERENCE = 1

FLAG_ACCURATE_SINCOS = 1
FLAG_PHASE_MFMA = 2

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


class Plan(ctypes.Structure):
    """idgx_plan (include/idgx.h)."""

    _fields_ = [
        ("nr_subgrids", ctypes.c_int32),
        ("subgrid_size", ctypes.c_int32),
        ("grid_size", ctypes.c_int32),
        ("nr_channels", ctypes.c_int32),
        ("nr_stations", ctypes.c_int32),
        ("nr_aterm_slots", ctypes.c_int32),
        ("nr_rows", ctypes.c_int64),
        ("max_nr_timesteps", ctypes.c_int32),
        ("image_size", ctypes.c_float),
        ("w_step_in_lambda", ctypes.c_float),
        ("precision", ctypes.c_int32),
        ("aterm_convention", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


FFT_FORWARD = 0
FFT_BACKWARD = 1


class Timing(ctypes.Structure):
    """idgx_timing: per-iteration microseconds and MB/s, as the reference's benchmark_result."""
    _fields_ = [("in_transfer_time", ctypes.c_double), ("out_transfer_time", ctypes.c_double),
                ("in_bandwidth", ctypes.c_double), ("out_bandwidth", ctypes.c_double),
                ("compute_time", ctypes.c_double)]


class Planner(ctypes.Structure):
    """idgx_planner (include/idgx.h)."""

    _fields_ = [("subgrid_size", ctypes.c_int32), ("grid_size", ctypes.c_int32), ("support", ctypes.c_int32),
                ("nr_times", ctypes.c_int32), ("nr_channels", ctypes.c_int32), ("aterm_interval", ctypes.c_int32),
                ("image_size", ctypes.c_double), ("w_step_in_lambda", ctypes.c_double)]


class AdderPlan(ctypes.Structure):
    """idgx_adder_plan (include/idgx.h)."""

    _fields_ = [("tile_size", ctypes.c_int32), ("nr_tiles", ctypes.c_int32), ("color_begin", ctypes.c_int32 * 5),
                ("z_min", ctypes.c_int32), ("nr_layers", ctypes.c_int32)]


SHAPE_TILES_PER_CLASS = -2  # gridder single launch: every item with its class's tiles per wave
STAGE_GRID = 1
STAGE_DEGRID = 2


class BatchClass(ctypes.Structure):
    """idgx_batch_class (include/idgx.h)."""

    _fields_ = [("plan", Plan), ("h_metadata", ctypes.c_void_p), ("h_channels", ctypes.c_void_p),
                ("h_vis_offset", ctypes.c_void_p), ("nr_vis_index", ctypes.c_int64)]


class BatchClassData(ctypes.Structure):
    _fields_ = [("d_metadata", ctypes.c_void_p), ("d_channels", ctypes.c_void_p), ("d_vis_index", ctypes.c_void_p),
                ("d_vis_offset", ctypes.c_void_p), ("d_spheroidal", ctypes.c_void_p), ("d_aterms", ctypes.c_void_p),
                ("d_subgrids", ctypes.c_void_p)]


class BatchData(ctypes.Structure):
    _fields_ = [("d_uvw", ctypes.c_void_p), ("d_wavenumbers", ctypes.c_void_p), ("d_visibilities", ctypes.c_void_p),
                ("d_visibilities_out", ctypes.c_void_p), ("cls", BatchClassData * 4)]


class IdgxError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what}: {strerror(status)} ({status})")
        self.status = status


_lib = None


def build():
    """Compile libidgx for gfx950 in-tree (make -C <package>)."""
    subprocess.check_call(["make", "-s", "-C", PKG, "-j8"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libidgx is not built ({LIB_PATH} missing): run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.idgx_check_plan.argtypes = [ctypes.POINTER(Plan), vp, vp]
        L.idgx_grid.argtypes = [ctypes.POINTER(Plan)] + [vp] * 9
        L.idgx_host_grid.argtypes = [ctypes.POINTER(Plan)] + [vp] * 8 + [ctypes.c_int, ctypes.POINTER(Timing)]
        L.idgx_host_degrid.argtypes = [ctypes.POINTER(Plan)] + [vp] * 8 + [ctypes.c_int, ctypes.POINTER(Timing)]
        L.idgx_degrid.argtypes = [ctypes.POINTER(Plan)] + [vp] * 9
        L.idgx_subgrid_fft.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, vp]
        L.idgx_adder_plan_build.argtypes = [ctypes.POINTER(Plan), vp, ctypes.POINTER(AdderPlan), vp, vp]
        L.idgx_add_subgrids.argtypes = [ctypes.POINTER(Plan), ctypes.POINTER(AdderPlan)] + [vp] * 6
        L.idgx_adder_plan_build_layered.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(AdderPlan), vp, vp]
        L.idgx_split_subgrids_layered.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.idgx_grid_fft.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int, vp]
        L.idgx_grid_fft_noalloc.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int, vp]
        L.idgx_grid_fft_box.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int32), vp]
        L.idgx_grid_box.argtypes = [ctypes.POINTER(Plan), vp, ctypes.POINTER(ctypes.c_int32)]
        L.idgx_grid_clear_box.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), vp]
        L.idgx_fft2d.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int, vp]
        L.idgx_fft2d_own_kernels.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.idgx_plan_workspace_bytes.argtypes = [ctypes.c_int64]
        L.idgx_plan_workspace_bytes.restype = ctypes.c_size_t
        L.idgx_plan_segments.argtypes = [ctypes.POINTER(Planner)] + [vp] * 7 + [ctypes.c_int64, vp, vp, vp,
                                                                                ctypes.c_int64,
                                                                                ctypes.POINTER(ctypes.c_int64), vp]
        L.idgx_wstack_to_image.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.idgx_wstack_from_image.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.idgx_grid_to_image.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int32), vp, vp]
        L.idgx_image_to_grid.argtypes = [ctypes.POINTER(Plan), vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int32), vp, vp]
        L.idgx_split_subgrids.argtypes = [ctypes.POINTER(Plan)] + [vp] * 4
        L.idgx_batch_create.argtypes = [ctypes.POINTER(BatchClass), ctypes.c_int, ctypes.POINTER(vp)]
        L.idgx_batch_destroy.argtypes = [vp]
        L.idgx_batch_destroy.restype = None
        L.idgx_batch_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.idgx_batch_get_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.idgx_batch_set_class_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.idgx_batch_get_class_shape.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]
        L.idgx_batch_set_single_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.idgx_batch_last_launches.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.idgx_batch_get_single_launch.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.idgx_batch_shape_info.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
        L.idgx_batch_plain_shape_info.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 2
        L.idgx_plain_kernel_shape.argtypes = [ctypes.POINTER(Plan), ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 2
        L.idgx_batch_autotune.argtypes = [vp, ctypes.POINTER(BatchData), ctypes.c_int, vp,
                                          ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.idgx_batch_grid.argtypes = [vp, ctypes.POINTER(BatchData), vp]
        L.idgx_batch_degrid.argtypes = [vp, ctypes.POINTER(BatchData), vp]
        L.idgx_batch_capture.argtypes = [vp, ctypes.POINTER(BatchData), ctypes.c_int, vp]
        L.idgx_batch_launch.argtypes = [vp, vp]
        L.idgx_batch_items.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 4
        L.idgx_strerror.restype = ctypes.c_char_p
        L.idgx_strerror.argtypes = [ctypes.c_int]
        L.idgx_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def strerror(status):
    return lib().idgx_strerror(int(status)).decode()


def exported_functions():
    """Names of the functions include/idgx.h declares."""
    import re

    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(idgx_\w+)\s*\(", text, re.M)))


def _dptr(x):
    """Device pointer of a torch tensor, an int address, or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    return ctypes.c_void_p(x.data_ptr())


def _hptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def make_plan(*, nr_subgrids, subgrid_size, grid_size, nr_channels, nr_stations, nr_aterm_slots, nr_rows,
              max_nr_timesteps, image_size, w_step=0.0, precision=PLAIN_FP32, aterm_convention=ATERM_ADJOINT,
              flags=0):
    if isinstance(precision, str):
        precision = PRECISIONS[precision]
    return Plan(int(nr_subgrids), int(subgrid_size), int(grid_size), int(nr_channels), int(nr_stations),
                int(nr_aterm_slots), int(nr_rows), int(max_nr_timesteps), float(image_size), float(w_step),
                int(precision), int(aterm_convention), int(flags))


def check_plan(plan, metadata, channels=None):
    """Host-side validation of a plan against its metadata (numpy arrays); returns the status code."""
    assert metadata.dtype == METADATA_DTYPE and metadata.flags.c_contiguous
    if channels is not None:
        assert channels.dtype == CHANNELS_DTYPE and channels.flags.c_contiguous
    return lib().idgx_check_plan(ctypes.byref(plan), _hptr(metadata), _hptr(channels))


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def grid(plan, d_metadata, d_channels, d_uvw, d_wavenumbers, d_visibilities, d_spheroidal, d_aterms, d_subgrids,
         stream=None):
    """idgx_grid on device buffers; raises IdgxError on a non-zero status."""
    rc = lib().idgx_grid(ctypes.byref(plan), _dptr(d_metadata), _dptr(d_channels), _dptr(d_uvw),
                         _dptr(d_wavenumbers), _dptr(d_visibilities), _dptr(d_spheroidal), _dptr(d_aterms),
                         _dptr(d_subgrids), _stream(stream))
    if rc != OK:
        raise IdgxError(rc, "idgx_grid")


def degrid(plan, d_metadata, d_channels, d_uvw, d_wavenumbers, d_subgrids, d_spheroidal, d_aterms, d_visibilities,
           stream=None):
    """idgx_degrid on device buffers; raises IdgxError on a non-zero status."""
    rc = lib().idgx_degrid(ctypes.byref(plan), _dptr(d_metadata), _dptr(d_channels), _dptr(d_uvw),
                           _dptr(d_wavenumbers), _dptr(d_subgrids), _dptr(d_spheroidal), _dptr(d_aterms),
                           _dptr(d_visibilities), _stream(stream))
    if rc != OK:
        raise IdgxError(rc, "idgx_degrid")


def _host_inputs(w):
    arrays = {k: np.ascontiguousarray(w[k]) for k in ("metadata", "uvw", "wavenumbers", "spheroidal", "aterms")}
    arrays["channels"] = None if w.get("channels") is None else np.ascontiguousarray(w["channels"])
    assert arrays["metadata"].dtype == METADATA_DTYPE
    assert arrays["uvw"].dtype == np.float32 and arrays["wavenumbers"].dtype == np.float32
    assert arrays["spheroidal"].dtype == np.float32 and arrays["aterms"].dtype == np.complex64
    return arrays


def host_grid(plan, w, iterations=1):
    """idgx_host_grid: the gridder from host arrays of workload w (metadata, uvw, wavenumbers, visibilities,
    spheroidal, aterms); returns (subgrids [n][4][S][S] complex64, Timing)."""
    h = _host_inputs(w)
    vis = np.ascontiguousarray(w["visibilities"], dtype=np.complex64)
    S = plan.subgrid_size
    out = np.zeros((max(plan.nr_subgrids, 1), 4, S, S), np.complex64)
    t = Timing()
    rc = lib().idgx_host_grid(ctypes.byref(plan), _hptr(h["metadata"]), _hptr(h["channels"]), _hptr(h["uvw"]),
                              _hptr(h["wavenumbers"]), _hptr(vis), _hptr(h["spheroidal"]), _hptr(h["aterms"]), _hptr(out), int(iterations),
                              ctypes.byref(t))
    if rc != OK:
        raise IdgxError(rc, "idgx_host_grid")
    return out[: plan.nr_subgrids], t


def host_degrid(plan, w, subgrids, iterations=1):
    """idgx_host_degrid: the degridder from host arrays; returns (visibilities [rows][C][4] complex64, Timing)."""
    h = _host_inputs(w)
    sub = np.ascontiguousarray(subgrids, dtype=np.complex64)
    out = np.zeros((plan.nr_rows, plan.nr_channels, 4), np.complex64)
    t = Timing()
    rc = lib().idgx_host_degrid(ctypes.byref(plan), _hptr(h["metadata"]), _hptr(h["channels"]), _hptr(h["uvw"]),
                                _hptr(h["wavenumbers"]), _hptr(sub), _hptr(h["spheroidal"]), _hptr(h["aterms"]),
                                _hptr(out), int(iterations), ctypes.byref(t))
    if rc != OK:
        raise IdgxError(rc, "idgx_host_degrid")
    return out, t


def subgrid_fft(plan, d_subgrids, direction, stream=None):
    """idgx_subgrid_fft: in-place forward (image -> uv) or backward transform of every subgrid."""
    rc = lib().idgx_subgrid_fft(ctypes.byref(plan), _dptr(d_subgrids), int(direction), _stream(stream))
    if rc != OK:
        raise IdgxError(rc, "idgx_subgrid_fft")


def adder_plan_build(plan, metadata):
    """Host-side conflict-free adder order: returns (AdderPlan, order int32[n], tile_start int32[n + 1])."""
    n = int(plan.nr_subgrids)
    ap = AdderPlan()
    order = np.zeros(max(n, 1), np.int32)
    tile_start = np.zeros(n + 1, np.int32)
    rc = lib().idgx_adder_plan_build(ctypes.byref(plan), _hptr(np.ascontiguousarray(metadata)), ctypes.byref(ap),
                                     _hptr(order), _hptr(tile_start))
    if rc != OK:
        raise IdgxError(rc, "idgx_adder_plan_build")
    return ap, order[:n], tile_start[:ap.nr_tiles + 1]


def adder_plan_build_layered(plan, metadata, z_min, nr_layers):
    """idgx_adder_plan_build_layered: the adder order for w-stacking onto nr_layers grids (layer = z - z_min)."""
    n = int(plan.nr_subgrids)
    ap = AdderPlan()
    order = np.zeros(max(n, 1), np.int32)
    tile_start = np.zeros(n + 1, np.int32)
    rc = lib().idgx_adder_plan_build_layered(ctypes.byref(plan), _hptr(np.ascontiguousarray(metadata)),
                                             int(z_min), int(nr_layers), ctypes.byref(ap), _hptr(order),
                                             _hptr(tile_start))
    if rc != OK:
        raise IdgxError(rc, "idgx_adder_plan_build_layered")
    return ap, order[:n], tile_start[:ap.nr_tiles + 1]


def add_subgrids(plan, adder_plan, d_metadata, d_order, d_tile_start, d_subgrids, d_grid, stream=None):
    rc = lib().idgx_add_subgrids(ctypes.byref(plan), ctypes.byref(adder_plan), _dptr(d_metadata), _dptr(d_order),
                                 _dptr(d_tile_start), _dptr(d_subgrids), _dptr(d_grid), _stream(stream))
    if rc != OK:
        raise IdgxError(rc, "idgx_add_subgrids")


def split_subgrids(plan, d_metadata, d_grid, d_subgrids, stream=None):
    rc = lib().idgx_split_subgrids(ctypes.byref(plan), _dptr(d_metadata), _dptr(d_grid), _dptr(d_subgrids),
                                   _stream(stream))
    if rc != OK:
        raise IdgxError(rc, "idgx_split_subgrids")


def split_subgrids_layered(plan, d_metadata, z_min, nr_layers, d_grid, d_subgrids, stream=None):
    _check(lib().idgx_split_subgrids_layered(ctypes.byref(plan), _dptr(d_metadata), int(z_min), int(nr_layers),
                                             _dptr(d_grid), _dptr(d_subgrids), _stream(stream)),
           "idgx_split_subgrids_layered")


def grid_fft(plan, d_grid, nr_layers, direction, stream=None):
    """idgx_grid_fft: in-place centred 2-D transform of a master grid complex64 [nr_layers][4][G][G]."""
    _check(lib().idgx_grid_fft(ctypes.byref(plan), _dptr(d_grid), int(nr_layers), int(direction), _stream(stream)),
           "idgx_grid_fft")


def grid_fft_noalloc(plan, d_grid, nr_layers, direction, stream=None):
    """idgx_grid_fft without any device allocation (G = 1024, 2048 or 4096; the scratch is a fixed array)."""
    _check(lib().idgx_grid_fft_noalloc(ctypes.byref(plan), _dptr(d_grid), nr_layers, direction, _stream(stream)),
           "idgx_grid_fft_noalloc")


def grid_box(plan, metadata):
    """idgx_grid_box: (x0, x1, y0, y1), the bounding box of the plan's subgrids on the master grid."""
    assert metadata.dtype == METADATA_DTYPE and metadata.flags.c_contiguous
    box = (ctypes.c_int32 * 4)()
    _check(lib().idgx_grid_box(ctypes.byref(plan), _hptr(metadata), box), "idgx_grid_box")
    return tuple(box)


def grid_fft_box(plan, d_grid, nr_layers, direction, box, stream=None):
    """idgx_grid_fft_box: the master-grid transform restricted to box = (x0, x1, y0, y1): BACKWARD takes the grid
    as zero outside the box, FORWARD produces only the box's cells."""
    b = (ctypes.c_int32 * 4)(*[int(v) for v in box])
    _check(lib().idgx_grid_fft_box(ctypes.byref(plan), _dptr(d_grid), int(nr_layers), int(direction), b,
                                   _stream(stream)), "idgx_grid_fft_box")


def grid_clear_box(plan, d_grid, nr_layers, box, stream=None):
    """idgx_grid_clear_box: zero the cells a boxed BACKWARD transform reads."""
    b = (ctypes.c_int32 * 4)(*[int(v) for v in box])
    _check(lib().idgx_grid_clear_box(ctypes.byref(plan), _dptr(d_grid), int(nr_layers), b, _stream(stream)),
           "idgx_grid_clear_box")


FFT_FP32 = 0
FFT_FP64 = 1


def fft2d(d_in, d_out, nx, ny, batch, sign, precision=FFT_FP32, stream=None):
    """idgx_fft2d: plain batched 2-D DFT of complex64 (or complex128) planes [batch][ny][nx], unnormalised."""
    _check(lib().idgx_fft2d(_dptr(d_in), _dptr(d_out), int(nx), int(ny), int(batch), int(sign), int(precision),
                            _stream(stream)), "idgx_fft2d")


def fft2d_own_kernels(nx, ny, precision=FFT_FP32):
    """True when idgx_fft2d runs this library's own kernels for the shape (False: rocFFT)."""
    r = lib().idgx_fft2d_own_kernels(int(nx), int(ny), int(precision))
    if r < 0:
        raise IdgxError(-r, "idgx_fft2d_own_kernels")
    return r == 1


def wstack_to_image(plan, d_layers, z_min, nr_layers, d_image, stream=None):
    _check(lib().idgx_wstack_to_image(ctypes.byref(plan), _dptr(d_layers), int(z_min), int(nr_layers),
                                      _dptr(d_image), _stream(stream)), "idgx_wstack_to_image")


def wstack_from_image(plan, d_image, z_min, nr_layers, d_layers, stream=None):
    _check(lib().idgx_wstack_from_image(ctypes.byref(plan), _dptr(d_image), int(z_min), int(nr_layers),
                                        _dptr(d_layers), _stream(stream)), "idgx_wstack_from_image")


def grid_to_image(plan, d_grid, z_min, nr_layers, box, d_image, stream=None):
    """idgx_grid_to_image: the boxed transform of the layered grid to the image and the w-screens with the sum over
    layers in one call (d_grid is left undefined)."""
    b = (ctypes.c_int32 * 4)(*[int(v) for v in box])
    _check(lib().idgx_grid_to_image(ctypes.byref(plan), _dptr(d_grid), int(z_min), int(nr_layers), b,
                                    _dptr(d_image), _stream(stream)), "idgx_grid_to_image")


def image_to_grid(plan, d_image, z_min, nr_layers, box, d_grid, stream=None):
    """idgx_image_to_grid: the w-screens from the image and the boxed transform to the layered grid in one call
    (the box's cells of every layer; the rest of the grid undefined)."""
    b = (ctypes.c_int32 * 4)(*[int(v) for v in box])
    _check(lib().idgx_image_to_grid(ctypes.byref(plan), _dptr(d_image), int(z_min), int(nr_layers), b,
                                    _dptr(d_grid), _stream(stream)), "idgx_image_to_grid")


def plan_segments_device(uvw, frequencies, station1, station2, seg_bl, seg_t0, seg_t1, *, subgrid_size, grid_size,
                         image_size, nr_times, support=8, aterm_interval=0, w_step=0.0, device="cuda", stream=None):
    """idgx_plan_segments: the subgrid plan computed on the GPU from host arrays (uvw float32 [rows][3] in metres,
    frequencies float64 [C]); returns host (metadata, channels) arrays, bit-identical to
    observation.plan_segments for segments sorted by (baseline, t0)."""
    import torch

    def up(a, dt):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=dt).reshape(-1)).to(device)

    nseg = int(np.asarray(seg_bl).size)
    pl = Planner(subgrid_size=int(subgrid_size), grid_size=int(grid_size), support=int(support),
                 nr_times=int(nr_times), nr_channels=int(np.asarray(frequencies).size),
                 aterm_interval=int(aterm_interval or 0), image_size=float(image_size),
                 w_step_in_lambda=float(w_step or 0.0))
    d = dict(uvw=up(uvw, np.float32), f=up(frequencies, np.float64),
             s1=up(np.asarray(station1, np.uint32).view(np.int32), np.int32),
             s2=up(np.asarray(station2, np.uint32).view(np.int32), np.int32),
             bl=up(seg_bl, np.int32), t0=up(seg_t0, np.int32), t1=up(seg_t1, np.int32))
    ws = torch.zeros(max(1, int(lib().idgx_plan_workspace_bytes(nseg))), dtype=torch.uint8, device=device)
    cap = nseg * pl.nr_channels
    md = torch.zeros(max(1, cap) * METADATA_DTYPE.itemsize, dtype=torch.uint8, device=device)
    ch = torch.zeros(max(1, cap) * CHANNELS_DTYPE.itemsize, dtype=torch.uint8, device=device)
    n = ctypes.c_int64(0)
    _check(lib().idgx_plan_segments(ctypes.byref(pl), _dptr(d["uvw"]), _dptr(d["f"]), _dptr(d["s1"]), _dptr(d["s2"]),
                                    _dptr(d["bl"]), _dptr(d["t0"]), _dptr(d["t1"]), nseg, _dptr(ws), _dptr(md),
                                    _dptr(ch), cap, ctypes.byref(n), _stream(stream)), "idgx_plan_segments")
    k = int(n.value)
    mdh = md.cpu().numpy().view(METADATA_DTYPE)[:k].copy()
    chh = ch.cpu().numpy().view(CHANNELS_DTYPE)[:k].copy()
    return mdh, chh


def plain_kernel_shape(plan, has_channel_windows):
    """(gridder pixels per thread, degridder channels per thread) of the plain FP32 one-shot kernels for `plan`."""
    px, nch = ctypes.c_int(), ctypes.c_int()
    _check(lib().idgx_plain_kernel_shape(ctypes.byref(plan), int(bool(has_channel_windows)), ctypes.byref(px),
                                         ctypes.byref(nch)), "idgx_plain_kernel_shape")
    return px.value, nch.value


def _check(rc, what):
    if rc != OK:
        raise IdgxError(rc, what)


class Batch:
    """A plan's persistent work-queue execution (idgx_batch_*).  `classes` is a list of (plan, metadata, channels,
    vis_offset) with host numpy arrays (channels / vis_offset may be None)."""

    def __init__(self, classes):
        arr = (BatchClass * len(classes))()
        self._keep = []
        for k, (plan, md, ch, off) in enumerate(classes):
            md = np.ascontiguousarray(md)
            ch = None if ch is None else np.ascontiguousarray(ch)
            off = None if off is None else np.ascontiguousarray(off, dtype=np.int64)
            self._keep += [md, ch, off]
            arr[k].plan = plan
            arr[k].h_metadata = md.ctypes.data
            arr[k].h_channels = None if ch is None else ch.ctypes.data
            arr[k].h_vis_offset = None if off is None else off.ctypes.data
            arr[k].nr_vis_index = 0 if off is None else int(off[-1])
        h = ctypes.c_void_p()
        _check(lib().idgx_batch_create(arr, len(classes), ctypes.byref(h)), "idgx_batch_create")
        self.handle = h
        self.nr_classes = len(classes)
        self.precision = int(classes[0][0].precision)

    def close(self):
        if self.handle:
            lib().idgx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def data(d_uvw, d_wavenumbers, d_visibilities, d_visibilities_out, classes):
        """classes: list of dicts with d_metadata, d_channels, d_vis_index, d_vis_offset, d_spheroidal, d_aterms,
        d_subgrids (device tensors or None)."""
        bd = BatchData()
        bd.d_uvw, bd.d_wavenumbers = _dptr(d_uvw), _dptr(d_wavenumbers)
        bd.d_visibilities, bd.d_visibilities_out = _dptr(d_visibilities), _dptr(d_visibilities_out)
        for k, c in enumerate(classes):
            for f in ("d_metadata", "d_channels", "d_vis_index", "d_vis_offset", "d_spheroidal", "d_aterms",
                      "d_subgrids"):
                setattr(bd.cls[k], f, _dptr(c.get(f)))
        return bd

    def grid(self, data, stream=None):
        _check(lib().idgx_batch_grid(self.handle, ctypes.byref(data), _stream(stream)), "idgx_batch_grid")

    def degrid(self, data, stream=None):
        _check(lib().idgx_batch_degrid(self.handle, ctypes.byref(data), _stream(stream)), "idgx_batch_degrid")

    def set_shape(self, gridder_shape, degridder_shape):
        _check(lib().idgx_batch_set_shape(self.handle, gridder_shape, degridder_shape), "idgx_batch_set_shape")

    def shape(self):
        g, d = ctypes.c_int(), ctypes.c_int()
        _check(lib().idgx_batch_get_shape(self.handle, ctypes.byref(g), ctypes.byref(d)), "idgx_batch_get_shape")
        return g.value, d.value

    def set_class_shape(self, cls, gridder_shape, degridder_shape):
        _check(lib().idgx_batch_set_class_shape(self.handle, cls, gridder_shape, degridder_shape),
               "idgx_batch_set_class_shape")

    def class_shape(self, cls):
        g, d = ctypes.c_int(), ctypes.c_int()
        _check(lib().idgx_batch_get_class_shape(self.handle, cls, ctypes.byref(g), ctypes.byref(d)),
               "idgx_batch_get_class_shape")
        return g.value, d.value

    def set_single_launch(self, gridder_shape, degridder_shape):
        """One launch per direction over the merged queue of all classes (shape -1: per-class launches)."""
        _check(lib().idgx_batch_set_single_launch(self.handle, gridder_shape, degridder_shape),
               "idgx_batch_set_single_launch")

    def single_launch(self):
        g, d = ctypes.c_int(), ctypes.c_int()
        _check(lib().idgx_batch_get_single_launch(self.handle, ctypes.byref(g), ctypes.byref(d)),
               "idgx_batch_get_single_launch")
        return g.value, d.value

    def last_launches(self):
        """(gridder, degridder) persistent launches the latest grid / degrid call issued."""
        g, d = ctypes.c_int(), ctypes.c_int()
        _check(lib().idgx_batch_last_launches(self.handle, ctypes.byref(g), ctypes.byref(d)),
               "idgx_batch_last_launches")
        return g.value, d.value

    @staticmethod
    def shapes():
        """Shapes of the mixed forms: (tiles per wave, LDS batch, phases on the matrix core)."""
        out = []
        for i in range(lib().idgx_batch_shape_count()):
            t, b, pm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _check(lib().idgx_batch_shape_info(i, ctypes.byref(t), ctypes.byref(b), ctypes.byref(pm)),
                   "idgx_batch_shape_info")
            out.append((t.value, b.value, pm.value))
        return out

    @staticmethod
    def plain_shapes():
        """Shapes of plain FP32 batches: (gridder pixels per thread, degridder visibilities per thread)."""
        out = []
        for i in range(lib().idgx_batch_plain_shape_count()):
            px, nv = ctypes.c_int(), ctypes.c_int()
            _check(lib().idgx_batch_plain_shape_info(i, ctypes.byref(px), ctypes.byref(nv)),
                   "idgx_batch_plain_shape_info")
            out.append((px.value, nv.value))
        return out

    def shape_table(self):
        """The shape table this batch's shape indices refer to."""
        return Batch.plain_shapes() if self.precision == PLAIN_FP32 else Batch.shapes()

    def autotune(self, data, repeats=2, stream=None):
        """Times every shape per class and direction on `data` (overwriting its outputs) and keeps each class's
        fastest.  A direction in single-launch mode then also times its per-class launches together and every shape
        of the merged launch, and keeps the fastest mode (possibly per-class, as single_launch() then reports with
        -1).  Returns (gridder ms, degridder ms) of what was kept."""
        g, d = ctypes.c_float(), ctypes.c_float()
        _check(lib().idgx_batch_autotune(self.handle, ctypes.byref(data), repeats, _stream(stream), ctypes.byref(g),
                                         ctypes.byref(d)), "idgx_batch_autotune")
        return g.value, d.value

    def capture(self, data, stages, stream):
        _check(lib().idgx_batch_capture(self.handle, ctypes.byref(data), stages, _stream(stream)), "idgx_batch_capture")

    def launch(self, stream):
        _check(lib().idgx_batch_launch(self.handle, _stream(stream)), "idgx_batch_launch")

    def items(self):
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().idgx_batch_items(self.handle, *[ctypes.byref(x) for x in v]), "idgx_batch_items")
        return tuple(x.value for x in v)
