"""Host-side checks (CPU): the product's reference-workload generator against the golden vectors, the C ABI's
exported symbols, and plan validation error codes."""
import ctypes

import numpy as np
import pytest

import idgx
from helpers import golden_names, load_golden
from idgx import observation, workload


@pytest.mark.parametrize("name", golden_names())
def test_reference_workload_bit_exact(name):
    w = load_golden(name)
    sh = w["shape"]
    g = workload.reference_workload(sh["nstations"], sh["ntimeslots"], sh["ntimesteps_per_subgrid"], sh["nchannels"],
                                    sh["subgrid_size"], sh["grid_size"])
    for key in ("uvw", "frequencies", "wavenumbers", "visibilities", "spheroidal", "aterms", "metadata", "baselines"):
        a = np.ascontiguousarray(g[key])
        b = np.ascontiguousarray(w[key]).reshape(a.shape) if key != "metadata" else w[key]
        assert a.tobytes() == b.tobytes(), key
    ramp = workload.reference_subgrids(w["nr_subgrids"], sh["subgrid_size"])
    assert ramp.tobytes() == w["ref_degrid_subgrids"].tobytes()


def test_c_abi_exports_every_declared_symbol():
    names = idgx.exported_functions()
    assert {"idgx_grid", "idgx_degrid", "idgx_check_plan", "idgx_strerror"} <= set(names)
    lib = idgx.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert idgx.lib().idgx_version().startswith(b"idgx")


def test_metadata_layout_is_36_bytes():
    assert idgx.METADATA_DTYPE.itemsize == 36
    assert ctypes.sizeof(idgx.Plan) == 56


def _plan(w, **kw):
    args = dict(nr_subgrids=w["nr_subgrids"], subgrid_size=w["subgrid_size"], grid_size=w["grid_size"],
                nr_channels=w["wavenumbers"].shape[0], nr_stations=w["nr_stations"],
                nr_aterm_slots=w["nr_aterm_slots"], nr_rows=w["uvw"].shape[0],
                max_nr_timesteps=w["max_nr_timesteps"], image_size=w["image_size"])
    args.update(kw)
    return idgx.make_plan(**args)


def test_check_plan_accepts_valid_and_rejects_out_of_bounds_metadata():
    w = observation.make_observation("small", visibilities=False)
    md, ch = w["metadata"], w["channels"]
    assert idgx.check_plan(_plan(w), md, ch) == idgx.OK
    bad_cases = {
        "station": ("station2", w["nr_stations"]),
        "rows": ("time_offset", w["uvw"].shape[0]),
        "negative rows": ("time_offset", -1),
        "aterm slot": ("aterm_index", 1),
        "timesteps": ("nr_timesteps", w["max_nr_timesteps"] + 1),
    }
    for what, (field, value) in bad_cases.items():
        m = md.copy()
        m[field][3] = value
        assert idgx.check_plan(_plan(w), m, ch) == idgx.ERR_INVALID_METADATA, what
    c = ch.copy()
    c["channel_begin"][2] = 10
    assert idgx.check_plan(_plan(w), md, c) == idgx.ERR_INVALID_METADATA


def test_check_plan_rejects_bad_plans():
    w = observation.make_observation("small", visibilities=False)
    md, ch = w["metadata"], w["channels"]
    assert idgx.check_plan(_plan(w, subgrid_size=7), md, ch) == idgx.ERR_UNSUPPORTED
    assert idgx.check_plan(_plan(w, subgrid_size=256), md, ch) == idgx.ERR_UNSUPPORTED
    assert idgx.check_plan(_plan(w, nr_channels=0), md, ch) == idgx.ERR_INVALID_ARGUMENT
    assert idgx.check_plan(_plan(w, image_size=0.0), md, ch) == idgx.ERR_INVALID_ARGUMENT
    assert idgx.check_plan(_plan(w, precision=9), md, ch) == idgx.ERR_INVALID_ARGUMENT
    assert idgx.check_plan(_plan(w, aterm_convention=5), md, ch) == idgx.ERR_INVALID_ARGUMENT
    assert idgx.check_plan(_plan(w, flags=8), md, ch) == idgx.ERR_INVALID_ARGUMENT


def test_grid_entry_points_reject_null_arguments_without_touching_the_device():
    w = observation.make_observation("small", visibilities=False)
    p = _plan(w)
    assert idgx.lib().idgx_grid(ctypes.byref(p), *([None] * 9)) == idgx.ERR_INVALID_ARGUMENT
    assert idgx.lib().idgx_degrid(ctypes.byref(p), *([None] * 9)) == idgx.ERR_INVALID_ARGUMENT
    assert idgx.lib().idgx_grid(None, *([None] * 9)) == idgx.ERR_INVALID_ARGUMENT
    assert idgx.strerror(idgx.ERR_INVALID_METADATA).startswith("subgrid metadata")


def test_host_buffer_calls_validate_on_the_host_before_touching_the_device():
    """idgx_host_grid / idgx_host_degrid reject null buffers, a zero iteration count and out-of-bounds metadata
    with a status code, before any device call (this runs without a GPU)."""
    w = observation.make_observation("small", visibilities=False)
    p = _plan(w)
    L = idgx.lib()
    t = idgx.Timing()
    assert L.idgx_host_grid(None, *([None] * 8), 1, ctypes.byref(t)) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_host_degrid(ctypes.byref(p), *([None] * 8), 1, None) == idgx.ERR_INVALID_ARGUMENT
    md, ch = w["metadata"].copy(), w["channels"]
    uvw, wn = np.ascontiguousarray(w["uvw"]), np.ascontiguousarray(w["wavenumbers"])
    sph, at = np.ascontiguousarray(w["spheroidal"]), np.ascontiguousarray(w["aterms"])
    vis = np.zeros((w["nr_rows"], wn.size, 4), np.complex64)
    sub = np.zeros((w["nr_subgrids"], 4, w["subgrid_size"], w["subgrid_size"]), np.complex64)
    hp = idgx._hptr
    args = [hp(md), hp(ch), hp(uvw), hp(wn), hp(vis), hp(sph), hp(at), hp(sub)]
    assert L.idgx_host_grid(ctypes.byref(p), *args, 0, None) == idgx.ERR_INVALID_ARGUMENT
    md["station1"][5] = w["nr_stations"]
    assert L.idgx_host_grid(ctypes.byref(p), *args, 1, None) == idgx.ERR_INVALID_METADATA
    assert L.idgx_host_degrid(ctypes.byref(p), *args, 1, None) == idgx.ERR_INVALID_METADATA


def test_grid_box_bounds_every_subgrid_clamped_to_the_grid():
    """idgx_grid_box (host only): the bounding box of the subgrids' cells, clamped to [0, G)."""
    w = observation.make_observation("small", visibilities=False)
    p = _plan(w)
    md = w["metadata"].copy()
    S, G = w["subgrid_size"], w["grid_size"]
    box = idgx.grid_box(p, md)
    assert box == (md["x"].min(), md["x"].max() + S, md["y"].min(), md["y"].max() + S)
    md["x"][0], md["y"][1] = -5, G - 3  # overhanging subgrids: the box stops at the grid's edges
    box = idgx.grid_box(p, md)
    assert box[0] == 0 and box[3] == G
    assert idgx.grid_box(_plan(w, nr_subgrids=0), md[:0].copy()) == (0, 0, 0, 0)


def test_boxed_grid_calls_reject_bad_boxes_without_touching_the_device():
    """idgx_grid_fft_box / idgx_grid_clear_box validate the box on the host (this runs without a GPU)."""
    w = observation.make_observation("small", visibilities=False)
    p = _plan(w)
    G = w["grid_size"]
    L = idgx.lib()
    fake = ctypes.c_void_p(1)  # never dereferenced: every call below fails its checks first
    for box in [(-1, 10, 0, 10), (10, 5, 0, 10), (0, G + 1, 0, 10), (0, 10, 20, 10), (0, 10, 0, G + 1)]:
        b = (ctypes.c_int32 * 4)(*box)
        assert L.idgx_grid_fft_box(ctypes.byref(p), fake, 1, idgx.FFT_FORWARD, b, None) == idgx.ERR_INVALID_ARGUMENT
        assert L.idgx_grid_clear_box(ctypes.byref(p), fake, 1, b, None) == idgx.ERR_INVALID_ARGUMENT
    good = (ctypes.c_int32 * 4)(0, 10, 0, 10)
    assert L.idgx_grid_fft_box(ctypes.byref(p), fake, 0, idgx.FFT_FORWARD, good, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_grid_fft_box(ctypes.byref(p), fake, 1, 7, good, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_grid_fft_box(ctypes.byref(p), None, 1, idgx.FFT_FORWARD, good, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_grid_fft_box(ctypes.byref(p), fake, 1, idgx.FFT_FORWARD, None, None) == idgx.ERR_INVALID_ARGUMENT


def test_grid_to_image_calls_reject_bad_arguments_without_touching_the_device():
    """idgx_grid_to_image / idgx_image_to_grid check the box, the layer count and the buffers on the host: aliased,
    missing or bad ones fail before any device work (this runs without a GPU)."""
    w = observation.make_observation("small", visibilities=False)
    p = _plan(w)
    G = w["grid_size"]
    L = idgx.lib()
    grid, image = ctypes.c_void_p(1), ctypes.c_void_p(2)  # never dereferenced
    good = (ctypes.c_int32 * 4)(0, 10, 0, 10)
    for fn in (L.idgx_grid_to_image, L.idgx_image_to_grid):
        for box in [(-1, 10, 0, 10), (10, 5, 0, 10), (0, G + 1, 0, 10), (0, 10, 20, 10), (0, 10, 0, G + 1)]:
            b = (ctypes.c_int32 * 4)(*box)
            assert fn(ctypes.byref(p), grid, -1, 3, b, image, None) == idgx.ERR_INVALID_ARGUMENT
        assert fn(ctypes.byref(p), grid, -1, 0, good, image, None) == idgx.ERR_INVALID_ARGUMENT
        assert fn(ctypes.byref(p), grid, -1, 3, None, image, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_grid_to_image(ctypes.byref(p), None, 0, 1, good, image, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_grid_to_image(ctypes.byref(p), grid, 0, 1, good, None, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_grid_to_image(ctypes.byref(p), grid, 0, 1, good, grid, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_image_to_grid(ctypes.byref(p), None, 0, 1, good, grid, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_image_to_grid(ctypes.byref(p), image, 0, 1, good, None, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_image_to_grid(ctypes.byref(p), image, 0, 1, good, image, None) == idgx.ERR_INVALID_ARGUMENT


def test_planner_covers_observation_and_keeps_subgrids_on_grid():
    w = observation.make_observation("ska_low", visibilities=False)
    md, ch = w["metadata"], w["channels"]
    S, G = w["subgrid_size"], w["grid_size"]
    assert (md["x"] >= 0).all() and (md["x"] + S <= G).all() and (md["y"] >= 0).all() and (md["y"] + S <= G).all()
    # Every (row, channel) sample is covered at most once, and almost all are covered.
    C = w["wavenumbers"].shape[0]
    cover = np.zeros((w["nr_rows"], C), np.int8)
    for s in range(0, md.size, 37):
        r0 = md["time_offset"][s]
        cover[r0:r0 + md["nr_timesteps"][s], ch["channel_begin"][s]:ch["channel_begin"][s] + ch["nr_channels"][s]] += 1
    assert cover.max() <= 1
    assert w["nr_visibilities"] > 0.99 * w["nr_rows"] * C
    # Samples of each subgrid lie within the usable footprint (S - support cells) around its centre.
    freqs = w["frequencies"].astype(np.float64)
    rows = md["time_offset"][:, None] + np.arange(128)[None, :]
    for s in range(0, md.size, 997):
        uvw = w["uvw"][rows[s, :md["nr_timesteps"][s]]].astype(np.float64)
        f = freqs[ch["channel_begin"][s]:ch["channel_begin"][s] + ch["nr_channels"][s]]
        gu = uvw[:, 0:1] * f[None, :] / observation.SPEED_OF_LIGHT * float(w["image_size"]) + G / 2
        gv = uvw[:, 1:2] * f[None, :] / observation.SPEED_OF_LIGHT * float(w["image_size"]) + G / 2
        assert gu.min() >= md["x"][s] + 3.5 and gu.max() <= md["x"][s] + S - 3.5
        assert gv.min() >= md["y"][s] + 3.5 and gv.max() <= md["y"][s] + S - 3.5


REFERENCE_INC = "/root/reference/gridding/inc"


@pytest.mark.skipif(not __import__("os").path.isdir(REFERENCE_INC), reason="reference sources not present here")
def test_gpu_launcher_is_source_compatible_with_reference_headers(tmp_path):
    """The drop-in TU compiles against the reference's own gridding_benchmark.h / types.hpp (build check only)."""
    import os
    import subprocess

    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ska-spc-tender-benchmarks_amd")
    root = os.path.dirname(pkg)
    cmd = ["g++", "-std=gnu++17", "-c", f"-I{REFERENCE_INC}",
           "-I/root/reference/benchmarks-common/inc", f"-I{root}/include",
           os.path.join(pkg, "host", "src", "launcher_hip.cpp"), "-o", str(tmp_path / "l.o")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_adder_plan_is_conflict_free_and_complete():
    """Every subgrid is listed once; within a colour, subgrids of different tiles never touch the same cell."""
    w = observation.make_observation("ska_low", visibilities=False)
    p = _plan(w)
    ap, order, tile_start = idgx.adder_plan_build(p, w["metadata"])
    md, S = w["metadata"], w["subgrid_size"]
    assert sorted(order.tolist()) == list(range(md.size))
    assert tile_start[0] == 0 and tile_start[-1] == md.size and np.all(np.diff(tile_start) > 0)
    cb = list(ap.color_begin)
    assert cb[0] == 0 and cb[4] == ap.nr_tiles and all(a <= b for a, b in zip(cb, cb[1:]))
    for c in range(4):
        # footprint bounding box of each tile of colour c; boxes of distinct tiles must not intersect
        boxes = []
        for t in range(cb[c], cb[c + 1]):
            s = order[tile_start[t]:tile_start[t + 1]]
            boxes.append((md["x"][s].min(), md["x"][s].max() + S, md["y"][s].min(), md["y"][s].max() + S))
        boxes = np.array(boxes)
        for i in range(len(boxes)):
            x0, x1, y0, y1 = boxes[i]
            others = np.delete(boxes, i, axis=0)
            overlap = (others[:, 0] < x1) & (x0 < others[:, 1]) & (others[:, 2] < y1) & (y0 < others[:, 3])
            assert not overlap.any()


def test_master_grid_and_layered_entry_points_check_arguments_on_the_host():
    """Argument checks that return before any device work: grid FFT sizes, layered adder plan z range."""
    import ctypes

    import idgx
    from idgx import METADATA_DTYPE

    md = np.zeros(3, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    md["z"] = [0, 2, 5]
    plan = idgx.make_plan(nr_subgrids=3, subgrid_size=16, grid_size=130, nr_channels=1, nr_stations=2,
                          nr_aterm_slots=1, nr_rows=3, max_nr_timesteps=1, image_size=0.1)
    L = idgx.lib()
    dummy = ctypes.c_void_p(1)
    assert L.idgx_grid_fft(ctypes.byref(plan), dummy, 1, idgx.FFT_FORWARD, None) == idgx.ERR_UNSUPPORTED  # 130 % 4
    plan.grid_size = 128
    assert L.idgx_grid_fft(ctypes.byref(plan), dummy, 0, idgx.FFT_FORWARD, None) == idgx.ERR_INVALID_ARGUMENT
    assert L.idgx_grid_fft(ctypes.byref(plan), dummy, 1, 7, None) == idgx.ERR_INVALID_ARGUMENT
    with pytest.raises(idgx.IdgxError):
        idgx.adder_plan_build_layered(plan, md, 0, 5)  # z = 5 is outside layers 0 .. 4
    ap, order, tile_start = idgx.adder_plan_build_layered(plan, md, 0, 6)
    assert (ap.z_min, ap.nr_layers, ap.nr_tiles) == (0, 6, 3) and sorted(order) == [0, 1, 2]
    ap1, _, _ = idgx.adder_plan_build(plan, md)
    assert (ap1.nr_layers, ap1.nr_tiles) == (1, 1)  # single grid: z ignored, one tile
