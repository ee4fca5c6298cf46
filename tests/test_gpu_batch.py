"""Batched execution (idgx_batch_*): persistent work-queue kernels over whole plans, visibility selections
(flagging), mixed subgrid sizes in one batch (a launch and a tile shape per class, or one launch over the merged
queue of all classes), hipGraph capture and the autotuner."""
import numpy as np
import pytest

from helpers import oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2
from idgx import SHAPE_TILES_PER_CLASS, STAGE_DEGRID, STAGE_GRID, Batch, observation
from idgx.device import DeviceBatch, DeviceWorkload, selection_lists, upload

pytestmark = pytest.mark.gpu


def _oneshot(w, precision):
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    sub = dw.subgrids()
    dw.degrid(plan)
    return sub, dw.visibilities_out()


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_batch_equals_oneshot_kernels_bitwise(precision):
    """Same item code and default tile shape as the one-shot kernels: identical bits, in any item order.  The
    batched degridder stages pixels prepared once per subgrid: identical words for BF16; for FP16 the power-of-two
    scale is chosen per subgrid instead of per LDS batch, so the results agree to rounding."""
    full = observation.make_observation("ska_low", visibilities=False)
    w = observation.subset(full, np.arange(5, full["nr_subgrids"], 311))
    w["visibilities"] = observation.random_visibilities(w["nr_rows"], w["wavenumbers"].shape[0], seed=3)
    sub1, vis1 = _oneshot(w, precision)
    db = DeviceBatch([w], precision=precision)
    assert Batch.shapes()[db.batch.shape()[0]] == (8, 256, 0) and Batch.shapes()[db.batch.shape()[1]] == (4, 256, 0)
    db.grid()
    db.degrid()
    assert db.subgrids().tobytes() == sub1.tobytes()
    if precision == "bf16x3":
        assert db.visibilities_out().tobytes() == vis1.tobytes()
    else:
        e = rel_l2(db.visibilities_out(), vis1)
        print(f"fp16x3 batched vs one-shot degridder: {e:.3e}")
        assert e <= 2e-6


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_batch_degridder_prepared_pixels_vs_fp64(precision):
    """Prepared pixels with non-identity A-terms and the reference taper, both A-term conventions."""
    from helpers import load_golden
    from idgx import ATERM_REFERENCE
    w = load_golden("ref_example")
    sub = w["ref_degrid_subgrids"]
    for conv in (0, ATERM_REFERENCE):
        db = DeviceBatch([w], precision=precision, aterm_convention=conv)
        db.classes[0]["d_subgrids"].copy_(upload(sub))
        db.degrid()
        ww = w if conv == 0 else dict(w, aterms=np.ascontiguousarray(np.conj(w["aterms"][..., [0, 2, 1, 3]])))
        v64 = oracle_degrid(ww, sub, fp64=True)
        e = rel_l2(db.visibilities_out(), v64)
        print(f"{precision} convention {conv}: batched degridder vs FP64 {e:.3e}")
        assert e <= 5e-5


def test_batch_with_flag_selection_vs_fp64():
    w = observation.make_observation("small")
    rng = np.random.default_rng(6)
    frac = rng.uniform(0.0, 0.6, w["nr_rows"] // 128)  # per baseline
    flags = rng.random((w["nr_rows"], w["wavenumbers"].shape[0])) < np.repeat(frac, 128)[:, None]
    vi, vo = selection_lists(w, flags)
    w["vis_index"], w["vis_offset"] = vi, vo
    db = DeviceBatch([w], precision="bf16x3")
    db.grid()
    got = db.subgrids()
    w0 = dict(w, visibilities=np.where(flags[..., None], 0, w["visibilities"]).astype(np.complex64))
    g64 = oracle_grid(w0, fp64=True)
    e = per_subgrid_rel_l2(got, g64)
    print(f"flagged ({flags.mean():.2f} of visibilities): gridder worst subgrid vs FP64 {e.max():.3e}")
    assert e.max() <= 5e-5
    db.degrid()
    vis = db.visibilities_out()
    v64 = oracle_degrid(w0, got, fp64=True)
    assert not np.any(vis[flags]), "flagged visibilities must not be written"
    ev = rel_l2(vis[~flags], v64[~flags])
    print(f"flagged degridder vs FP64 {ev:.3e}")
    assert ev <= 5e-5


def _mixed_size_classes(w):
    """Three classes (S = 24, 32, 48) from one observation, sharing the visibility rows."""
    classes = []
    for k, S in enumerate((24, 32, 48)):
        idx = np.arange(k, w["nr_subgrids"], 3)
        c = observation.subset(w, idx)
        c["visibilities"] = w["visibilities"][c["source_rows"]]
        c.update(subgrid_size=S, spheroidal=observation.prolate_spheroidal_taper(S),
                 aterms=observation.identity_aterms(1, w["nr_stations"], S))
        classes.append(c)
    # Classes share the visibility rows: give every class all rows, with its own time offsets.
    rows = np.concatenate([c["source_rows"] for c in classes])
    base = 0
    shared = []
    for c in classes:
        md = c["metadata"].copy()
        md["time_offset"] += base
        base += c["nr_rows"]
        shared.append(dict(c, metadata=md))
    uvw, vis = w["uvw"][rows], w["visibilities"][rows]
    for c in shared:
        c.update(uvw=uvw, visibilities=vis, nr_rows=rows.size)
    return shared


def test_mixed_subgrid_sizes_in_one_batch():
    """Three subgrid sizes in one batch, by default one persistent launch per direction over all of them (gridder
    items with their class's tiles): the same bits as the one-shot kernels of each class."""
    w = observation.make_observation("small")
    shared = _mixed_size_classes(w)
    db = DeviceBatch(shared, precision="bf16x3")
    assert db.batch.single_launch()[0] == SHAPE_TILES_PER_CLASS
    db.grid()
    assert db.batch.last_launches()[0] == 1
    for k, c in enumerate(shared):
        ref, _ = _oneshot(c, "bf16x3")
        assert db.subgrids(k).tobytes() == ref.tobytes(), f"class S={c['subgrid_size']}"
    db.degrid()
    vis_b = db.visibilities_out()
    for k, c in enumerate(shared):
        dw = DeviceWorkload(c)
        dw.d_subgrids.copy_(upload(db.subgrids(k)))
        dw.degrid(dw.plan(precision="bf16x3"))
        r0 = int(c["metadata"]["time_offset"][0])
        r1 = r0 + int(c["metadata"]["nr_timesteps"].sum())
        assert dw.visibilities_out()[r0:r1].tobytes() == vis_b[r0:r1].tobytes()


def test_graph_replays_match_eager():
    import torch

    w = observation.make_observation("small")
    db = DeviceBatch([w], precision="bf16x3")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        db.grid(stream)
        db.degrid(stream)
    stream.synchronize()
    sub0, vis0 = db.subgrids(), db.visibilities_out()
    db.batch.capture(db.data, STAGE_GRID | STAGE_DEGRID, stream)
    for r in range(3):
        db.classes[0]["d_subgrids"].zero_()
        db.d_vis_out.zero_()
        torch.cuda.synchronize()
        db.batch.launch(stream)
        stream.synchronize()
        assert db.subgrids().tobytes() == sub0.tobytes(), f"replay {r}"
        assert db.visibilities_out().tobytes() == vis0.tobytes(), f"replay {r}"
        x = torch.ones(1000, device="cuda").sum()  # eager work between replays
        assert float(x) == 1000.0


def test_graph_of_a_mixed_size_batch_replays_its_per_class_launches():
    """Three classes with different tile shapes: the captured graph holds the counter reset, the degridder pixel
    preparation and one persistent launch per class and direction; every replay equals the eager run bit for bit."""
    import torch

    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision="fp16x3")
    db.batch.set_single_launch(-1, -1)
    shapes = Batch.shapes()
    for k, (g, d) in enumerate([((3, 256, 0), (4, 256, 0)), ((8, 256, 0), (6, 256, 0)), ((6, 256, 0), (2, 256, 0))]):
        db.batch.set_class_shape(k, shapes.index(g), shapes.index(d))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        db.grid(stream)
        db.degrid(stream)
    stream.synchronize()
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()
    db.batch.capture(db.data, STAGE_GRID | STAGE_DEGRID, stream)
    for r in range(3):
        for c in db.classes:
            c["d_subgrids"].zero_()
        db.d_vis_out.zero_()
        torch.cuda.synchronize()
        db.batch.launch(stream)
        stream.synchronize()
        for k in range(3):
            assert db.subgrids(k).tobytes() == sub0[k].tobytes(), f"replay {r}, class {k}"
        assert db.visibilities_out().tobytes() == vis0.tobytes(), f"replay {r}"
        x = torch.ones(1000, device="cuda").sum()  # eager work between replays
        assert float(x) == 1000.0


def test_autotuner_picks_a_shape_and_keeps_results():
    w = observation.make_observation("small")
    db = DeviceBatch([w], precision="bf16x3")
    db.grid()
    db.degrid()
    sub0, vis0 = db.subgrids(), db.visibilities_out()
    tg, td = db.batch.autotune(db.data, repeats=2)
    gs, ds = db.batch.shape()
    shapes = Batch.shapes()
    print(f"autotune: gridder shape {shapes[gs]} {tg:.3f} ms, degridder shape {shapes[ds]} {td:.3f} ms")
    assert 0 <= gs < len(shapes) and 0 <= ds < len(shapes) and tg > 0 and td > 0
    db.grid()
    db.degrid()
    assert rel_l2(db.subgrids(), sub0) < 1e-6
    assert rel_l2(db.visibilities_out(), vis0) < 1e-6


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_every_tile_shape_gives_the_same_results(precision):
    """All compiled tile shapes (tiles per wave, LDS depth, phases on the matrix core) agree with the default ones
    to rounding, for subgrid sizes whose rows are not a multiple of the tile (S = 24, 48) and for S = 32."""
    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision=precision)
    db.batch.set_single_launch(-1, -1)  # each compiled shape on its own launches
    db.grid()
    db.degrid()
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()
    for k in range(len(Batch.shapes())):
        db.batch.set_shape(k, k)
        for c in db.classes:
            c["d_subgrids"].zero_()
        db.d_vis_out.zero_()
        db.grid()
        eg = max(rel_l2(db.subgrids(j), sub0[j]) for j in range(3))
        db.degrid()
        ev = rel_l2(db.visibilities_out(), vis0)
        print(f"{precision} shape {Batch.shapes()[k]}: gridder {eg:.2e}, degridder {ev:.2e}")
        assert eg < 2e-6 and ev < 2e-6


def test_classes_run_with_their_own_tile_shapes():
    """Each class keeps its own shapes (set one by one, or by the autotuner) and the results do not depend on
    them beyond rounding."""
    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision="fp16x3")
    db.batch.set_single_launch(-1, -1)
    db.grid()
    db.degrid()
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()
    shapes = Batch.shapes()
    picks = [(shapes.index((3, 256, 0)), shapes.index((6, 256, 0))),
             (shapes.index((8, 256, 0)), shapes.index((4, 256, 0))),
             (shapes.index((6, 256, 0)), shapes.index((2, 256, 0)))]
    for k, (g, d) in enumerate(picks):
        db.batch.set_class_shape(k, g, d)
    assert [db.batch.class_shape(k) for k in range(3)] == picks
    assert db.batch.shape() == picks[0]
    for c in db.classes:
        c["d_subgrids"].zero_()
    db.d_vis_out.zero_()
    db.grid()
    db.degrid()
    for k in range(3):
        assert rel_l2(db.subgrids(k), sub0[k]) < 2e-6, f"class {k}"
    assert rel_l2(db.visibilities_out(), vis0) < 2e-6
    tg, td = db.batch.autotune(db.data, repeats=1)
    tuned = [db.batch.class_shape(k) for k in range(3)]
    print("autotuned per class:", [(shapes[g], shapes[d]) for g, d in tuned], f"{tg:.3f} / {td:.3f} ms")
    assert tg > 0 and td > 0
    db.grid()
    db.degrid()
    for k in range(3):
        assert rel_l2(db.subgrids(k), sub0[k]) < 2e-6, f"class {k}"
    assert rel_l2(db.visibilities_out(), vis0) < 2e-6
    with pytest.raises(Exception):
        db.batch.set_class_shape(3, 0, 0)


def _zero_outputs(db):
    for c in db.classes:
        c["d_subgrids"].zero_()
    db.d_vis_out.zero_()


@pytest.mark.parametrize("gshape,dshape", [((4, 256, 0), (4, 256, 0)), ((3, 256, 0), (8, 256, 1))])
def test_single_launch_over_all_classes_equals_per_class_launches_bitwise(gshape, dshape):
    """Single-launch mode: one persistent launch per direction over the merged queue of the S = 24/32/48 classes
    gives the per-class launches' results bit for bit when every class uses the same shape (each item writes its
    own outputs, so the order of the merged queue does not matter)."""
    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision="fp16x3")
    shapes = Batch.shapes()
    g, d = shapes.index(gshape), shapes.index(dshape)
    default_tiles = tuple(shapes[db.batch.class_shape(k)[0]][0] for k in range(3))
    assert db.batch.single_launch() == (SHAPE_TILES_PER_CLASS, shapes.index((4, 256, 0)))  # several classes
    assert default_tiles == (3, 8, 6)
    db.batch.set_single_launch(-1, -1)
    db.batch.set_shape(g, d)
    assert db.batch.single_launch() == (-1, -1)
    items_per_class = db.batch.items()
    db.grid()
    db.degrid()
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()

    assert db.batch.last_launches() == (3, 3)
    db.batch.set_single_launch(g, d)
    assert db.batch.single_launch() == (g, d)
    gi, di, gw, dw = db.batch.items()
    assert (gi, di) == items_per_class[:2] and gw > 0 and dw > 0
    _zero_outputs(db)
    db.grid()
    db.degrid()
    assert db.batch.last_launches() == (1, 1)
    for k in range(3):
        assert db.subgrids(k).tobytes() == sub0[k].tobytes(), f"class {k}"
    assert db.visibilities_out().tobytes() == vis0.tobytes()

    db.batch.set_single_launch(-1, -1)  # back to per-class launches; their queues were kept
    assert db.batch.single_launch() == (-1, -1)
    _zero_outputs(db)
    db.grid()
    db.degrid()
    for k in range(3):
        assert db.subgrids(k).tobytes() == sub0[k].tobytes(), f"class {k}"
    assert db.visibilities_out().tobytes() == vis0.tobytes()


def test_single_launch_matches_fp64_with_flagged_classes():
    """The merged queue with visibility selections (flagging) on every class: gridder and degridder within the
    mixed-form tolerance of the FP64 oracle, and flagged visibilities left unwritten."""
    w = observation.make_observation("small")
    classes = _mixed_size_classes(w)
    rng = np.random.default_rng(5)
    flags = rng.random((classes[0]["nr_rows"], w["wavenumbers"].shape[0])) < 0.3
    for c in classes:
        c["vis_index"], c["vis_offset"] = selection_lists(c, flags)
    db = DeviceBatch(classes, precision="fp16x3")
    db.batch.set_single_launch(Batch.shapes().index((4, 256, 0)), Batch.shapes().index((4, 256, 0)))
    db.grid()
    db.degrid()
    vis = db.visibilities_out()
    assert not np.any(vis[flags]), "flagged visibilities must not be written"
    for k, c in enumerate(classes):
        c0 = dict(c, visibilities=np.where(flags[..., None], 0, c["visibilities"]).astype(np.complex64))
        got = db.subgrids(k)
        e = per_subgrid_rel_l2(got, oracle_grid(c0, fp64=True)).max()
        v64 = oracle_degrid(c0, got, fp64=True)
        r0 = int(c["metadata"]["time_offset"][0])
        r1 = r0 + int(c["metadata"]["nr_timesteps"].sum())
        keep = ~flags[r0:r1]
        ev = rel_l2(vis[r0:r1][keep], v64[r0:r1][keep])
        print(f"class S={c['subgrid_size']}: gridder worst subgrid vs FP64 {e:.2e}, degridder {ev:.2e}")
        assert e <= 5e-5 and ev <= 5e-5


def test_graph_of_a_single_launch_batch_replays_match_eager():
    import torch

    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision="fp16x3")
    shapes = Batch.shapes()
    db.batch.set_single_launch(shapes.index((4, 256, 0)), shapes.index((6, 256, 0)))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        db.grid(stream)
        db.degrid(stream)
    stream.synchronize()
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()
    db.batch.capture(db.data, STAGE_GRID | STAGE_DEGRID, stream)
    for r in range(3):
        _zero_outputs(db)
        torch.cuda.synchronize()
        db.batch.launch(stream)
        stream.synchronize()
        for k in range(3):
            assert db.subgrids(k).tobytes() == sub0[k].tobytes(), f"replay {r}, class {k}"
        assert db.visibilities_out().tobytes() == vis0.tobytes(), f"replay {r}"
        x = torch.ones(1000, device="cuda").sum()  # eager work between replays
        assert float(x) == 1000.0


def test_autotuner_tunes_the_single_launch_shape():
    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision="fp16x3")
    db.grid()
    db.degrid()
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()
    db.batch.set_single_launch(0, 0)
    tg, td = db.batch.autotune(db.data, repeats=1)
    g, d = db.batch.single_launch()
    shapes = Batch.shapes()
    names = {SHAPE_TILES_PER_CLASS: "tiles per class", -1: "per-class launches"}
    print(f"single launch autotuned: gridder {names.get(g) or shapes[g]} {tg:.3f} ms, "
          f"degridder {names.get(d) or shapes[d]} {td:.3f} ms")
    # The fastest of the merged launch's shapes and the per-class launches (-1) run together.
    assert g in (SHAPE_TILES_PER_CLASS, -1) or 0 <= g < len(shapes)
    assert -1 <= d < len(shapes) and tg > 0 and td > 0
    _zero_outputs(db)
    db.grid()
    db.degrid()
    for k in range(3):
        assert rel_l2(db.subgrids(k), sub0[k]) < 2e-6, f"class {k}"
    assert rel_l2(db.visibilities_out(), vis0) < 2e-6


def test_single_launch_with_each_class_tiles_equals_per_class_launches_bitwise():
    """Gridder single launch with IDGX_SHAPE_TILES_PER_CLASS: one persistent kernel in which every item takes its
    class's tiles per wave (3 / 8 / 6 for S = 24 / 32 / 48) gives the per-class launches' bits; changing a class's
    shape afterwards rebuilds the merged queue."""
    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision="fp16x3")
    db.batch.set_single_launch(-1, -1)
    shapes = Batch.shapes()
    g_picks = [shapes.index((3, 256, 0)), shapes.index((8, 256, 0)), shapes.index((6, 256, 0))]
    d = shapes.index((4, 256, 0))
    for k, g in enumerate(g_picks):
        db.batch.set_class_shape(k, g, d)
    db.grid()
    db.degrid()
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()
    assert db.batch.last_launches() == (3, 3)
    db.batch.set_single_launch(SHAPE_TILES_PER_CLASS, d)
    assert db.batch.single_launch() == (SHAPE_TILES_PER_CLASS, d)
    _zero_outputs(db)
    db.grid()
    db.degrid()
    assert db.batch.last_launches() == (1, 1)
    for k in range(3):
        assert db.subgrids(k).tobytes() == sub0[k].tobytes(), f"class {k}"
    assert db.visibilities_out().tobytes() == vis0.tobytes()
    # Class 1 switches to 4 tiles: per-class launches and the single launch follow it together.
    db.batch.set_class_shape(1, shapes.index((4, 256, 0)), d)
    db.batch.set_single_launch(-1, -1)
    _zero_outputs(db)
    db.grid()
    sub1 = [db.subgrids(k) for k in range(3)]
    db.batch.set_single_launch(SHAPE_TILES_PER_CLASS, d)
    _zero_outputs(db)
    db.grid()
    for k in range(3):
        assert db.subgrids(k).tobytes() == sub1[k].tobytes(), f"class {k} after the shape change"
    assert rel_l2(db.subgrids(1), sub0[1]) < 2e-6


def test_graph_of_the_default_single_launch_replays_match_eager():
    """A mixed-size batch in its default mode (one launch per direction, gridder items with their class's tiles)
    captured in a hipGraph: every replay equals the eager run bit for bit, one persistent launch per direction."""
    import torch

    w = observation.make_observation("small")
    db = DeviceBatch(_mixed_size_classes(w), precision="fp16x3")
    assert db.batch.single_launch()[0] == SHAPE_TILES_PER_CLASS
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        db.grid(stream)
        db.degrid(stream)
    stream.synchronize()
    assert db.batch.last_launches() == (1, 1)
    sub0 = [db.subgrids(k) for k in range(3)]
    vis0 = db.visibilities_out()
    db.batch.capture(db.data, STAGE_GRID | STAGE_DEGRID, stream)
    assert db.batch.last_launches() == (1, 1)
    for r in range(3):
        _zero_outputs(db)
        torch.cuda.synchronize()
        db.batch.launch(stream)
        stream.synchronize()
        for k in range(3):
            assert db.subgrids(k).tobytes() == sub0[k].tobytes(), f"replay {r}, class {k}"
        assert db.visibilities_out().tobytes() == vis0.tobytes(), f"replay {r}"
        x = torch.ones(1000, device="cuda").sum()  # eager work between replays
        assert float(x) == 1000.0
