"""Plain FP32 batches (idgx_batch_* with IDGX_PLAIN_FP32): the persistent forms of the plain kernels over whole plans,
with visibility selections (flagged samples skipped), several subgrid classes in one launch per direction, hipGraph
capture and the autotuner.

The persistent items evaluate every phase and sum in the same order as the one-shot kernels (idgx_grid / idgx_degrid)
wherever a subgrid's channel window is at most 64 channels wide, and skipping a flagged sample only leaves out an
exact zero term.  So on those plans the batch must equal the one-shot kernels bit for bit (the one-shot gridder fed
the visibilities with the flagged samples zeroed), besides meeting the 1e-5 bar against the CPU restatement of the
reference path."""
import numpy as np
import pytest

from helpers import oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2
from idgx import SHAPE_TILES_PER_CLASS, STAGE_DEGRID, STAGE_GRID, Batch, observation
from idgx.device import DeviceBatch, DeviceWorkload

pytestmark = pytest.mark.gpu

BAR = 1e-5


def _oneshot(w, d_subgrids=None):
    """Plain one-shot gridder output, and the one-shot degridder run on `d_subgrids` (default: its own output)."""
    dw = DeviceWorkload(w)
    plan = dw.plan(precision="fp32")
    dw.grid(plan)
    sub = dw.subgrids()
    if d_subgrids is not None:
        dw.d_subgrids.copy_(d_subgrids)
    dw.degrid(plan)
    return sub, dw.visibilities_out()


def _window_mask(c, n_rows, C):
    m = np.zeros((n_rows, C), bool)
    md, ch = c["metadata"], c["channels"]
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        c0, nc = int(ch["channel_begin"][s]), int(ch["nr_channels"][s])
        m[r0:r0 + nt, c0:c0 + nc] = True
    return m


def test_plain_shape_table():
    shapes = Batch.plain_shapes()
    assert len(shapes) >= 3
    assert {px for px, _ in shapes} == {2, 3, 4} and {nv for _, nv in shapes} == {1, 2, 4}


@pytest.fixture(scope="module")
def ska_low_subset():
    full = observation.make_observation("ska_low", visibilities=False)
    w = observation.subset(full, np.arange(3, full["nr_subgrids"], 401))
    w["visibilities"] = observation.random_visibilities(w["nr_rows"], w["wavenumbers"].shape[0], seed=11)
    return w


def test_plain_batch_equals_oneshot_bitwise_in_every_shape(ska_low_subset):
    """Dense channel windows (at most 64 channels): every shape of the persistent kernels gives the one-shot
    kernels' bits, gridder and degridder (the degridder on the same input subgrids)."""
    w = ska_low_subset
    assert int(w["channels"]["nr_channels"].max()) <= 64
    sub1, vis1 = _oneshot(w)
    db = DeviceBatch([w], precision="fp32")
    for shape in range(len(Batch.plain_shapes())):
        db.batch.set_shape(shape, shape)
        db.grid()
        got = db.subgrids()
        db.degrid()
        vis = db.visibilities_out()
        print(f"shape {Batch.plain_shapes()[shape]}: gridder equal {np.array_equal(got, sub1)}, degridder equal "
              f"{np.array_equal(vis, vis1)}")
        assert np.array_equal(got, sub1)
        assert np.array_equal(vis, vis1)
    # and against the CPU restatement of the reference path
    e = per_subgrid_rel_l2(sub1, oracle_grid(w))
    print(f"SKA-Low subset ({w['nr_subgrids']} subgrids): plain gridder worst subgrid vs CPU restatement {e.max():.3e}")
    assert e.max() <= BAR


@pytest.fixture(scope="module")
def uneven():
    return observation.make_batched_observation("batched_uneven_small")


def test_plain_batch_with_flags_vs_cpu_and_oneshot(uneven):
    """Three subgrid classes (S = 24 / 32 / 48) in one persistent launch per direction, ragged timesteps and ~30 %
    of the samples flagged through selection lists: gridder and degridder against the CPU restatement with the
    flagged samples zeroed (<= 1e-5 per subgrid / per class), bit for bit against the one-shot kernels on the zeroed
    visibilities, and flagged samples left unwritten."""
    bo = uneven
    flags, C = bo["flags"], bo["wavenumbers"].shape[0]
    vis0 = np.where(flags[..., None], 0, bo["visibilities"]).astype(np.complex64)
    db = DeviceBatch(bo["classes"], precision="fp32")
    assert db.batch.single_launch() == (SHAPE_TILES_PER_CLASS, 0)
    db.grid()
    db.degrid()
    assert db.batch.last_launches() == (1, 1)
    vis = db.visibilities_out()
    assert not np.any(vis[flags]), "flagged visibilities must not be written"
    for k, c in enumerate(bo["classes"]):
        got = db.subgrids(k)
        ref = oracle_grid(dict(c, visibilities=vis0))
        eg = per_subgrid_rel_l2(got, ref)
        sel = _window_mask(c, bo["nr_rows"], C) & ~flags
        vref = oracle_degrid(dict(c, visibilities=vis0), got)
        ev = rel_l2(vis[sel], vref[sel])
        one_sub, one_vis = _oneshot(dict(c, visibilities=vis0), d_subgrids=db.classes[k]["d_subgrids"])
        print(f"class S={c['subgrid_size']} ({c['nr_subgrids']} subgrids, {int(sel.sum())} selected): gridder worst "
              f"subgrid vs CPU {eg.max():.3e}, degridder {ev:.3e}; equal to one-shot: "
              f"{np.array_equal(got, one_sub)} / {np.array_equal(vis[sel], one_vis[sel])}")
        assert eg.max() <= BAR and ev <= BAR
        assert np.array_equal(got, one_sub)
        assert np.array_equal(vis[sel], one_vis[sel])


def test_plain_batch_single_launch_graph_and_autotune(uneven):
    """Per-class launches equal the single launch bit for bit; a captured graph replays the eager result three times
    with eager work in between; the autotuner picks shapes of the plain table and the results do not change."""
    bo = uneven
    db = DeviceBatch(bo["classes"], precision="fp32")
    db.grid()
    db.degrid()
    ref_sub = [db.subgrids(k) for k in range(len(bo["classes"]))]
    ref_vis = db.visibilities_out()
    db.batch.set_single_launch(-1, -1)
    db.grid()
    db.degrid()
    assert db.batch.last_launches() == (3, 3)
    for k in range(len(bo["classes"])):
        assert np.array_equal(db.subgrids(k), ref_sub[k])
    assert np.array_equal(db.visibilities_out(), ref_vis)
    db.batch.set_single_launch(SHAPE_TILES_PER_CLASS, 0)

    import torch

    stream = torch.cuda.Stream()
    db.batch.capture(db.data, STAGE_GRID | STAGE_DEGRID, stream)
    for r in range(3):
        for c in db.classes:
            c["d_subgrids"].zero_()
        db.d_vis_out.zero_()
        torch.cuda.synchronize()
        db.batch.launch(stream)
        stream.synchronize()
        for k in range(len(bo["classes"])):
            assert np.array_equal(db.subgrids(k), ref_sub[k]), f"replay {r} class {k}"
        assert np.array_equal(db.visibilities_out(), ref_vis), f"replay {r}"

    tg, td = db.batch.autotune(db.data, repeats=1)
    table = db.batch.shape_table()
    assert table == Batch.plain_shapes()
    chosen = [db.batch.class_shape(k) for k in range(len(bo["classes"]))]
    print(f"autotuned plain shapes per class {[(table[g], table[d]) for g, d in chosen]}, single launch "
          f"{db.batch.single_launch()}: gridder {tg:.3f} ms, degridder {td:.3f} ms")
    db.grid()
    db.degrid()
    for k in range(len(bo["classes"])):
        assert np.array_equal(db.subgrids(k), ref_sub[k])
    assert np.array_equal(db.visibilities_out(), ref_vis)
