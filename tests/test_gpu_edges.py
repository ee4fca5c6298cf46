"""Edge cases of the gridder and degridder entry points: the largest subgrid size a plan accepts (S = 128), empty
plans, and subgrids that carry no visibilities (zero timesteps, or an empty channel window).

Bars: the plain form against the FP32 restatement of the reference CPU path (relative L2 <= 1e-5 per subgrid);
the split mixed forms against the FP64 oracle (<= 5e-5 per subgrid); subgrids without visibilities are exactly
zero after the gridder and leave the degridder's output untouched."""
import numpy as np
import pytest

from helpers import oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2
from idgx import Batch, host_grid, make_plan, observation, grid, degrid
from idgx.device import DeviceBatch, DeviceWorkload, upload
from idgx.workload import reference_workload

pytestmark = pytest.mark.gpu

PARITY = 1e-5
SPLIT_BAR = 5e-5


def _vis_errors(w, got, ref):
    md, ch = w["metadata"], w.get("channels")
    C = w["wavenumbers"].shape[0]
    errs = []
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        c0, nc = (0, C) if ch is None else (int(ch["channel_begin"][s]), int(ch["nr_channels"][s]))
        if nt and nc:
            errs.append(rel_l2(got[r0:r0 + nt, c0:c0 + nc], ref[r0:r0 + nt, c0:c0 + nc]))
    return np.array(errs)


@pytest.fixture(scope="module")
def s128():
    w = observation.make_observation("small", subgrid_size=128, timesteps_per_subgrid=32)
    w = observation.subset(w, np.arange(0, w["nr_subgrids"], max(1, w["nr_subgrids"] // 6)))
    rng = np.random.default_rng(128)
    sub = (rng.standard_normal((w["nr_subgrids"], 4, 128, 128)) +
           1j * rng.standard_normal((w["nr_subgrids"], 4, 128, 128))).astype(np.complex64)
    return w, sub


def test_largest_subgrid_size_plain_matches_reference_restatement(s128):
    w, sub = s128
    dw = DeviceWorkload(w)
    plan = dw.plan()
    dw.grid(plan)
    eg = per_subgrid_rel_l2(dw.subgrids(), oracle_grid(w)).max()
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    ed = _vis_errors(w, dw.visibilities_out(), oracle_degrid(w, sub)).max()
    print(f"S=128 ({w['nr_subgrids']} subgrids) plain: gridder worst {eg:.2e}, degridder worst {ed:.2e}")
    assert eg <= PARITY and ed <= PARITY


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_largest_subgrid_size_mixed_one_shot_and_batch_vs_fp64(s128, precision):
    w, sub = s128
    g64, v64 = oracle_grid(w, fp64=True), oracle_degrid(w, sub, fp64=True)
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    eg = per_subgrid_rel_l2(dw.subgrids(), g64).max()
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    ed = _vis_errors(w, dw.visibilities_out(), v64).max()
    db = DeviceBatch([w], precision=precision)
    db.grid()
    ebg = per_subgrid_rel_l2(db.subgrids(), g64).max()
    db.classes[0]["d_subgrids"].copy_(upload(sub))
    db.degrid()
    ebd = _vis_errors(w, db.visibilities_out(), v64).max()
    print(f"S=128 {precision}: one-shot gridder {eg:.2e} degridder {ed:.2e}; batch gridder {ebg:.2e} "
          f"degridder {ebd:.2e}")
    assert max(eg, ed, ebg, ebd) <= SPLIT_BAR


def test_empty_plan_leaves_outputs_untouched():
    """A plan with no subgrids is a valid no-op for the one-shot calls, the batch and the host-buffer calls."""
    import torch

    w = observation.make_observation("small")
    dw = DeviceWorkload(w)
    p = make_plan(nr_subgrids=0, subgrid_size=32, grid_size=w["grid_size"], nr_channels=dw.C,
                  nr_stations=w["nr_stations"], nr_aterm_slots=w["nr_aterm_slots"], nr_rows=dw.nr_rows,
                  max_nr_timesteps=w["max_nr_timesteps"], image_size=w["image_size"], precision="fp16x3")
    dw.d_subgrids.fill_(7)
    dw.d_vis_out.fill_(9)
    for precision in (0, p.precision):
        p.precision = precision
        grid(p, dw.d_metadata, dw.d_channels, dw.d_uvw, dw.d_wavenumbers, dw.d_visibilities, dw.d_spheroidal,
             dw.d_aterms, dw.d_subgrids)
        degrid(p, dw.d_metadata, dw.d_channels, dw.d_uvw, dw.d_wavenumbers, dw.d_subgrids, dw.d_spheroidal,
               dw.d_aterms, dw.d_vis_out)
    torch.cuda.synchronize()
    assert bool((dw.d_subgrids == 7).all()) and bool((dw.d_vis_out == 9).all())
    empty = dict(w, metadata=w["metadata"][:0].copy(), channels=w["channels"][:0].copy(), nr_subgrids=0)
    db = DeviceBatch([empty], precision="fp16x3")
    db.grid()
    db.degrid()
    assert db.batch.items()[:2] == (0, 0)
    sub, t = host_grid(p, empty)
    assert sub.shape == (0, 4, 32, 32) and t.compute_time >= 0


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_subgrids_without_visibilities(precision):
    """Subgrid 0 gets zero timesteps and subgrid 1 an empty channel window: the gridder writes exact zeros for both
    (one-shot and batch), the degridder leaves their visibilities untouched, and every other subgrid still meets
    its bar."""
    w = observation.make_observation("small")
    w = observation.subset(w, np.arange(0, w["nr_subgrids"], 4))
    md, ch = w["metadata"].copy(), w["channels"].copy()
    # The samples the two subgrids would have covered (no other subgrid covers them).
    windows = [(int(md["time_offset"][s]), int(md["nr_timesteps"][s]), int(ch["channel_begin"][s]),
                int(ch["nr_channels"][s])) for s in (0, 1)]
    md["nr_timesteps"][0] = 0
    ch["nr_channels"][1] = 0
    w = dict(w, metadata=md, channels=ch)
    rng = np.random.default_rng(3)
    sub = (rng.standard_normal((w["nr_subgrids"], 4, 32, 32)) +
           1j * rng.standard_normal((w["nr_subgrids"], 4, 32, 32))).astype(np.complex64)
    fp64 = precision != "fp32"
    bar = SPLIT_BAR if fp64 else PARITY
    g_ref, v_ref = oracle_grid(w, fp64=fp64), oracle_degrid(w, sub, fp64=fp64)
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.d_subgrids.fill_(1)
    dw.grid(plan)
    got = dw.subgrids()
    assert not got[:2].any(), "subgrids without visibilities must be zero"
    assert per_subgrid_rel_l2(got[2:], g_ref[2:]).max() <= bar
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    vis = dw.visibilities_out()
    for r0, nt, c0, nc in windows:
        assert nt > 0 and nc > 0 and not vis[r0:r0 + nt, c0:c0 + nc].any()
    assert _vis_errors(w, vis, v_ref).max() <= bar
    if precision != "fp32":
        db = DeviceBatch([w], precision=precision)
        db.classes[0]["d_subgrids"].fill_(1)
        db.grid()
        bsub = db.subgrids()
        assert not bsub[:2].any()
        assert per_subgrid_rel_l2(bsub[2:], g_ref[2:]).max() <= bar
        db.classes[0]["d_subgrids"].copy_(upload(sub))
        db.degrid()
        bvis = db.visibilities_out()
        for r0, nt, c0, nc in windows:
            assert not bvis[r0:r0 + nt, c0:c0 + nc].any()
        assert _vis_errors(w, bvis, v_ref).max() <= bar


@pytest.mark.parametrize("C", [12, 10])
def test_plain_degridder_channel_chunks_on_ragged_plans_without_windows(C):
    """Without channel windows the plain degridder takes 4 channels per thread when the channel count is a multiple
    of 4 (C = 12) and 2 otherwise (C = 10).  Both on subgrids of 0, 37, 64 and 128 timesteps: every covered
    visibility meets the bar against the FP32 restatement of the reference CPU path, and the rows of the subgrid
    without timesteps stay zero."""
    w = reference_workload(4, 2, 128, C, 32, 1024)
    md = w["metadata"].copy()
    md["nr_timesteps"][0] = 0
    md["nr_timesteps"][1] = 37
    md["nr_timesteps"][2::3] = 64
    w = dict(w, metadata=md)
    rng = np.random.default_rng(C)
    n = w["nr_subgrids"]
    sub = (rng.standard_normal((n, 4, 32, 32)) + 1j * rng.standard_normal((n, 4, 32, 32))).astype(np.complex64)
    dw = DeviceWorkload(w)
    plan = dw.plan()
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    vis = dw.visibilities_out()
    errs = _vis_errors(w, vis, oracle_degrid(w, sub))
    print(f"C={C}: {errs.size} subgrids with visibilities, worst rel-L2 {errs.max():.2e}")
    r0 = int(md["time_offset"][0])
    assert not vis[r0:r0 + 128].any()
    assert errs.max() <= PARITY


@pytest.mark.parametrize("S", [24, 40, 48])
def test_plain_pixels_per_thread_by_subgrid_size(S):
    """The plain gridder takes 3 pixels per thread where that leaves fewer idle lanes (S = 24, 48) and 4 otherwise
    (S = 40); both directions meet the bar against the FP32 restatement of the reference CPU path."""
    w = observation.make_observation("small", subgrid_size=S, timesteps_per_subgrid=64)
    w = observation.subset(w, np.arange(0, w["nr_subgrids"], 3))
    rng = np.random.default_rng(S)
    sub = (rng.standard_normal((w["nr_subgrids"], 4, S, S)) +
           1j * rng.standard_normal((w["nr_subgrids"], 4, S, S))).astype(np.complex64)
    dw = DeviceWorkload(w)
    plan = dw.plan()
    dw.grid(plan)
    eg = per_subgrid_rel_l2(dw.subgrids(), oracle_grid(w)).max()
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    ed = _vis_errors(w, dw.visibilities_out(), oracle_degrid(w, sub)).max()
    print(f"S={S} ({w['nr_subgrids']} subgrids) plain: gridder worst {eg:.2e}, degridder worst {ed:.2e}")
    assert eg <= PARITY and ed <= PARITY
