"""The batched uneven and full-cycle configurations end to end on the GPU, against the FP64 oracle.

batched_uneven_small: three subgrid classes (S = 24 / 32 / 48 by baseline-length tercile) in one persistent launch,
ragged per-baseline time chunks, ~30 % of the visibilities flagged through selection lists.
full_cycle: S = 64 on a 16384^2 grid, direction-dependent Jones A-terms per station changing every 64 timesteps
(two A-term slots; the planner cuts subgrids at the boundary), full w-term.
Bar: relative L2 <= 5e-5 vs FP64 per subgrid (gridder) and per class (degridder) for the hi/lo forms.
"""
import numpy as np
import pytest

from helpers import oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2
from idgx import FFT_BACKWARD, FFT_FORWARD, observation
from idgx.device import DeviceBatch, DeviceWorkload, upload

pytestmark = pytest.mark.gpu

BAR = 5e-5


def _window_mask(c, n_rows, C):
    """Rows x channels covered by the subgrids of class c."""
    m = np.zeros((n_rows, C), bool)
    md, ch = c["metadata"], c["channels"]
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        c0, nc = int(ch["channel_begin"][s]), int(ch["nr_channels"][s])
        m[r0:r0 + nt, c0:c0 + nc] = True
    return m


@pytest.fixture(scope="module")
def uneven():
    return observation.make_batched_observation("batched_uneven_small")


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_batched_uneven_vs_fp64(uneven, precision):
    bo = uneven
    flags, C = bo["flags"], bo["wavenumbers"].shape[0]
    vis0 = np.where(flags[..., None], 0, bo["visibilities"]).astype(np.complex64)
    db = DeviceBatch(bo["classes"], precision=precision)
    tg, td = db.batch.autotune(db.data, repeats=1)
    db.grid()
    db.degrid()
    vis = db.visibilities_out()
    assert not np.any(vis[flags]), "flagged visibilities must not be written"
    for k, c in enumerate(bo["classes"]):
        got = db.subgrids(k)
        g64 = oracle_grid(dict(c, visibilities=vis0), fp64=True)
        eg = per_subgrid_rel_l2(got, g64)
        v64 = oracle_degrid(dict(c, visibilities=vis0), got, fp64=True)
        sel = _window_mask(c, bo["nr_rows"], C) & ~flags
        ev = rel_l2(vis[sel], v64[sel])
        print(f"{precision} class S={c['subgrid_size']} ({c['nr_subgrids']} subgrids, {c['nr_visibilities']} vis): "
              f"gridder worst subgrid {eg.max():.3e}, degridder {ev:.3e}")
        assert eg.max() <= BAR and ev <= BAR


@pytest.fixture(scope="module")
def cycle():
    return observation.make_observation("full_cycle")


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_full_cycle_jones_vs_fp64(cycle, precision):
    w = cycle
    assert w["nr_aterm_slots"] == 2 and set(np.unique(w["metadata"]["aterm_index"])) == {0, 1}
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    got = dw.subgrids()
    g64 = oracle_grid(w, fp64=True)
    eg = per_subgrid_rel_l2(got, g64)
    dw.degrid(plan)
    ev = rel_l2(dw.visibilities_out(), oracle_degrid(w, got, fp64=True))
    print(f"full_cycle {precision}: gridder worst subgrid vs FP64 {eg.max():.3e}, degridder {ev:.3e}")
    assert eg.max() <= BAR and ev <= BAR


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_full_cycle_adjoint_on_the_16k_grid(cycle, precision):
    """<degrid(fft_b(split(G))), V> = <G, add(fft_f(grid(V)))> with the 4 x 16384^2 master grid in HBM; G is
    random on the bounding box of the subgrids (the only part the splitter reads) and zero elsewhere."""
    import torch

    w = cycle
    S, Gs = int(w["subgrid_size"]), int(w["grid_size"])
    md = w["metadata"]
    x0, x1 = int(md["x"].min()), int(md["x"].max()) + S
    y0, y1 = int(md["y"].min()), int(md["y"].max()) + S
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    dw.fft(plan, FFT_FORWARD)
    d_grid = dw.grid_buffer()
    dw.add(plan, d_grid)
    g = d_grid.view(torch.complex64).view(4, Gs, Gs)
    gV = g[:, y0:y1, x0:x1].cpu().numpy().astype(np.complex128)
    outside = float(g.abs().sum()) - float(np.abs(gV).sum())
    rng = np.random.default_rng(31)
    box = (rng.standard_normal(gV.shape) + 1j * rng.standard_normal(gV.shape)).astype(np.complex64)
    g.zero_()
    g[:, y0:y1, x0:x1] = torch.from_numpy(box).to(g.device)
    dw.split(plan, d_grid)
    dw.fft(plan, FFT_BACKWARD)
    dw.degrid(plan)
    dG = dw.visibilities_out().astype(np.complex128)
    lhs = np.vdot(dG, w["visibilities"].astype(np.complex128))
    rhs = np.vdot(box.astype(np.complex128), gV)
    print(f"full_cycle adjoint {precision}: rel diff {abs(lhs - rhs) / abs(rhs):.2e}; box {y1 - y0}x{x1 - x0}")
    assert abs(outside) <= 1e-3 * float(np.abs(gV).sum()), "the adder wrote outside the subgrid footprints"
    assert abs(lhs - rhs) <= 2e-5 * abs(rhs)


def test_device_selection_lists_match_the_host_lists(uneven):
    import torch

    from idgx.device import device_selection_lists

    bo = uneven
    d_flags = torch.from_numpy(bo["flags"]).cuda()
    for c in bo["classes"]:
        d_vi, d_off, h_off = device_selection_lists(c, d_flags, chunk_entries=4096)
        assert np.array_equal(h_off, c["vis_offset"]) and np.array_equal(d_off.cpu().numpy(), c["vis_offset"])
        assert np.array_equal(d_vi.cpu().numpy().view(np.uint32), c["vis_index"])


# ---- config 4 on its own workload: the full 256-station plan, every k-th subgrid of each class ----------------------

@pytest.fixture(scope="module")
def c4():
    """The full batched_uneven plan (256 stations, 512 x 1 s, 128 channels 150-200 MHz, S = 24 / 32 / 48 by
    baseline-length tercile, time chunks log-uniform in [16, 256], G = 8192), and a subset of about 60 subgrids per
    class spread evenly over each class's plan, with seeded visibilities and per-baseline flags on their rows."""
    bo = observation.make_batched_observation("batched_uneven", visibilities=False, flags=False)
    idx = [np.unique(np.linspace(0, int(c["nr_subgrids"]) - 1, 60).astype(np.int64)) for c in bo["classes"]]
    sb = observation.subset_batched(bo, idx, seed=9)
    return bo, sb


def _c4_describe(sb):
    for c in sb["classes"]:
        ch = c["channels"]
        print(f"  S={c['subgrid_size']}: {c['nr_subgrids']} subgrids, T {int(c['metadata']['nr_timesteps'].min())}.."
              f"{int(c['metadata']['nr_timesteps'].max())}, window {int(ch['nr_channels'].min())}.."
              f"{int(ch['nr_channels'].max())} channels, {c['nr_visibilities']} selected visibilities")


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_c4_subset_single_launch_vs_fp64(c4, precision):
    """The default single persistent launch per direction over the three classes of the real plan against the FP64
    oracle: worst subgrid <= 5e-5 (gridder), per class <= 5e-5 (degridder); flagged samples left unwritten."""
    _, sb = c4
    _c4_describe(sb)
    flags, C = sb["flags"], sb["wavenumbers"].shape[0]
    vis0 = np.where(flags[..., None], 0, sb["visibilities"]).astype(np.complex64)
    db = DeviceBatch(sb["classes"], precision=precision)
    db.grid()
    db.degrid()
    assert db.batch.last_launches() == (1, 1)
    vis = db.visibilities_out()
    assert not np.any(vis[flags]), "flagged visibilities must not be written"
    for k, c in enumerate(sb["classes"]):
        got = db.subgrids(k)
        eg = per_subgrid_rel_l2(got, oracle_grid(dict(c, visibilities=vis0), fp64=True))
        v64 = oracle_degrid(dict(c, visibilities=vis0), got, fp64=True)
        sel = _window_mask(c, sb["nr_rows"], C) & ~flags
        ev = rel_l2(vis[sel], v64[sel])
        print(f"C4 {precision} S={c['subgrid_size']}: gridder worst subgrid vs FP64 {eg.max():.3e}, degridder {ev:.3e}")
        assert eg.max() <= BAR and ev <= BAR


def test_c4_subset_plain_vs_cpu_restatement(c4):
    """Plain FP32 on the real plan's classes, against the CPU restatement of the reference path with the flagged
    samples zeroed (<= 1e-5 per subgrid / per class): the plain batch (one launch per direction, flagged samples
    skipped) and the one-shot kernels per class (every sample of each window)."""
    _, sb = c4
    flags, C = sb["flags"], sb["wavenumbers"].shape[0]
    vis0 = np.where(flags[..., None], 0, sb["visibilities"]).astype(np.complex64)
    db = DeviceBatch(sb["classes"], precision="fp32")
    db.grid()
    db.degrid()
    vis = db.visibilities_out()
    assert not np.any(vis[flags]), "flagged visibilities must not be written"
    for k, c in enumerate(sb["classes"]):
        cw = dict(c, visibilities=vis0)
        got = db.subgrids(k)
        ref = oracle_grid(cw)
        eg = per_subgrid_rel_l2(got, ref)
        sel = _window_mask(c, sb["nr_rows"], C) & ~flags
        ev = rel_l2(vis[sel], oracle_degrid(cw, got)[sel])
        dw = DeviceWorkload(cw)
        plan = dw.plan(precision="fp32")
        dw.grid(plan)
        one = dw.subgrids()
        eo = per_subgrid_rel_l2(one, ref)
        dw.degrid(plan)
        eov = rel_l2(dw.visibilities_out()[sel], oracle_degrid(cw, one)[sel])
        print(f"C4 plain S={c['subgrid_size']}: batch gridder worst subgrid vs CPU {eg.max():.3e}, degridder "
              f"{ev:.3e}; one-shot {eo.max():.3e} / {eov:.3e}")
        assert max(eg.max(), ev, eo.max(), eov) <= 1e-5


def test_c4_device_planner_matches_the_host_planner(c4):
    """idgx_plan_segments on the full C4 geometry (every ragged time segment of every class) gives the host
    planner's metadata and channel windows bit for bit."""
    from idgx import plan_segments_device

    bo, _ = c4
    cfg = bo["config"]
    total = 0
    for c in bo["classes"]:
        md_d, ch_d = plan_segments_device(bo["uvw"], c["frequencies64"], *c["stations"], *c["segments"],
                                          subgrid_size=int(c["subgrid_size"]), grid_size=cfg["grid_size"],
                                          image_size=cfg["image_size"], nr_times=cfg["n_times"])
        assert md_d.size == c["metadata"].size
        assert md_d.tobytes() == c["metadata"].tobytes() and ch_d.tobytes() == c["channels"].tobytes()
        total += md_d.size
    print(f"C4 device plan: {total} subgrids in {len(bo['classes'])} classes, bit-identical to the host plan")
