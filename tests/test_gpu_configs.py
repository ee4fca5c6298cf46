"""The batched uneven and full-cycle configurations end to end on the GPU, against the FP64 oracle.

batched_uneven_small: three subgrid classes (S = 24 / 32 / 48 by baseline-length tercile) in one persistent launch,
ragged per-baseline time chunks, ~30 % of the visibilities flagged through selection lists.
full_cycle: the 256-station layout, S = 64 on a 16384^2 grid, direction-dependent Jones A-terms per station changing
every 64 timesteps (two A-term slots; the planner cuts subgrids at the boundary), |w| up to ~5300 wavelengths on six
w-layers.
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


def _max_w_lambda(w):
    """Largest |w| of any sample the plan grids, in wavelengths (each subgrid's rows at its window's top channel)."""
    md, ch = w["metadata"], w["channels"]
    f = w["frequencies"].astype(np.float64)
    out = 0.0
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        fmax = f[int(ch["channel_begin"][s]) + int(ch["nr_channels"][s]) - 1]
        out = max(out, float(np.abs(w["uvw"][r0:r0 + nt, 2]).max()) * fmax / observation.SPEED_OF_LIGHT)
    return out


@pytest.fixture(scope="module")
def cycle_subset(cycle):
    """About 150 subgrids of the full-cycle plan spread over it, plus the 16 whose samples reach the largest |w|, so
    the subset spans the whole w range (every w-layer) at a size the FP64 oracle checks in seconds."""
    w = cycle
    md = w["metadata"]
    n_times = w["config"]["n_times"]
    wabs = np.abs(w["uvw"][:, 2]).reshape(-1, n_times)
    bl, t0 = md["time_offset"] // n_times, md["time_offset"] % n_times
    wmax = np.array([wabs[b, t:t + 64].max() for b, t in zip(bl, t0)])
    idx = np.unique(np.concatenate([np.linspace(0, md.size - 1, 150).astype(np.int64), np.argsort(wmax)[-16:]]))
    return observation.subset(w, idx)


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_full_cycle_jones_vs_fp64(cycle_subset, precision):
    """Full-cycle subgrids (S = 64, Jones A-terms in two slots, w-layers, |w| up to ~5300 wavelengths) against the
    FP64 oracle: the mixed form <= 5e-5 worst subgrid / whole degridder output; the plain form against the CPU
    restatement of the reference path (<= 1e-5), its FP64 error stated."""
    w = cycle_subset
    wmax = _max_w_lambda(w)
    assert w["nr_aterm_slots"] == 2 and set(np.unique(w["metadata"]["aterm_index"])) == {0, 1}
    assert wmax >= 4000.0
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    got = dw.subgrids()
    eg = per_subgrid_rel_l2(got, oracle_grid(w, fp64=True))
    dw.degrid(plan)
    vis = dw.visibilities_out()
    ev = rel_l2(vis, oracle_degrid(w, got, fp64=True))
    z = w["metadata"]["z"]
    print(f"full_cycle {precision}: {w['nr_subgrids']} subgrids, max |w| {wmax:.0f} wavelengths, w-layers "
          f"{int(z.min())}..{int(z.max())}: gridder worst subgrid vs FP64 {eg.max():.3e}, degridder {ev:.3e}")
    if precision == "fp32":
        ec = per_subgrid_rel_l2(got, oracle_grid(w))
        evc = rel_l2(vis, oracle_degrid(w, got))
        print(f"full_cycle fp32 vs the CPU restatement: gridder worst subgrid {ec.max():.3e}, degridder {evc:.3e}")
        assert ec.max() <= 1e-5 and evc <= 1e-5
        assert eg.max() <= 1e-3  # FP32 phases on 40 km baselines: stated, not held to the mixed forms' bar
    else:
        assert eg.max() <= BAR and ev <= BAR


def _vdot64(a, b):
    """<a, b> = sum conj(a) b of two complex64 device tensors (flat float32 views), accumulated in FP64 a plane at a
    time."""
    import torch

    a, b = a.view(torch.complex64), b.view(torch.complex64)
    acc = 0j
    step = 1 << 28
    for i in range(0, a.numel(), step):
        acc += complex(torch.vdot(a[i:i + step].to(torch.complex128), b[i:i + step].to(torch.complex128)).item())
    return acc


@pytest.mark.parametrize("precision,fused", [("fp32", False), ("fp16x3", False), ("fp16x3", True)])
def test_full_cycle_adjoint_on_the_16k_grid(cycle, precision, fused):
    """The whole full-cycle plan (102,901 subgrids, |w| to ~5300 wavelengths) through the w-stacked imaging cycle on
    six 16384^2 w-layers in HBM, and back: with M = to_image . grid_fft(BACKWARD) . layered add . subgrid FFT .
    gridder and M^H = degridder . inverse subgrid FFT . layered split . grid_fft(FORWARD) . from_image,
    <M^H I, V> = <I, M V> for a random image I, inner products accumulated in FP64 on the device.  Bar: 2e-5 for the
    mixed form (FP64 phase reduction); 2e-4 for plain FP32, whose phases round as the reference's do (a large
    phase offset minus k (u l + v m + w n), each rounded to FP32) and so differ from FP64 by up to 1.4e-4 per
    subgrid on this plan's 40 km baselines (test_full_cycle_jones_vs_fp64), in the gridder and degridder
    differently: the identity cannot hold closer than that.  fused: both image-side steps in one call each on the
    subgrids' box (idgx_grid_to_image, idgx_image_to_grid: screens and layer sum inside the row pass)."""
    import torch

    from idgx import FFT_BACKWARD, FFT_FORWARD, grid_box

    w = cycle
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    z_min, nl = dw.layers()
    G = int(w["grid_size"])
    box = grid_box(plan, w["metadata"])
    print(f"full_cycle adjoint {precision}{' fused' if fused else ''}: {w['nr_subgrids']} subgrids, {nl} w-layers "
          f"from z = {z_min}, G = {G}, box {box}", flush=True)
    dw.grid(plan)
    dw.fft(plan, FFT_FORWARD)
    d_layers = dw.layer_buffer(nl)
    dw.add_layered(plan, d_layers)
    d_image = dw.grid_buffer()
    if fused:
        dw.grid_to_image(plan, d_layers, d_image, box)
    else:
        dw.grid_fft(plan, d_layers, nl, FFT_BACKWARD)
        dw.to_image(plan, d_layers, d_image)
    gen = torch.Generator(device=d_image.device)
    gen.manual_seed(41)
    d_rand = torch.randn(4 * G * G * 2, generator=gen, device=d_image.device, dtype=torch.float32)
    rhs = _vdot64(d_rand, d_image.view(torch.float32))
    del d_image
    if fused:
        dw.image_to_grid(plan, d_rand, d_layers, box)
    else:
        dw.from_image(plan, d_rand, d_layers)
        dw.grid_fft(plan, d_layers, nl, FFT_FORWARD)
    dw.split_layered(plan, d_layers)
    del d_layers
    dw.fft(plan, FFT_BACKWARD)
    dw.degrid(plan)
    lhs = _vdot64(dw.d_vis_out.view(torch.float32), dw.d_visibilities.view(torch.float32))
    rel = abs(lhs - rhs) / abs(rhs)
    print(f"full_cycle adjoint {precision}{' fused' if fused else ''}: <M^H I, V> = {lhs:.6e}, <I, M V> = "
          f"{rhs:.6e}, rel diff {rel:.2e}")
    assert rel <= (2e-4 if precision == "fp32" else 2e-5)


def test_device_selection_lists_match_the_host_lists(uneven):
    import torch

    from idgx.device import device_selection_lists

    bo = uneven
    d_flags = torch.from_numpy(bo["flags"]).cuda()
    for c in bo["classes"]:
        d_vi, d_off, h_off = device_selection_lists(c, d_flags, chunk_entries=4096)
        assert np.array_equal(h_off, c["vis_offset"]) and np.array_equal(d_off.cpu().numpy(), c["vis_offset"])
        assert np.array_equal(d_vi.cpu().numpy().view(np.uint32), c["vis_index"])


# ---- the batched uneven workload itself: the full 256-station plan, every k-th subgrid of each class -------------

@pytest.fixture(scope="module")
def uneven_plan():
    """The full batched_uneven plan (256 stations, 512 x 1 s, 128 channels 150-200 MHz, S = 24 / 32 / 48 by
    baseline-length tercile, time chunks log-uniform in [16, 256], G = 8192), and a subset of about 60 subgrids per
    class spread evenly over each class's plan, with seeded visibilities and per-baseline flags on their rows."""
    bo = observation.make_batched_observation("batched_uneven", visibilities=False, flags=False)
    idx = [np.unique(np.linspace(0, int(c["nr_subgrids"]) - 1, 60).astype(np.int64)) for c in bo["classes"]]
    sb = observation.subset_batched(bo, idx, seed=9)
    return bo, sb


def _describe_classes(sb):
    for c in sb["classes"]:
        ch = c["channels"]
        print(f"  S={c['subgrid_size']}: {c['nr_subgrids']} subgrids, T {int(c['metadata']['nr_timesteps'].min())}.."
              f"{int(c['metadata']['nr_timesteps'].max())}, window {int(ch['nr_channels'].min())}.."
              f"{int(ch['nr_channels'].max())} channels, {c['nr_visibilities']} selected visibilities")


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_batched_uneven_subset_single_launch_vs_fp64(uneven_plan, precision):
    """The default single persistent launch per direction over the three classes of the real plan against the FP64
    oracle: worst subgrid <= 5e-5 (gridder), per class <= 5e-5 (degridder); flagged samples left unwritten."""
    _, sb = uneven_plan
    _describe_classes(sb)
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
        print(f"batched_uneven {precision} S={c['subgrid_size']}: gridder worst subgrid vs FP64 {eg.max():.3e}, degridder {ev:.3e}")
        assert eg.max() <= BAR and ev <= BAR


def test_batched_uneven_subset_plain_vs_cpu_restatement(uneven_plan):
    """Plain FP32 on the real plan's classes, against the CPU restatement of the reference path with the flagged
    samples zeroed (<= 1e-5 per subgrid / per class): the plain batch (one launch per direction, flagged samples
    skipped) and the one-shot kernels per class (every sample of each window)."""
    _, sb = uneven_plan
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
        print(f"batched_uneven plain S={c['subgrid_size']}: batch gridder worst subgrid vs CPU {eg.max():.3e}, degridder "
              f"{ev:.3e}; one-shot {eo.max():.3e} / {eov:.3e}")
        assert max(eg.max(), ev, eo.max(), eov) <= 1e-5


def test_batched_uneven_device_planner_matches_the_host_planner(uneven_plan):
    """idgx_plan_segments on the full batched_uneven geometry (every ragged time segment of every class) gives the host
    planner's metadata and channel windows bit for bit."""
    from idgx import plan_segments_device

    bo, _ = uneven_plan
    cfg = bo["config"]
    total = 0
    for c in bo["classes"]:
        md_d, ch_d = plan_segments_device(bo["uvw"], c["frequencies64"], *c["stations"], *c["segments"],
                                          subgrid_size=int(c["subgrid_size"]), grid_size=cfg["grid_size"],
                                          image_size=cfg["image_size"], nr_times=cfg["n_times"])
        assert md_d.size == c["metadata"].size
        assert md_d.tobytes() == c["metadata"].tobytes() and ch_d.tobytes() == c["channels"].tobytes()
        total += md_d.size
    print(f"batched_uneven device plan: {total} subgrids in {len(bo['classes'])} classes, bit-identical to the host plan")
