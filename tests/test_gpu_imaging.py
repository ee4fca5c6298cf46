"""Master-grid transform and w-stacking, the step after the adder (idgx_grid_fft, layered adder / splitter,
idgx_wstack_to_image / _from_image): against the FP64 brute-force DFT and placement oracles, the adjoint identity
of the whole imaging chain, and a point source imaged through gridder -> subgrid FFT -> layered adder -> master
FFT -> w-stacking vs the direct FP64 DFT of its visibilities."""
import numpy as np
import pytest

import oracle
from helpers import rel_l2
from idgx import FFT_BACKWARD, FFT_FORWARD, METADATA_DTYPE, CHANNELS_DTYPE, grid_box, grid_clear_box, grid_fft, \
    grid_fft_box, observation
from idgx.device import DeviceWorkload, download, upload

pytestmark = pytest.mark.gpu


def _centred(x, sign):
    """sum_X x[X] exp(sign 2 pi i (K - G/2)(X - G/2) / G) in both dimensions (G % 4 == 0), in FP64."""
    G = x.shape[-1]
    c = (-1.0) ** np.arange(G)
    return oracle.dft2d_fp64(x * c[:, None] * c[None, :], sign) * c[:, None] * c[None, :]


def _grid_workload(G, md, S=16):
    return dict(metadata=md, channels=None, subgrid_size=S, grid_size=G, nr_stations=2, nr_aterm_slots=1,
                image_size=np.float32(0.1), w_step=50.0, wavenumbers=np.full(1, 3.0, np.float32),
                uvw=np.zeros((max(1, int(md["nr_timesteps"].sum())), 3), np.float32),
                visibilities=np.zeros((max(1, int(md["nr_timesteps"].sum())), 1, 4), np.complex64),
                spheroidal=np.ones((S, S), np.float32), aterms=observation.identity_aterms(1, 2, S),
                max_nr_timesteps=int(md["nr_timesteps"].max()), nr_subgrids=md.size)


def test_grid_fft_matches_the_centred_dft_and_round_trips():
    rng = np.random.default_rng(1)
    G, L = 64, 2
    x = (rng.standard_normal((L, 4, G, G)) + 1j * rng.standard_normal((L, 4, G, G))).astype(np.complex64)
    md = np.zeros(1, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    dw = DeviceWorkload(_grid_workload(G, md))
    plan = dw.plan()
    d = upload(x)
    dw.grid_fft(plan, d, L, FFT_FORWARD)
    fwd = download(d, np.complex64, x.shape)
    e_f = max(rel_l2(fwd[l, p], _centred(x[l, p].astype(np.complex128), +1)) for l in range(L) for p in range(4))
    dw.grid_fft(plan, d, L, FFT_BACKWARD)
    rt = download(d, np.complex64, x.shape) / (G * G)
    d2 = upload(x)
    dw.grid_fft(plan, d2, L, FFT_BACKWARD)
    bwd = download(d2, np.complex64, x.shape)
    e_b = max(rel_l2(bwd[l, p], _centred(x[l, p].astype(np.complex128), -1)) for l in range(L) for p in range(4))
    print(f"grid FFT vs FP64 DFT: forward {e_f:.2e}, backward {e_b:.2e}, round trip {rel_l2(rt, x):.2e}")
    assert e_f < 1e-6 and e_b < 1e-6 and rel_l2(rt, x) < 1e-6


def _centred_np(x, sign):
    """The centred transform in FP64 through numpy's FFT between the same sign flips (G % 4 == 0)."""
    G = x.shape[-1]
    cc = np.outer((-1.0) ** np.arange(G), (-1.0) ** np.arange(G))
    y = np.fft.ifft2(x * cc) * (G * G) if sign > 0 else np.fft.fft2(x * cc)
    return y * cc


def _fft_plan(G):
    md = np.zeros(1, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    return DeviceWorkload(_grid_workload(G, md)).plan()


@pytest.mark.parametrize("G", [2048, 1028])
def test_boxed_grid_fft_equals_the_full_transform(G):
    """idgx_grid_fft_box: BACKWARD on a grid that is zero outside a box (not aligned to the 64-column tiles) gives
    the full transform's image bit for bit; FORWARD gives the full transform's cells inside the box to FP32
    rounding (its row transforms run first).
    The shared scratch holds an unrelated transform's data first, so a read of a column the box leaves out would
    show.  G = 1028 runs the rocFFT path, which ignores the box.  Whole-grid and empty boxes too, and a FORWARD box
    that covers most of the grid (the full transform runs)."""
    rng = np.random.default_rng(7)
    plan = _fft_plan(G)
    shape = (2, 4, G, G)
    junk = upload((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64))
    grid_fft(plan, junk, 2, FFT_BACKWARD)  # leaves unrelated data in the scratch
    box = (300, 701, 1000, 1403) if G == 2048 else (10, 90, 500, 530)
    x = np.zeros(shape, np.complex64)
    x[..., box[2]:box[3], box[0]:box[1]] = (rng.standard_normal((2, 4, box[3] - box[2], box[1] - box[0])) +
                                            1j * rng.standard_normal((2, 4, box[3] - box[2], box[1] - box[0])))
    d_box, d_full = upload(x), upload(x)
    grid_fft_box(plan, d_box, 2, FFT_BACKWARD, box)
    grid_fft(plan, d_full, 2, FFT_BACKWARD)
    img = download(d_full, np.complex64, shape)
    assert download(d_box, np.complex64, shape).tobytes() == img.tobytes()
    assert rel_l2(img[1, 2], _centred_np(x[1, 2].astype(np.complex128), -1)) < 1e-6
    if G == 2048:
        # The box's rows only: junk in the other rows is not read (so idgx_grid_clear_box need not zero them).
        xj = x.copy()
        xj[..., :box[2], :] = 5 + 5j
        xj[..., box[3]:, :] = -3 + 1j
        d_j = upload(xj)
        grid_fft_box(plan, d_j, 2, FFT_BACKWARD, box)
        assert download(d_j, np.complex64, shape).tobytes() == img.tobytes()
    # FORWARD from the dense image: the box's cells match, the rest is not produced.
    d_box, d_full = upload(img), upload(img)
    grid_fft_box(plan, d_box, 2, FFT_FORWARD, box)
    grid_fft(plan, d_full, 2, FFT_FORWARD)
    rows, cols = slice(box[2], box[3]), slice(box[0], box[1])
    got, full = download(d_box, np.complex64, shape), download(d_full, np.complex64, shape)
    e_box = rel_l2(got[..., rows, cols], full[..., rows, cols])
    print(f"G={G} boxed FORWARD vs the full transform on the box: {e_box:.2e}")
    assert e_box < 1e-6
    assert rel_l2(got[0, 0][rows, cols] / np.float32(G * G), x[0, 0][rows, cols]) < 1e-6
    # Whole-grid box: the full transform; empty box backward: zeros in, zeros out.
    d_box, d_full = upload(x), upload(x)
    grid_fft_box(plan, d_box, 2, FFT_BACKWARD, (0, G, 0, G))
    grid_fft(plan, d_full, 2, FFT_BACKWARD)
    assert download(d_box, np.complex64, shape).tobytes() == download(d_full, np.complex64, shape).tobytes()
    d0 = upload(np.zeros(shape, np.complex64))
    grid_fft_box(plan, d0, 2, FFT_BACKWARD, (0, 0, 0, 0))
    assert not download(d0, np.complex64, shape).any()
    # FORWARD on a box that covers most of the grid runs the full transform (cheaper than rows first): bit for bit.
    big = (0, G - 1, 40, G - 30)
    d_box, d_full = upload(img), upload(img)
    grid_fft_box(plan, d_box, 2, FFT_FORWARD, big)
    grid_fft(plan, d_full, 2, FFT_FORWARD)
    assert download(d_box, np.complex64, shape).tobytes() == download(d_full, np.complex64, shape).tobytes()


def test_boxed_imaging_cycle_equals_the_full_cycle():
    """Two imaging cycles in a row (clear, add, master FFT to the image and back, split) on the subgrids' bounding
    box: idgx_grid_clear_box + idgx_grid_fft_box give the split subgrids of the full clear and full transforms to
    FP32 rounding, the second cycle starting from the grid the boxed FORWARD transform left undefined outside the
    box."""
    import torch

    w = observation.make_observation("small")
    dw = DeviceWorkload(w)
    plan = dw.plan(precision="fp16x3")
    dw.grid(plan)
    dw.fft(plan, FFT_FORWARD, dw.d_subgrids)
    d_uv = dw.d_subgrids.clone()
    box = grid_box(plan, w["metadata"])
    G = int(w["grid_size"])
    assert 0 < box[1] - box[0] < G and 0 < box[3] - box[2] < G

    def cycle(boxed, d_grid):
        if boxed:
            grid_clear_box(plan, d_grid, 1, box)
        else:
            d_grid.zero_()
        dw.add(plan, d_grid, d_uv)
        for direction in (FFT_BACKWARD, FFT_FORWARD):
            if boxed:
                grid_fft_box(plan, d_grid, 1, direction, box)
            else:
                grid_fft(plan, d_grid, 1, direction)
        out = torch.empty_like(d_uv)
        dw.split(plan, d_grid, out)
        return out.cpu().numpy().view(np.complex64)

    d_full, d_box = dw.grid_buffer(), dw.grid_buffer()
    d_box.fill_(7)  # not zero anywhere: the boxed clear must zero what the transform to the image reads
    for r in range(2):
        e = rel_l2(cycle(True, d_box), cycle(False, d_full))
        print(f"boxed cycle {r} vs the full cycle: {e:.2e}")
        assert e < 1e-6, f"cycle {r}"


@pytest.mark.parametrize("G", [1024, 2048, 4096, 8192])
def test_three_pass_grid_fft_matches_numpy_fp64(G):
    """G = 1024 .. 16384 run the hand-written three-pass transform (column passes N1 x N2 = 32x32, 32x64, 64x64,
    64x128): every correlation plane transformed, planes 0 and 3 compared with numpy's FP64 FFT both ways, and the
    round trip."""
    rng = np.random.default_rng(G)
    x = (rng.standard_normal((4, G, G), np.float32) + 1j * rng.standard_normal((4, G, G), np.float32)).astype(
        np.complex64)
    plan = _fft_plan(G)
    d = upload(x)
    grid_fft(plan, d, 1, FFT_FORWARD)
    fwd = download(d, np.complex64, x.shape)
    e_f = max(rel_l2(fwd[p], _centred_np(x[p].astype(np.complex128), +1)) for p in (0, 3))
    grid_fft(plan, d, 1, FFT_BACKWARD)
    e_rt = rel_l2(download(d, np.complex64, x.shape) / np.float32(G * G), x)
    d2 = upload(x)
    grid_fft(plan, d2, 1, FFT_BACKWARD)
    bwd = download(d2, np.complex64, x.shape)
    e_b = max(rel_l2(bwd[p], _centred_np(x[p].astype(np.complex128), -1)) for p in (0, 3))
    print(f"G={G} three-pass grid FFT vs numpy FP64: forward {e_f:.2e}, backward {e_b:.2e}, round trip {e_rt:.2e}")
    assert e_f < 1e-6 and e_b < 1e-6 and e_rt < 1e-6


@pytest.mark.parametrize("G", [256, 512, 3072, 1028])
def test_other_size_grid_fft_matches_numpy_fp64(G):
    """Sizes outside 1024 .. 16384 run the plain 2-D transform between the two sign flips: the LDS kernels (256: one
    row pass and one column pass; 512 likewise), the radix-3 three-pass set (3072 = 3 x 1024) and rocFFT (1028).
    Every correlation plane transformed, planes 0 and 3 compared with numpy's FP64 FFT both ways, and the round
    trip."""
    from idgx import FFT_FP32, fft2d_own_kernels

    assert fft2d_own_kernels(G, G, FFT_FP32) == (G != 1028)
    rng = np.random.default_rng(G)
    x = (rng.standard_normal((4, G, G), np.float32) + 1j * rng.standard_normal((4, G, G), np.float32)).astype(
        np.complex64)
    plan = _fft_plan(G)
    d = upload(x)
    grid_fft(plan, d, 1, FFT_FORWARD)
    fwd = download(d, np.complex64, x.shape)
    e_f = max(rel_l2(fwd[p], _centred_np(x[p].astype(np.complex128), +1)) for p in (0, 3))
    grid_fft(plan, d, 1, FFT_BACKWARD)
    e_rt = rel_l2(download(d, np.complex64, x.shape) / np.float32(G * G), x)
    d2 = upload(x)
    grid_fft(plan, d2, 1, FFT_BACKWARD)
    bwd = download(d2, np.complex64, x.shape)
    e_b = max(rel_l2(bwd[p], _centred_np(x[p].astype(np.complex128), -1)) for p in (0, 3))
    print(f"G={G} grid FFT vs numpy FP64: forward {e_f:.2e}, backward {e_b:.2e}, round trip {e_rt:.2e}")
    assert e_f < 1e-6 and e_b < 1e-6 and e_rt < 1e-6


def test_three_pass_grid_fft_at_16384():
    """The 16384^2 master grid (8.6 GB, the full-cycle configuration): four output rows of plane 0 against the
    FP64 centred DFT (column sums by matrix-vector product, then numpy's row FFT), Parseval on every plane, and the
    round trip, all on the device."""
    import torch

    G = 16384
    torch.manual_seed(5)
    x = torch.randn((4, G, G, 2), device="cuda", dtype=torch.float32)
    x0 = x[0].cpu().numpy().view(np.complex64)[..., 0]
    e_in = [float(torch.linalg.vector_norm(x[p], dtype=torch.float64)) ** 2 for p in range(4)]
    d = x.clone()
    grid_fft(_fft_plan(G), d, 1, FFT_FORWARD)
    torch.cuda.synchronize()
    c = (-1.0) ** np.arange(G)
    worst_row = 0.0
    for L in (0, 1, G // 2, 12345):
        w = np.exp(2j * np.pi * (((L - G // 2) * (np.arange(G) - G // 2)) % G) / G)
        v = w @ x0  # sum over rows Y of x[Y, X] exp(+2 pi i (L - G/2)(Y - G/2) / G)
        ref = np.fft.ifft(v * c) * G * c
        got = d[0, L].cpu().numpy().view(np.complex64)[..., 0]
        worst_row = max(worst_row, rel_l2(got, ref))
    e_out = [float(torch.linalg.vector_norm(d[p], dtype=torch.float64)) ** 2 for p in range(4)]
    parseval = max(abs(e_out[p] / (e_in[p] * G * G) - 1.0) for p in range(4))
    grid_fft(_fft_plan(G), d, 1, FFT_BACKWARD)
    d /= G * G
    rt = float(torch.linalg.vector_norm(d - x, dtype=torch.float64) / torch.linalg.vector_norm(x, dtype=torch.float64))
    print(f"G=16384 three-pass grid FFT: worst row vs FP64 {worst_row:.2e}, Parseval {parseval:.2e}, round trip {rt:.2e}")
    assert worst_row < 1e-6 and parseval < 1e-5 and rt < 1e-6


def test_layered_adder_and_splitter_match_the_placement_oracle_per_layer():
    rng = np.random.default_rng(5)
    S, G, n = 16, 96, 240
    md = np.zeros(n, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    md["x"] = rng.integers(-S // 2, G - S // 2, n)
    md["y"] = rng.integers(-S // 2, G - S // 2, n)
    md["z"] = rng.integers(-2, 2, n)  # layers -2 .. 1
    md["x"][:5], md["y"][:5] = 30, 30  # a pile in one tile, spread over layers
    sub = (rng.integers(-8, 9, (n, 4, S, S)) + 1j * rng.integers(-8, 9, (n, 4, S, S))).astype(np.complex64)
    dw = DeviceWorkload(_grid_workload(G, md, S))
    plan = dw.plan()
    z_min, nl = dw.layers()
    assert (z_min, nl) == (-2, 4)
    dw.d_subgrids.copy_(upload(sub))
    d_layers = dw.layer_buffer(nl)
    dw.add_layered(plan, d_layers)
    got = dw.grid_array(d_layers, nl)
    for k in range(nl):
        sel = md["z"] == z_min + k
        np.testing.assert_array_equal(got[k], oracle.adder_fp64(md[sel], sub[sel], G).astype(np.complex64))
    grids = (rng.standard_normal((nl, 4, G, G)) + 1j * rng.standard_normal((nl, 4, G, G))).astype(np.complex64)
    dw.split_layered(plan, upload(grids))
    split = dw.subgrids()
    for k in range(nl):
        sel = md["z"] == z_min + k
        np.testing.assert_array_equal(split[sel], oracle.splitter_fp64(md[sel], grids[k], S).astype(np.complex64))


def test_wstack_screens_match_their_formula_and_are_adjoint():
    rng = np.random.default_rng(9)
    G, z_min, nl, w_step, im = 64, -3, 5, 170.0, 0.1
    md = np.zeros(1, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    w = _grid_workload(G, md)
    w.update(w_step=w_step, image_size=np.float32(im))
    dw = DeviceWorkload(w)
    plan = dw.plan()
    layers = (rng.standard_normal((nl, 4, G, G)) + 1j * rng.standard_normal((nl, 4, G, G))).astype(np.complex64)
    image = (rng.standard_normal((4, G, G)) + 1j * rng.standard_normal((4, G, G))).astype(np.complex64)
    from idgx import wstack_from_image, wstack_to_image
    d_img = dw.layer_buffer(1)
    wstack_to_image(plan, upload(layers), z_min, nl, d_img)
    got = download(d_img, np.complex64, (4, G, G)).astype(np.complex128)
    lm = (np.arange(G) - G / 2) * float(np.float32(im)) / G
    r2 = lm[None, :] ** 2 + lm[:, None] ** 2
    n = r2 / (1 + np.sqrt(1 - r2))
    wz = (z_min + np.arange(nl) + 0.5) * w_step
    screens = np.exp(-2j * np.pi * wz[:, None, None] * n[None])
    ref = (layers.astype(np.complex128) * screens[:, None]).sum(0)
    d_lay = dw.layer_buffer(nl)
    wstack_from_image(plan, upload(image), z_min, nl, d_lay)
    back = download(d_lay, np.complex64, (nl, 4, G, G)).astype(np.complex128)
    lhs = np.vdot(back, layers.astype(np.complex128))
    rhs = np.vdot(image.astype(np.complex128), got)
    print(f"w-stack to_image vs formula {rel_l2(got, ref):.2e}; adjoint rel diff {abs(lhs - rhs) / abs(rhs):.2e}")
    assert rel_l2(got, ref) < 1e-6 and abs(lhs - rhs) <= 1e-6 * abs(rhs)


@pytest.mark.parametrize("G", [1024, 2048, 1028])
def test_grid_to_image_and_back_match_the_separate_calls(G):
    """idgx_grid_to_image / idgx_image_to_grid, whose row pass applies the w-screens and sums the layers, against
    idgx_grid_fft_box followed by idgx_wstack_to_image, and idgx_wstack_from_image followed by idgx_grid_fft_box
    FORWARD (both pinned to FP64 above).  w_z n reaches about 16 turns; the box is off the 64-column tiles.  One
    correlation is also checked against the FP64 formula directly (numpy FFTs).  Then an empty box (a zero image),
    the whole grid, and a smaller box after a larger one (the kept scratch is reused).  G = 1028 runs the two calls
    (no three-pass kernels there): bit for bit the separate chain."""
    from idgx import grid_to_image, image_to_grid, wstack_from_image, wstack_to_image

    rng = np.random.default_rng(11)
    z_min, nl, w_step, im = -2, 5 if G < 2048 else 3, 2600.0, 0.1
    md = np.zeros(1, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    w = _grid_workload(G, md)
    w.update(w_step=w_step, image_size=np.float32(im))
    dw = DeviceWorkload(w)
    plan = dw.plan()
    shape = (nl, 4, G, G)

    def layers_in(box):
        x = np.zeros(shape, np.complex64)
        h, wd = box[3] - box[2], box[1] - box[0]
        x[..., box[2]:box[3], box[0]:box[1]] = (rng.standard_normal((nl, 4, h, wd)) +
                                                1j * rng.standard_normal((nl, 4, h, wd)))
        return x

    def to_image_both(x, box):
        d_sep, d_img_sep = upload(x), dw.layer_buffer(1)
        grid_fft_box(plan, d_sep, nl, FFT_BACKWARD, box)
        wstack_to_image(plan, d_sep, z_min, nl, d_img_sep)
        d_fused, d_img = upload(x), dw.layer_buffer(1)
        d_img.fill_(3)  # every image cell must be written
        grid_to_image(plan, d_fused, z_min, nl, box, d_img)
        return download(d_img, np.complex64, (4, G, G)), download(d_img_sep, np.complex64, (4, G, G))

    big = (300, 701, 100, 523) if G >= 2048 else (150, 351, 50, 261)
    x = layers_in(big)
    img, img_sep = to_image_both(x, big)
    e_to = rel_l2(img, img_sep)
    if G != 1028:
        # Rows outside the box are not read: junk there leaves the image bit for bit as it was.
        xj = x.copy()
        xj[..., :big[2], :] = 5 + 5j
        xj[..., big[3]:, :] = -3 + 1j
        d_j, d_img_j = upload(xj), dw.layer_buffer(1)
        grid_to_image(plan, d_j, z_min, nl, big, d_img_j)
        assert download(d_img_j, np.complex64, (4, G, G)).tobytes() == img.tobytes()
    # From the image: the box's cells of every layer.
    X = (rng.standard_normal((4, G, G)) + 1j * rng.standard_normal((4, G, G))).astype(np.complex64)
    d_sep = dw.layer_buffer(nl)
    wstack_from_image(plan, upload(X), z_min, nl, d_sep)
    grid_fft_box(plan, d_sep, nl, FFT_FORWARD, big)
    d_fused = dw.layer_buffer(nl)
    image_to_grid(plan, upload(X), z_min, nl, big, d_fused)
    rows, cols = slice(big[2], big[3]), slice(big[0], big[1])
    got = download(d_fused, np.complex64, shape)[..., rows, cols]
    sep = download(d_sep, np.complex64, shape)[..., rows, cols]
    e_from = rel_l2(got, sep)
    print(f"G={G}: grid_to_image vs the separate calls {e_to:.2e}; image_to_grid {e_from:.2e}")
    if G == 1028:
        assert e_to == 0.0 and e_from == 0.0
        return
    assert e_to < 1e-6 and e_from < 1e-6
    if G == 1024:
        lm = (np.arange(G) - G / 2) * float(np.float32(im)) / G
        r2 = lm[None, :] ** 2 + lm[:, None] ** 2
        n = r2 / (1 + np.sqrt(1 - r2))
        wz = (z_min + np.arange(nl) + 0.5) * w_step
        assert (wz[-1] * n).max() > 10  # phases of many turns
        ref = sum(np.exp(-2j * np.pi * wz[z] * n) * _centred_np(x[z, 1].astype(np.complex128), -1)
                  for z in range(nl))
        e64 = rel_l2(img[1], ref)
        print(f"G={G}: grid_to_image vs FP64 {e64:.2e}")
        assert e64 < 1e-6
    # Empty box: zero layers give a zero image (every cell written); from the image, the grid is not touched.
    img0, _ = to_image_both(np.zeros(shape, np.complex64), (0, 0, 0, 0))
    assert not img0.any()
    d_g = dw.layer_buffer(nl)
    d_g.fill_(7)
    image_to_grid(plan, upload(X), z_min, nl, (0, 0, 0, 0), d_g)
    assert bool((d_g == 7).all())
    # The whole grid, then a box smaller than the first (the scratch grown for the whole grid is reused).
    for box in ((0, G, 0, G), (64, 129, 7, 90)):
        x = layers_in(box)
        img, img_sep = to_image_both(x, box)
        e = rel_l2(img, img_sep)
        print(f"G={G} box {box}: grid_to_image vs the separate calls {e:.2e}")
        assert e < 1e-6


@pytest.fixture(scope="module")
def wstack():
    return observation.make_observation("wstack")


@pytest.mark.parametrize("precision,fused", [("fp32", False), ("fp16x3", False), ("fp16x3", True)])
def test_imaging_chain_with_wstacking_is_adjoint(wstack, precision, fused):
    """<degrid(fft_b(split(fft_f(from_image(X))))), V> = <X, to_image(fft_b(add(fft_f(grid(V)))))>; fused: the
    transforms and screens of each side in one call on the subgrids' box (idgx_grid_to_image / idgx_image_to_grid)."""
    w = wstack
    G = int(w["grid_size"])
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    z_min, nl = dw.layers()
    assert nl >= 5
    box = grid_box(plan, w["metadata"])
    dw.grid(plan)
    dw.fft(plan, FFT_FORWARD)
    d_layers = dw.layer_buffer(nl)
    dw.add_layered(plan, d_layers)
    d_img = dw.layer_buffer(1)
    if fused:
        dw.grid_to_image(plan, d_layers, d_img, box)
    else:
        dw.grid_fft(plan, d_layers, nl, FFT_BACKWARD)
        dw.to_image(plan, d_layers, d_img)
    img = download(d_img, np.complex64, (4, G, G)).astype(np.complex128)
    rng = np.random.default_rng(17)
    X = (rng.standard_normal((4, G, G)) + 1j * rng.standard_normal((4, G, G))).astype(np.complex64)
    if fused:
        dw.image_to_grid(plan, upload(X), d_layers, box)
    else:
        dw.from_image(plan, upload(X), d_layers)
        dw.grid_fft(plan, d_layers, nl, FFT_FORWARD)
    dw.split_layered(plan, d_layers)
    dw.fft(plan, FFT_BACKWARD)
    dw.degrid(plan)
    dX = dw.visibilities_out().astype(np.complex128)
    lhs = np.vdot(dX, w["visibilities"].astype(np.complex128))
    rhs = np.vdot(X.astype(np.complex128), img)
    print(f"w-stacked chain adjoint {precision}{' fused' if fused else ''} ({nl} layers, G = {G}): rel diff "
          f"{abs(lhs - rhs) / abs(rhs):.2e}")
    assert abs(lhs - rhs) <= 2e-5 * abs(rhs)


def _dirty_image(w, vis, pixels):
    """Direct FP64 DFT dirty image sum_v V exp(-i k (u l + v m + w n)) at master pixels (X, Y)."""
    G, im = int(w["grid_size"]), float(w["image_size"])
    uvw = w["uvw"].astype(np.float64)
    k = w["wavenumbers"].astype(np.float64)
    out = []
    for X, Y in pixels:
        l, m = (X - G / 2) * im / G, (Y - G / 2) * im / G
        n = (l * l + m * m) / (1 + np.sqrt(1 - l * l - m * m))
        ph = -(uvw[:, 0:1] * l + uvw[:, 1:2] * m + uvw[:, 2:3] * n) * k[None, :]
        out.append((vis * np.exp(1j * ph)[..., None]).sum((0, 1)))
    return np.array(out)


def _image(dw, plan, w, fused=False):
    G = int(w["grid_size"])
    z_min, nl = dw.layers()
    dw.grid(plan)
    dw.fft(plan, FFT_FORWARD)
    d_layers = dw.layer_buffer(nl)
    dw.add_layered(plan, d_layers)
    d_img = dw.layer_buffer(1)
    if fused:
        dw.grid_to_image(plan, d_layers, d_img, grid_box(plan, w["metadata"]))
    else:
        dw.grid_fft(plan, d_layers, nl, FFT_BACKWARD)
        dw.to_image(plan, d_layers, d_img)
    return download(d_img, np.complex64, (4, G, G))


@pytest.mark.parametrize("precision,fused", [("fp32", False), ("fp16x3", False), ("fp16x3", True)])
def test_point_source_through_the_wstacked_chain_matches_the_direct_dft(wstack, precision, fused):
    """A point source at (l0, m0) with the full w-term, imaged with w-stacking, divided by the chain's response to
    a visibility at the uv origin (the taper correction), equals the direct FP64 DFT dirty image to 5e-5 of its
    peak over the central half of the field (IDG is exact up to the taper's aliasing); the peak sits at the source
    pixel.  fused: idgx_grid_to_image on the subgrids' box."""
    w = dict(wstack)
    G, im = int(w["grid_size"]), float(w["image_size"])
    X0, Y0 = G // 2 + 301, G // 2 - 197
    l0, m0 = (X0 - G / 2) * im / G, (Y0 - G / 2) * im / G
    n0 = (l0 * l0 + m0 * m0) / (1 + np.sqrt(1 - l0 * l0 - m0 * m0))
    uvw = w["uvw"].astype(np.float64)
    k = w["wavenumbers"].astype(np.float64)
    ph = (uvw[:, 0:1] * l0 + uvw[:, 1:2] * m0 + uvw[:, 2:3] * n0) * k[None, :]
    pol = np.array([1.0, 0.0, 0.0, 1.0])
    w["visibilities"] = (np.exp(1j * ph)[..., None] * pol).astype(np.complex64)
    dw = DeviceWorkload(w)
    img = _image(dw, dw.plan(precision=precision), w, fused)[0]
    # Response to one visibility at the uv origin, in a subgrid centred on the grid (layer 0).
    md = np.zeros(1, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    S = int(w["subgrid_size"])
    md["x"] = md["y"] = G // 2 - S // 2
    ch = np.zeros(1, CHANNELS_DTYPE)
    ch["nr_channels"] = 1
    d = dict(w, metadata=md, channels=ch, uvw=np.zeros((1, 3), np.float32), nr_subgrids=1, max_nr_timesteps=1,
             visibilities=(np.ones((1, w["wavenumbers"].shape[0], 1)) * pol).astype(np.complex64))
    dd = DeviceWorkload(d)
    resp = _image(dd, dd.plan(precision=precision), d, fused)[0]
    pix = [(X, Y) for X in range(G // 4, 3 * G // 4, 61) for Y in range(G // 4, 3 * G // 4, 59)] + \
          [(X0 + dx, Y0 + dy) for dx in (-1, 0, 1) for dy in (-1, 0, 1)]
    ref = _dirty_image(w, w["visibilities"].astype(np.complex128), pix)[:, 0]
    got = np.array([img[Y, X] / resp[Y, X] for X, Y in pix])
    peak = np.abs(ref).max()
    err = np.abs(got - ref).max() / peak
    c0, c1 = G // 4, 3 * G // 4
    Yp, Xp = np.unravel_index(np.argmax(np.abs(img / resp)[c0:c1, c0:c1]), (c1 - c0, c1 - c0))
    Yp, Xp = Yp + c0, Xp + c0
    print(f"{precision}{' fused' if fused else ''}: point source through the w-stacked chain vs direct DFT: max "
          f"error {err:.2e} of the peak; "
          f"peak at ({Xp}, {Yp}), source at ({X0}, {Y0})")
    assert err < 5e-5 and (Xp, Yp) == (X0, Y0)
