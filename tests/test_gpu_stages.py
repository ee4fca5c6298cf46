"""Subgrid FFT, adder and splitter on the GPU, and the whole imaging chain.

The transforms have no reference counterpart (the reference stops at image-domain subgrids); they are pinned by
the oracle's brute-force DFT (itself pinned to numpy.fft) with the transform's defining shift factors, by round
trips, by a closed form through the gridder, by the oracle's brute-force adder/splitter, and by adjoint
identities.
"""
import numpy as np
import pytest

import oracle
from helpers import rel_l2
from idgx import FFT_BACKWARD, FFT_FORWARD, METADATA_DTYPE, make_plan, observation
from idgx.device import DeviceWorkload, upload

pytestmark = pytest.mark.gpu


def _workload(S, G, md, n_stations=2):
    return dict(metadata=md, channels=None, subgrid_size=S, grid_size=G, nr_stations=n_stations, nr_aterm_slots=1,
                image_size=np.float32(0.05), w_step=0.0, wavenumbers=np.full(1, 3.0, np.float32),
                uvw=np.zeros((max(1, int(md["nr_timesteps"].sum())), 3), np.float32),
                visibilities=np.zeros((max(1, int(md["nr_timesteps"].sum())), 1, 4), np.complex64),
                spheroidal=np.ones((S, S), np.float32), aterms=observation.identity_aterms(1, n_stations, S),
                max_nr_timesteps=int(md["nr_timesteps"].max()), nr_subgrids=md.size)


def _spec_forward(x):
    """The forward transform by definition, evaluated with the oracle's brute-force DFT."""
    S = x.shape[-1]
    n = np.arange(S)
    pre = (-1.0) ** n
    post = np.exp(1j * np.pi * (n - S / 2) * (1 + 1 / S))
    y = oracle.dft2d_fp64(x * pre[:, None] * pre[None, :], +1)
    return y * post[:, None] * post[None, :]


def _spec_backward(x):
    S = x.shape[-1]
    n = np.arange(S)
    pre = np.conj(np.exp(1j * np.pi * (n - S / 2) * (1 + 1 / S)))
    post = (-1.0) ** n
    y = oracle.dft2d_fp64(x * pre[:, None] * pre[None, :], -1)
    return y * post[:, None] * post[None, :]


def test_spec_is_the_half_pixel_centred_transform():
    """The diagonal-factor form used above equals the defining sum exp(+2 pi i (k - S/2)(x - S/2 + 1/2)/S)."""
    S = 12
    rng = np.random.default_rng(0)
    x = rng.standard_normal((S, S)) + 1j * rng.standard_normal((S, S))
    k = np.arange(S)[:, None] - S / 2
    j = np.arange(S)[None, :] - S / 2 + 0.5
    F = np.exp(2j * np.pi * k * j / S)
    np.testing.assert_allclose(_spec_forward(x[None])[0], F @ x @ F.T, atol=1e-9)
    np.testing.assert_allclose(_spec_backward(x[None])[0], F.conj().T @ x @ F.conj(), atol=1e-9)


@pytest.mark.parametrize("S", [8, 12, 16, 24, 32, 48, 64])
def test_subgrid_fft_matches_dft(S):
    rng = np.random.default_rng(S)
    n = 5
    x = (rng.standard_normal((n, 4, S, S)) + 1j * rng.standard_normal((n, 4, S, S))).astype(np.complex64)
    md = np.zeros(n, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    dw = DeviceWorkload(_workload(S, 256, md))
    plan = dw.plan()
    dw.d_subgrids.copy_(upload(x))
    dw.fft(plan, FFT_FORWARD)
    fwd = dw.subgrids()
    e_f = rel_l2(fwd, _spec_forward(x.astype(np.complex128)))
    dw.d_subgrids.copy_(upload(x))
    dw.fft(plan, FFT_BACKWARD)
    e_b = rel_l2(dw.subgrids(), _spec_backward(x.astype(np.complex128)))
    dw.d_subgrids.copy_(upload(fwd))
    dw.fft(plan, FFT_BACKWARD)
    e_rt = rel_l2(dw.subgrids() / (S * S), x)
    print(f"S={S}: forward {e_f:.2e} backward {e_b:.2e} round trip {e_rt:.2e}")
    assert max(e_f, e_b, e_rt) < 2e-6


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_gridder_then_fft_puts_a_visibility_on_its_cell(precision):
    """A visibility exactly j cells from the subgrid centre grids (unit taper, identity A-terms) and transforms
    into S^2 V at cell S/2 + j and zero elsewhere."""
    S, G, im = 32, 512, 0.05
    md = np.zeros(1, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    md["x"], md["y"] = 200, 180
    k = 3.0
    ju, jv = 5, -3
    u_c = (md["x"][0] + S / 2 - G / 2) / im
    v_c = (md["y"][0] + S / 2 - G / 2) / im
    w = _workload(S, G, md)
    w["image_size"] = np.float32(im)
    w["wavenumbers"] = np.full(1, k, np.float32)
    # u in metres with k u / 2 pi = u_c + ju / image_size; the float rounding leaves ~1e-5 of a cell.
    w["uvw"] = np.array([[(u_c + ju / im) * 2 * np.pi / k, (v_c + jv / im) * 2 * np.pi / k, 0.0]], np.float32)
    V = np.array([1 + 2j, -0.5j, 0.25, 3 - 1j], np.complex64)
    w["visibilities"] = V.reshape(1, 1, 4)
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    dw.fft(plan, FFT_FORWARD)
    g = dw.subgrids()[0]
    peak = g[:, S // 2 + jv, S // 2 + ju]
    np.testing.assert_allclose(peak, S * S * V, rtol=2e-3)
    rest = g.copy()
    rest[:, S // 2 + jv, S // 2 + ju] = 0
    assert np.abs(rest).max() < 2e-3 * S * S


def _overlapping_subgrids(rng, n, S, G):
    md = np.zeros(n, METADATA_DTYPE)
    md["nr_timesteps"] = 1
    md["station2"] = 1
    md["x"] = rng.integers(-S // 2, G - S // 2, n)
    md["y"] = rng.integers(-S // 2, G - S // 2, n)
    md["x"][:6] = 40  # a pile of identical subgrids in one tile
    md["y"][:6] = 40
    return md


@pytest.mark.parametrize("S,G", [(16, 128), (24, 200), (48, 256), (64, 320)])
def test_adder_and_splitter_match_oracle_exactly_and_are_adjoint(S, G):
    """Integer-valued data (exact sums), piles of subgrids in one tile, subgrids hanging over every edge and a few
    far outside the grid (their tile index is clamped by the plan)."""
    rng = np.random.default_rng(2)
    n = 300
    md = _overlapping_subgrids(rng, n, S, G)
    md["x"][6:9] = [-3 * S, G + 2 * S, -S + 3]
    md["y"][6:9] = [5, -4 * S, G - 2]
    sub = (rng.integers(-8, 9, (n, 4, S, S)) + 1j * rng.integers(-8, 9, (n, 4, S, S))).astype(np.complex64)
    dw = DeviceWorkload(_workload(S, G, md))
    plan = dw.plan()
    dw.d_subgrids.copy_(upload(sub))
    d_grid = dw.grid_buffer()
    dw.add(plan, d_grid)
    got = dw.grid_array(d_grid)
    np.testing.assert_array_equal(got, oracle.adder_fp64(md, sub, G).astype(np.complex64))
    gridv = (rng.standard_normal((4, G, G)) + 1j * rng.standard_normal((4, G, G))).astype(np.complex64)
    d_grid2 = upload(gridv)
    dw.split(plan, d_grid2)
    split = dw.subgrids()
    np.testing.assert_array_equal(split, oracle.splitter_fp64(md, gridv, S).astype(np.complex64))
    lhs = np.vdot(split.astype(np.complex128), sub.astype(np.complex128))
    rhs = np.vdot(gridv.astype(np.complex128), got.astype(np.complex128))
    assert abs(lhs - rhs) <= 1e-9 * abs(rhs)


def test_adder_is_deterministic_with_real_valued_data():
    rng = np.random.default_rng(3)
    S, G, n = 32, 256, 800
    md = _overlapping_subgrids(rng, n, S, G)
    sub = (rng.standard_normal((n, 4, S, S)) + 1j * rng.standard_normal((n, 4, S, S))).astype(np.complex64)
    dw = DeviceWorkload(_workload(S, G, md))
    plan = dw.plan()
    dw.d_subgrids.copy_(upload(sub))
    a, b = dw.grid_buffer(), dw.grid_buffer()
    dw.add(plan, a)
    dw.add(plan, b)
    ga, gb = dw.grid_array(a), dw.grid_array(b)
    assert ga.tobytes() == gb.tobytes()
    assert rel_l2(ga, oracle.adder_fp64(md, sub, G)) < 1e-6


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_imaging_chain_adjoint(precision):
    """<degrid(fft_b(split(G))), V> = <G, add(fft_f(grid(V)))> on the small configuration with Jones A-terms."""
    w = observation.make_observation("small")
    rng = np.random.default_rng(12)
    S, Gs = w["subgrid_size"], w["grid_size"]
    a = rng.standard_normal(w["aterms"].shape) + 1j * rng.standard_normal(w["aterms"].shape)
    w["aterms"] = (w["aterms"] + 0.2 * a).astype(np.complex64)
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    dw.fft(plan, FFT_FORWARD)
    d_grid = dw.grid_buffer()
    dw.add(plan, d_grid)
    gV = dw.grid_array(d_grid).astype(np.complex128)
    Gm = (rng.standard_normal((4, Gs, Gs)) + 1j * rng.standard_normal((4, Gs, Gs))).astype(np.complex64)
    dw.split(plan, upload(Gm))
    dw.fft(plan, FFT_BACKWARD)
    dw.degrid(plan)
    dG = dw.visibilities_out().astype(np.complex128)
    lhs = np.vdot(dG, w["visibilities"].astype(np.complex128))
    rhs = np.vdot(Gm.astype(np.complex128), gV)
    print(f"chain adjoint {precision}: rel diff {abs(lhs - rhs) / abs(rhs):.2e}")
    assert abs(lhs - rhs) <= 2e-5 * abs(rhs)
