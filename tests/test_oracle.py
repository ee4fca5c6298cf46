"""Pins of the oracle (CPU, no GPU needed).

- The FP32 restatement reproduces the compiled reference's CPU gridder and degridder bit for bit on every golden
  vector (tests/golden/, produced by tests/golden/make_golden.py from the reference sources).
- The FP64 oracle gridder agrees with the bit-exact FP32 restatement where FP32 phase rounding is small (short
  baselines), and reduces to textbook closed forms: a visibility at the subgrid's phase centre grids to a constant
  subgrid; one cell off-centre grids to the DFT basis vector of that cell.
- The FP64 degridder is the exact adjoint of the FP64 gridder: <degrid(G), V> = <G, grid(V)> with random
  non-identity A-terms, taper, subgrids and visibilities.
- The brute-force DFT matches numpy.fft; the adder/splitter match direct slicing and are mutual adjoints.
"""
import numpy as np
import pytest

import oracle
from helpers import golden_names, load_golden, oracle_degrid, oracle_grid, rel_l2
from idgx import METADATA_DTYPE, observation


@pytest.mark.parametrize("name", golden_names())
def test_fp32_gridder_bit_exact_vs_reference(name):
    w = load_golden(name)
    got = oracle_grid(w, nthreads=4)
    assert np.array_equal(got.view(np.uint32), w["ref_grid_subgrids"].view(np.uint32))


@pytest.mark.parametrize("name", golden_names())
def test_fp32_degridder_bit_exact_vs_reference(name):
    w = load_golden(name)
    got = oracle_degrid(w, w["ref_degrid_subgrids"], convention=oracle.ATERM_REFERENCE, nthreads=4)
    assert np.array_equal(got.view(np.uint32), w["ref_degrid_visibilities"].view(np.uint32))


def test_fp32_restatement_is_contraction_sensitive():
    """A plausible mistake (a fused multiply-add in the phase) must be visible at the bit level: perturbing one
    wavenumber by one ulp changes the output, so bit-exactness above is not vacuous."""
    w = load_golden("ref_s32")
    w["wavenumbers"] = w["wavenumbers"].copy()
    w["wavenumbers"][3] = np.nextafter(w["wavenumbers"][3], np.float32(10))
    got = oracle_grid(w, nthreads=4)
    assert not np.array_equal(got, w["ref_grid_subgrids"])


def _one_subgrid(S=16, G=256, image_size=0.08, T=1, C=1, wavenumber=3.0):
    md = np.zeros(1, METADATA_DTYPE)
    md["nr_timesteps"] = T
    md["station1"], md["station2"] = 0, 1
    md["x"], md["y"] = 100, 90
    w = dict(metadata=md, channels=None, subgrid_size=S, grid_size=G, nr_stations=2, nr_aterm_slots=1,
             image_size=np.float32(image_size), w_step=0.0, wavenumbers=np.full(C, wavenumber, np.float32),
             spheroidal=np.ones((S, S), np.float32), aterms=observation.identity_aterms(1, 2, S))
    return w


def _centre_uv(w):
    """uv (metres) whose k*u equals the subgrid's phase-centre offset (pure double arithmetic)."""
    S, G, im = w["subgrid_size"], w["grid_size"], float(w["image_size"])
    k = float(w["wavenumbers"][0])
    md = w["metadata"][0]
    u_off = (md["x"] + S / 2 - G / 2) * 2 * np.pi / im
    v_off = (md["y"] + S / 2 - G / 2) * 2 * np.pi / im
    return u_off / k, v_off / k


def test_fp64_gridder_closed_form_phase_centre():
    w = _one_subgrid()
    u, v = _centre_uv(w)
    w["uvw"] = np.array([[u, v, 0.0]], np.float32)
    vis = np.zeros((1, 1, 4), np.complex64)
    vis[0, 0] = [1 + 2j, 0.5j, -1, 3]
    w["visibilities"] = vis
    got = oracle_grid(w, fp64=True)[0]
    # u, v were rounded to float: the residual phase is below 1e-3 rad at this size.
    for pol in range(4):
        np.testing.assert_allclose(got[pol], np.full((16, 16), vis[0, 0, pol]), atol=2e-3 * abs(vis[0, 0, pol]) + 1e-12)


def test_fp64_gridder_closed_form_one_cell_off_centre():
    """k*u = u_off + 2 pi / image_size puts the sample one cell right of the centre: the subgrid is the DFT basis
    vector exp(-2 pi i (x + 0.5 - S/2) / S), whose 2-D FFT is a single spike."""
    w = _one_subgrid(S=16)
    S, im = 16, float(w["image_size"])
    u, v = _centre_uv(w)
    k = float(w["wavenumbers"][0])
    w["uvw"] = np.array([[u + 2 * np.pi / im / k, v, 0.0]], np.float64).astype(np.float32)
    vis = np.zeros((1, 1, 4), np.complex64)
    vis[0, 0, 0] = 1
    w["visibilities"] = vis
    got = oracle_grid(w, fp64=True)[0, 0]
    x = np.arange(S)
    expect = np.exp(-2j * np.pi * (x + 0.5 - S / 2) / S)[None, :] * np.ones((S, 1))
    assert rel_l2(got, expect) < 5e-3
    spectrum = np.abs(np.fft.fft2(got))
    assert spectrum.argmax() == S - 1  # e^{-2 pi i x/S} is frequency -1 -> bin S-1 of row 0
    assert spectrum.max() > 0.99 * S * S


def test_fp64_gridder_agrees_with_bit_exact_fp32_on_short_baselines():
    w = observation.make_observation("small")
    g32 = oracle_grid(w, nthreads=8)
    g64 = oracle_grid(w, fp64=True, nthreads=8)
    assert rel_l2(g32, g64) < 1e-5


def test_fp64_degridder_agrees_with_fp32_restatement_on_short_baselines():
    w = observation.make_observation("small")
    rng = np.random.default_rng(3)
    sub = (rng.standard_normal((w["nr_subgrids"], 4, 32, 32)) + 1j * rng.standard_normal((w["nr_subgrids"], 4, 32, 32))
           ).astype(np.complex64)
    v32 = oracle_degrid(w, sub, nthreads=8)
    v64 = oracle_degrid(w, sub, fp64=True, nthreads=8)
    assert rel_l2(v32, v64) < 1e-5


def _random_jones(rng, n_slots, n_st, S):
    a = rng.standard_normal((n_slots, n_st, S, S, 4)) + 1j * rng.standard_normal((n_slots, n_st, S, S, 4))
    return (0.3 * a + np.array([1, 0, 0, 1])).astype(np.complex64)


def test_fp64_degridder_is_adjoint_of_gridder():
    rng = np.random.default_rng(11)
    S, C, T = 16, 5, 7
    n_st = 3
    md = np.zeros(3, METADATA_DTYPE)
    md["time_offset"] = [0, T, 2 * T]
    md["nr_timesteps"] = T
    md["station1"] = [0, 0, 1]
    md["station2"] = [1, 2, 2]
    md["aterm_index"] = [0, 1, 1]
    md["x"] = [40, 60, 70]
    md["y"] = [50, 45, 80]
    md["z"] = [0, 1, 2]
    ch = np.zeros(3, oracle.CHANNELS_DTYPE)
    ch["channel_begin"] = [0, 1, 2]
    ch["nr_channels"] = [5, 3, 2]
    w = dict(metadata=md, channels=ch, subgrid_size=S, grid_size=128, nr_stations=n_st, nr_aterm_slots=2,
             image_size=np.float32(0.05), w_step=7.5,
             wavenumbers=np.linspace(3.0, 3.5, C).astype(np.float32),
             uvw=(rng.standard_normal((3 * T, 3)) * 300).astype(np.float32),
             spheroidal=rng.random((S, S)).astype(np.float32),
             aterms=_random_jones(rng, 2, n_st, S))
    V = (rng.standard_normal((3 * T, C, 4)) + 1j * rng.standard_normal((3 * T, C, 4))).astype(np.complex64)
    G = (rng.standard_normal((3, 4, S, S)) + 1j * rng.standard_normal((3, 4, S, S))).astype(np.complex64)
    w["visibilities"] = V
    gV = oracle_grid(w, fp64=True)
    dG = oracle_degrid(w, G, fp64=True)
    mask = np.zeros((3 * T, C), bool)
    for s in range(3):
        mask[md["time_offset"][s]:md["time_offset"][s] + T, ch["channel_begin"][s]:ch["channel_begin"][s] + ch["nr_channels"][s]] = True
    lhs = np.vdot(dG[mask], V.astype(np.complex128)[mask])
    rhs = np.vdot(G.astype(np.complex128), gV)
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs)
    # A transposed A-term in the degridder breaks the identity: the check is not vacuous.
    wt = dict(w, aterms=np.ascontiguousarray(w["aterms"][..., [0, 2, 1, 3]]))
    dGt = oracle_degrid(wt, G, fp64=True)
    assert abs(np.vdot(dGt[mask], V.astype(np.complex128)[mask]) - rhs) > 1e-6 * abs(rhs)


def test_dft2d_matches_numpy_fft():
    rng = np.random.default_rng(5)
    for S in (8, 24, 32):
        x = rng.standard_normal((3, S, S)) + 1j * rng.standard_normal((3, S, S))
        np.testing.assert_allclose(oracle.dft2d_fp64(x, -1), np.fft.fft2(x), rtol=0, atol=1e-9 * S * S)
        np.testing.assert_allclose(oracle.dft2d_fp64(x, +1), np.fft.ifft2(x) * S * S, rtol=0, atol=1e-9 * S * S)


def test_adder_and_splitter_match_slicing_and_are_adjoint():
    rng = np.random.default_rng(9)
    S, G, n = 8, 40, 6
    md = np.zeros(n, METADATA_DTYPE)
    md["x"] = [0, 5, 32, 36, -3, 10]  # includes subgrids overhanging the edges
    md["y"] = [0, 7, 30, 10, 2, 36]
    sub = (rng.standard_normal((n, 4, S, S)) + 1j * rng.standard_normal((n, 4, S, S))).astype(np.complex64)
    grid = oracle.adder_fp64(md, sub, G)
    expect = np.zeros((4, G + 2 * S, G + 2 * S), np.complex128)
    for s in range(n):
        y0, x0 = md["y"][s] + S, md["x"][s] + S
        expect[:, y0:y0 + S, x0:x0 + S] += sub[s]
    np.testing.assert_allclose(grid, expect[:, S:S + G, S:S + G], atol=1e-12)
    g = rng.standard_normal((4, G, G)) + 1j * rng.standard_normal((4, G, G))
    split = oracle.splitter_fp64(md, g, S)
    np.testing.assert_allclose(np.vdot(split, sub.astype(np.complex128)), np.vdot(g, grid), rtol=1e-12)
