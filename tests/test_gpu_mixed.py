"""Mixed-precision gridder/degridder (FP64 phase reduction, FP32 phasors, BF16/FP16 MFMA contraction) against the
FP64 oracle.

Bars (relative L2 vs the FP64 oracle, whole output and worst subgrid):
  bf16x3, fp16x3 (hi/lo split, three products): <= 5e-5
  bf16, fp16 (single product): reported; bounded by their unit roundoff (bf16 2^-9, fp16 2^-12)
"""
import numpy as np
import pytest

from helpers import load_golden, oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2
from idgx import ATERM_ADJOINT, ATERM_REFERENCE, observation
from idgx.device import DeviceWorkload, upload

pytestmark = pytest.mark.gpu

SPLIT_BAR = 5e-5
SINGLE_BARS = {"bf16": 8e-3, "fp16": 1e-3}
PRECISIONS = ["bf16x3", "fp16x3", "bf16", "fp16"]


def _bar(precision):
    return SPLIT_BAR if precision.endswith("x3") else SINGLE_BARS[precision]


def _vis_errors(w, got, ref):
    md, ch = w["metadata"], w.get("channels")
    C = w["wavenumbers"].shape[0]
    errs = []
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        c0, nc = (0, C) if ch is None else (int(ch["channel_begin"][s]), int(ch["nr_channels"][s]))
        errs.append(rel_l2(got[r0:r0 + nt, c0:c0 + nc], ref[r0:r0 + nt, c0:c0 + nc]))
    return np.array(errs)


@pytest.fixture(scope="module")
def small():
    w = observation.make_observation("small")
    rng = np.random.default_rng(21)
    sub = (rng.standard_normal((w["nr_subgrids"], 4, 32, 32)) +
           1j * rng.standard_normal((w["nr_subgrids"], 4, 32, 32))).astype(np.complex64)
    return w, oracle_grid(w, fp64=True), sub, oracle_degrid(w, sub, fp64=True)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_small_config_mixed_vs_fp64(small, precision):
    w, g64, sub, v64 = small
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    got = dw.subgrids()
    eg = rel_l2(got, g64)
    egs = per_subgrid_rel_l2(got, g64).max()
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    vis = dw.visibilities_out()
    ed = rel_l2(vis, v64)
    eds = _vis_errors(w, vis, v64).max()
    print(f"small {precision}: gridder vs FP64 {eg:.3e} (worst subgrid {egs:.3e}); degridder {ed:.3e} "
          f"(worst {eds:.3e})")
    bar = _bar(precision)
    assert egs <= bar and eds <= bar


@pytest.mark.parametrize("S", [8, 16, 20, 24, 40, 48])
def test_one_shot_tile_choice_per_subgrid_size_vs_fp64(S):
    """The one-shot gridder picks its tiles per wave from the subgrid size (S = 8, 20: 2; 16: 4; 24: 3; 40, 48: 6,
    with strips 8, 16, 4, 8, 8, 16 columns wide); every choice meets the split-form bar against FP64."""
    w = observation.make_observation("small", subgrid_size=S, timesteps_per_subgrid=64)
    w = observation.subset(w, np.arange(0, w["nr_subgrids"], 2))
    rng = np.random.default_rng(S)
    sub = (rng.standard_normal((w["nr_subgrids"], 4, S, S)) +
           1j * rng.standard_normal((w["nr_subgrids"], 4, S, S))).astype(np.complex64)
    dw = DeviceWorkload(w)
    plan = dw.plan(precision="fp16x3")
    dw.grid(plan)
    egs = per_subgrid_rel_l2(dw.subgrids(), oracle_grid(w, fp64=True)).max()
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    eds = _vis_errors(w, dw.visibilities_out(), oracle_degrid(w, sub, fp64=True)).max()
    print(f"S={S} ({w['nr_subgrids']} subgrids) fp16x3 one-shot: gridder worst {egs:.2e}, degridder worst {eds:.2e}")
    assert egs <= SPLIT_BAR and eds <= SPLIT_BAR


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_ska_low_subset_mixed_meets_5e5_where_fp32_does_not(precision):
    full = observation.make_observation("ska_low", visibilities=False)
    idx = np.arange(3, full["nr_subgrids"], 457)
    w = observation.subset(full, idx)
    w["visibilities"] = observation.random_visibilities(w["nr_rows"], w["wavenumbers"].shape[0], seed=9)
    g64 = oracle_grid(w, fp64=True)
    dw = DeviceWorkload(w)
    dw.grid(dw.plan(precision=precision))
    got = dw.subgrids()
    dw.grid(dw.plan(precision="fp32"))
    got32 = dw.subgrids()
    e_mixed = per_subgrid_rel_l2(got, g64)
    e_plain = per_subgrid_rel_l2(got32, g64)
    print(f"ska_low subset ({idx.size} subgrids) gridder vs FP64: {precision} whole {rel_l2(got, g64):.3e} worst "
          f"{e_mixed.max():.3e}; plain fp32 whole {rel_l2(got32, g64):.3e} worst {e_plain.max():.3e}")
    assert rel_l2(got, g64) <= SPLIT_BAR and e_mixed.max() <= SPLIT_BAR
    rng = np.random.default_rng(4)
    sub = (rng.standard_normal(got.shape) + 1j * rng.standard_normal(got.shape)).astype(np.complex64)
    v64 = oracle_degrid(w, sub, fp64=True)
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(dw.plan(precision=precision))
    vis = dw.visibilities_out()
    ev = _vis_errors(w, vis, v64)
    print(f"ska_low subset degridder vs FP64: {precision} whole {rel_l2(vis, v64):.3e} worst {ev.max():.3e}")
    assert rel_l2(vis, v64) <= SPLIT_BAR and ev.max() <= SPLIT_BAR


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_golden_aterms_and_taper_mixed(precision):
    """Non-identity A-terms and the reference taper: the mixed gridder vs FP64, and the mixed degridder in the
    reference A-term convention vs FP64 of A1^H P A2 (hermitian A-terms through the adjoint oracle)."""
    w = load_golden("ref_example")
    g64 = oracle_grid(w, fp64=True)
    dw = DeviceWorkload(w)
    dw.grid(dw.plan(precision=precision))
    eg = per_subgrid_rel_l2(dw.subgrids(), g64).max()
    wh = dict(w, aterms=np.ascontiguousarray(np.conj(w["aterms"][..., [0, 2, 1, 3]])))
    v64 = oracle_degrid(wh, w["ref_degrid_subgrids"], fp64=True)
    dw.d_subgrids.copy_(upload(w["ref_degrid_subgrids"]))
    dw.degrid(dw.plan(precision=precision, aterm_convention=ATERM_REFERENCE))
    ev = rel_l2(dw.visibilities_out(), v64)
    print(f"golden ref_example {precision}: gridder worst subgrid vs FP64 {eg:.3e}; degridder {ev:.3e}")
    assert eg <= SPLIT_BAR and ev <= SPLIT_BAR


def test_mixed_adjoint_dot_product():
    """<degrid(G), V> = <G, grid(V)> for the bf16x3 kernels, to the mixed-precision tolerance."""
    w = observation.make_observation("small")
    rng = np.random.default_rng(8)
    S = 32
    G = (rng.standard_normal((w["nr_subgrids"], 4, S, S)) + 1j * rng.standard_normal((w["nr_subgrids"], 4, S, S))
         ).astype(np.complex64)
    dw = DeviceWorkload(w)
    plan = dw.plan(precision="bf16x3", aterm_convention=ATERM_ADJOINT)
    dw.grid(plan)
    gV = dw.subgrids().astype(np.complex128)
    dw.d_subgrids.copy_(upload(G))
    dw.degrid(plan)
    dG = dw.visibilities_out().astype(np.complex128)
    lhs = np.vdot(dG, w["visibilities"].astype(np.complex128))
    rhs = np.vdot(G.astype(np.complex128), gV)
    print(f"adjoint: |lhs - rhs| / |rhs| = {abs(lhs - rhs) / abs(rhs):.3e}")
    assert abs(lhs - rhs) <= 5e-5 * abs(rhs)


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_mixed_forms_handle_large_dynamic_range(precision):
    """Visibilities and pixels far beyond the FP16 range (65504): the per-batch power-of-two scaling keeps the
    FP16 forms exact to the split precision; BF16 has the FP32 exponent range anyway."""
    w = observation.make_observation("small")
    w["visibilities"] = (w["visibilities"] * np.float32(3e6)).astype(np.complex64)
    w["visibilities"][::7] *= np.float32(1e-4)  # and a spread of magnitudes
    dw = DeviceWorkload(w)
    plan = dw.plan(precision=precision)
    dw.grid(plan)
    got = dw.subgrids()
    g64 = oracle_grid(w, fp64=True)
    eg = per_subgrid_rel_l2(got, g64).max()
    rng = np.random.default_rng(2)
    sub = (rng.standard_normal(got.shape) + 1j * rng.standard_normal(got.shape)).astype(np.complex64) * np.float32(2e7)
    dw.d_subgrids.copy_(upload(sub))
    dw.degrid(plan)
    ev = rel_l2(dw.visibilities_out(), oracle_degrid(w, sub, fp64=True))
    print(f"{precision} large range: gridder worst subgrid {eg:.3e}, degridder {ev:.3e}")
    assert np.isfinite(got).all() and eg <= SPLIT_BAR and ev <= SPLIT_BAR
