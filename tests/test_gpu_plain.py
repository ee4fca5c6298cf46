"""Plain FP32 HIP gridder/degridder vs the reference CPU path (golden vectors) and the oracle.

Bar: relative L2 error <= 1e-5 per subgrid (gridder) and per subgrid's visibilities (degridder) against the
reference's own CPU output; FP64-oracle error is computed and bounded by the FP32 phase-rounding floor.
"""
import numpy as np
import pytest

import oracle
from helpers import covered_mask, golden_names, load_golden, oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2
from idgx import ATERM_ADJOINT, ATERM_REFERENCE, FLAG_ACCURATE_SINCOS
from idgx.device import DeviceWorkload
from idgx import observation

pytestmark = pytest.mark.gpu

PARITY = 1e-5  # relative L2 per subgrid, plain FP32 vs the reference CPU path
SINCOS_FLAGS = [0, FLAG_ACCURATE_SINCOS]


def _per_subgrid_vis_error(w, got, ref):
    md, ch = w["metadata"], w.get("channels")
    C = w["wavenumbers"].shape[0]
    errs = []
    for s in range(md.size):
        r0 = int(md["time_offset"][s])
        nt = int(md["nr_timesteps"][s])
        c0, nc = (0, C) if ch is None else (int(ch["channel_begin"][s]), int(ch["nr_channels"][s]))
        errs.append(rel_l2(got[r0:r0 + nt, c0:c0 + nc], ref[r0:r0 + nt, c0:c0 + nc]))
    return np.array(errs)


@pytest.mark.parametrize("flags", SINCOS_FLAGS)
@pytest.mark.parametrize("name", golden_names())
def test_grid_matches_reference_cpu(name, flags):
    w = load_golden(name)
    dw = DeviceWorkload(w)
    dw.grid(dw.plan(flags=flags))
    got = dw.subgrids()
    err = per_subgrid_rel_l2(got, w["ref_grid_subgrids"])
    print(f"{name} flags={flags}: max per-subgrid rel-L2 vs reference CPU {err.max():.3e}")
    assert err.max() <= PARITY


@pytest.mark.parametrize("flags", SINCOS_FLAGS)
@pytest.mark.parametrize("name", golden_names())
def test_degrid_matches_reference_cpu(name, flags):
    w = load_golden(name)
    w["visibilities"] = np.zeros_like(w["visibilities"])
    dw = DeviceWorkload(w)
    dw.d_subgrids.copy_(__import__("idgx.device", fromlist=["upload"]).upload(w["ref_degrid_subgrids"]))
    dw.degrid(dw.plan(aterm_convention=ATERM_REFERENCE, flags=flags))
    got = dw.visibilities_out()
    err = _per_subgrid_vis_error(w, got, w["ref_degrid_visibilities"])
    print(f"{name} flags={flags}: max per-subgrid rel-L2 vs reference CPU {err.max():.3e}")
    assert err.max() <= PARITY


@pytest.mark.parametrize("name", golden_names())
def test_degrid_adjoint_convention_with_hermitian_aterms(name):
    """A1' P A2'^H with A' = A^H is A1^H P A2: the adjoint path reproduces the reference convention."""
    w = load_golden(name)
    a = w["aterms"]
    w["aterms"] = np.ascontiguousarray(np.conj(a[..., [0, 2, 1, 3]]))
    dw = DeviceWorkload(w)
    from idgx.device import upload

    dw.d_subgrids.copy_(upload(w["ref_degrid_subgrids"]))
    dw.degrid(dw.plan(aterm_convention=ATERM_ADJOINT))
    err = _per_subgrid_vis_error(w, dw.visibilities_out(), w["ref_degrid_visibilities"])
    assert err.max() <= PARITY


@pytest.fixture(scope="module")
def small():
    return observation.make_observation("small")


@pytest.mark.parametrize("flags", SINCOS_FLAGS)
def test_small_config_round_trip(small, flags):
    w = small
    dw = DeviceWorkload(w)
    plan = dw.plan(flags=flags)
    dw.grid(plan)
    got = dw.subgrids()
    ref32 = oracle_grid(w)
    ref64 = oracle_grid(w, fp64=True)
    e32 = per_subgrid_rel_l2(got, ref32)
    e64 = rel_l2(got, ref64)
    print(f"small grid flags={flags}: vs CPU fp32 max {e32.max():.3e}; vs FP64 oracle {e64:.3e}; "
          f"CPU fp32 vs FP64 {rel_l2(ref32, ref64):.3e}")
    assert e32.max() <= PARITY
    assert e64 <= 2 * rel_l2(ref32, ref64) + PARITY

    dw.degrid(plan)  # degrid the GPU's own subgrids (adjoint convention; identity A-terms)
    vis = dw.visibilities_out()
    vref32 = oracle_degrid(w, got)
    vref64 = oracle_degrid(w, got, fp64=True)
    ev = _per_subgrid_vis_error(w, vis, vref32)
    print(f"small degrid flags={flags}: vs CPU fp32 max {ev.max():.3e}; vs FP64 {rel_l2(vis, vref64):.3e}")
    assert ev.max() <= PARITY


def test_ska_low_subset_channel_windows():
    """A deterministic subset of the SKA-Low-like plan (every 601st subgrid): channel windows, w-term, long
    baselines."""
    full = observation.make_observation("ska_low", visibilities=False)
    idx = np.arange(0, full["nr_subgrids"], 601)
    w = observation.subset(full, idx)
    w["visibilities"] = observation.random_visibilities(w["nr_rows"], w["wavenumbers"].shape[0], seed=7)
    dw = DeviceWorkload(w)
    plan = dw.plan()
    dw.grid(plan)
    got = dw.subgrids()
    ref32 = oracle_grid(w)
    e32 = per_subgrid_rel_l2(got, ref32)
    ref64 = oracle_grid(w, fp64=True)
    print(f"ska_low subset ({idx.size} subgrids) grid: vs CPU fp32 max {e32.max():.3e}; vs FP64 {rel_l2(got, ref64):.3e}"
          f"; CPU fp32 vs FP64 {rel_l2(ref32, ref64):.3e}")
    assert e32.max() <= PARITY
    dw.degrid(plan)
    vis = dw.visibilities_out()
    vref = oracle_degrid(w, got)
    mask = covered_mask(w)
    ev = _per_subgrid_vis_error(w, vis, vref)
    print(f"ska_low subset degrid: vs CPU fp32 max {ev.max():.3e}")
    assert ev.max() <= PARITY
    assert not np.any(vis[~mask]), "entries outside the plan must stay untouched (zero)"
