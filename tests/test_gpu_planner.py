"""The device-side planner (idgx_plan_segments) against the host planner (observation.plan_segments): bit-identical
metadata and channel windows on every configuration's geometry (fixed and ragged time segments, A-term slot cuts,
w-layers, channel grouping, off-grid drops)."""
import numpy as np
import pytest

from idgx import observation, plan_segments_device

pytestmark = pytest.mark.gpu


def _geometry(cfg):
    rng = np.random.default_rng(cfg["seed"])
    enu = observation.station_layout(rng, cfg["n_stations"], cfg["radius_m"], cfg["core_fraction"],
                                     cfg["core_radius_m"])
    s1, s2 = observation.baseline_pairs(cfg["n_stations"])
    uvw = observation.earth_rotation_uvw(enu, s1, s2, cfg["n_times"], cfg["dec_deg"])
    if cfg.get("zero_w"):
        uvw[..., 2] = 0.0
    freqs = np.linspace(cfg["f_lo"], cfg["f_hi"], cfg["n_channels"])
    return uvw.astype(np.float32), freqs, s1, s2


def _compare(uvw32, freqs, s1, s2, bl, t0, t1, **kw):
    md_h, ch_h = observation.plan_segments(uvw32.astype(np.float64), freqs, s1, s2, bl, t0, t1, **kw)
    md_d, ch_d = plan_segments_device(uvw32.reshape(-1, 3), freqs, s1, s2, bl, t0, t1,
                                      nr_times=uvw32.shape[1], subgrid_size=kw["subgrid_size"],
                                      grid_size=kw["grid_size"], image_size=kw["image_size"],
                                      aterm_interval=kw.get("aterm_interval") or 0, w_step=kw.get("w_step") or 0.0)
    assert md_d.size == md_h.size and md_h.size > 0
    assert md_d.tobytes() == md_h.tobytes() and ch_d.tobytes() == ch_h.tobytes()
    return md_h.size


@pytest.mark.parametrize("name", ["small", "ska_low", "full_cycle", "wstack"])
def test_device_planner_matches_the_host_planner(name):
    cfg = observation.CONFIGS[name]
    uvw32, freqs, s1, s2 = _geometry(cfg)
    bl, t0, t1 = observation.time_segments(len(s1), cfg["n_times"], cfg["timesteps_per_subgrid"],
                                           cfg.get("aterm_interval"))
    order = np.lexsort((t0, bl))  # segments sorted by (baseline, t0)
    n = _compare(uvw32, freqs, s1, s2, bl[order], t0[order], t1[order], subgrid_size=cfg["subgrid_size"],
                 grid_size=cfg["grid_size"], image_size=cfg["image_size"], aterm_interval=cfg.get("aterm_interval"),
                 w_step=cfg.get("w_step"))
    print(f"{name}: {n} subgrids, device plan bit-identical to the host plan")


def test_device_planner_matches_on_ragged_segments_and_three_subgrid_sizes():
    cfg = observation.BATCHED_CONFIGS["batched_uneven_small"]
    uvw32, freqs, s1, s2 = _geometry(cfg)
    seg_rng = np.random.default_rng(3)
    for S in cfg["subgrid_sizes"]:
        bl, t0, t1 = observation.ragged_time_segments(seg_rng, np.arange(len(s1)), cfg["n_times"], *cfg["t_range"],
                                                      aterm_interval=40)
        n = _compare(uvw32, freqs, s1, s2, bl, t0, t1, subgrid_size=S, grid_size=cfg["grid_size"],
                     image_size=cfg["image_size"], aterm_interval=40, w_step=75.0)
        print(f"ragged S={S}: {n} subgrids, device plan bit-identical to the host plan")
