"""Seeded synthetic interferometer observations and the IDG subgrid planner.

The reference benchmark only ships a toy generator (random ellipses, random subgrid corners; reproduced exactly
in workload.reference_workload).  The workloads below are the SKA-Low-like configurations the hot path is sized
for.  Conventions:

* Array site latitude -26.824722 deg (the SKA-Low site), stations on a plane (ENU, up = 0).
* Earth-rotation synthesis at declination dec with 1 s timesteps centred on transit; uvw in metres.
* Channels np.linspace(f_lo, f_hi, C); wavenumbers 2 pi f / c (float32).
* Visibilities complex64 with real and imaginary parts ~ N(0, 1).
* Identity A-terms unless Jones terms are requested; a prolate-spheroidal taper.
* Grid coordinate of a uv point (wavelengths): g = u * image_size + G / 2, so a subgrid with corner x0 has its
  phase centre at u = (x0 + S/2 - G/2) / image_size, which is the reference's u_offset
  (gridding/src/gridding_benchmark.cpp:74-81).

Planner: for each baseline the timesteps are cut into chunks of `timesteps_per_subgrid` (and at A-term slot
boundaries), and each chunk's channels are grouped greedily so that the uv footprint of every (timestep,
channel) sample of the group spans at most S - support grid cells in u and in v.  The subgrid is centred on the
footprint; groups whose subgrid would leave the grid are dropped (their visibilities are not gridded and are not
counted in the throughput).
"""
import math

import numpy as np

from . import CHANNELS_DTYPE, METADATA_DTYPE

SPEED_OF_LIGHT = 299792458.0
SITE_LATITUDE_DEG = -26.824722
EARTH_ROTATION_RAD_S = 7.2921159e-5


def station_layout(rng, n_stations, radius_m, core_fraction=0.0, core_radius_m=0.0):
    """ENU positions (m): core_fraction of the stations uniform (by area) in the core disk, the rest uniform in the
    full disk."""
    n_core = int(round(core_fraction * n_stations))
    r = np.empty(n_stations)
    r[:n_core] = core_radius_m * np.sqrt(rng.random(n_core))
    r[n_core:] = radius_m * np.sqrt(rng.random(n_stations - n_core))
    phi = 2 * np.pi * rng.random(n_stations)
    return np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(n_stations)], axis=1)


def baseline_pairs(n_stations):
    s1, s2 = np.triu_indices(n_stations, k=1)
    return s1.astype(np.uint32), s2.astype(np.uint32)


def earth_rotation_uvw(enu, s1, s2, n_times, dec_deg, dt_s=1.0, latitude_deg=SITE_LATITUDE_DEG):
    """uvw (m), float64 [n_baselines][n_times][3], hour angles centred on transit."""
    lat = math.radians(latitude_deg)
    dec = math.radians(dec_deg)
    e, n, up = enu[:, 0], enu[:, 1], enu[:, 2]
    # ENU -> equatorial (X towards the meridian at the equator, Y east, Z north pole).
    X = -math.sin(lat) * n + math.cos(lat) * up
    Y = e
    Z = math.cos(lat) * n + math.sin(lat) * up
    bx, by, bz = X[s2] - X[s1], Y[s2] - Y[s1], Z[s2] - Z[s1]
    ha = (np.arange(n_times) - (n_times - 1) / 2.0) * dt_s * EARTH_ROTATION_RAD_S
    sh, ch = np.sin(ha)[None, :], np.cos(ha)[None, :]
    bx, by, bz = bx[:, None], by[:, None], bz[:, None]
    u = sh * bx + ch * by
    v = -math.sin(dec) * ch * bx + math.sin(dec) * sh * by + math.cos(dec) * bz
    w = math.cos(dec) * ch * bx - math.cos(dec) * sh * by + math.sin(dec) * bz
    return np.stack([u, v, w], axis=-1)


def prolate_spheroidal_taper(S, c=3.0 * math.pi):
    """Separable 2-D taper from the zeroth prolate spheroidal angular function of bandwidth c, normalised to 1
    at the subgrid centre, evaluated at nu = |x + 0.5 - S/2| / (S/2)."""
    from scipy.special import pro_ang1

    nu = np.abs(np.arange(S) + 0.5 - S / 2.0) / (S / 2.0)
    val = pro_ang1(0, 0, c, np.minimum(nu, 1.0 - 1e-9))[0]
    val = val / pro_ang1(0, 0, c, 0.0)[0]
    val = np.clip(val, 0.0, None)
    return np.outer(val, val).astype(np.float32)


def identity_aterms(n_slots, n_stations, S):
    a = np.zeros((n_slots, n_stations, S, S, 4), np.complex64)
    a[..., 0] = 1
    a[..., 3] = 1
    return a


def plan_subgrids(uvw_m, frequencies, s1, s2, *, subgrid_size, grid_size, image_size, timesteps_per_subgrid,
                  support=8, aterm_interval=None):
    """IDG plan: returns (metadata, channels) for rows laid out as row = baseline * n_times + t."""
    n_bl, n_times, _ = uvw_m.shape
    S, G = subgrid_size, grid_size
    width = S - support
    scale = frequencies.astype(np.float64) / SPEED_OF_LIGHT * image_size  # cells per metre, per channel
    # Time chunks (cut at A-term slot boundaries too).
    bounds = list(range(0, n_times, timesteps_per_subgrid))
    if aterm_interval:
        bounds = sorted(set(bounds) | set(range(0, n_times, aterm_interval)))
    bounds.append(n_times)
    meta_parts, chan_parts = [], []
    for t0, t1 in zip(bounds[:-1], bounds[1:]):
        seg = uvw_m[:, t0:t1, :]
        umin, umax = seg[..., 0].min(1), seg[..., 0].max(1)
        vmin, vmax = seg[..., 1].min(1), seg[..., 1].max(1)
        start = np.zeros(n_bl, np.int64)
        all_bl = np.arange(n_bl)

        def footprint(b, a, z):
            # grid-cell extent of the samples of baselines b with channels in [a, z]
            fa, fz = scale[a], scale[z]
            lo_u = np.minimum(umin[b] * fa, umin[b] * fz)
            hi_u = np.maximum(umax[b] * fa, umax[b] * fz)
            lo_v = np.minimum(vmin[b] * fa, vmin[b] * fz)
            hi_v = np.maximum(vmax[b] * fa, vmax[b] * fz)
            return lo_u, hi_u, lo_v, hi_v

        groups = []  # (baseline idx array, c_begin, c_end_inclusive)
        C = len(frequencies)
        for c in range(1, C + 1):
            if c < C:
                lo_u, hi_u, lo_v, hi_v = footprint(all_bl, start, np.full(n_bl, c))
                fits = (hi_u - lo_u <= width) & (hi_v - lo_v <= width)
                close = ~fits
            else:
                close = np.ones(n_bl, bool)
            idx = np.nonzero(close)[0]
            if idx.size:
                groups.append((idx, start[idx].copy(), np.full(idx.size, c - 1)))
                start[idx] = c
        bl = np.concatenate([g[0] for g in groups])
        ca = np.concatenate([g[1] for g in groups])
        cb = np.concatenate([g[2] for g in groups])
        lo_u, hi_u, lo_v, hi_v = footprint(bl, ca, cb)
        x0 = np.floor((lo_u + hi_u) / 2 + G / 2 - S / 2 + 0.5).astype(np.int64)
        y0 = np.floor((lo_v + hi_v) / 2 + G / 2 - S / 2 + 0.5).astype(np.int64)
        keep = (x0 >= 0) & (x0 + S <= G) & (y0 >= 0) & (y0 + S <= G)
        bl, ca, cb, x0, y0 = bl[keep], ca[keep], cb[keep], x0[keep], y0[keep]
        md = np.zeros(bl.size, METADATA_DTYPE)
        md["baseline_offset"] = 0
        md["time_offset"] = bl * n_times + t0
        md["nr_timesteps"] = t1 - t0
        md["aterm_index"] = (t0 // aterm_interval) if aterm_interval else 0
        md["station1"] = s1[bl]
        md["station2"] = s2[bl]
        md["x"] = x0
        md["y"] = y0
        md["z"] = 0
        ch = np.zeros(bl.size, CHANNELS_DTYPE)
        ch["channel_begin"] = ca
        ch["nr_channels"] = cb - ca + 1
        meta_parts.append(md)
        chan_parts.append(ch)
    md = np.concatenate(meta_parts)
    ch = np.concatenate(chan_parts)
    order = np.lexsort((ch["channel_begin"], md["time_offset"]))
    return np.ascontiguousarray(md[order]), np.ascontiguousarray(ch[order])


# name -> parameters of the seeded synthetic configurations
CONFIGS = {
    # Small plain FP32 round trip: 10 stations in a 1 km disk, 128 x 1 s, 16 ch 150-165 MHz, w = 0.
    "small": dict(n_stations=10, radius_m=1000.0, core_fraction=0.0, core_radius_m=0.0, n_times=128,
                  f_lo=150e6, f_hi=165e6, n_channels=16, subgrid_size=32, timesteps_per_subgrid=128,
                  grid_size=1024, image_size=0.05, dec_deg=-30.0, zero_w=True, seed=1),
    # SKA-Low-like: 128 stations, half in a 1 km core, the rest in a 20 km disk; 256 x 1 s; 64 ch 150-180 MHz;
    # full w-term (|w| <~ 2000 wavelengths at this site and declination).
    "ska_low": dict(n_stations=128, radius_m=20000.0, core_fraction=0.5, core_radius_m=1000.0, n_times=256,
                    f_lo=150e6, f_hi=180e6, n_channels=64, subgrid_size=32, timesteps_per_subgrid=128,
                    grid_size=4096, image_size=0.1, dec_deg=-30.0, zero_w=False, seed=2),
}


def make_observation(name=None, *, visibilities=True, **overrides):
    """Geometry, plan and (optionally) host visibilities of a named configuration.

    Returns the dict the tests and bench.py feed to the C ABI (same keys as workload.reference_workload)."""
    cfg = dict(CONFIGS[name]) if name else {}
    cfg.update(overrides)
    rng = np.random.default_rng(cfg["seed"])
    enu = station_layout(rng, cfg["n_stations"], cfg["radius_m"], cfg["core_fraction"], cfg["core_radius_m"])
    s1, s2 = baseline_pairs(cfg["n_stations"])
    uvw = earth_rotation_uvw(enu, s1, s2, cfg["n_times"], cfg["dec_deg"])
    if cfg.get("zero_w"):
        uvw[..., 2] = 0.0
    uvw32 = uvw.astype(np.float32)
    freqs = np.linspace(cfg["f_lo"], cfg["f_hi"], cfg["n_channels"])
    wavenumbers = (2 * np.pi * freqs / SPEED_OF_LIGHT).astype(np.float32)
    S = cfg["subgrid_size"]
    md, ch = plan_subgrids(uvw32.astype(np.float64), freqs, s1, s2, subgrid_size=S, grid_size=cfg["grid_size"],
                           image_size=cfg["image_size"], timesteps_per_subgrid=cfg["timesteps_per_subgrid"],
                           aterm_interval=cfg.get("aterm_interval"))
    n_rows = len(s1) * cfg["n_times"]
    w = dict(
        config=cfg,
        uvw=np.ascontiguousarray(uvw32.reshape(n_rows, 3)),
        frequencies=freqs.astype(np.float32),
        wavenumbers=wavenumbers,
        spheroidal=prolate_spheroidal_taper(S),
        aterms=identity_aterms(1, cfg["n_stations"], S),
        metadata=md,
        channels=ch,
        nr_subgrids=md.size,
        subgrid_size=S,
        grid_size=cfg["grid_size"],
        nr_stations=cfg["n_stations"],
        nr_aterm_slots=1,
        image_size=np.float32(cfg["image_size"]),
        w_step=0.0,
        max_nr_timesteps=int(md["nr_timesteps"].max()) if md.size else 0,
        nr_rows=n_rows,
        nr_visibilities=int((md["nr_timesteps"].astype(np.int64) * ch["nr_channels"]).sum()),
    )
    if visibilities:
        w["visibilities"] = random_visibilities(n_rows, cfg["n_channels"], cfg["seed"])
    return w


def random_visibilities(n_rows, n_channels, seed):
    """complex64 [n_rows][C][4], real and imaginary parts ~ N(0, 1)."""
    rng = np.random.default_rng(seed + 1000)
    v = rng.standard_normal((n_rows, n_channels, 4, 2), dtype=np.float32)
    return np.ascontiguousarray(v).view(np.complex64).reshape(n_rows, n_channels, 4)


def subset(w, subgrid_indices):
    """A self-contained workload holding only the given subgrids: rows are gathered and time offsets rewritten,
    so the oracle and the device see identical, compact inputs."""
    idx = np.asarray(subgrid_indices)
    md = w["metadata"][idx].copy()
    ch = None if w["channels"] is None else w["channels"][idx].copy()
    rows = []
    new_md = md.copy()
    cursor = 0
    for k in range(md.size):
        r0 = int(md["time_offset"][k]) + int(md["baseline_offset"][k]) - int(w["metadata"]["baseline_offset"][0])
        nt = int(md["nr_timesteps"][k])
        rows.append(np.arange(r0, r0 + nt))
        new_md["baseline_offset"][k] = 0
        new_md["time_offset"][k] = cursor
        cursor += nt
    rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    out = dict(w)
    out.update(metadata=np.ascontiguousarray(new_md), channels=None if ch is None else np.ascontiguousarray(ch),
               uvw=np.ascontiguousarray(w["uvw"][rows]), nr_subgrids=md.size, nr_rows=rows.size,
               max_nr_timesteps=int(md["nr_timesteps"].max()) if md.size else 0, source_rows=rows)
    if "visibilities" in w and w["visibilities"] is not None:
        out["visibilities"] = np.ascontiguousarray(w["visibilities"][rows])
    if ch is not None:
        out["nr_visibilities"] = int((new_md["nr_timesteps"].astype(np.int64) * ch["nr_channels"]).sum())
    return out
