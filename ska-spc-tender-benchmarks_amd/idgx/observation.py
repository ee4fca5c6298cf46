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
import time

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


def jones_aterms(rng, n_slots, n_stations, S, sigma=0.1):
    """Direction-dependent Jones matrices [slot][station][y][x][xx, xy, yx, yy]: J = I + E with E a smooth
    (quadratic in the pixel coordinates) complex perturbation whose entries have standard deviation ~sigma, each
    pixel's matrix renormalised to |det J| = 1."""
    t = (np.arange(S) + 0.5) / (S / 2.0) - 1.0  # pixel centres in [-1, 1]
    yy, xx = np.meshgrid(t, t, indexing="ij")
    basis = np.stack([np.ones_like(xx), xx, yy, xx * yy, xx * xx - 1 / 3, yy * yy - 1 / 3])  # [6][S][S]
    nb = basis.shape[0]
    coef = (rng.standard_normal((n_slots, n_stations, 4, nb, 2)) @ np.array([1.0, 1j])) * (sigma / np.sqrt(nb))
    e = np.einsum("abkn,nyx->abyxk", coef, basis)
    j = e.copy()
    j[..., 0] += 1.0
    j[..., 3] += 1.0
    det = j[..., 0] * j[..., 3] - j[..., 1] * j[..., 2]
    j /= np.sqrt(np.abs(det))[..., None]
    return np.ascontiguousarray(j.astype(np.complex64))


def time_segments(n_baselines, n_times, timesteps_per_subgrid, aterm_interval=None):
    """Fixed time chunks for every baseline, cut at A-term slot boundaries too: (baseline, t0, t1) arrays."""
    bounds = set(range(0, n_times, timesteps_per_subgrid))
    if aterm_interval:
        bounds |= set(range(0, n_times, aterm_interval))
    bounds = sorted(bounds) + [n_times]
    t0 = np.repeat(np.asarray(bounds[:-1], np.int64)[:, None], n_baselines, axis=1)
    t1 = np.repeat(np.asarray(bounds[1:], np.int64)[:, None], n_baselines, axis=1)
    bl = np.repeat(np.arange(n_baselines, dtype=np.int64)[None, :], len(bounds) - 1, axis=0)
    return bl.ravel(), t0.ravel(), t1.ravel()


def ragged_time_segments(rng, baselines, n_times, t_min, t_max, aterm_interval=None):
    """Per-baseline time chunks with lengths log-uniform in [t_min, t_max] (the last one truncated at n_times), cut
    at A-term slot boundaries too: (baseline, t0, t1) arrays, baseline-major."""
    baselines = np.asarray(baselines, np.int64)
    k = -(-n_times // t_min) + 1
    lengths = np.rint(np.exp(rng.uniform(np.log(t_min), np.log(t_max), (baselines.size, k)))).astype(np.int64)
    ends = np.minimum(np.cumsum(lengths, axis=1), n_times)
    starts = np.concatenate([np.zeros((baselines.size, 1), np.int64), ends[:, :-1]], axis=1)
    valid = ends > starts
    bl = np.repeat(baselines[:, None], k, axis=1)[valid]
    t0, t1 = starts[valid], ends[valid]
    if aterm_interval:
        cuts = np.arange(aterm_interval, n_times, aterm_interval)
        pieces = []
        for b_, a_, z_ in zip(bl, t0, t1):
            edges = [a_] + [int(c) for c in cuts if a_ < c < z_] + [z_]
            pieces += [(b_, e0, e1) for e0, e1 in zip(edges[:-1], edges[1:])]
        bl, t0, t1 = (np.asarray(x, np.int64) for x in zip(*pieces))
    return bl, t0, t1


def plan_segments(uvw_m, frequencies, s1, s2, seg_bl, seg_t0, seg_t1, *, subgrid_size, grid_size, image_size,
                  support=8, aterm_interval=None, w_step=None):
    """IDG plan over time segments: segment i covers rows seg_bl[i] * n_times + [seg_t0[i], seg_t1[i]) (rows laid
    out as row = baseline * n_times + t); a segment must not cross an A-term slot boundary.  Its channels are
    grouped greedily so that the uv footprint of every (timestep, channel) sample of a group spans at most
    S - support grid cells in u and in v; each group becomes one subgrid centred on its footprint, dropped if it
    would leave the grid.  With w_step (wavelengths), the subgrid's w-layer z = floor(w / w_step) puts the middle of
    its w range (over its timesteps, at its centre frequency) within w_step / 2 of its phase reference
    (z + 0.5) w_step.
    Returns (metadata, channels) sorted by (first row, first channel)."""
    n_bl, n_times, _ = uvw_m.shape
    S, G = subgrid_size, grid_size
    width = S - support
    scale = frequencies.astype(np.float64) / SPEED_OF_LIGHT * image_size  # cells per metre, per channel
    seg_bl, seg_t0, seg_t1 = (np.asarray(x, np.int64) for x in (seg_bl, seg_t0, seg_t1))
    n_seg = seg_bl.size
    # Per-segment extent of u and v (metres): reductions over [start, end) of the flattened rows.
    start = seg_bl * n_times + seg_t0
    end = seg_bl * n_times + seg_t1
    bounds = np.stack([start, end], axis=1).ravel()

    def seg_reduce(op, comp):
        a = np.append(uvw_m[..., comp].reshape(-1), 0.0)
        return op.reduceat(a, bounds)[::2] if n_seg else np.zeros(0)

    umin, umax = seg_reduce(np.minimum, 0), seg_reduce(np.maximum, 0)
    vmin, vmax = seg_reduce(np.minimum, 1), seg_reduce(np.maximum, 1)
    wmin, wmax = (seg_reduce(np.minimum, 2), seg_reduce(np.maximum, 2)) if w_step else (None, None)

    def footprint(i, a, z):
        # grid-cell extent of the samples of segments i with channels in [a, z]
        fa, fz = scale[a], scale[z]
        lo_u = np.minimum(umin[i] * fa, umin[i] * fz)
        hi_u = np.maximum(umax[i] * fa, umax[i] * fz)
        lo_v = np.minimum(vmin[i] * fa, vmin[i] * fz)
        hi_v = np.maximum(vmax[i] * fa, vmax[i] * fz)
        return lo_u, hi_u, lo_v, hi_v

    groups = []  # (segment idx array, c_begin, c_end_inclusive)
    first = np.zeros(n_seg, np.int64)
    every = np.arange(n_seg)
    C = len(frequencies)
    for c in range(1, C + 1):
        if c < C:
            lo_u, hi_u, lo_v, hi_v = footprint(every, first, np.full(n_seg, c))
            close = ~((hi_u - lo_u <= width) & (hi_v - lo_v <= width))
        else:
            close = np.ones(n_seg, bool)
        idx = np.nonzero(close)[0]
        if idx.size:
            groups.append((idx, first[idx].copy(), np.full(idx.size, c - 1)))
            first[idx] = c
    seg = np.concatenate([g[0] for g in groups]) if groups else np.zeros(0, np.int64)
    ca = np.concatenate([g[1] for g in groups]) if groups else np.zeros(0, np.int64)
    cb = np.concatenate([g[2] for g in groups]) if groups else np.zeros(0, np.int64)
    lo_u, hi_u, lo_v, hi_v = footprint(seg, ca, cb)
    x0 = np.floor((lo_u + hi_u) / 2 + G / 2 - S / 2 + 0.5).astype(np.int64)
    y0 = np.floor((lo_v + hi_v) / 2 + G / 2 - S / 2 + 0.5).astype(np.int64)
    keep = (x0 >= 0) & (x0 + S <= G) & (y0 >= 0) & (y0 + S <= G)
    seg, ca, cb, x0, y0 = seg[keep], ca[keep], cb[keep], x0[keep], y0[keep]
    bl, t0, t1 = seg_bl[seg], seg_t0[seg], seg_t1[seg]
    md = np.zeros(seg.size, METADATA_DTYPE)
    md["baseline_offset"] = 0
    md["time_offset"] = bl * n_times + t0
    md["nr_timesteps"] = t1 - t0
    md["aterm_index"] = (t0 // aterm_interval) if aterm_interval else 0
    md["station1"] = s1[bl]
    md["station2"] = s2[bl]
    md["x"] = x0
    md["y"] = y0
    md["z"] = 0
    if w_step:
        f_c = 0.5 * (frequencies[ca] + frequencies[cb])
        w_mid = 0.5 * (wmin[seg] + wmax[seg])
        md["z"] = np.floor(w_mid * f_c / SPEED_OF_LIGHT / w_step).astype(np.int64)
    ch = np.zeros(seg.size, CHANNELS_DTYPE)
    ch["channel_begin"] = ca
    ch["nr_channels"] = cb - ca + 1
    order = np.lexsort((ch["channel_begin"], md["time_offset"]))
    return np.ascontiguousarray(md[order]), np.ascontiguousarray(ch[order])


def plan_subgrids(uvw_m, frequencies, s1, s2, *, subgrid_size, grid_size, image_size, timesteps_per_subgrid,
                  support=8, aterm_interval=None, w_step=None):
    """IDG plan with the same time chunks for every baseline: returns (metadata, channels) for rows laid out as
    row = baseline * n_times + t."""
    bl, t0, t1 = time_segments(uvw_m.shape[0], uvw_m.shape[1], timesteps_per_subgrid, aterm_interval)
    return plan_segments(uvw_m, frequencies, s1, s2, bl, t0, t1, subgrid_size=subgrid_size, grid_size=grid_size,
                         image_size=image_size, support=support, aterm_interval=aterm_interval, w_step=w_step)


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
    # Full imaging cycle: the 256-station layout of the batched configuration (same layout parameters and seed: half
    # the stations in a 1 km core, the rest in a 20 km disk) with the small configuration's sampling (128 x 1 s, 16
    # channels 150-165 MHz), pointed at declination -41 deg so that |w| reaches 5267 wavelengths (the 10-station
    # 1 km layout cannot exceed ~1100); S=64 on a 16384^2 grid, image 0.1 rad, Jones A-terms per station changing
    # every 64 timesteps (the planner cuts subgrids there), w-layers 2000 wavelengths apart (6 layers).
    "full_cycle": dict(n_stations=256, radius_m=20000.0, core_fraction=0.5, core_radius_m=1000.0, n_times=128,
                       f_lo=150e6, f_hi=165e6, n_channels=16, subgrid_size=64, timesteps_per_subgrid=64,
                       grid_size=16384, image_size=0.1, dec_deg=-41.0, zero_w=False, aterm_interval=64,
                       jones=True, w_step=2000.0, seed=4),
    # W-stacking: the small layout pointed 27 degrees from zenith (|w| up to ~500 wavelengths), 32-timestep
    # subgrids, w-layers 100 wavelengths apart.
    "wstack": dict(n_stations=10, radius_m=1000.0, core_fraction=0.0, core_radius_m=0.0, n_times=128,
                   f_lo=150e6, f_hi=165e6, n_channels=16, subgrid_size=32, timesteps_per_subgrid=32,
                   grid_size=2048, image_size=0.1, dec_deg=0.0, zero_w=False, w_step=100.0, seed=7),
}

# Batched uneven workloads (make_batched_observation): baselines split into length terciles with subgrid sizes
# 24 / 32 / 48, per-baseline time chunks log-uniform in t_range, per-baseline flag fraction U(0, flag_max).
BATCHED_CONFIGS = {
    # 256 stations (half in a 1 km core, the rest in a 20 km disk), 512 x 1 s, 128 ch 150-200 MHz, G = 8192.
    "batched_uneven": dict(n_stations=256, radius_m=20000.0, core_fraction=0.5, core_radius_m=1000.0, n_times=512,
                           f_lo=150e6, f_hi=200e6, n_channels=128, subgrid_sizes=(24, 32, 48), t_range=(16, 256),
                           grid_size=8192, image_size=0.1, dec_deg=-30.0, zero_w=False, flag_max=0.6, seed=4),
    # The same construction at a size the oracle checks in seconds.
    "batched_uneven_small": dict(n_stations=12, radius_m=3000.0, core_fraction=0.5, core_radius_m=500.0,
                                 n_times=96, f_lo=150e6, f_hi=180e6, n_channels=16, subgrid_sizes=(24, 32, 48),
                                 t_range=(8, 64), grid_size=2048, image_size=0.1, dec_deg=-30.0, zero_w=False,
                                 flag_max=0.6, seed=5),
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
                           aterm_interval=cfg.get("aterm_interval"), w_step=cfg.get("w_step"))
    n_rows = len(s1) * cfg["n_times"]
    n_slots = -(-cfg["n_times"] // cfg["aterm_interval"]) if cfg.get("aterm_interval") else 1
    aterms = jones_aterms(np.random.default_rng(cfg["seed"] + 7), n_slots, cfg["n_stations"], S) \
        if cfg.get("jones") else identity_aterms(n_slots, cfg["n_stations"], S)
    w = dict(
        config=cfg,
        uvw=np.ascontiguousarray(uvw32.reshape(n_rows, 3)),
        frequencies=freqs.astype(np.float32),
        wavenumbers=wavenumbers,
        spheroidal=prolate_spheroidal_taper(S),
        aterms=aterms,
        metadata=md,
        channels=ch,
        nr_subgrids=md.size,
        subgrid_size=S,
        grid_size=cfg["grid_size"],
        nr_stations=cfg["n_stations"],
        nr_aterm_slots=n_slots,
        image_size=np.float32(cfg["image_size"]),
        w_step=float(cfg.get("w_step") or 0.0),
        max_nr_timesteps=int(md["nr_timesteps"].max()) if md.size else 0,
        nr_rows=n_rows,
        nr_visibilities=int((md["nr_timesteps"].astype(np.int64) * ch["nr_channels"]).sum()),
    )
    if visibilities:
        w["visibilities"] = random_visibilities(n_rows, cfg["n_channels"], cfg["seed"])
    return w


def baseline_flags(rng_fraction, rng_flags, n_baselines, n_times, n_channels, flag_max, rows_per_chunk=1 << 16):
    """Boolean flags [rows][C] (True = flagged): baseline b flags each (timestep, channel) sample with
    probability f_b ~ U(0, flag_max).  Generated in row chunks (float32 uniforms) to bound the memory."""
    frac = rng_fraction.uniform(0.0, flag_max, n_baselines).astype(np.float32)
    n_rows = n_baselines * n_times
    flags = np.empty((n_rows, n_channels), bool)
    for r0 in range(0, n_rows, rows_per_chunk):
        r1 = min(n_rows, r0 + rows_per_chunk)
        f = frac[np.arange(r0, r1) // n_times][:, None]
        flags[r0:r1] = rng_flags.random((r1 - r0, n_channels), dtype=np.float32) < f
    return flags


def make_batched_observation(name=None, *, visibilities=True, flags=True, **overrides):
    """A batched uneven workload: one subgrid class per baseline-length tercile (subgrid sizes 24 / 32 / 48 by
    default), ragged per-baseline time chunks, per-baseline flagging expressed as visibility selection lists.

    Returns dict(config, classes=[workload dicts sharing uvw / wavenumbers / visibilities rows, each with its
    'vis_index' / 'vis_offset'], flags, nr_rows, nr_visibilities = selected visibilities)."""
    from .device import selection_lists

    cfg = dict(BATCHED_CONFIGS[name]) if name else {}
    cfg.update(overrides)
    rng = np.random.default_rng(cfg["seed"])
    enu = station_layout(rng, cfg["n_stations"], cfg["radius_m"], cfg["core_fraction"], cfg["core_radius_m"])
    s1, s2 = baseline_pairs(cfg["n_stations"])
    n_bl, n_times, C = len(s1), cfg["n_times"], cfg["n_channels"]
    uvw = earth_rotation_uvw(enu, s1, s2, n_times, cfg["dec_deg"])
    if cfg.get("zero_w"):
        uvw[..., 2] = 0.0
    uvw32 = uvw.astype(np.float32)
    del uvw
    uvw64 = uvw32.astype(np.float64)
    freqs = np.linspace(cfg["f_lo"], cfg["f_hi"], C)
    wavenumbers = (2 * np.pi * freqs / SPEED_OF_LIGHT).astype(np.float32)
    n_rows = n_bl * n_times
    shared = dict(uvw=np.ascontiguousarray(uvw32.reshape(n_rows, 3)), frequencies=freqs.astype(np.float32),
                  wavenumbers=wavenumbers, nr_rows=n_rows, grid_size=cfg["grid_size"],
                  nr_stations=cfg["n_stations"], image_size=np.float32(cfg["image_size"]), w_step=0.0)
    if visibilities:
        shared["visibilities"] = random_visibilities(n_rows, C, cfg["seed"])
    fl = baseline_flags(np.random.default_rng(cfg["seed"] + 2), np.random.default_rng(cfg["seed"] + 3), n_bl, n_times,
                        C, cfg["flag_max"]) if flags else None
    length = np.linalg.norm(enu[s2] - enu[s1], axis=1)
    terciles = np.array_split(np.argsort(length, kind="stable"), len(cfg["subgrid_sizes"]))
    seg_rng = np.random.default_rng(cfg["seed"] + 1)
    classes = []
    for k, S in enumerate(cfg["subgrid_sizes"]):
        bl, t0, t1 = ragged_time_segments(seg_rng, np.sort(terciles[k]), n_times, *cfg["t_range"])
        t_plan = time.perf_counter()
        md, ch = plan_segments(uvw64, freqs, s1, s2, bl, t0, t1, subgrid_size=S, grid_size=cfg["grid_size"],
                               image_size=cfg["image_size"])
        t_plan = time.perf_counter() - t_plan
        c = dict(shared, config=cfg, metadata=md, channels=ch, nr_subgrids=md.size, subgrid_size=S,
                 segments=(bl, t0, t1), host_plan_s=t_plan, frequencies64=freqs, stations=(s1, s2),
                 spheroidal=prolate_spheroidal_taper(S), aterms=identity_aterms(1, cfg["n_stations"], S),
                 nr_aterm_slots=1, max_nr_timesteps=int(md["nr_timesteps"].max()) if md.size else 0)
        if fl is not None:
            c["vis_index"], c["vis_offset"] = selection_lists(c, fl)
            c["nr_visibilities"] = int(c["vis_offset"][-1])
        else:
            c["nr_visibilities"] = int((md["nr_timesteps"].astype(np.int64) * ch["nr_channels"]).sum())
        classes.append(c)
    return dict(shared, config=cfg, classes=classes, flags=fl,
                nr_visibilities=sum(c["nr_visibilities"] for c in classes))


def subset_batched(bo, indices, *, seed=0):
    """Subgrids indices[k] of every class k of a batched observation (built with visibilities=False, flags=False) as
    a compact batched observation: the rows they read gathered into one row space (class by class), seeded N(0, 1)
    visibilities for those rows, flags drawn for those rows only (per-baseline fraction f_b ~ U(0, flag_max) from
    the configuration's seed, as baseline_flags draws them, per-sample uniforms from `seed`) and each class's
    selection lists.  The plan records, channel windows and uvw rows are the full plan's, so the subset exercises
    the full workload's geometry (subgrid sizes, windows, ragged timesteps, long baselines) at a size the oracle
    checks in seconds.  'source_rows' maps the rows back to the full observation."""
    from .device import selection_lists

    cfg = bo["config"]
    n_times, C = cfg["n_times"], bo["wavenumbers"].shape[0]
    parts, cursor = [], 0
    for k, c in enumerate(bo["classes"]):
        sw = subset(dict(c, visibilities=None), indices[k])
        sw["metadata"]["time_offset"] += cursor
        cursor += sw["nr_rows"]
        parts.append(sw)
    rows = np.concatenate([sw["source_rows"] for sw in parts]) if parts else np.zeros(0, np.int64)
    n_rows = rows.size
    frac = np.random.default_rng(cfg["seed"] + 2).uniform(0.0, cfg["flag_max"], len(bo["classes"][0]["stations"][0]))
    frac = frac.astype(np.float32)
    rng = np.random.default_rng(seed + 17)
    flags = rng.random((n_rows, C), dtype=np.float32) < frac[rows // n_times][:, None]
    shared = dict(uvw=np.ascontiguousarray(bo["uvw"][rows]), frequencies=bo["frequencies"],
                  wavenumbers=bo["wavenumbers"], nr_rows=n_rows, grid_size=bo["grid_size"],
                  nr_stations=bo["nr_stations"], image_size=bo["image_size"], w_step=bo["w_step"],
                  visibilities=random_visibilities(n_rows, C, cfg["seed"] + seed))
    classes = []
    for sw in parts:
        c = dict(sw, **shared)
        c["max_nr_timesteps"] = int(c["metadata"]["nr_timesteps"].max()) if c["metadata"].size else 0
        c["vis_index"], c["vis_offset"] = selection_lists(c, flags)
        c["nr_visibilities"] = int(c["vis_offset"][-1])
        classes.append(c)
    return dict(shared, config=cfg, classes=classes, flags=flags, source_rows=rows,
                nr_visibilities=sum(c["nr_visibilities"] for c in classes))


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
