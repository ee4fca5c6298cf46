"""Workload generators and the planner (host only): time segments, the batched uneven configuration, selection
lists, Jones A-terms and the A-term-aware full-cycle plan."""
import numpy as np

from idgx import observation
from idgx.device import selection_lists


def _selection_by_loop(w, flags):
    """Straightforward per-subgrid restatement of the selection lists (np.nonzero over each window)."""
    md, ch = w["metadata"], w["channels"]
    index, offset = [np.zeros(0, np.uint32)], [0]
    for s in range(md.size):
        r0, T = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        c0, nc = int(ch["channel_begin"][s]), int(ch["nr_channels"][s])
        t, c = np.nonzero(~flags[r0:r0 + T, c0:c0 + nc])
        index.append((t.astype(np.uint32) << 16) | c.astype(np.uint32))
        offset.append(offset[-1] + t.size)
    return np.concatenate(index), np.asarray(offset, np.int64)


def test_fixed_time_segments_tile_every_baseline_and_respect_aterm_slots():
    bl, t0, t1 = observation.time_segments(5, 300, 128, aterm_interval=64)
    for b in range(5):
        sel = bl == b
        edges = np.concatenate([t0[sel], t1[sel][-1:]])
        assert list(edges) == [0, 64, 128, 192, 256, 300]
    assert ((t0 // 64) == ((t1 - 1) // 64)).all()


def test_ragged_segments_are_log_uniform_chunks_that_tile_the_observation():
    rng = np.random.default_rng(3)
    bl, t0, t1 = observation.ragged_time_segments(rng, np.arange(400), 512, 16, 256, aterm_interval=100)
    L = t1 - t0
    assert (L >= 1).all() and (L <= 256).all()
    for b in (0, 17, 399):
        sel = bl == b
        assert t0[sel][0] == 0 and t1[sel][-1] == 512 and (t0[sel][1:] == t1[sel][:-1]).all()
    assert ((t0 // 100) == ((t1 - 1) // 100)).all()
    # without cuts, the untruncated chunk lengths are log-uniform in [16, 256]: log2 is uniform in [4, 8]
    bl, t0, t1 = observation.ragged_time_segments(np.random.default_rng(5), np.arange(200), 8192, 16, 256)
    lg = np.log2((t1 - t0)[t1 < 8192])
    assert lg.min() >= 4 - 0.05 and lg.max() <= 8 + 0.01
    assert abs(np.median(lg) - 6) < 0.1 and abs(np.mean(lg < 5) - 0.25) < 0.03


def test_selection_lists_match_the_per_subgrid_restatement():
    w = observation.make_observation("small", visibilities=False)
    flags = np.random.default_rng(4).random((w["nr_rows"], w["wavenumbers"].shape[0])) < 0.35
    for chunk in (1 << 26, 3000):
        vi, vo = selection_lists(w, flags, chunk_entries=chunk)
        ri, ro = _selection_by_loop(w, flags)
        assert np.array_equal(vi, ri) and np.array_equal(vo, ro)


def test_batched_uneven_small_structure():
    bo = observation.make_batched_observation("batched_uneven_small")
    cfg = bo["config"]
    C = bo["wavenumbers"].shape[0]
    assert [c["subgrid_size"] for c in bo["classes"]] == [24, 32, 48]
    # flags: per-baseline fractions in [0, 0.6): the overall fraction is near 0.3
    assert 0.2 < bo["flags"].mean() < 0.4
    cover = np.zeros((bo["nr_rows"], C), np.int8)
    stations = set()
    for c in bo["classes"]:
        md, ch, S = c["metadata"], c["channels"], c["subgrid_size"]
        assert (md["x"] >= 0).all() and (md["x"] + S <= cfg["grid_size"]).all()
        assert (md["nr_timesteps"] <= cfg["t_range"][1]).all()
        for s in range(md.size):
            r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
            c0, nc = int(ch["channel_begin"][s]), int(ch["nr_channels"][s])
            cover[r0:r0 + nt, c0:c0 + nc] += 1
        vi, vo = _selection_by_loop(c, bo["flags"])
        assert np.array_equal(c["vis_index"], vi) and np.array_equal(c["vis_offset"], vo)
        stations |= {(int(a), int(b)) for a, b in zip(md["station1"], md["station2"])}
    assert cover.max() == 1  # classes and subgrids never overlap
    assert cover.mean() > 0.99
    assert len(stations) == bo["classes"][0]["nr_stations"] * (bo["classes"][0]["nr_stations"] - 1) // 2
    assert bo["nr_visibilities"] == int((cover.astype(bool) & ~bo["flags"]).sum())


def test_jones_aterms_are_smooth_unimodular_and_not_identity():
    a = observation.jones_aterms(np.random.default_rng(1), 2, 5, 32)
    J = a.reshape(2, 5, 32, 32, 2, 2).astype(np.complex128)
    assert np.allclose(np.abs(np.linalg.det(J)), 1.0, atol=1e-5)
    off = np.abs(J - np.eye(2)).mean()
    assert 0.02 < off < 0.3
    # smooth: neighbouring pixels differ far less than the perturbation itself
    assert np.abs(np.diff(J, axis=3)).max() < 0.2 * np.abs(J - np.eye(2)).max()


def test_full_cycle_plan_cuts_at_aterm_boundaries():
    w = observation.make_observation("full_cycle", visibilities=False)
    md = w["metadata"]
    t_local = md["time_offset"] % w["config"]["n_times"]
    assert w["nr_aterm_slots"] == 2 and w["aterms"].shape == (2, 256, 64, 64, 4)
    assert (md["aterm_index"] == t_local // 64).all() and (md["nr_timesteps"] == 64).all()
    assert (md["x"] >= 0).all() and (md["x"] + 64 <= 16384).all()


def test_full_cycle_reaches_its_w_range_with_w_layers():
    """The full-cycle configuration's |w| reaches +-5000 wavelengths (at the top channel), every subgrid's w-layer
    puts the middle of its w range within w_step / 2 of its phase reference (z + 0.5) w_step, and the layers span
    the w range."""
    w = observation.make_observation("full_cycle", visibilities=False)
    md, ch = w["metadata"], w["channels"]
    f = w["frequencies"].astype(np.float64)
    w_lambda = np.abs(w["uvw"][:, 2].astype(np.float64)) * f.max() / observation.SPEED_OF_LIGHT
    assert w_lambda.max() >= 5000.0
    step = w["w_step"]
    n_times = w["config"]["n_times"]
    uvw = w["uvw"].reshape(-1, n_times, 3).astype(np.float64)
    bl, t0 = md["time_offset"] // n_times, md["time_offset"] % n_times
    for s in np.linspace(0, md.size - 1, 500).astype(int):
        ws = uvw[bl[s], t0[s]:t0[s] + md["nr_timesteps"][s], 2]
        fc = 0.5 * (f[ch["channel_begin"][s]] + f[ch["channel_begin"][s] + ch["nr_channels"][s] - 1])
        mid = 0.5 * (ws.min() + ws.max()) * fc / observation.SPEED_OF_LIGHT
        assert abs(mid - (md["z"][s] + 0.5) * step) <= 0.5 * step + 1e-6
    assert md["z"].min() * step <= -5000.0 + step and (md["z"].max() + 1) * step >= 5000.0 - step


def test_earth_rotation_uvw_invariants():
    """uvw of earth-rotation synthesis: a rotation of the baseline, so |uvw| = |baseline| at every hour angle; at
    transit u = east, v = north cos(dec - lat) - up sin(dec - lat), w = north sin(dec - lat) + up cos(dec - lat)
    (the source's direction cosines in the local frame); a baseline's (u, v) for a source at the pole turns by the
    earth's rotation per second."""
    rng = np.random.default_rng(11)
    enu = rng.normal(0.0, 3000.0, (9, 3))
    s1, s2 = observation.baseline_pairs(9)
    b = enu[s2] - enu[s1]
    lat = np.radians(observation.SITE_LATITUDE_DEG)
    for dec_deg in (-30.0, -41.0, 10.0, -89.0):
        uvw = observation.earth_rotation_uvw(enu, s1, s2, 101, dec_deg)
        assert np.allclose(np.linalg.norm(uvw, axis=-1), np.linalg.norm(b, axis=-1)[:, None], rtol=1e-12)
        mid = uvw[:, 50]  # hour angle 0: the middle of 101 steps centred on transit
        d = np.radians(dec_deg) - lat
        assert np.allclose(mid[:, 0], b[:, 0], atol=1e-9)
        assert np.allclose(mid[:, 1], b[:, 1] * np.cos(d) - b[:, 2] * np.sin(d), atol=1e-9)
        assert np.allclose(mid[:, 2], b[:, 1] * np.sin(d) + b[:, 2] * np.cos(d), atol=1e-9)
    pole = observation.earth_rotation_uvw(enu, s1, s2, 2, -90.0)
    ang = np.arctan2(pole[:, 1, 1], pole[:, 1, 0]) - np.arctan2(pole[:, 0, 1], pole[:, 0, 0])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    assert np.allclose(np.abs(ang), observation.EARTH_ROTATION_RAD_S, rtol=1e-9)
    assert np.allclose(pole[:, 0, 2], pole[:, 1, 2])  # w is constant for a source at the pole


def test_prolate_spheroidal_taper_invariants():
    """The taper is the outer product of a 1-D profile that is symmetric about the subgrid centre, largest at the
    two central pixels (1 at nu = 0 by normalisation), decreasing towards the edge and non-negative; and that
    profile is the zeroth prolate spheroidal wave function of bandwidth c = 3 pi: an eigenfunction of the finite
    Fourier transform on [-1, 1], int exp(i c x t) psi(t) dt = lambda psi(x) (checked by Gauss-Legendre
    quadrature on the profile's own formula)."""
    from scipy.special import pro_ang1

    for S in (16, 24, 32, 48, 64):
        t = observation.prolate_spheroidal_taper(S).astype(np.float64)
        prof = t[S // 2]
        assert np.allclose(t, np.outer(prof, prof) / prof[S // 2], rtol=1e-6)
        assert np.allclose(prof, prof[::-1], rtol=1e-6)
        half = prof[S // 2:]
        assert (np.diff(half) <= 1e-7).all() and (prof >= 0).all()
        assert 0.95 < half[0] <= 1.0 and half[-1] < 0.05
    c = 3.0 * np.pi
    xq, wq = np.polynomial.legendre.leggauss(200)

    def psi(x):
        return pro_ang1(0, 0, c, np.clip(x, -1 + 1e-12, 1 - 1e-12))[0] / pro_ang1(0, 0, c, 0.0)[0]

    x = np.linspace(-0.95, 0.95, 39)
    ft = (np.exp(1j * c * np.outer(x, xq)) * (wq * psi(xq))).sum(axis=1)
    lam = ft[19] / psi(x[19])
    assert np.allclose(ft, lam * psi(x), atol=1e-8 * abs(lam))
    assert np.isclose(lam.imag, 0.0, atol=1e-10) and lam.real > 0
