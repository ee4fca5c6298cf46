"""Device residency for a workload: upload the host arrays once, then run the C-ABI gridder/degridder on them.

torch provides the HBM allocations and the stream (plumbing only); every result is computed by libidgx.
"""
import numpy as np

from . import METADATA_DTYPE, CHANNELS_DTYPE, ATERM_ADJOINT, PLAIN_FP32, make_plan, check_plan, grid, degrid, \
    IdgxError, OK, FFT_FORWARD, FFT_BACKWARD, subgrid_fft, adder_plan_build, add_subgrids, split_subgrids, \
    adder_plan_build_layered, split_subgrids_layered, grid_fft, wstack_to_image, wstack_from_image, grid_to_image, \
    image_to_grid


def _torch():
    import torch

    return torch


def upload(a, device="cuda"):
    """Copy a numpy array to a flat uint8 device tensor (keeps any dtype, including complex and structs)."""
    torch = _torch()
    flat = np.ascontiguousarray(a).reshape(-1).view(np.uint8)
    return torch.from_numpy(flat.copy()).to(device)


def download(t, dtype, shape):
    return t.cpu().numpy().view(dtype).reshape(shape)


class DeviceWorkload:
    """Inputs of one workload resident in HBM, plus output buffers for both directions."""

    def __init__(self, w, device="cuda", visibilities_on_device=None):
        torch = _torch()
        self.w = w
        self.device = device
        self.S = int(w["subgrid_size"])
        self.n = int(w["nr_subgrids"])
        self.C = int(w["wavenumbers"].shape[0])
        self.nr_rows = int(w["uvw"].shape[0])
        self.d_metadata = upload(w["metadata"], device)
        self.d_channels = None if w.get("channels") is None else upload(w["channels"], device)
        self.d_uvw = upload(w["uvw"], device)
        self.d_wavenumbers = upload(w["wavenumbers"], device)
        self.d_spheroidal = upload(w["spheroidal"], device)
        self.d_aterms = upload(w["aterms"], device)
        if visibilities_on_device is not None:
            self.d_visibilities = visibilities_on_device
        else:
            self.d_visibilities = upload(w["visibilities"], device)
        self.d_subgrids = torch.zeros(max(self.n, 1) * 4 * self.S * self.S * 8, dtype=torch.uint8, device=device)
        self.d_vis_out = torch.zeros(self.nr_rows * self.C * 4 * 8, dtype=torch.uint8, device=device)

    def plan(self, precision=PLAIN_FP32, aterm_convention=ATERM_ADJOINT, flags=0):
        w = self.w
        p = make_plan(nr_subgrids=self.n, subgrid_size=self.S, grid_size=w["grid_size"], nr_channels=self.C,
                      nr_stations=w["nr_stations"], nr_aterm_slots=w["nr_aterm_slots"], nr_rows=self.nr_rows,
                      max_nr_timesteps=w["max_nr_timesteps"], image_size=w["image_size"], w_step=w["w_step"],
                      precision=precision, aterm_convention=aterm_convention, flags=flags)
        rc = check_plan(p, w["metadata"], w.get("channels"))
        if rc != OK:
            raise IdgxError(rc, "idgx_check_plan")
        return p

    def grid(self, plan, stream=None, subgrids_in=None):
        """Run the gridder; the subgrids land in self.d_subgrids."""
        grid(plan, self.d_metadata, self.d_channels, self.d_uvw, self.d_wavenumbers, self.d_visibilities,
             self.d_spheroidal, self.d_aterms, self.d_subgrids, stream)

    def degrid(self, plan, d_subgrids=None, stream=None):
        """Run the degridder on d_subgrids (default: self.d_subgrids); visibilities land in self.d_vis_out."""
        degrid(plan, self.d_metadata, self.d_channels, self.d_uvw, self.d_wavenumbers,
               self.d_subgrids if d_subgrids is None else d_subgrids, self.d_spheroidal, self.d_aterms,
               self.d_vis_out, stream)

    def subgrids(self):
        return download(self.d_subgrids, np.complex64, (max(self.n, 1), 4, self.S, self.S))[: self.n]

    def visibilities_out(self):
        return download(self.d_vis_out, np.complex64, (self.nr_rows, self.C, 4))

    # ---- stages around the gridder / degridder -------------------------------------------------------------
    def fft(self, plan, direction=FFT_FORWARD, d_subgrids=None, stream=None):
        subgrid_fft(plan, self.d_subgrids if d_subgrids is None else d_subgrids, direction, stream)

    def grid_buffer(self):
        """A zeroed complex64 [4][G][G] master grid in HBM (flat bytes)."""
        torch = _torch()
        G = int(self.w["grid_size"])
        return torch.zeros(4 * G * G * 8, dtype=torch.uint8, device=self.device)

    def add(self, plan, d_grid, d_subgrids=None, stream=None):
        """Adds the (uv) subgrids onto d_grid in the conflict-free tile order."""
        if not hasattr(self, "_adder"):
            ap, order, tile_start = adder_plan_build(plan, self.w["metadata"])
            self._adder = (ap, upload(order, self.device), upload(tile_start, self.device))
        ap, d_order, d_tile_start = self._adder
        add_subgrids(plan, ap, self.d_metadata, d_order, d_tile_start,
                     self.d_subgrids if d_subgrids is None else d_subgrids, d_grid, stream)

    def split(self, plan, d_grid, d_subgrids=None, stream=None):
        split_subgrids(plan, self.d_metadata, d_grid, self.d_subgrids if d_subgrids is None else d_subgrids, stream)

    def grid_array(self, d_grid, nr_layers=1):
        G = int(self.w["grid_size"])
        shape = (4, G, G) if nr_layers == 1 else (nr_layers, 4, G, G)
        return download(d_grid, np.complex64, shape)

    # ---- w-stacking: layers z_min .. z_min + nr_layers - 1 of the plan's metadata z ------------------------------
    def layers(self):
        """(z_min, nr_layers) spanned by the plan's w-layers."""
        z = self.w["metadata"]["z"]
        return (int(z.min()), int(z.max()) - int(z.min()) + 1) if z.size else (0, 1)

    def layer_buffer(self, nr_layers):
        """A zeroed complex64 [nr_layers][4][G][G] stack of master grids in HBM (flat bytes)."""
        torch = _torch()
        G = int(self.w["grid_size"])
        return torch.zeros(nr_layers * 4 * G * G * 8, dtype=torch.uint8, device=self.device)

    def add_layered(self, plan, d_layers, d_subgrids=None, stream=None):
        z_min, nl = self.layers()
        if getattr(self, "_adder_layered", None) is None:
            ap, order, tile_start = adder_plan_build_layered(plan, self.w["metadata"], z_min, nl)
            self._adder_layered = (ap, upload(order, self.device), upload(tile_start, self.device))
        ap, d_order, d_tile_start = self._adder_layered
        add_subgrids(plan, ap, self.d_metadata, d_order, d_tile_start,
                     self.d_subgrids if d_subgrids is None else d_subgrids, d_layers, stream)

    def split_layered(self, plan, d_layers, d_subgrids=None, stream=None):
        z_min, nl = self.layers()
        split_subgrids_layered(plan, self.d_metadata, z_min, nl, d_layers,
                               self.d_subgrids if d_subgrids is None else d_subgrids, stream)

    def grid_fft(self, plan, d_grid, nr_layers, direction, stream=None):
        grid_fft(plan, d_grid, nr_layers, direction, stream)

    def to_image(self, plan, d_layers, d_image, stream=None):
        z_min, nl = self.layers()
        wstack_to_image(plan, d_layers, z_min, nl, d_image, stream)

    def from_image(self, plan, d_image, d_layers, stream=None):
        z_min, nl = self.layers()
        wstack_from_image(plan, d_image, z_min, nl, d_layers, stream)

    def grid_to_image(self, plan, d_layers, d_image, box, stream=None):
        """The boxed transform of the layers and the w-screens in one call (idgx_grid_to_image); d_layers is left
        undefined."""
        z_min, nl = self.layers()
        grid_to_image(plan, d_layers, z_min, nl, box, d_image, stream)

    def image_to_grid(self, plan, d_image, d_layers, box, stream=None):
        """The w-screens and the boxed transform to the layers in one call (idgx_image_to_grid): the box's cells."""
        z_min, nl = self.layers()
        image_to_grid(plan, d_image, z_min, nl, box, d_layers, stream)


def selection_lists(w, flags, chunk_entries=1 << 26):
    """Visibility selection of every subgrid of workload w from a boolean flag mask [rows][C] (True = flagged):
    returns (vis_index uint32 [(t << 16) | c], vis_offset int64 [n + 1]) relative to each subgrid's window, entries
    of a subgrid in (t, c) row-major order.  Vectorised over chunks of subgrids of about chunk_entries window
    entries each."""
    md, ch = w["metadata"], w.get("channels")
    C = w["wavenumbers"].shape[0]
    n = md.size
    r0 = md["time_offset"].astype(np.int64) + md["baseline_offset"].astype(np.int64) - \
        (int(md["baseline_offset"][0]) if n else 0)
    T = md["nr_timesteps"].astype(np.int64)
    c0 = np.zeros(n, np.int64) if ch is None else ch["channel_begin"].astype(np.int64)
    nc = np.full(n, C, np.int64) if ch is None else ch["nr_channels"].astype(np.int64)
    window = T * nc
    flat = flags.reshape(-1)
    index, counts = [], np.zeros(n, np.int64)
    s0 = 0
    cum = np.cumsum(window)
    while s0 < n:
        base = cum[s0 - 1] if s0 else 0
        s1 = max(s0 + 1, int(np.searchsorted(cum, base + chunk_entries, side="right")))
        s1 = min(s1, n)
        sub = np.arange(s0, s1)
        # one entry per (subgrid, t): its subgrid, t, and channel count
        st_s = np.repeat(sub, T[sub])
        st_t = np.arange(st_s.size) - np.repeat(np.cumsum(T[sub]) - T[sub], T[sub])
        # one entry per (subgrid, t, c)
        e_st = np.repeat(np.arange(st_s.size), nc[st_s])
        e_c = np.arange(e_st.size) - np.repeat(np.cumsum(nc[st_s]) - nc[st_s], nc[st_s])
        e_s, e_t = st_s[e_st], st_t[e_st]
        keep = ~flat[(r0[e_s] + e_t) * C + c0[e_s] + e_c]
        index.append(((e_t[keep].astype(np.uint32) << 16) | e_c[keep].astype(np.uint32)))
        counts[s0:s1] = np.bincount(e_s[keep] - s0, minlength=s1 - s0)
        s0 = s1
    vi = np.concatenate(index) if index else np.zeros(0, np.uint32)
    offset = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return np.ascontiguousarray(vi.astype(np.uint32)), offset


def device_flags(n_baselines, n_times, n_channels, flag_max, seed, device="cuda", rows_per_chunk=1 << 20):
    """Boolean flags [rows][C] generated in HBM: baseline b flags each (timestep, channel) sample with probability
    f_b ~ U(0, flag_max) (f_b drawn as on the host, numpy seed + 2; the per-sample uniforms from a torch generator
    seeded with seed + 3)."""
    torch = _torch()
    frac = np.random.default_rng(seed + 2).uniform(0.0, flag_max, n_baselines).astype(np.float32)
    d_frac = torch.from_numpy(frac).to(device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed + 3)
    n_rows = n_baselines * n_times
    flags = torch.empty((n_rows, n_channels), dtype=torch.bool, device=device)
    for r0 in range(0, n_rows, rows_per_chunk):
        r1 = min(n_rows, r0 + rows_per_chunk)
        f = d_frac[torch.arange(r0, r1, device=device) // n_times][:, None]
        flags[r0:r1] = torch.rand((r1 - r0, n_channels), generator=gen, device=device) < f
    return flags


def device_selection_lists(w, d_flags, chunk_entries=1 << 27):
    """selection_lists computed in HBM from a device flag mask [rows][C]: returns (d_vis_index int32 tensor holding
    the uint32 entries, d_vis_offset int64 tensor, vis_offset numpy int64)."""
    torch = _torch()
    dev = d_flags.device
    md, ch = w["metadata"], w.get("channels")
    C = int(d_flags.shape[1])
    n = md.size
    r0 = md["time_offset"].astype(np.int64) + md["baseline_offset"].astype(np.int64) - \
        (int(md["baseline_offset"][0]) if n else 0)
    T = md["nr_timesteps"].astype(np.int64)
    c0 = np.zeros(n, np.int64) if ch is None else ch["channel_begin"].astype(np.int64)
    nc = np.full(n, C, np.int64) if ch is None else ch["nr_channels"].astype(np.int64)
    d_r0, d_T, d_c0, d_nc = (torch.from_numpy(x).to(dev) for x in (r0, T, c0, nc))
    cum = np.cumsum(T * nc)
    flat = d_flags.reshape(-1)
    out, counts = [], torch.zeros(n, dtype=torch.int64, device=dev)
    s0 = 0
    while s0 < n:
        base = cum[s0 - 1] if s0 else 0
        s1 = min(n, max(s0 + 1, int(np.searchsorted(cum, base + chunk_entries, side="right"))))
        sub = torch.arange(s0, s1, device=dev)
        Ts = d_T[sub]
        st_s = torch.repeat_interleave(sub, Ts)
        st_t = torch.arange(st_s.numel(), device=dev) - torch.repeat_interleave(torch.cumsum(Ts, 0) - Ts, Ts)
        ncs = d_nc[st_s]
        e_st = torch.repeat_interleave(torch.arange(st_s.numel(), device=dev), ncs)
        e_c = torch.arange(e_st.numel(), device=dev) - torch.repeat_interleave(torch.cumsum(ncs, 0) - ncs, ncs)
        e_s, e_t = st_s[e_st], st_t[e_st]
        keep = ~flat[(d_r0[e_s] + e_t) * C + d_c0[e_s] + e_c]
        entry = (e_t[keep] << 16) | e_c[keep]
        out.append(torch.where(entry >= (1 << 31), entry - (1 << 32), entry).to(torch.int32))
        counts[s0:s1] = torch.bincount(e_s[keep] - s0, minlength=s1 - s0)
        s0 = s1
    d_vi = torch.cat(out) if out else torch.zeros(0, dtype=torch.int32, device=dev)
    d_off = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(counts, 0)])
    return d_vi, d_off, d_off.cpu().numpy()


class DeviceBatch:
    """One or more subgrid classes of an observation resident in HBM, run through the batch API (persistent
    work-queue kernels).  `classes` are workload dicts sharing uvw / wavenumbers / visibilities rows; each may
    carry 'vis_index' / 'vis_offset' selection lists (host 'vis_offset' is always needed; 'd_vis_index' /
    'd_vis_offset' device tensors, when present, are used instead of uploading the host lists)."""

    def __init__(self, classes, precision="bf16x3", device="cuda", visibilities_on_device=None, aterm_convention=0):
        from . import Batch

        torch = _torch()
        self.device = device
        w0 = classes[0]
        self.C = int(w0["wavenumbers"].shape[0])
        self.nr_rows = int(w0["uvw"].shape[0])
        self.d_uvw = upload(w0["uvw"], device)
        self.d_wavenumbers = upload(w0["wavenumbers"], device)
        self.d_visibilities = visibilities_on_device if visibilities_on_device is not None else \
            upload(w0["visibilities"], device)
        self.d_vis_out = torch.zeros(self.nr_rows * self.C * 4 * 8, dtype=torch.uint8, device=device)
        self.classes = []
        specs, cls_data = [], []
        for w in classes:
            S, n = int(w["subgrid_size"]), int(w["nr_subgrids"])
            c = dict(w=w, S=S, n=n,
                     d_metadata=upload(w["metadata"], device),
                     d_channels=None if w.get("channels") is None else upload(w["channels"], device),
                     d_vis_index=w["d_vis_index"] if w.get("d_vis_index") is not None else
                     None if w.get("vis_index") is None else upload(w["vis_index"], device),
                     d_vis_offset=w["d_vis_offset"] if w.get("d_vis_offset") is not None else
                     None if w.get("vis_offset") is None else upload(w["vis_offset"], device),
                     d_spheroidal=upload(w["spheroidal"], device), d_aterms=upload(w["aterms"], device),
                     d_subgrids=torch.zeros(max(n, 1) * 4 * S * S * 8, dtype=torch.uint8, device=device))
            plan = make_plan(nr_subgrids=n, subgrid_size=S, grid_size=w["grid_size"], nr_channels=self.C,
                             nr_stations=w["nr_stations"], nr_aterm_slots=w["nr_aterm_slots"], nr_rows=self.nr_rows,
                             max_nr_timesteps=w["max_nr_timesteps"], image_size=w["image_size"], w_step=w["w_step"],
                             precision=precision, aterm_convention=aterm_convention)
            c["plan"] = plan
            self.classes.append(c)
            specs.append((plan, w["metadata"], w.get("channels"), w.get("vis_offset")))
        self.batch = Batch(specs)
        self.data = Batch.data(self.d_uvw, self.d_wavenumbers, self.d_visibilities, self.d_vis_out, self.classes)

    def rebatch(self, precision):
        """Replace the batch by one of another precision over the same HBM buffers (the old batch, and its device
        memory, is released first: a mixed batch holds prepared pixels as large as the subgrids)."""
        from . import Batch, PRECISIONS

        self.batch.close()
        specs = []
        for c in self.classes:
            plan = type(c["plan"]).from_buffer_copy(c["plan"])
            plan.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
            c["plan"] = plan
            w = c["w"]
            specs.append((plan, w["metadata"], w.get("channels"), w.get("vis_offset")))
        self.batch = Batch(specs)

    def grid(self, stream=None):
        self.batch.grid(self.data, stream)

    def degrid(self, stream=None):
        self.batch.degrid(self.data, stream)

    def subgrids(self, k=0):
        c = self.classes[k]
        return download(c["d_subgrids"], np.complex64, (max(c["n"], 1), 4, c["S"], c["S"]))[: c["n"]]

    def visibilities_out(self):
        return download(self.d_vis_out, np.complex64, (self.nr_rows, self.C, 4))
