"""Shared helpers of the test suite: fixture loading, error metrics and oracle calls on a workload dict."""
import json
import os

import numpy as np

import oracle
from idgx import METADATA_DTYPE

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_names():
    with open(os.path.join(GOLDEN, "MANIFEST.json")) as f:
        return sorted(json.load(f))


def load_golden(name):
    """A fixture as a workload dict (same keys as idgx.workload.reference_workload) plus the reference outputs."""
    with open(os.path.join(GOLDEN, "MANIFEST.json")) as f:
        man = json.load(f)[name]
    sh = man["shape"]
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    T, C, S = sh["ntimesteps_per_subgrid"], sh["nchannels"], sh["subgrid_size"]
    md = np.ascontiguousarray(z["metadata"]).view(METADATA_DTYPE)
    nsub = md.size
    w = dict(
        shape=sh,
        uvw=np.ascontiguousarray(z["uvw"].reshape(nsub * T, 3)),
        frequencies=z["frequencies"],
        wavenumbers=z["wavenumbers"],
        visibilities=np.ascontiguousarray(z["visibilities"].reshape(nsub * T, C, 4)),
        baselines=z["baselines"],
        spheroidal=z["spheroidal"],
        aterms=z["aterms"],
        metadata=md,
        channels=None,
        nr_subgrids=nsub,
        subgrid_size=S,
        grid_size=sh["grid_size"],
        nr_stations=sh["nstations"],
        nr_aterm_slots=sh["ntimeslots"],
        image_size=np.float32(man["image_size"]),
        w_step=float(man["w_step_in_lambda"]),
        max_nr_timesteps=T,
        nr_rows=nsub * T,
        ref_grid_subgrids=z["grid_subgrids"],
        ref_degrid_subgrids=z["degrid_subgrids"],
        ref_degrid_visibilities=np.ascontiguousarray(z["degrid_visibilities"].reshape(nsub * T, C, 4)),
    )
    return w


def rel_l2(a, b):
    b = np.asarray(b)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - b) / (den if den > 0 else 1.0))


def per_subgrid_rel_l2(a, b):
    """Relative L2 error of each leading-axis slice."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    return num / np.where(den > 0, den, 1.0)


def _kw(w):
    return dict(grid_size=int(w["grid_size"]), nr_stations=int(w["nr_stations"]), image_size=np.float32(w["image_size"]),
                w_step=float(w["w_step"]), channels=w.get("channels"))


def oracle_grid(w, fp64=False, nthreads=0):
    f = oracle.grid_fp64 if fp64 else oracle.grid_fp32
    return f(w["metadata"], w["uvw"], w["wavenumbers"], w["visibilities"], w["spheroidal"], w["aterms"],
             subgrid_size=int(w["subgrid_size"]), nthreads=nthreads, **_kw(w))


def oracle_degrid(w, subgrids, fp64=False, convention=oracle.ATERM_ADJOINT, nthreads=0):
    kw = _kw(w)
    if fp64:
        return oracle.degrid_fp64(w["metadata"], w["uvw"], w["wavenumbers"], subgrids, w["spheroidal"], w["aterms"],
                                  nr_rows=w["uvw"].shape[0], nthreads=nthreads, **kw)
    return oracle.degrid_fp32(w["metadata"], w["uvw"], w["wavenumbers"], subgrids, w["spheroidal"], w["aterms"],
                              nr_rows=w["uvw"].shape[0], convention=convention, nthreads=nthreads, **kw)


def covered_mask(w):
    """Boolean [rows][C] of the visibility entries the plan covers."""
    C = w["wavenumbers"].shape[0]
    mask = np.zeros((w["uvw"].shape[0], C), bool)
    md, ch = w["metadata"], w.get("channels")
    for s in range(md.size):
        r0 = int(md["time_offset"][s]) + int(md["baseline_offset"][s]) - int(md["baseline_offset"][0])
        c0, nc = (0, C) if ch is None else (int(ch["channel_begin"][s]), int(ch["nr_channels"][s]))
        mask[r0:r0 + int(md["nr_timesteps"][s]), c0:c0 + nc] = True
    return mask
