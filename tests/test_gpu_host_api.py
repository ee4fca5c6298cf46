"""Host-buffer entry points (idgx_host_grid / idgx_host_degrid): the reference launcher's contract in one C call.

They must give exactly the device-pointer paths' results (same kernels on the same inputs), hence the reference
CPU path's within the plain-form bar, and report positive per-iteration timings."""
import numpy as np
import pytest

from helpers import golden_names, load_golden, per_subgrid_rel_l2, rel_l2
from idgx import ATERM_REFERENCE, host_degrid, host_grid, observation
from idgx.device import DeviceBatch, DeviceWorkload, upload

pytestmark = pytest.mark.gpu

PARITY = 1e-5  # relative L2 per subgrid, plain FP32 vs the reference CPU path


def _timing_ok(t):
    return t.in_transfer_time > 0 and t.out_transfer_time > 0 and t.compute_time > 0 and \
        t.in_bandwidth > 0 and t.out_bandwidth > 0


@pytest.mark.parametrize("name", golden_names())
def test_host_grid_equals_device_path_and_reference(name):
    w = load_golden(name)
    dw = DeviceWorkload(w)
    plan = dw.plan()
    dw.grid(plan)
    sub, t = host_grid(plan, w, iterations=2)
    assert sub.tobytes() == dw.subgrids().tobytes()
    assert per_subgrid_rel_l2(sub, w["ref_grid_subgrids"]).max() <= PARITY
    print(f"{name}: in {t.in_transfer_time:.1f} us ({t.in_bandwidth:.0f} MB/s), compute {t.compute_time:.1f} us, "
          f"out {t.out_transfer_time:.1f} us ({t.out_bandwidth:.0f} MB/s)")
    assert _timing_ok(t)


@pytest.mark.parametrize("name", golden_names())
def test_host_degrid_equals_device_path_and_reference(name):
    w = load_golden(name)
    w["visibilities"] = np.zeros_like(w["visibilities"])
    dw = DeviceWorkload(w)
    dw.d_subgrids.copy_(upload(w["ref_degrid_subgrids"]))
    plan = dw.plan(aterm_convention=ATERM_REFERENCE)
    dw.degrid(plan)
    vis, t = host_degrid(plan, w, w["ref_degrid_subgrids"], iterations=3)
    assert vis.tobytes() == dw.visibilities_out().tobytes()
    md = w["metadata"]
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        assert rel_l2(vis[r0:r0 + nt], w["ref_degrid_visibilities"][r0:r0 + nt]) <= PARITY
    assert _timing_ok(t)


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_host_calls_of_the_mixed_forms_equal_the_batch(precision):
    """Mixed precisions run as a one-class batch inside the call: same bits as idgx_batch_* on device buffers,
    including per-subgrid channel windows."""
    w = observation.make_observation("small")
    assert w.get("channels") is not None
    db = DeviceBatch([w], precision=precision)
    db.grid()
    db.degrid()
    plan = DeviceWorkload(w).plan(precision=precision)
    sub, tg = host_grid(plan, w)
    assert sub.tobytes() == db.subgrids().tobytes()
    vis, td = host_degrid(plan, w, sub)
    assert vis.tobytes() == db.visibilities_out().tobytes()
    assert _timing_ok(tg) and _timing_ok(td)
