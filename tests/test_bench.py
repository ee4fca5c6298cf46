"""bench.py's bookkeeping (no GPU): the kernel names it looks up in the committed profiles follow the library's own
launch rule (idgx_plain_kernel_shape), and its roofline record is built from those profiles."""
import importlib.util
import os

import pytest

from idgx import make_plan, plain_kernel_shape

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _plan(S, C=64):
    return make_plan(nr_subgrids=1, subgrid_size=S, grid_size=4096, nr_channels=C, nr_stations=2, nr_aterm_slots=1,
                     nr_rows=1, max_nr_timesteps=1, image_size=0.1)


def test_plain_kernel_shape_rule():
    # Gridder: 3 pixels per thread where 768-pixel blocks leave fewer idle lanes than 1024-pixel ones.
    for S in range(8, 129, 2):
        P = S * S
        want = 3 if -(-P // 768) * 768 < -(-P // 1024) * 1024 else 4
        assert plain_kernel_shape(_plan(S), False)[0] == want, S
    assert [plain_kernel_shape(_plan(S), False)[0] for S in (24, 32, 48)] == [3, 4, 3]
    # Degridder: 4 channels per thread only without channel windows and with a channel count divisible by 4.
    assert plain_kernel_shape(_plan(32, 16), False)[1] == 4
    assert plain_kernel_shape(_plan(32, 16), True)[1] == 2
    assert plain_kernel_shape(_plan(32, 18), False)[1] == 2


def test_kernel_keys_resolve_in_the_committed_profiles(bench):
    shape = plain_kernel_shape(_plan(32), True)  # the SKA-Low-like plan: S = 32 with channel windows
    for direction in ("gridder", "degridder"):
        key = bench.kernel_key("fp32", "one-shot", direction, plain_shape=shape)
        assert bench.measured_traffic("ska_low", key) is not None, key
        assert bench.pipe_profile("ska_low", key) is not None, key
    assert bench.kernel_key("fp32", "one-shot", "degridder", plain_shape=shape) == "degridder_plain<2, 1>"
    assert bench.kernel_key("fp32", "batched", "gridder", "tiles per class") == "gridder_plain_queue_tiles<1>"
    assert bench.kernel_key("fp32", "batched", "degridder", (4, 2)) == "degridder_plain_queue<2, 1>"
    key = bench.kernel_key("fp16x3", "batched", "degridder", (4, 256, 0))
    assert key == "degridder_mfma_queue<1, true, 4, 256, false>"
    assert bench.measured_traffic("ska_low", key) is not None


def test_batched_uneven_keys_resolve_in_the_committed_profiles(bench):
    # the default single launch of the batched uneven plan: the gridder with each class's own tiles per wave
    key = bench.kernel_key("fp16x3", "batched", "gridder", "tiles per class")
    assert key == "gridder_mfma_queue_tiles<1, true>"
    assert bench.measured_traffic("batched_uneven", key) is not None
    pipes = bench.pipe_profile("batched_uneven", key)
    assert pipes is not None and pipes["VALUBusy"] > 90


def test_roofline_record_is_valu_bound_with_counters(bench):
    key = "degridder_mfma_queue<1, true, 4, 256, false>"
    sincos = 136.3e9
    r = bench.roofline_record("ska_low", "fp16x3", key, "degridder", 4.7e15, sincos, 0.042)
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s" and r["peak"] == bench.FP32_PEAK_TFLOPS
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"])
    assert r["valu"]["busy_pct"] > 90 and 5 < r["valu"]["insts_per_pair"] < 20
    assert 0 < r["mfma"]["frac"] < 1 and 0 < r["transcendental"]["frac"] < 1
    assert r["traffic"] is not None
