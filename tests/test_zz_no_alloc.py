"""The allocation-free master-grid transform (idgx_grid_fft_noalloc) against idgx_grid_fft through the C ABI: the same
kernels with a fixed scratch array instead of a plane allocated on first use, so the results must agree bit for bit,
on two calls in a row (the second smaller), and the call must not change the device's free memory."""
import numpy as np
import pytest

from idgx import ERR_UNSUPPORTED, FFT_BACKWARD, FFT_FORWARD, IdgxError, grid_fft, grid_fft_noalloc, make_plan

pytestmark = pytest.mark.gpu


def _plan(G):
    return make_plan(nr_subgrids=0, subgrid_size=32, grid_size=G, nr_channels=1, nr_stations=1, nr_aterm_slots=1,
                     nr_rows=0, max_nr_timesteps=0, image_size=0.1)


def _grid(G, layers, seed):
    import torch

    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    return torch.randn(layers * 4 * G * G * 2, generator=gen, device="cuda", dtype=torch.float32)


def test_grid_fft_noalloc_matches_grid_fft_on_two_calls():
    import torch

    # First call: G = 4096, one plane per group (the largest the fixed scratch holds); second, smaller: G = 1024 with
    # two w-layers (eight planes per group).
    for G, layers, direction, seed in ((4096, 1, FFT_BACKWARD, 1), (1024, 2, FFT_FORWARD, 2)):
        print(f"noalloc: G={G}, {layers} layer(s), direction {direction}", flush=True)
        src = _grid(G, layers, seed)
        a, b = src.clone(), src.clone()
        torch.cuda.synchronize()
        grid_fft(_plan(G), a, layers, direction)
        free0 = torch.cuda.mem_get_info()[0]
        grid_fft_noalloc(_plan(G), b, layers, direction)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), f"G={G}: results differ"
        assert not torch.equal(a, src), "the transform ran"
        assert free1 == free0, f"G={G}: free device memory changed by {free0 - free1} bytes"
        del src, a, b


def test_grid_fft_noalloc_rejects_grids_beyond_its_scratch():
    import torch

    d = torch.zeros(16, device="cuda")
    for G in (8192, 16384, 64, 1000):
        with pytest.raises(IdgxError) as e:
            grid_fft_noalloc(_plan(G), d, 1, FFT_FORWARD)
        assert e.value.status == ERR_UNSUPPORTED
    assert np.all(d.cpu().numpy() == 0)
