"""remap and fourier_interpolation_correction on the GPU against the oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("remap", ["h2hc", "hc2h", "h2h", "hc2hc"])
@pytest.mark.parametrize("dtype", [np.complex64, np.float32, np.complex128])
@pytest.mark.parametrize("shape", [(2, 16, 16, 16), (1, 9, 7, 12), (3, 1, 30, 30)])
def test_remap_matches_the_oracle(remap, dtype, shape):
    import noa_amd.xform as nx
    rng = np.random.default_rng(71)
    rshape = (shape[0], shape[1], shape[2], shape[3] // 2 + 1)
    x = rng.standard_normal(rshape).astype(dtype)
    if np.iscomplexobj(x):
        x = x + 1j * rng.standard_normal(rshape).astype(x.real.dtype)
    want = np.zeros_like(x)
    O.remap_hx2hx(remap, x, want, shape)
    out = torch.zeros(rshape, dtype={np.complex64: torch.complex64, np.float32: torch.float32,
                                     np.complex128: torch.complex128}[dtype], device="cuda")
    nx.remap(remap, torch.from_numpy(x).cuda(), out, shape)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_remap_rejects_in_place_layout_change():
    import noa_amd.xform as nx
    t = torch.zeros((1, 8, 8, 5), dtype=torch.complex64, device="cuda")
    with pytest.raises(nx.NoaError, match="in-place"):
        nx.remap("h2hc", t, t, (1, 8, 8, 8))


@pytest.mark.parametrize("post", [True, False])
@pytest.mark.parametrize("in_dtype, out_dtype", [(np.float32, np.float32), (np.float64, np.float64),
                                                 (np.float32, np.float64)])
def test_gridding_correction_matches_the_oracle(post, in_dtype, out_dtype):
    import noa_amd.xform as nx
    rng = np.random.default_rng(72)
    shape = (2, 20, 18, 17)
    x = rng.standard_normal((1,) + shape[1:]).astype(in_dtype)      # broadcast over the batch
    want = np.zeros(shape, out_dtype)
    O.fourier_interpolation_correction(x, want, post_correction=post)
    out = torch.zeros(shape, dtype=torch.float64 if out_dtype == np.float64 else torch.float32, device="cuda")
    nx.fourier_interpolation_correction(torch.from_numpy(x).cuda(), out, post_correction=post)
    torch.cuda.synchronize()
    tol = 1e-13 if (in_dtype == np.float64 or out_dtype == np.float64) else 1e-6
    assert rel_l2(out.cpu().numpy(), want) < tol


def test_gridding_correction_in_place():
    import noa_amd.xform as nx
    rng = np.random.default_rng(73)
    x = rng.standard_normal((1, 16, 16, 16)).astype(np.float32)
    want = np.zeros_like(x)
    O.fourier_interpolation_correction(x, want)
    t = torch.from_numpy(x).cuda()
    nx.fourier_interpolation_correction(t, t)
    torch.cuda.synchronize()
    assert rel_l2(t.cpu().numpy(), want) < 1e-6
    with pytest.raises(nx.NoaError, match="not supported"):
        nx.fourier_interpolation_correction(t, t, interp=nx.Interp.CUBIC)
