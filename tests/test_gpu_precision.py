"""Extraction from FP16 / BF16 volume storage (FP32 math): exactness and error against the FP64 oracle.

The half-precision storage formats are an extension (noa's f16 path computes in f16, Interpolation.hpp:445-448,
and noa has no bf16): the volume is stored as {f16, f16} or {bf16, bf16} pairs and converted to FP32 at load.
"""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import random_rotations, rel_l2

pytestmark = pytest.mark.gpu


def _volume(n, rng):
    """CN(0,1) x exp(-20 |f|^2) on a centred rfft grid (the extraction workload's spectrum, smaller)."""
    f = (np.arange(n) - n // 2) / n
    fx = np.arange(n // 2 + 1) / n
    env = np.exp(-20 * (f[:, None, None] ** 2 + f[None, :, None] ** 2 + fx[None, None, :] ** 2))
    v = (rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))) * env
    return v.astype(np.complex64)[None]


def _extract_gpu(vol, n, rot):
    import noa_amd.xform as nx
    b = rot.shape[0]
    out = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device="cuda")
    nx.extract_central_slices_3d("hc2hc", vol, None, (1, n, n, n), out, None, (b, 1, n, n), None,
                                 nx.Mat33Array(rot, "cuda"))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("layout", ["default", "bricks", "linear"])
@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_half_storage_equals_fp32_math_on_the_rounded_volume(monkeypatch, storage, layout):
    """Bit-identical to the c32 extraction of the rounded volume, with the gather from the row-major volume or from
    its block copy (forced either way)."""
    import noa_amd.xform as nx
    print(f"case half storage {storage} layout={layout}", flush=True)
    if layout != "default":
        monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", layout)
    rng = np.random.default_rng(31)
    n = 48
    vol = torch.from_numpy(_volume(n, rng)).cuda()
    rot = random_rotations(16, rng)
    if storage == "fp16":
        half = vol.to(torch.complex32)
        rounded = half.to(torch.complex64)
    else:
        half = nx.ComplexBF16.from_complex(vol)
        rounded = torch.view_as_complex(half.pairs.to(torch.float32).contiguous())
    got = _extract_gpu(half, n, rot)
    want = _extract_gpu(rounded.contiguous(), n, rot)
    np.testing.assert_array_equal(got, want)


def test_storage_error_against_the_fp64_oracle():
    """rel-L2 of each storage against the FP64 oracle (double volume, double rotations): FP32 ~1e-7, FP16 ~1e-3,
    BF16 ~1e-2 (8-bit significand). Pixels within 1e-3 of the 0.5 cutoff circle are left out: there FP32 and FP64
    coordinates can fall on different sides of the cutoff, and the pixel is the value or 0."""
    import noa_amd.xform as nx
    rng = np.random.default_rng(32)
    n = 48
    vol32 = _volume(n, rng)
    rot = random_rotations(16, rng, dtype=np.float64)
    ref = np.zeros((16, 1, n, n // 2 + 1), np.complex128)
    O.extract_central_slices_3d("hc2hc", vol32.astype(np.complex128), None, (1, n, n, n), ref, None, (16, 1, n, n),
                                None, rot)
    dev = torch.from_numpy(vol32).cuda()
    errs = {}
    fy = (np.arange(n) - n // 2)[:, None] / n
    fx = np.arange(n // 2 + 1)[None, :] / n
    keep = np.broadcast_to(np.abs(np.sqrt(fy ** 2 + fx ** 2) - 0.5) > 1e-3, (16, 1, n, n // 2 + 1))
    for name, v in (("fp32", dev), ("fp16", dev.to(torch.complex32)), ("bf16", nx.ComplexBF16.from_complex(dev))):
        got = _extract_gpu(v, n, rot.astype(np.float32))
        errs[name] = rel_l2(got.astype(np.complex128)[keep], ref[keep])
    assert errs["fp32"] < 1e-5
    assert errs["fp16"] < 2e-3
    assert errs["bf16"] < 2e-2
    assert errs["fp32"] < errs["fp16"] < errs["bf16"]
