"""phase_shift_2d and the fused shifted extraction on the GPU against the oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import random_rotations, random_spectrum, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("remap", ["h2h", "hc2hc", "h2hc", "hc2h"])
@pytest.mark.parametrize("case", ["batched", "single", "half", "zero", "c64", "no-input", "cutoff"])
def test_phase_shift_matches_the_oracle(remap, case):
    import noa_amd.xform as nx
    rng = np.random.default_rng(51)
    b, h, w = 5, 30, 24
    dtype = np.complex128 if case == "c64" else np.complex64
    spec = random_spectrum((b, 1, h, w // 2 + 1), rng).astype(dtype)
    shifts = rng.uniform(-6, 6, (b, 2)).astype(np.float64 if case == "c64" else np.float32)
    cutoff = 0.3 if case == "cutoff" else 1.0
    if case in ("single", "no-input"):
        shifts = shifts[0]
    elif case == "half":
        shifts = np.array([h / 2, -w / 2], np.float32)
    elif case == "zero":
        shifts = np.zeros(2, np.float32)
    want = np.zeros_like(spec)
    O.phase_shift_2d(remap, None if case == "no-input" else spec, want, (b, 1, h, w), shifts, fftfreq_cutoff=cutoff)
    out = torch.zeros(spec.shape, dtype=torch.complex128 if case == "c64" else torch.complex64, device="cuda")
    inp = None if case == "no-input" else torch.from_numpy(spec).cuda()
    sh = torch.from_numpy(shifts).cuda() if shifts.ndim == 2 else shifts
    nx.phase_shift_2d(remap, inp, out, (b, 1, h, w), sh, fftfreq_cutoff=cutoff)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    if case in ("half", "zero"):
        np.testing.assert_array_equal(got, want)
    else:
        assert rel_l2(got, want) < (1e-12 if case == "c64" else 3e-7)


def test_phase_shift_in_place():
    import noa_amd.xform as nx
    rng = np.random.default_rng(52)
    spec = random_spectrum((3, 1, 32, 17), rng)
    shifts = rng.uniform(-4, 4, (3, 2)).astype(np.float32)
    want = np.zeros_like(spec)
    O.phase_shift_2d("hc2hc", spec, want, (3, 1, 32, 32), shifts)
    t = torch.from_numpy(spec).cuda()
    nx.phase_shift_2d("hc2hc", t, t, (3, 1, 32, 32), torch.from_numpy(shifts).cuda())
    torch.cuda.synchronize()
    assert rel_l2(t.cpu().numpy(), want) < 3e-7
    with pytest.raises(nx.NoaError, match="In-place remapping"):
        nx.phase_shift_2d("h2hc", t, t, (3, 1, 32, 32), torch.from_numpy(shifts).cuda())


@pytest.mark.parametrize("remap", ["hc2hc", "h2h"])
def test_fused_shifted_extraction_equals_extract_then_shift(remap):
    """The fused epilogue equals extraction followed by phase_shift_2d on the GPU exactly, and the oracle's
    extraction + phase shift within the sin/cos rounding."""
    import noa_amd.xform as nx
    rng = np.random.default_rng(53)
    n, b = 48, 12
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    shifts = rng.uniform(-5, 5, (b, 2)).astype(np.float32)
    layout = "hc2hc" if remap.endswith("hc") else "h2h"
    dvol = torch.from_numpy(vol).cuda()
    drot = torch.from_numpy(rot).cuda()
    dsh = torch.from_numpy(shifts).cuda()
    fused = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device="cuda")
    nx.extract_central_slices_3d_shifted(remap, dvol, None, (1, n, n, n), fused, None, (b, 1, n, n), None, drot,
                                         shifts=dsh)
    plain = torch.empty_like(fused)
    nx.extract_central_slices_3d(remap, dvol, None, (1, n, n, n), plain, None, (b, 1, n, n), None, drot)
    nx.phase_shift_2d(layout, plain, plain, (b, 1, n, n), dsh)
    torch.cuda.synchronize()
    assert torch.equal(fused, plain)
    want = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
    O.extract_central_slices_3d(remap, vol, None, (1, n, n, n), want, None, (b, 1, n, n), None, rot)
    O.phase_shift_2d(layout, want.copy(), want, (b, 1, n, n), shifts)
    assert rel_l2(fused.cpu().numpy(), want) < 3e-7
