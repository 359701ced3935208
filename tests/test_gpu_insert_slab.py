"""The slab insertion kernel (one power-of-two size, Mat33<f32>, LINEAR, cutoff <= 0.5) against the oracle.

Bars: rel-L2 1e-6 of values and weights against noa's per-voxel gather (the kernel evaluates noa's coordinate
expressions bit for bit; the window polynomial differs from noa's sin/cos form by < 4e-7 and the sums are regrouped),
bit-identical run to run, and within 1e-6 of the one-wave kernel (NOA_HIP_INSERT_KERNEL=wave) while not
bit-identical to it (the two kernels are distinct code paths).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import random_rotations, random_spectrum, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(autouse=True)
def slab_kernel(monkeypatch):
    """The slab kernel is the default where it applies; NOA_HIP_INSERT_KERNEL=slab makes these tests independent of
    the tuning table."""
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", "slab")


def nx():
    import noa_amd.xform as m
    return m


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def run_gpu(remap, sl, w, shape_s, n, inv, out_dtype, wvol=True):
    vol = torch.zeros((1, n, n, n // 2 + 1), dtype=out_dtype, device=DEV)
    wv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV) if wvol else None
    nx().insert_central_slices_3d(remap, sl, w, shape_s, vol, wv, (1, n, n, n), None, inv)
    torch.cuda.synchronize()
    return vol.cpu().numpy(), (wv.cpu().numpy() if wvol else None)


def run_oracle(remap, sl, w, shape_s, n, inv, out_dtype, wvol=True):
    vol = np.zeros((1, n, n, n // 2 + 1), out_dtype)
    wv = np.zeros((1, n, n, n // 2 + 1), np.float32) if wvol else None
    O.insert_central_slices_3d(remap, sl, w, shape_s, vol, wv, (1, n, n, n), None, inv, n_threads=8)
    return vol, wv


CASES = [
    # remap, slice dtype, weights, output dtype
    ("h2h", np.complex64, True, np.complex64),
    ("hc2hc", np.complex64, True, np.complex64),
    ("hc2h", np.complex64, True, np.complex64),
    ("h2hc", np.complex64, True, np.complex64),
    ("h2h", np.float32, False, np.float32),
    ("hc2hc", np.complex64, False, np.float32),
]


@pytest.mark.parametrize("remap,tin,has_w,tout", CASES)
def test_slab_matches_the_oracle(remap, tin, has_w, tout):
    print(f"case slab {remap} {np.dtype(tin)} weights={has_w} -> {np.dtype(tout)}", flush=True)
    rng = np.random.default_rng(300)
    n, b = 64, 200
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng, tin)
    w = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32) if has_w else None
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    want_v, want_w = run_oracle(remap, sl, w, (b, 1, n, n), n, inv, tout)
    got_v, got_w = run_gpu(remap, dev(sl), dev(w) if has_w else None, (b, 1, n, n), n, dev(inv),
                           getattr(torch, np.dtype(tout).name))
    assert rel_l2(got_v, want_v) < 1e-6
    assert rel_l2(got_w, want_w) < 1e-6


@pytest.mark.parametrize("paired", ["0", "1"])
@pytest.mark.parametrize("remap", ["h2h", "hc2hc"])
def test_slab_axis_poses_reach_the_slice_edges(monkeypatch, paired, remap):
    """Identity and 90-degree poses put voxels exactly on slice samples, on the last row and on x = n/2, where the
    x + 1 tap lies past the end of the slice: the tap must contribute nothing and the valid one its full value,
    whether the two taps of a row are separate loads or one merged load."""
    print(f"case slab axis poses {remap} paired={paired}", flush=True)
    monkeypatch.setenv("NOA_HIP_INSERT_PAIRED", paired)
    rng = np.random.default_rng(301)
    n = 32
    poses = [np.eye(3)]
    for fn in (nx().rotate_z, nx().rotate_y, nx().rotate_x):
        for deg in (90, 180, 270):
            poses.append(fn(np.deg2rad(deg)))
    rot = np.round(np.stack(poses)).astype(np.float32)      # exact 0/+-1 entries
    b = rot.shape[0]
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng, np.complex64)
    w = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)
    inv = np.ascontiguousarray(np.transpose(rot, (0, 2, 1)))
    want_v, want_w = run_oracle(remap, sl, w, (b, 1, n, n), n, inv, np.complex64)
    got_v, got_w = run_gpu(remap, dev(sl), dev(w), (b, 1, n, n), n, dev(inv), torch.complex64)
    assert rel_l2(got_v, want_v) < 1e-6
    assert rel_l2(got_w, want_w) < 1e-6
    # Per voxel as well: the edge voxels are few, a wrong tap there would hide in a global norm.
    np.testing.assert_allclose(got_v, want_v, rtol=0, atol=2e-6 * np.abs(want_v).max())
    np.testing.assert_allclose(got_w, want_w, rtol=0, atol=2e-6 * np.abs(want_w).max())


def test_slab_is_deterministic_and_distinct_from_the_wave_kernel(monkeypatch):
    print("case slab determinism and wave-kernel agreement", flush=True)
    rng = np.random.default_rng(302)
    n, b = 128, 1500
    sl = dev(random_spectrum((b, 1, n, n // 2 + 1), rng, np.complex64))
    w = dev(rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32))
    inv = dev(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    a_v, a_w = run_gpu("h2h", sl, w, (b, 1, n, n), n, inv, torch.complex64)
    b_v, b_w = run_gpu("h2h", sl, w, (b, 1, n, n), n, inv, torch.complex64)
    np.testing.assert_array_equal(a_v, b_v)
    np.testing.assert_array_equal(a_w, b_w)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", "wave")
    c_v, c_w = run_gpu("h2h", sl, w, (b, 1, n, n), n, inv, torch.complex64)
    assert rel_l2(a_v, c_v) < 1e-6 and rel_l2(a_w, c_w) < 1e-6
    assert not np.array_equal(a_v, c_v), "the slab path ran the one-wave kernel"


def test_slab_strided_slices_and_broadcast_rotation():
    """Slices and weights cut from wider rows (row pitch > n/2 + 1, same pitch for both), and one rotation for every
    slice (a host Mat33)."""
    print("case slab strided slices, broadcast rotation", flush=True)
    rng = np.random.default_rng(303)
    n, b = 64, 40
    wide = random_spectrum((b, 1, n, n // 2 + 9), rng, np.complex64)
    wwide = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 9)).astype(np.float32)
    sl, w = wide[..., : n // 2 + 1], wwide[..., : n // 2 + 1]
    rot = random_rotations(1, rng)[0]
    inv = np.ascontiguousarray(rot.T)
    want_v, want_w = run_oracle("hc2hc", sl, w, (b, 1, n, n), n, inv, np.complex64)
    gs = torch.from_numpy(wide).to(DEV)[..., : n // 2 + 1]
    gw = torch.from_numpy(wwide).to(DEV)[..., : n // 2 + 1]
    got_v, got_w = run_gpu("hc2hc", gs, gw, (b, 1, n, n), n, inv, torch.complex64)
    assert rel_l2(got_v, want_v) < 1e-6
    assert rel_l2(got_w, want_w) < 1e-6


def test_slab_adds_to_the_volume_and_leaves_voxels_beyond_the_cutoff():
    """Insertion accumulates into the volume; voxels beyond the cutoff keep their value (noa returns early)."""
    print("case slab accumulate / cutoff", flush=True)
    rng = np.random.default_rng(304)
    n, b = 64, 50
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng, np.complex64)
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    start = random_spectrum((1, n, n, n // 2 + 1), rng, np.complex64)
    want = start.copy()
    O.insert_central_slices_3d("h2h", sl, None, (b, 1, n, n), want, None, (1, n, n, n), None, inv,
                               fftfreq_cutoff=0.4)
    vol = dev(start)
    opts = nx().InsertCentralSlicesOptions(fftfreq_cutoff=0.4)
    nx().insert_central_slices_3d("h2h", dev(sl), None, (b, 1, n, n), vol, None, (1, n, n, n), None, dev(inv), opts)
    torch.cuda.synchronize()
    got = vol.cpu().numpy()
    assert rel_l2(got, want) < 1e-6
    fr = np.fft.fftfreq(n)
    r = np.sqrt(fr[:, None, None] ** 2 + fr[None, :, None] ** 2 + (np.arange(n // 2 + 1) / n)[None, None, :] ** 2)
    outside = np.broadcast_to(r > 0.4 + 1e-6, got.shape)
    np.testing.assert_array_equal(got[outside], start[outside])
