"""The tiled rasterisation (LDS accumulators per 16^3 volume tile, each tile visited by the slices whose plane
crosses it, one device-wide atomic flush per work item).

It takes launches of 64 slices or more into a cubic volume with FP32 values and weights, Mat33<f32> or
Quaternion<f32> rotations, Mat22 inverse scalings or none, no Ewald curvature and cutoff <= 0.5; NOA_HIP_RASTERIZE_LAYOUT=tiles|direct forces either path. Bars: rel-L2 1e-6 of values and weights
against the oracle (noa's per-pixel scatter; the sums are regrouped, by LDS and device atomics in both paths) and
against the one-lane-per-pixel kernel.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import random_rotations, random_spectrum, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


def nx():
    import noa_amd.xform as m
    return m


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def axis_poses():
    m = nx()
    poses = [np.eye(3)]
    for fn in (m.rotate_z, m.rotate_y, m.rotate_x):
        for deg in (90, 180, 270):
            poses.append(fn(np.deg2rad(deg)))
    return np.round(np.stack(poses)).astype(np.float32)


def gpu(layout, monkeypatch, remap, sl, w, b, n, rot, out_dtype=torch.complex64, wvol=True, cutoff=0.5, start=None,
        wstart=None):
    monkeypatch.setenv("NOA_HIP_RASTERIZE_LAYOUT", layout)
    wr = n // 2 + 1
    vol = dev(start) if start is not None else torch.zeros((1, n, n, wr), dtype=out_dtype, device=DEV)
    wv = None
    if wvol:
        wv = dev(wstart) if wstart is not None else torch.zeros((1, n, n, wr), dtype=torch.float32, device=DEV)
    opts = nx().RasterizeCentralSlicesOptions(fftfreq_cutoff=cutoff)
    nx().rasterize_central_slices_3d(remap, sl, w, (b, 1, n, n), vol, wv, (1, n, n, n), None, rot, opts)
    torch.cuda.synchronize()
    return vol.cpu().numpy(), (wv.cpu().numpy() if wvol else None)


def oracle(remap, sl, w, b, n, rot, out_dtype=np.complex64, wvol=True, cutoff=0.5, start=None, wstart=None):
    wr = n // 2 + 1
    vol = start.copy() if start is not None else np.zeros((1, n, n, wr), out_dtype)
    wv = None
    if wvol:
        wv = wstart.copy() if wstart is not None else np.zeros((1, n, n, wr), np.float32)
    O.rasterize_central_slices_3d(remap, sl, w, (b, 1, n, n), vol, wv, (1, n, n, n), None, rot,
                                  fftfreq_cutoff=cutoff, n_threads=8)
    return vol, wv


@pytest.mark.parametrize("remap,n,b,cutoff", [
    ("h2h", 64, 300, 0.5),
    ("hc2hc", 64, 300, 0.5),
    ("hc2h", 33, 120, 0.5),        # odd size
    ("h2hc", 48, 200, 0.31),       # smaller cutoff
    ("hc2hc", 128, 96, 0.5),
])
def test_tiles_match_the_oracle_and_the_direct_kernel(monkeypatch, remap, n, b, cutoff):
    print(f"case rasterize tiles {remap} n={n} b={b} cutoff={cutoff}", flush=True)
    rng = np.random.default_rng(1700 + n)
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)
    rot = random_rotations(b, rng)
    rot[: len(axis_poses())] = axis_poses()     # taps exactly on lattice points, tile faces and the x = 0 plane
    want_v, want_w = oracle(remap, sl, w, b, n, rot, cutoff=cutoff)
    got_v, got_w = gpu("tiles", monkeypatch, remap, dev(sl), dev(w), b, n, dev(rot), cutoff=cutoff)
    assert rel_l2(got_v, want_v) < 1e-6
    assert rel_l2(got_w, want_w) < 1e-6
    ref_v, ref_w = gpu("direct", monkeypatch, remap, dev(sl), dev(w), b, n, dev(rot), cutoff=cutoff)
    assert rel_l2(got_v, ref_v) < 1e-6
    assert rel_l2(got_w, ref_w) < 1e-6


def test_tiles_axis_poses_per_voxel(monkeypatch):
    """Axis poses only: every pixel lands on a lattice point (fraction 0), so each voxel receives whole values and
    the x = 0 duplicates carry the conjugate; checked per voxel, where a misplaced tap cannot hide in a norm."""
    print("case rasterize tiles axis poses", flush=True)
    rng = np.random.default_rng(1710)
    n = 32
    rot = np.concatenate([axis_poses()] * 7)                  # 70 slices: the tiled path by default
    b = rot.shape[0]
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)
    for remap in ("h2h", "hc2hc"):
        want_v, want_w = oracle(remap, sl, w, b, n, rot)
        got_v, got_w = gpu("tiles", monkeypatch, remap, dev(sl), dev(w), b, n, dev(rot))
        np.testing.assert_allclose(got_v, want_v, rtol=0, atol=2e-6 * np.abs(want_v).max())
        np.testing.assert_allclose(got_w, want_w, rtol=0, atol=2e-6 * np.abs(want_w).max())


@pytest.mark.parametrize("kind", ["real", "c2r", "value", "no_weights", "weight_value"])
def test_tiles_value_types_and_weights(monkeypatch, kind):
    print(f"case rasterize tiles {kind}", flush=True)
    rng = np.random.default_rng(1720)
    n, b = 48, 128
    rot = random_rotations(b, rng)
    spec = random_spectrum((b, 1, n, n // 2 + 1), rng)
    wts = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)
    out_np, out_t, wvol = np.complex64, torch.complex64, True
    sl_np, sl_dev, w_np, w_dev = spec, dev(spec), wts, dev(wts)
    if kind == "real":
        sl_np = rng.standard_normal((b, 1, n, n // 2 + 1)).astype(np.float32)
        sl_dev, out_np, out_t = dev(sl_np), np.float32, torch.float32
    elif kind == "c2r":
        out_np, out_t = np.float32, torch.float32
    elif kind == "value":
        sl_np = sl_dev = 2.5
        out_np, out_t = np.float32, torch.float32
    elif kind == "no_weights":
        w_np = w_dev = None
        wvol = False
    elif kind == "weight_value":
        w_np = w_dev = 0.75
    want_v, want_w = oracle("hc2hc", sl_np, w_np, b, n, rot, out_dtype=out_np, wvol=wvol)
    got_v, got_w = gpu("tiles", monkeypatch, "hc2hc", sl_dev, w_dev, b, n, dev(rot), out_dtype=out_t, wvol=wvol)
    assert rel_l2(got_v, want_v) < 1e-6
    if wvol:
        assert rel_l2(got_w, want_w) < 1e-6


def test_tiles_strided_slices_accumulate_and_keep_voxels_beyond_the_cutoff(monkeypatch):
    """Slices cut from wider rows, a volume that already holds values (rasterisation adds to it), and voxels that no
    tap reaches keep their value."""
    print("case rasterize tiles strided/accumulate", flush=True)
    rng = np.random.default_rng(1730)
    n, b = 64, 150
    wide = random_spectrum((b, 1, n, n // 2 + 7), rng)
    sl = wide[..., : n // 2 + 1]
    rot = random_rotations(b, rng)
    start = random_spectrum((1, n, n, n // 2 + 1), rng)
    wstart = rng.uniform(0, 1, (1, n, n, n // 2 + 1)).astype(np.float32)
    want_v, want_w = oracle("h2h", sl, None, b, n, rot, cutoff=0.4, start=start, wstart=wstart)
    gs = torch.from_numpy(wide).to(DEV)[..., : n // 2 + 1]
    got_v, got_w = gpu("tiles", monkeypatch, "h2h", gs, None, b, n, dev(rot), cutoff=0.4, start=start, wstart=wstart)
    assert rel_l2(got_v, want_v) < 1e-6
    assert rel_l2(got_w, want_w) < 1e-6
    untouched = (want_v == start) & (want_w == wstart)
    np.testing.assert_array_equal(got_v[untouched], start[untouched])


def test_tiles_with_a_degenerate_transform(monkeypatch):
    """A singular matrix collapses the plane: every pixel of that slice is a candidate of every tile, and each is
    still rasterised once."""
    print("case rasterize tiles degenerate transform", flush=True)
    rng = np.random.default_rng(1740)
    n, b = 32, 70
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    rot[3] = np.array([[0.3, 0.5, 0.5], [0.1, 0.5, 0.5], [0.2, -0.5, -0.5]], np.float32)
    rot[5] = np.zeros((3, 3), np.float32)
    want_v, want_w = oracle("h2h", sl, None, b, n, rot)
    got_v, got_w = gpu("tiles", monkeypatch, "h2h", dev(sl), None, b, n, dev(rot))
    assert rel_l2(got_v, want_v) < 1e-6
    assert rel_l2(got_w, want_w) < 1e-6


def test_default_selection_by_slice_count(monkeypatch):
    """Without the switch, 64 slices or more take the tiled path and fewer the direct kernel; both agree."""
    print("case rasterize default selection", flush=True)
    monkeypatch.delenv("NOA_HIP_RASTERIZE_LAYOUT", raising=False)
    rng = np.random.default_rng(1750)
    n = 64
    for b in (63, 64):
        sl = random_spectrum((b, 1, n, n // 2 + 1), rng)
        rot = random_rotations(b, rng)
        vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        nx().rasterize_central_slices_3d("hc2hc", dev(sl), None, (b, 1, n, n), vol, None, (1, n, n, n), None, dev(rot))
        torch.cuda.synchronize()
        want, _ = oracle("hc2hc", sl, None, b, n, rot, wvol=False)
        assert rel_l2(vol.cpu().numpy(), want) < 1e-6


@pytest.mark.parametrize("remap,n,b", [("h2h", 64, 200), ("hc2hc", 48, 96)])
def test_tiles_take_quaternions(monkeypatch, remap, n, b):
    """Quaternion<f32> rotations on the tiled path (the plane geometry from their matrix, noa's quaternion rotation
    per pixel): against the oracle with the same quaternions and against the direct kernel, rel-L2 1e-6."""
    print(f"case rasterize tiles quaternions {remap} n={n} b={b}", flush=True)
    rng = np.random.default_rng(1800 + n)
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)
    rot = random_rotations(b, rng)
    rot[: len(axis_poses())] = axis_poses()
    q = nx().quaternion_from_matrix(rot.astype(np.float64)).astype(np.float32)
    want_v = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    want_w = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.rasterize_central_slices_3d(remap, sl, w, (b, 1, n, n), want_v, want_w, (1, n, n, n), None, q, quaternion=True,
                                  n_threads=8)
    got_v, got_w = gpu("tiles", monkeypatch, remap, dev(sl), dev(w), b, n, nx().Quaternions(dev(q)))
    dir_v, dir_w = gpu("direct", monkeypatch, remap, dev(sl), dev(w), b, n, nx().Quaternions(dev(q)))
    assert rel_l2(got_v, want_v) < 1e-6 and rel_l2(got_w, want_w) < 1e-6
    assert rel_l2(got_v, dir_v) < 1e-6 and rel_l2(got_w, dir_w) < 1e-6


def _scalings(b, rng):
    """Anisotropic inverse scalings (magnification 0.9-1.1 along a random in-plane axis): R(t) diag(s) R(-t)."""
    t = rng.uniform(0, np.pi, b)
    s = rng.uniform(0.9, 1.1, (b, 2))
    c, d = np.cos(t), np.sin(t)
    rt = np.stack([np.stack([c, -d], -1), np.stack([d, c], -1)], -2)
    return np.ascontiguousarray(rt @ (s[:, :, None] * np.transpose(rt, (0, 2, 1)))).astype(np.float32)


@pytest.mark.parametrize("remap,n,b,quat", [("h2h", 64, 200, False), ("hc2hc", 48, 96, False), ("h2h", 64, 120, True)])
def test_tiles_take_inverse_scalings(monkeypatch, remap, n, b, quat):
    """Mat22 inverse scalings on the tiled path (the plane's pixel columns scaled, noa's general expression per
    pixel without Ewald curvature): against the oracle with the same scalings and against the direct kernel,
    rel-L2 1e-6."""
    print(f"case rasterize tiles scalings {remap} n={n} b={b} quat={quat}", flush=True)
    rng = np.random.default_rng(1900 + n + b)
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)
    rot = random_rotations(b, rng)
    rot[: len(axis_poses())] = axis_poses()
    scale = _scalings(b, rng)
    if quat:
        rot = nx().quaternion_from_matrix(rot.astype(np.float64)).astype(np.float32)
    want_v = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    want_w = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.rasterize_central_slices_3d(remap, sl, w, (b, 1, n, n), want_v, want_w, (1, n, n, n), scale, rot,
                                  quaternion=quat, n_threads=8)
    got, direct = {}, {}
    for layout, out in (("tiles", got), ("direct", direct)):
        monkeypatch.setenv("NOA_HIP_RASTERIZE_LAYOUT", layout)
        vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        wv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
        r = nx().Quaternions(dev(rot)) if quat else dev(rot)
        nx().rasterize_central_slices_3d(remap, dev(sl), dev(w), (b, 1, n, n), vol, wv, (1, n, n, n), dev(scale), r,
                                         nx().RasterizeCentralSlicesOptions())
        torch.cuda.synchronize()
        out["v"], out["w"] = vol.cpu().numpy(), wv.cpu().numpy()
    assert np.abs(want_v).max() > 0
    assert rel_l2(got["v"], want_v) < 1e-6 and rel_l2(got["w"], want_w) < 1e-6
    assert rel_l2(got["v"], direct["v"]) < 1e-6 and rel_l2(got["w"], direct["w"]) < 1e-6
