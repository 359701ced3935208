"""The tiled extraction (volume tiles staged in LDS, each visited by the slices whose plane crosses it).

It serves many slices from a cubic volume with cutoff <= 0.5 and Mat33<f32> rotations; NOA_HIP_EXTRACT_LAYOUT=tiles
forces it, =linear forces the row-major gather. Bars: bit-identical to the gather and to the oracle without shifts;
with the fused phase shift bit-identical to the gather and within 3e-7 rel-L2 of the oracle (sin/cos of two
libraries). Every pixel must be written: outputs start as NaN.
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


def run(layout, monkeypatch, remap, vol, n, rot, b, shifts=None, cutoff=0.5, row_pad=0):
    monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", layout)
    wr = n // 2 + 1
    full = torch.full((b, 1, n, wr + row_pad), float("nan"), dtype=torch.complex64, device=DEV)
    out = full[..., :wr]
    opts = nx().ExtractCentralSlicesOptions(fftfreq_cutoff=cutoff)
    if shifts is None:
        nx().extract_central_slices_3d(remap, vol, None, (1, n, n, n), out, None, (b, 1, n, n), None, rot, opts)
    else:
        nx().extract_central_slices_3d_shifted(remap, vol, None, (1, n, n, n), out, None, (b, 1, n, n), None, rot,
                                               shifts=shifts, options=opts)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def oracle(remap, vol, n, rot, b, shifts=None, cutoff=0.5):
    want = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
    O.extract_central_slices_3d(remap, vol, None, (1, n, n, n), want, None, (b, 1, n, n), None, rot,
                                fftfreq_cutoff=cutoff, n_threads=8)
    if shifts is not None:
        out_layout = "hc2hc" if remap.endswith("hc") else "h2h"
        O.phase_shift_2d(out_layout, want.copy(), want, (b, 1, n, n), shifts)
    return want


def axis_poses():
    """Identity and the 90/180/270-degree rotations about each axis: samples land exactly on lattice points and on
    tile boundaries."""
    m = nx()
    poses = [np.eye(3)]
    for fn in (m.rotate_z, m.rotate_y, m.rotate_x):
        for deg in (90, 180, 270):
            poses.append(fn(np.deg2rad(deg)))
    return np.round(np.stack(poses)).astype(np.float32)


@pytest.mark.parametrize("remap,n,b,cutoff,row_pad", [
    ("hc2hc", 64, 300, 0.5, 0),
    ("h2h", 48, 200, 0.5, 3),      # padded output rows
    ("hc2h", 33, 150, 0.5, 0),     # odd size
    ("h2hc", 64, 120, 0.31, 0),    # smaller cutoff: more pixels for the outside kernel
    ("h2h", 128, 64, 0.5, 0),
])
def test_tiles_equal_the_gather_and_the_oracle(monkeypatch, remap, n, b, cutoff, row_pad):
    print(f"case tiles {remap} n={n} b={b} cutoff={cutoff} pad={row_pad}", flush=True)
    rng = np.random.default_rng(1200 + n)
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    rot[: len(axis_poses())] = axis_poses()
    got = run("tiles", monkeypatch, remap, dev(vol), n, dev(rot), b, cutoff=cutoff, row_pad=row_pad)
    ref = run("linear", monkeypatch, remap, dev(vol), n, dev(rot), b, cutoff=cutoff, row_pad=row_pad)
    assert not np.isnan(got).any(), "pixels left unwritten"
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(got, oracle(remap, vol, n, rot, b, cutoff=cutoff))


@pytest.mark.parametrize("kind", ["batched", "broadcast", "half"])
def test_tiles_with_the_fused_phase_shift(monkeypatch, kind):
    print(f"case tiles shifted {kind}", flush=True)
    rng = np.random.default_rng(1300)
    n, b = 64, 160
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    if kind == "batched":
        sh = rng.uniform(-5, 5, (b, 2)).astype(np.float32)
        arg = dev(sh)
    elif kind == "broadcast":
        sh = np.array([1.75, -3.25], np.float32)
        arg = sh
    else:
        sh = np.array([n / 2, n / 2], np.float32)
        arg = sh
    got = run("tiles", monkeypatch, "hc2hc", dev(vol), n, dev(rot), b, shifts=arg)
    ref = run("linear", monkeypatch, "hc2hc", dev(vol), n, dev(rot), b, shifts=arg)
    assert not np.isnan(got).any()
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    # A single shift is passed as one (the half-shape one takes noa's exact +-1 pattern), batched ones per slice.
    want = oracle("hc2hc", vol, n, rot, b, shifts=sh)
    assert rel_l2(got, want) < 3e-7


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_tiles_from_16bit_storage(monkeypatch, storage):
    """FP16/BF16 pairs are converted to FP32 when the tile is staged: the same bits as the gather's conversion."""
    print(f"case tiles storage {storage}", flush=True)
    rng = np.random.default_rng(1400)
    n, b = 64, 200
    vol = dev(random_spectrum((1, n, n, n // 2 + 1), rng))
    v = vol.to(torch.complex32) if storage == "fp16" else nx().ComplexBF16.from_complex(vol)
    rot = dev(random_rotations(b, rng))
    got = run("tiles", monkeypatch, "hc2hc", v, n, rot, b)
    ref = run("linear", monkeypatch, "hc2hc", v, n, rot, b)
    assert not np.isnan(got).any()
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_tiles_with_a_degenerate_transform(monkeypatch):
    """A singular matrix collapses the plane onto a line: every pixel of such a slice is a candidate of every tile,
    and the result still equals the gather."""
    print("case tiles degenerate transform", flush=True)
    rng = np.random.default_rng(1500)
    n, b = 32, 6
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    rot[2] = np.array([[0.3, 0.5, 0.5], [0.1, 0.5, 0.5], [0.2, -0.5, -0.5]], np.float32)   # equal 2nd and 3rd columns
    rot[4] = np.zeros((3, 3), np.float32)
    got = run("tiles", monkeypatch, "h2h", dev(vol), n, dev(rot), b)
    ref = run("linear", monkeypatch, "h2h", dev(vol), n, dev(rot), b)
    assert not np.isnan(got).any()
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(got, oracle("h2h", vol, n, rot, b))


def test_tuning_table_selects_tiles(monkeypatch):
    """The launch tuning table's layout 3 selects the tiled path (the autotuner's candidate): same bits as forcing
    it, and as the default (block-copy) gather."""
    print("case tiles through the tuning table", flush=True)
    from noa_amd import autotune
    rng = np.random.default_rng(1600)
    n, b = 64, 16 * 64
    vol = dev(random_spectrum((1, n, n, n // 2 + 1), rng))
    rot = dev(random_rotations(b, rng))
    forced = run("tiles", monkeypatch, "hc2hc", vol, n, rot, b)
    monkeypatch.delenv("NOA_HIP_EXTRACT_LAYOUT")
    outs = []
    try:
        for values in ((3, 64, 0, 0), (0, 0, 0, 0)):
            autotune.apply([autotune.Entry(op="extract", box=n, values=values)])
            out = torch.full((b, 1, n, n // 2 + 1), float("nan"), dtype=torch.complex64, device=DEV)
            nx().extract_central_slices_3d("hc2hc", vol, None, (1, n, n, n), out, None, (b, 1, n, n), None, rot)
            torch.cuda.synchronize()
            outs.append(out.cpu().numpy())
    finally:
        autotune.apply([])
    for o in outs:
        np.testing.assert_array_equal(o.view(np.uint32), forced.view(np.uint32))
