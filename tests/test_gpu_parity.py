"""Parity of the HIP path (noa_amd.xform -> C ABI -> gfx950 kernels) with the CPU oracle on the same inputs.

Bars: extraction (LINEAR/NEAREST, no w-window) is bit-exact with the oracle (no transcendental functions,
no fused multiply-add on either side). Interpolated insertion sums in the reference's slice order and differs
only through sin/cos ulps: rel-L2 <= 1e-6. Rasterisation accumulates with float atomics in an unspecified
order: rel-L2 <= 1e-6 against the oracle's serial order (the bound on the atomic-order nondeterminism).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import h2hc, mat33_blocks, random_rotations, random_spectrum, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


def nx():
    import noa_amd.xform as m
    return m


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def cpu(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# ------------------------------------------------------------------------------------------------ extraction
EXTRACT_CASES = [
    # remap, in dtype, out dtype, interp
    ("h2h", np.complex64, np.complex64, 1),
    ("hc2hc", np.complex64, np.complex64, 1),
    ("hc2h", np.complex64, np.complex64, 1),
    ("h2hc", np.complex64, np.complex64, 1),
    ("hc2hc", np.complex64, np.float32, 1),
    ("h2h", np.float32, np.float32, 1),
    ("hc2hc", np.complex64, np.complex64, 0),
    ("h2h", np.complex128, np.complex128, 1),
]


@pytest.mark.parametrize("remap,tin,tout,interp", EXTRACT_CASES)
def test_extract_bit_exact(remap, tin, tout, interp):
    print(f"case extract {remap} {np.dtype(tin)}->{np.dtype(tout)} interp={interp}", flush=True)
    rng = np.random.default_rng(100)
    n, b = 48, 12
    vol = random_spectrum((1, n, n, n // 2 + 1), rng, tin)
    rdt = np.float64 if tin == np.complex128 else np.float32
    rot = random_rotations(b, rng, dtype=rdt)
    want = np.zeros((b, 1, n, n // 2 + 1), dtype=tout)
    O.extract_central_slices_3d(remap, vol, None, (1, n, n, n), want, None, (b, 1, n, n), None, rot, interp=interp)
    out = torch.zeros((b, 1, n, n // 2 + 1), dtype=getattr(torch, np.dtype(tout).name), device=DEV)
    nx().extract_central_slices_3d(remap, to_dev(vol), None, (1, n, n, n), out, None, (b, 1, n, n), None, to_dev(rot),
                                   nx().ExtractCentralSlicesOptions(interp=nx().Interp(interp)))
    np.testing.assert_array_equal(cpu(out), want)


def test_extract_weights_scaling_ews_target_quaternions():
    print("case extract weights/scaling/ews/target/quaternion", flush=True)
    rng = np.random.default_rng(101)
    n, b = 40, 9
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    wvol = rng.uniform(0, 2, (1, n, n, n // 2 + 1)).astype(np.float32)
    rot = random_rotations(b, rng)
    scale = np.stack([np.diag(rng.uniform(0.9, 1.1, 2)).astype(np.float32) for _ in range(b)])
    target = (1, 44, 44, 44)
    q = nx().quaternion_from_matrix(rot.astype(np.float64)).astype(np.float32)
    for rotation, oracle_rot, quat in ((to_dev(rot), rot, False), (nx().Quaternions(to_dev(q)), q, True)):
        want = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
        wwant = np.zeros((b, 1, n, n // 2 + 1), np.float32)
        O.extract_central_slices_3d("hc2h", vol, wvol, (1, n, n, n), want, wwant, (b, 1, n, n), scale, oracle_rot,
                                    fftfreq_cutoff=0.45, target_shape=target, ews_radius=(60.0, 70.0), quaternion=quat)
        out = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        wout = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.float32, device=DEV)
        opts = nx().ExtractCentralSlicesOptions(fftfreq_cutoff=0.45, target_shape=target, ews_radius=(60.0, 70.0))
        nx().extract_central_slices_3d("hc2h", to_dev(vol), to_dev(wvol), (1, n, n, n), out, wout, (b, 1, n, n),
                                       to_dev(scale), rotation, opts)
        np.testing.assert_array_equal(cpu(out), want)
        np.testing.assert_array_equal(cpu(wout), wwant)


def test_extract_w_windowed_sinc():
    print("case extract w-windowed sinc", flush=True)
    rng = np.random.default_rng(102)
    n, b = 32, 6
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    want = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((b, 1, n, n // 2 + 1), np.float32)
    O.extract_central_slices_3d("hc2hc", vol, None, (1, n, n, n), want, wwant, (b, 1, n, n), None, rot,
                                fftfreq_sinc=0.02, fftfreq_blackman=0.1)
    out = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    wout = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    opts = nx().ExtractCentralSlicesOptions(w_windowed_sinc=nx().WindowedSinc(0.02, 0.1))
    nx().extract_central_slices_3d("hc2hc", to_dev(vol), None, (1, n, n, n), out, wout, (b, 1, n, n), None,
                                   to_dev(rot), opts)
    assert rel_l2(cpu(out), want) < 1e-6
    assert rel_l2(cpu(wout), wwant) < 1e-6


def test_extract_odd_sizes_strided_output_single_rotation():
    print("case extract odd/strided/single", flush=True)
    rng = np.random.default_rng(103)
    n, h = 33, 35
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(1, rng)[0]
    want = np.zeros((4, 1, h, h // 2 + 1), np.complex64)
    O.extract_central_slices_3d("hc2hc", vol, None, (1, n, n, n), want, None, (4, 1, h, h), None, rot,
                                fftfreq_cutoff=0.7)
    big = torch.zeros((4, 1, h, h // 2 + 1 + 5), dtype=torch.complex64, device=DEV)
    out = big[..., : h // 2 + 1]
    nx().extract_central_slices_3d("hc2hc", to_dev(vol), None, (1, n, n, n), out, None, (4, 1, h, h), None, rot,
                                   nx().ExtractCentralSlicesOptions(fftfreq_cutoff=0.7))
    np.testing.assert_array_equal(cpu(out), want)


@pytest.mark.parametrize("remap,n,h,shifts", [("hc2hc", 48, 48, False), ("h2h", 33, 35, True), ("hc2h", 31, 27, False),
                                              ("h2hc", 24, 50, True)])
def test_extract_volume_layouts(monkeypatch, remap, n, h, shifts):
    """The c32 extraction gathers from a 4x2x4-block copy of the volume when many slices are extracted; forced both
    ways here: the same bits in either layout, equal to the oracle (bit-exact without shifts, within the sin/cos
    rounding of the fused phase shift with them), odd and even sizes, slices larger or smaller than the volume."""
    print(f"case extract layouts {remap} n={n} h={h} shifts={shifts}", flush=True)
    rng = np.random.default_rng(104)
    b = 20
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    sh = rng.uniform(-4, 4, (b, 2)).astype(np.float32)
    want = np.zeros((b, 1, h, h // 2 + 1), np.complex64)
    O.extract_central_slices_3d(remap, vol, None, (1, n, n, n), want, None, (b, 1, h, h), None, rot,
                                fftfreq_cutoff=0.6)
    if shifts:
        out_layout = "hc2hc" if remap.endswith("hc") else "h2h"
        O.phase_shift_2d(out_layout, want.copy(), want, (b, 1, h, h), sh)
    outs = {}
    for layout in ("bricks", "linear"):
        monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", layout)
        out = torch.zeros((b, 1, h, h // 2 + 1), dtype=torch.complex64, device=DEV)
        opts = nx().ExtractCentralSlicesOptions(fftfreq_cutoff=0.6)
        if shifts:
            nx().extract_central_slices_3d_shifted(remap, to_dev(vol), None, (1, n, n, n), out, None, (b, 1, h, h),
                                                   None, to_dev(rot), shifts=to_dev(sh), options=opts)
        else:
            nx().extract_central_slices_3d(remap, to_dev(vol), None, (1, n, n, n), out, None, (b, 1, h, h), None,
                                           to_dev(rot), opts)
        outs[layout] = cpu(out)
    np.testing.assert_array_equal(outs["bricks"], outs["linear"])
    if shifts:
        assert rel_l2(outs["bricks"], want) < 3e-7
    else:
        np.testing.assert_array_equal(outs["bricks"], want)


# ------------------------------------------------------------------------------------------------ rasterisation
@pytest.mark.parametrize("remap", ["h2h", "hc2hc", "hc2h", "h2hc"])
def test_rasterize_parity(remap):
    print(f"case rasterize {remap}", flush=True)
    rng = np.random.default_rng(110)
    n, b = 48, 24
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    rot = random_rotations(b, rng)
    vol = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wvol = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.rasterize_central_slices_3d(remap, s, w, (b, 1, n, n), vol, wvol, (1, n, n, n), None, rot)
    gv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    gw = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    nx().rasterize_central_slices_3d(remap, to_dev(s), to_dev(w), (b, 1, n, n), gv, gw, (1, n, n, n), None,
                                     to_dev(rot))
    assert rel_l2(cpu(gv), vol) < 1e-6
    assert rel_l2(cpu(gw), wvol) < 1e-6


def test_rasterize_real_value_input_and_f64():
    print("case rasterize value/f64", flush=True)
    rng = np.random.default_rng(111)
    n, b = 32, 10
    rot = random_rotations(b, rng, dtype=np.float64)
    vol = np.zeros((1, n, n, n // 2 + 1), np.float64)
    O.rasterize_central_slices_3d("hc2hc", 2.5, None, (b, 1, n, n), vol, None, (1, n, n, n), None, rot)
    gv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float64, device=DEV)
    nx().rasterize_central_slices_3d("hc2hc", 2.5, None, (b, 1, n, n), gv, None, (1, n, n, n), None, to_dev(rot))
    assert rel_l2(cpu(gv), vol) < 1e-12


# ------------------------------------------------------------------------------------------------ insertion
@pytest.mark.parametrize("remap", ["h2h", "hc2hc", "hc2h", "h2hc"])
def test_insert_parity(remap):
    print(f"case insert {remap}", flush=True)
    rng = np.random.default_rng(120)
    n, b = 40, 30
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)     # insertion accumulates: start from non-zero
    wvol = rng.uniform(0, 1, (1, n, n, n // 2 + 1)).astype(np.float32)
    gv, gw = to_dev(vol), to_dev(wvol)
    O.insert_central_slices_3d(remap, s, w, (b, 1, n, n), vol, wvol, (1, n, n, n), None, inv,
                               fftfreq_sinc=0.03, fftfreq_blackman=0.06)
    opts = nx().InsertCentralSlicesOptions(windowed_sinc=nx().WindowedSinc(0.03, 0.06))
    nx().insert_central_slices_3d(remap, to_dev(s), to_dev(w), (b, 1, n, n), gv, gw, (1, n, n, n), None, to_dev(inv),
                                  opts)
    assert rel_l2(cpu(gv), vol) < 1e-6
    assert rel_l2(cpu(gw), wvol) < 1e-6


@pytest.mark.parametrize("kernel", ["wave", "slab"])
def test_insert_work_orders_give_the_same_bits(monkeypatch, kernel):
    """The work plan's item order (0: expected work, 1: first slice, 2: XCD-aware Morton runs padded with no-op items)
    only changes when each item runs: every voxel's sum, and the split sub-bricks' reduction, are the same bits."""
    print(f"case insert work orders {kernel}", flush=True)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", kernel)
    monkeypatch.setenv("NOA_HIP_INSERT_LIST", "32")            # many split sub-bricks
    rng = np.random.default_rng(122)
    n, b = 64, 400
    s = to_dev(random_spectrum((b, 1, n, n // 2 + 1), rng))
    w = to_dev(rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32))
    inv = to_dev(np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1))))
    outs = []
    for order in ("0", "1", "2"):
        monkeypatch.setenv("NOA_HIP_INSERT_ORDER", order)
        gv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        gw = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
        nx().insert_central_slices_3d("h2h", s, w, (b, 1, n, n), gv, gw, (1, n, n, n), None, inv)
        outs.append((cpu(gv), cpu(gw)))
    for v, wt in outs[1:]:
        np.testing.assert_array_equal(v, outs[0][0])
        np.testing.assert_array_equal(wt, outs[0][1])


def test_insert_defaults_value_input_single_slice_broadcast():
    print("case insert value/broadcast", flush=True)
    rng = np.random.default_rng(121)
    n = 32
    inv = np.ascontiguousarray(np.transpose(random_rotations(7, rng), (0, 2, 1)))
    s = random_spectrum((1, 1, n, n // 2 + 1), rng)   # one slice, broadcast over 7 rotations
    vol = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wvol = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d("hc2hc", s, 1.0, (7, 1, n, n), vol, wvol, (1, n, n, n), None, inv)
    # noa broadcasts a B=1 input over the batch only through the interpolator (shape must still match).
    s7 = np.broadcast_to(s, (7, 1, n, n // 2 + 1))
    gv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    gw = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    nx().insert_central_slices_3d("hc2hc", to_dev(s7), 1.0, (7, 1, n, n), gv, gw, (1, n, n, n), None, to_dev(inv))
    vol7 = np.zeros_like(vol)
    wvol7 = np.zeros_like(wvol)
    O.insert_central_slices_3d("hc2hc", np.ascontiguousarray(s7), 1.0, (7, 1, n, n), vol7, wvol7, (1, n, n, n), None, inv)
    assert rel_l2(cpu(gv), vol7) < 1e-6
    assert rel_l2(cpu(gw), wvol7) < 1e-6
    assert rel_l2(vol7, vol) < 1e-6


# ------------------------------------------------------------------------------------------------ errors
def test_invalid_arguments_raise_before_launch():
    n = 16
    vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    out = torch.zeros((2, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    rot = torch.eye(3, device=DEV).repeat(2, 1, 1)
    with pytest.raises(nx().NoaError, match="shape"):
        nx().extract_central_slices_3d("h2h", vol, None, (1, n, n, n), out, None, (3, 1, n, n), None, rot)
    with pytest.raises(nx().NoaError, match="n_slices"):
        nx().extract_central_slices_3d("h2h", vol, None, (1, n, n, n), out, None, (2, 1, n, n), None,
                                       torch.eye(3, device=DEV).repeat(3, 1, 1))
    with pytest.raises(nx().NoaError, match="HX2HX"):
        nx().extract_central_slices_3d("f2f", vol, None, (1, n, n, n), out, None, (2, 1, n, n), None, rot)
    with pytest.raises(nx().NoaError, match="overlap"):
        nx().rasterize_central_slices_3d("h2h", vol[:, :1, :, :].reshape(1, 1, n, n // 2 + 1).expand(1, 1, n, n // 2 + 1),
                                         None, (1, 1, n, n), vol, None, (1, n, n, n), None, rot[:1])
    with pytest.raises(nx().NoaError, match="same device"):
        nx().extract_central_slices_3d("h2h", vol.cpu(), None, (1, n, n, n), out, None, (2, 1, n, n), None, rot)


# ------------------------------------------------------------------------------------------------ insertion paths
def _insert_dev(remap, s, w, shape, n, inv, scale=None, opts=None, out_dtype=torch.complex64, quat=False):
    vol = torch.zeros((1, n, n, n // 2 + 1), dtype=out_dtype, device=DEV)
    wvol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    rot = nx().Quaternions(to_dev(inv)) if quat else to_dev(inv)
    nx().insert_central_slices_3d(remap, to_dev(s), None if w is None else to_dev(w), shape, vol, wvol,
                                  (1, n, n, n), scale, rot, opts)
    return cpu(vol), cpu(wvol)


@pytest.mark.parametrize("remap,interp,fb", [("h2h", 1, -1.0), ("hc2hc", 1, 0.05), ("hc2h", 0, -1.0),
                                             ("h2hc", 1, 0.02)])
def test_insert_culled_equals_general_path(monkeypatch, remap, interp, fb):
    # NOA_HIP_INSERT_KERNEL=general selects the general (one lane per voxel, every slice) kernel; by default the
    # culled one-wave kernel runs (48 is not a power of two), also with a batched identity scaling. All sum the same
    # terms in the same order.
    print(f"case insert culled-vs-general {remap} interp={interp} fb={fb}", flush=True)
    rng = np.random.default_rng(130)
    n, b = 48, 40
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    opts = nx().InsertCentralSlicesOptions(interp=nx().Interp(interp), windowed_sinc=nx().WindowedSinc(-1, fb))
    eye = torch.eye(2, device=DEV).repeat(b, 1, 1)
    v1, w1 = _insert_dev(remap, s, w, (b, 1, n, n), n, inv, None, opts)
    v3, w3 = _insert_dev(remap, s, w, (b, 1, n, n), n, inv, eye, opts)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", "general")
    v2, w2 = _insert_dev(remap, s, w, (b, 1, n, n), n, inv, None, opts)
    assert rel_l2(v1, v2) < 1e-6 and rel_l2(w1, w2) < 1e-6
    np.testing.assert_array_equal(v3, v1)
    np.testing.assert_array_equal(w3, w1)
    want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d(remap, s, w, (b, 1, n, n), want, wwant, (1, n, n, n), None, inv, interp=interp,
                               fftfreq_blackman=fb)
    assert rel_l2(v1, want) < 1e-6 and rel_l2(w1, wwant) < 1e-6


def test_insert_culled_quaternion_real_and_abs2_outputs():
    print("case insert culled quaternion / f32 / c32->f32", flush=True)
    rng = np.random.default_rng(131)
    n, b = 40, 25
    fwd = random_rotations(b, rng, dtype=np.float64)
    inv = np.ascontiguousarray(np.transpose(fwd, (0, 2, 1)))
    q = nx().quaternion_from_matrix(inv).astype(np.float32)
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    O.insert_central_slices_3d("hc2hc", s, None, (b, 1, n, n), want, None, (1, n, n, n), None, q, quaternion=True)
    got, _ = _insert_dev("hc2hc", s, None, (b, 1, n, n), n, q, quat=True)
    assert rel_l2(got, want) < 1e-6
    # complex slices into a real volume: |z|^2 of the complex sum (cast_or_abs_squared)
    want_r = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d("h2h", s, None, (b, 1, n, n), want_r, None, (1, n, n, n), None, inv.astype(np.float32))
    got_r, _ = _insert_dev("h2h", s, None, (b, 1, n, n), n, inv.astype(np.float32), out_dtype=torch.float32)
    assert rel_l2(got_r, want_r) < 1e-6
    sr = rng.standard_normal((b, 1, n, n // 2 + 1)).astype(np.float32)
    want_f = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d("h2h", sr, None, (b, 1, n, n), want_f, None, (1, n, n, n), None, inv.astype(np.float32))
    got_f, _ = _insert_dev("h2h", sr, None, (b, 1, n, n), n, inv.astype(np.float32), out_dtype=torch.float32)
    assert rel_l2(got_f, want_f) < 1e-6


@pytest.mark.parametrize("remap", ["hc2hc", "h2h"])
def test_insert_many_slices_split_bricks(remap):
    # 1500 slices into 32^3: the bricks near the origin get several slice segments whose partial sums are
    # reduced in segment order; compared with the oracle's strict slice order and with itself (determinism).
    # Power-of-two size, Mat33 rotations and weights in and out: the compile-time configurations of the kernel.
    print(f"case insert split bricks {remap}", flush=True)
    rng = np.random.default_rng(133)
    n, b = 32, 1500
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d(remap, s, w, (b, 1, n, n), want, wwant, (1, n, n, n), None, inv)
    v1, w1 = _insert_dev(remap, s, w, (b, 1, n, n), n, inv)
    v2, w2 = _insert_dev(remap, s, w, (b, 1, n, n), n, inv)
    assert rel_l2(v1, want) < 1e-6 and rel_l2(w1, wwant) < 1e-6
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(w1, w2)


@pytest.mark.parametrize("kernel", ["slab", "wave"])
def test_insert_rows_that_are_not_unit_vectors(monkeypatch, kernel):
    # Scaled matrices (rows of norm 0.45 and 0.8): the slab |pz| <= f_b is thicker in samples than for a rotation, so
    # a column of a sub-brick holds more than four candidates and the slab kernel packs a slice's candidates in
    # several passes. Against the oracle, which evaluates noa's per-voxel expressions with the matrices as given.
    print(f"case insert scaled matrices {kernel}", flush=True)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", kernel)
    rng = np.random.default_rng(134)
    n, b = 64, 60
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    scale = np.where(np.arange(b) % 2 == 0, 0.45, 0.8).astype(np.float32)[:, None, None]
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)) * scale).astype(np.float32)
    want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d("h2h", s, w, (b, 1, n, n), want, wwant, (1, n, n, n), None, inv)
    v, wv = _insert_dev("h2h", s, w, (b, 1, n, n), n, inv)
    assert rel_l2(v, want) < 1e-6 and rel_l2(wv, wwant) < 1e-6


@pytest.mark.parametrize("kernel", ["slab", "wave"])
def test_insert_quaternions_on_both_culled_kernels(monkeypatch, kernel):
    # A power-of-two size: the slab kernel takes Quaternion<f32> through the quaternions' matrices (coordinates equal
    # to noa's quaternion rotation up to rounding), the one-wave kernel with noa's quaternion expressions.
    print(f"case insert quaternions {kernel}", flush=True)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", kernel)
    rng = np.random.default_rng(135)
    n, b = 64, 300
    fwd = random_rotations(b, rng, dtype=np.float64)
    inv = np.ascontiguousarray(np.transpose(fwd, (0, 2, 1)))
    q = nx().quaternion_from_matrix(inv).astype(np.float32)
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d("hc2hc", s, w, (b, 1, n, n), want, wwant, (1, n, n, n), None, q, quaternion=True,
                               n_threads=8)
    got, gotw = _insert_dev("hc2hc", s, w, (b, 1, n, n), n, q, quat=True)
    if kernel == "wave":
        assert rel_l2(got, want) < 1e-6 and rel_l2(gotw, wwant) < 1e-6
        return
    # The slab kernel equals the oracle fed the quaternions' float32 matrices; the oracle's own quaternion and matrix
    # paths for one rotation differ by rounding (rel-L2 4.2e-6 for these inputs, on the CPU), hence 1e-5 against it.
    mat = _quaternion_matrices(q)
    mwant = np.zeros_like(want)
    mwwant = np.zeros_like(wwant)
    O.insert_central_slices_3d("hc2hc", s, w, (b, 1, n, n), mwant, mwwant, (1, n, n, n), None, mat, n_threads=8)
    assert rel_l2(got, mwant) < 1e-6 and rel_l2(gotw, mwwant) < 1e-6
    assert rel_l2(got, want) < 1e-5 and rel_l2(gotw, wwant) < 1e-5


def _quaternion_matrices(q):
    """Rotation matrices of (z, y, x, w) unit quaternions in float32, in the order of operations of the kernels'
    conversion; the rotation itself is checked against the quaternion oracle."""
    f = np.float32
    z, y, x, w = (q[:, i].astype(f) for i in range(4))
    tx, ty, tz = f(2) * x, f(2) * y, f(2) * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz, tyy, tyz, tzz = tx * x, ty * x, tz * x, ty * y, tz * y, tz * z
    m = np.stack([f(1) - (txx + tyy), tyz + twx, txz - twy, tyz - twx, f(1) - (txx + tzz), txy + twz,
                  txz + twy, txy - twz, f(1) - (tyy + tzz)], 1)
    return np.ascontiguousarray(m.reshape(-1, 3, 3).astype(f))


def _scalings(b, rng):
    """Anisotropic inverse scalings (0.9-1.1 along a random in-plane axis): R(t) diag(s) R(-t)."""
    t = rng.uniform(0, np.pi, b)
    sv = rng.uniform(0.9, 1.1, (b, 2))
    c, d = np.cos(t), np.sin(t)
    rt = np.stack([np.stack([c, -d], -1), np.stack([d, c], -1)], -2)
    return np.ascontiguousarray(rt @ (sv[:, :, None] * np.transpose(rt, (0, 2, 1)))).astype(np.float32)


@pytest.mark.parametrize("remap,n,b,interp", [("h2h", 64, 200, 1), ("hc2hc", 48, 120, 1), ("hc2h", 40, 80, 0)])
def test_insert_inverse_scalings_on_the_culled_kernel(monkeypatch, remap, n, b, interp):
    # Mat22 inverse scalings: the one-wave kernel (fz and its culling are those of the rotation alone; the 2D
    # frequency is scaled) against the oracle and against the general kernel, rel-L2 1e-6.
    print(f"case insert scalings {remap} n={n} b={b} interp={interp}", flush=True)
    rng = np.random.default_rng(1700 + n)
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    scale = _scalings(b, rng)
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    opts = nx().InsertCentralSlicesOptions(interp=nx().Interp(interp))
    want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d(remap, s, w, (b, 1, n, n), want, wwant, (1, n, n, n), scale, inv, interp=interp,
                               n_threads=8)
    got, gotw = _insert_dev(remap, s, w, (b, 1, n, n), n, inv, torch.from_numpy(scale).to(DEV), opts)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", "general")
    gen, genw = _insert_dev(remap, s, w, (b, 1, n, n), n, inv, torch.from_numpy(scale).to(DEV), opts)
    assert np.abs(want).max() > 0
    assert rel_l2(got, want) < 1e-6 and rel_l2(gotw, wwant) < 1e-6
    assert rel_l2(got, gen) < 1e-6 and rel_l2(gotw, genw) < 1e-6


def test_insert_oversampled_and_odd_sizes():
    # Slices larger than the volume (footprints exceed the LDS staging: global-read fallback), odd volume size,
    # explicit target shape and a cutoff above 0.5.
    print("case insert oversampled/odd", flush=True)
    rng = np.random.default_rng(132)
    n, h, b = 33, 64, 12
    s = random_spectrum((b, 1, h, h // 2 + 1), rng)
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    for target, cutoff in (((0, 0, 0, 0), 0.5), ((1, 40, 40, 40), 0.6)):
        want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
        O.insert_central_slices_3d("hc2hc", s, None, (b, 1, h, h), want, None, (1, n, n, n), None, inv,
                                   fftfreq_cutoff=cutoff, target_shape=target)
        vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        opts = nx().InsertCentralSlicesOptions(fftfreq_cutoff=cutoff, target_shape=target)
        nx().insert_central_slices_3d("hc2hc", to_dev(s), None, (b, 1, h, h), vol, None, (1, n, n, n), None,
                                      to_dev(inv), opts)
        assert rel_l2(cpu(vol), want) < 1e-6


# ------------------------------------------------------------------------------------------------ windows > 2
# CUBIC (2), CUBIC_BSPLINE (3), LANCZOS4/6/8 (4, 5, 6): per-tap flip, 4-8 taps per dimension. CUBIC and the
# B-spline are plain arithmetic (bit-identical to the oracle); Lanczos evaluates sin/cos (ulps apart).
@pytest.mark.parametrize("interp", [2, 3, 4, 5, 6, 102])
@pytest.mark.parametrize("remap", ["hc2hc", "h2h"])
def test_extract_window_methods(interp, remap):
    print(f"case extract window interp={interp} {remap}", flush=True)
    rng = np.random.default_rng(140)
    n, b = 24, 4
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    rot = random_rotations(b, rng)
    want = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
    O.extract_central_slices_3d(remap, vol, None, (1, n, n, n), want, None, (b, 1, n, n), None, rot, interp=interp)
    out = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    opts = nx().ExtractCentralSlicesOptions(interp=nx().Interp(interp))
    nx().extract_central_slices_3d(remap, to_dev(vol), None, (1, n, n, n), out, None, (b, 1, n, n), None,
                                   to_dev(rot), opts)
    got = cpu(out)
    if interp % 100 in (2, 3):
        np.testing.assert_array_equal(got, want)
    else:
        assert rel_l2(got, want) < 1e-6


def test_extract_window_methods_weights_w_window_f64():
    print("case extract window cubic + weights + w-window / f64", flush=True)
    rng = np.random.default_rng(141)
    n, b = 20, 3
    vol = random_spectrum((1, n, n, n // 2 + 1), rng)
    wvol = rng.uniform(0.5, 2, (1, n, n, n // 2 + 1)).astype(np.float32)
    rot = random_rotations(b, rng)
    want = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((b, 1, n, n // 2 + 1), np.float32)
    O.extract_central_slices_3d("hc2hc", vol, wvol, (1, n, n, n), want, wwant, (b, 1, n, n), None, rot, interp=2,
                                fftfreq_sinc=0.1, fftfreq_blackman=0.15)
    out = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    wout = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    opts = nx().ExtractCentralSlicesOptions(interp=nx().Interp(2), w_windowed_sinc=nx().WindowedSinc(0.1, 0.15))
    nx().extract_central_slices_3d("hc2hc", to_dev(vol), to_dev(wvol), (1, n, n, n), out, wout, (b, 1, n, n), None,
                                   to_dev(rot), opts)
    assert rel_l2(cpu(out), want) < 1e-6 and rel_l2(cpu(wout), wwant) < 1e-6
    vol64 = vol.astype(np.complex128)
    rot64 = rot.astype(np.float64)
    want64 = np.zeros((b, 1, n, n // 2 + 1), np.complex128)
    O.extract_central_slices_3d("h2h", vol64, None, (1, n, n, n), want64, None, (b, 1, n, n), None, rot64, interp=3)
    out64 = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.complex128, device=DEV)
    nx().extract_central_slices_3d("h2h", to_dev(vol64), None, (1, n, n, n), out64, None, (b, 1, n, n), None,
                                   to_dev(rot64), nx().ExtractCentralSlicesOptions(interp=nx().Interp(3)))
    np.testing.assert_array_equal(cpu(out64), want64)


@pytest.mark.parametrize("interp", [2, 3, 5])
def test_insert_window_methods(interp):
    print(f"case insert window interp={interp}", flush=True)
    rng = np.random.default_rng(142)
    n, b = 24, 10
    s = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.5, 1.5, (b, 1, n, n // 2 + 1)).astype(np.float32)
    inv = np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))
    want = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wwant = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d("hc2h", s, w, (b, 1, n, n), want, wwant, (1, n, n, n), None, inv, interp=interp,
                               fftfreq_sinc=0.05, fftfreq_blackman=0.08)
    opts = nx().InsertCentralSlicesOptions(interp=nx().Interp(interp), windowed_sinc=nx().WindowedSinc(0.05, 0.08))
    v, wv = _insert_dev("hc2h", s, w, (b, 1, n, n), n, inv, None, opts)
    assert rel_l2(v, want) < 1e-6 and rel_l2(wv, wwant) < 1e-6


@pytest.mark.parametrize("interp", [2, 3, 6])
def test_insert_extract_window_methods(interp):
    print(f"case insert_extract window interp={interp}", flush=True)
    rng = np.random.default_rng(143)
    n, b_in, b_out = 24, 12, 3
    s = random_spectrum((b_in, 1, n, n // 2 + 1), rng)
    rin = random_rotations(b_in, rng)
    rout = random_rotations(b_out, rng)
    want = np.zeros((b_out, 1, n, n // 2 + 1), np.complex64)
    O.insert_and_extract_central_slices_3d("hc2hc", s, None, (b_in, 1, n, n), want, None, (b_out, 1, n, n), None,
                                           rin, None, rout, interp=interp, input_fftfreq_sinc=0.05,
                                           input_fftfreq_blackman=0.1)
    out = torch.zeros((b_out, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    opts = nx().InsertAndExtractCentralSlicesOptions(interp=nx().Interp(interp),
                                                     input_windowed_sinc=nx().WindowedSinc(0.05, 0.1))
    nx().insert_and_extract_central_slices_3d("hc2hc", to_dev(s), None, (b_in, 1, n, n), out, None, (b_out, 1, n, n),
                                              None, to_dev(rin), None, to_dev(rout), opts)
    assert rel_l2(cpu(out), want) < 1e-6
