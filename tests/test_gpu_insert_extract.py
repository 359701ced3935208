"""insert_and_extract_central_slices_3d on the GPU against the oracle, and the reference's own invariants.

Bar: the GPU sums the same terms in noa's slice order with the same single-precision expressions (coordinates,
cutoff and window tests are bit-identical); only sin/cos in the windowed sinc differ by ulps, so the outputs agree
to rel-L2 <= 1e-6 and per element to 1e-4 of the largest magnitude. A slice dropped by the tile culling at a
single pixel would break the per-element bound.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import allclose_abs_safe, random_rotations, random_spectrum, rel_l2

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


def close(got, want, rel=1e-6, elem=1e-4):
    scale = max(float(np.abs(want).max()), 1e-30)
    assert rel_l2(got, want) <= rel, rel_l2(got, want)
    assert float(np.abs(got.astype(np.complex128) - want).max()) <= elem * scale


def tilt_rotations(n, step_deg, offset_deg, sign, dtype=np.float32):
    """The reference tests' poses: rotations about y by sign * (i * step + offset) degrees (ZYX Euler)."""
    e = nx().euler2matrix
    return np.stack([e(np.deg2rad([0.0, sign * (i * step_deg + offset_deg), 0.0]), axes="zyx", dtype=dtype)
                     for i in range(n)]).astype(dtype)


CASES = [
    # remap, in dtype, out dtype, weights, interp, extra
    ("hc2hc", np.complex64, np.complex64, "none", 1, ""),
    ("h2h", np.complex64, np.complex64, "array", 1, ""),
    ("hc2h", np.complex64, np.complex64, "value", 1, ""),
    ("h2hc", np.complex64, np.complex64, "array", 1, "add"),
    ("hc2hc", np.complex64, np.float32, "array", 1, ""),
    ("h2h", np.float32, np.float32, "array", 1, "correct"),
    ("hc2hc", np.complex64, np.complex64, "array", 0, ""),
    ("hc2hc", np.complex128, np.complex128, "array", 1, ""),
    ("hc2hc", np.complex64, np.complex64, "array", 1, "quaternion"),
    ("hc2hc", np.complex64, np.complex64, "array", 1, "scale"),
    ("hc2hc", np.complex64, np.complex64, "array", 1, "ews"),
    ("hc2hc", np.complex64, np.complex64, "array", 1, "w-window"),
    ("h2h", np.complex64, np.complex64, "array", 1, "w-window add"),
    ("hc2hc", np.complex64, np.complex64, "none", 1, "value-input"),
    ("hc2hc", np.complex64, np.complex64, "array", 1, "single-output-pose"),
]


@pytest.mark.parametrize("remap,tin,tout,weights,interp,extra", CASES)
def test_insert_extract_matches_the_oracle(remap, tin, tout, weights, interp, extra):
    print(f"case insert_extract {remap} {np.dtype(tin)}->{np.dtype(tout)} w={weights} interp={interp} {extra}",
          flush=True)
    rng = np.random.default_rng(200)
    n, b_in, b_out = 40, 70, 6
    f64 = tin == np.complex128
    rdt = np.float64 if f64 else np.float32
    wdt = np.float64 if f64 else np.float32
    slices = random_spectrum((b_in, 1, n, n // 2 + 1), rng, tin)
    inp = complex(0.5, -1.5) if "value-input" in extra else slices
    w_in = {"none": None, "array": rng.uniform(0.1, 2, (b_in, 1, n, n // 2 + 1)).astype(wdt),
            "value": 0.75}[weights]
    rot_in = random_rotations(b_in, rng, dtype=rdt)
    rot_out = random_rotations(b_out, rng, dtype=rdt)
    if "single-output-pose" in extra:
        rot_out = rot_out[0]
    kw = dict(interp=interp, input_fftfreq_sinc=0.05, input_fftfreq_blackman=0.08, fftfreq_cutoff=0.45)
    opts = dict(interp=interp, input_windowed_sinc=nx().WindowedSinc(0.05, 0.08), fftfreq_cutoff=0.45)
    s_in = s_out = None
    if "scale" in extra:
        s_in = np.array([[1.02, 0.01], [-0.02, 0.97]], rdt)
        s_out = np.stack([np.array([[0.98, 0.0], [0.01, 1.03]], rdt)] * b_out)
    if "ews" in extra:
        kw["ews_radius"] = opts["ews_radius"] = (60.0, 55.0)
    if "w-window" in extra:
        kw.update(w_fftfreq_sinc=0.02, w_fftfreq_blackman=0.06)
        opts["w_windowed_sinc"] = nx().WindowedSinc(0.02, 0.06)
    if "add" in extra:
        kw["add_to_output"] = opts["add_to_output"] = True
    if "correct" in extra:
        kw["correct_weights"] = opts["correct_weights"] = True
    quat = "quaternion" in extra
    rin_arg = nx().quaternion_from_matrix(rot_in).astype(rdt) if quat else rot_in

    start = random_spectrum((b_out, 1, n, n // 2 + 1), rng, tout)
    want = start.copy()
    want_w = np.full((b_out, 1, n, n // 2 + 1), 0.25, wdt)
    O.insert_and_extract_central_slices_3d(remap, inp, w_in, (b_in, 1, n, n), want, want_w, (b_out, 1, n, n),
                                           s_in, rin_arg, None if s_out is None else s_out, rot_out,
                                           input_quaternion=quat, **kw)
    out = to_dev(start)
    out_w = torch.full((b_out, 1, n, n // 2 + 1), 0.25, dtype=getattr(torch, np.dtype(wdt).name), device=DEV)
    rin_dev = nx().Quaternions(to_dev(rin_arg)) if quat else to_dev(rot_in)
    rout_dev = rot_out if rot_out.ndim == 2 else to_dev(rot_out)
    nx().insert_and_extract_central_slices_3d(
        remap, inp if isinstance(inp, complex) else to_dev(slices),
        w_in if (w_in is None or isinstance(w_in, float)) else to_dev(w_in), (b_in, 1, n, n), out, out_w,
        (b_out, 1, n, n), s_in, rin_dev, None if s_out is None else to_dev(s_out), rout_dev,
        nx().InsertAndExtractCentralSlicesOptions(**opts))
    got, got_w = cpu(out), cpu(out_w)
    tol = 1e-12 if f64 else 1e-6
    close(got, want, rel=tol, elem=tol * 100)
    close(got_w, want_w, rel=tol, elem=tol * 100)


def test_culling_keeps_every_contribution_with_the_default_window():
    """Many slices and the narrowest window (1/n): the tile culling must not drop any pixel's contribution."""
    print("case insert_extract culling n=64 b_in=300", flush=True)
    rng = np.random.default_rng(201)
    n, b_in, b_out = 64, 300, 8
    slices = random_spectrum((b_in, 1, n, n // 2 + 1), rng)
    rot_in = random_rotations(b_in, rng)
    rot_out = random_rotations(b_out, rng)
    want = np.zeros((b_out, 1, n, n // 2 + 1), np.complex64)
    want_w = np.zeros((b_out, 1, n, n // 2 + 1), np.float32)
    O.insert_and_extract_central_slices_3d("hc2hc", slices, None, (b_in, 1, n, n), want, want_w, (b_out, 1, n, n),
                                           None, rot_in, None, rot_out)
    out = torch.zeros((b_out, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    out_w = torch.zeros((b_out, 1, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    nx().insert_and_extract_central_slices_3d("hc2hc", to_dev(slices), None, (b_in, 1, n, n), out, out_w,
                                              (b_out, 1, n, n), None, to_dev(rot_in), None, to_dev(rot_out))
    got, got_w = cpu(out), cpu(out_w)
    # The weight is the number of contributions (times their windows): a dropped slice shows here first.
    assert np.count_nonzero(want_w) > 1000
    np.testing.assert_allclose(got_w, want_w, rtol=1e-5, atol=1e-6)
    close(got, want)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
def test_reference_texture_remap_invariant(dtype):
    """TestXformFourierInsertInterpolateExtract.cpp:89-139: hc2hc from a Texture == hc2h followed by h2hc remap."""
    print(f"case insert_extract texture/remap {np.dtype(dtype)}", flush=True)
    x = nx()
    b_in, b_out, n = 20, 5, 256
    rin = tilt_rotations(b_in, 2, 0, -1)
    rout = tilt_rotations(b_out, 2, 1, 1)
    count = b_in * n * (n // 2 + 1)
    lin = np.linspace(-50, 50, count).reshape(b_in, 1, n, n // 2 + 1)   # noa::linspace, endpoint included
    data = (lin + 1j * lin).astype(dtype) if np.dtype(dtype).kind == "c" else lin.astype(dtype)
    inp = to_dev(data)
    tdt = getattr(torch, np.dtype(dtype).name)
    out0 = torch.empty((b_out, 1, n, n // 2 + 1), dtype=tdt, device=DEV)
    out1 = torch.empty_like(out0)
    out2 = torch.empty_like(out0)
    opts = x.InsertAndExtractCentralSlicesOptions(input_windowed_sinc=x.WindowedSinc(0.01), fftfreq_cutoff=0.8)
    x.insert_and_extract_central_slices_3d("hc2hc", x.Texture(inp, x.Interp.LINEAR), None, (b_in, 1, n, n), out0,
                                           None, (b_out, 1, n, n), None, to_dev(rin), None, to_dev(rout), opts)
    x.insert_and_extract_central_slices_3d("hc2h", inp, None, (b_in, 1, n, n), out1, None, (b_out, 1, n, n), None,
                                           to_dev(rin), None, to_dev(rout), opts)
    x.remap("h2hc", out1, out2, (b_out, 1, n, n))
    assert allclose_abs_safe(cpu(out0), cpu(out2), 5e-5)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reference_weights_invariant(dtype):
    """TestXformFourierInsertInterpolateExtract.cpp:141-182: with input weights equal to the (real) input values,
    the output weights equal the output values."""
    print(f"case insert_extract weights {np.dtype(dtype)}", flush=True)
    x = nx()
    rng = np.random.default_rng(202)
    b_in, b_out, n = 20, 5, 256
    rin = tilt_rotations(b_in, 2, 0, -1)
    rout = tilt_rotations(b_out, 2, 1, 1)
    data = rng.uniform(-10, 10, (b_in, 1, n, n // 2 + 1)).astype(dtype)
    tdt = getattr(torch, np.dtype(dtype).name)
    out0 = torch.empty((b_out, 1, n, n // 2 + 1), dtype=tdt, device=DEV)
    out1 = torch.empty_like(out0)
    opts = x.InsertAndExtractCentralSlicesOptions(input_windowed_sinc=x.WindowedSinc(0.01), fftfreq_cutoff=0.8)
    x.insert_and_extract_central_slices_3d("hc2hc", to_dev(data), to_dev(data.copy()), (b_in, 1, n, n), out0, out1,
                                           (b_out, 1, n, n), None, to_dev(rin), None, to_dev(rout), opts)
    assert allclose_abs_safe(cpu(out0), cpu(out1), 5e-5)


def test_reference_matrix_equals_quaternion():
    """TestXformFourierInsertInterpolateExtract.cpp:231-283: Mat33 and Quaternion input rotations give the same
    slices (compared below 0.49 cycle/pixel, as the reference's low-pass). The two round the rotated coordinates
    differently, so the bound is 5e-5 of the spectrum's largest magnitude (the oracle pin explains why)."""
    print("case insert_extract matrix/quaternion", flush=True)
    x = nx()
    rng = np.random.default_rng(203)
    n = 256
    m = x.euler2matrix(np.deg2rad([45.0, 0.0, 0.0]), axes="zyx", dtype=np.float32)[None]
    q = x.quaternion_from_matrix(m).astype(np.float32)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    real = rng.uniform(-50, 50, (n, n))
    spec = np.fft.fftshift(np.fft.rfft2(real), axes=0).astype(np.complex64)[None, None]   # h2hc of r2c
    rout = np.eye(3, dtype=np.float32)[None]
    opts = x.InsertAndExtractCentralSlicesOptions(input_windowed_sinc=x.WindowedSinc(0.01), fftfreq_cutoff=0.8)
    out0 = torch.empty((1, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    out1 = torch.empty_like(out0)
    x.insert_and_extract_central_slices_3d("hc2hc", to_dev(spec), None, (1, 1, n, n), out0, None, (1, 1, n, n), None,
                                           to_dev(m), None, to_dev(rout), opts)
    x.insert_and_extract_central_slices_3d("hc2hc", to_dev(spec), None, (1, 1, n, n), out1, None, (1, 1, n, n), None,
                                           x.Quaternions(to_dev(q)), None, to_dev(rout), opts)
    fy = (np.arange(n) - n // 2)[:, None] / n
    fx = np.arange(n // 2 + 1)[None, :] / n
    keep = np.sqrt(fy ** 2 + fx ** 2) <= 0.49
    a, b = cpu(out0)[0, 0], cpu(out1)[0, 0]
    a0, b0 = np.where(keep, a, 0), np.where(keep, b, 0)
    assert np.abs(a0 - b0).max() <= 5e-5 * np.abs(a0).max()
    assert rel_l2(b0, a0) < 1e-5
