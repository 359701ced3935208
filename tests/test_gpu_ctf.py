"""ctf_anisotropic (and its fused weighted form), ctf_isotropic over 1, 2 and 3 dimensions, and both CTFs on full
layouts and strided arrays, on the GPU against the oracle.

Bar: the same single-precision expressions in the same order; atan2, cos, sin and exp differ by ulps between the
device and the host, and the phase reaches a few hundred radians at high frequency, so the CTF agrees to ~1e-5
absolute (|CTF| <= 1): rel-L2 <= 1e-5 and per element 1e-4. FP64 CTFs agree to 1e-12.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import random_spectrum, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


def sig():
    import noa_amd.signal as m
    return m


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def cpu(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _ctfs(b, f64=False):
    """Per-slice CTFs as in the CTF-weighted insertion workload: 300 kV, Cs 2.7 mm, amplitude 0.07, 1 A/px,
    defocus 0.5-3 um, astigmatism 0-0.1 um at angles 0-pi; one with a phase plate and B-factor."""
    rng = np.random.default_rng(60)
    out = []
    for i in range(b):
        extra = dict(phase_shift=0.4, bfactor=-30.0, scale=0.8) if i == 0 else {}
        out.append(sig().CTFAnisotropic((1.0, 1.05), (rng.uniform(0.5, 3.0), rng.uniform(0, 0.1), rng.uniform(0, np.pi)),
                                        300, 0.07, 2.7, dtype=np.float64 if f64 else np.float32, **extra))
    return out


def test_python_ctf_objects_equal_the_oracles():
    """The host-side constructor gives noa's members (compared with the oracle's restatement byte for byte, except
    k3 whose float atan may differ by an ulp between numpy and libm)."""
    for f64 in (False, True):
        c = sig().CTFAnisotropic((1.0, 1.05), (1.7, 0.05, 0.3), 300, 0.07, 2.7, 0.4, -30.0, 0.8,
                                 dtype=np.float64 if f64 else np.float32)
        ref = O.make_ctf((1.0, 1.05), (1.7, 0.05, 0.3), 300, 0.07, 2.7, 0.4, -30.0, 0.8, f64=f64)
        dt = np.float64 if f64 else np.float32
        a, b = np.frombuffer(c.to_bytes(), dt), np.frombuffer(ref, dt)
        np.testing.assert_array_equal(a[:-1], b[:-1])
        assert a[-1] == pytest.approx(b[-1], rel=2e-7)


CASES = [
    # remap, input dtype (None: write the CTF), output dtype, batched, abs, squared, f64 ctf, range stop
    ("h2h", None, np.float32, True, False, False, False, -1.0),
    ("hc2hc", None, np.float32, False, False, True, False, -1.0),
    ("h2hc", None, np.complex64, True, True, False, False, -1.0),
    ("hc2h", np.complex64, np.complex64, True, False, False, False, -1.0),
    ("h2h", np.complex64, np.float32, True, False, False, False, -1.0),
    ("hc2hc", np.float32, np.float32, True, False, True, False, 0.3),
    ("h2h", np.complex128, np.complex128, True, False, False, True, -1.0),
    ("hc2hc", None, np.float64, False, True, False, True, -1.0),
]


@pytest.mark.parametrize("remap,tin,tout,batched,ctf_abs,squared,f64,stop", CASES)
def test_ctf_matches_the_oracle(remap, tin, tout, batched, ctf_abs, squared, f64, stop):
    print(f"case ctf {remap} {tin}->{np.dtype(tout)} batched={batched} abs={ctf_abs} sq={squared} f64={f64}", flush=True)
    rng = np.random.default_rng(61)
    b, h, w = 5, 64, 60
    ctfs = _ctfs(b, f64)
    arg = ctfs if batched else ctfs[0]
    raw = [c.to_bytes() for c in ctfs] if batched else ctfs[0].to_bytes()
    x = None if tin is None else random_spectrum((b, 1, h, w // 2 + 1), rng, tin)
    want = np.zeros((b, 1, h, w // 2 + 1), tout)
    O.ctf_anisotropic_2d(remap, x, want, (b, 1, h, w), raw, fftfreq_range=(0.0, stop, True), ctf_abs=ctf_abs,
                         ctf_squared=squared)
    out = torch.zeros(want.shape, dtype=getattr(torch, np.dtype(tout).name), device=DEV)
    sig().ctf_anisotropic(remap, None if x is None else to_dev(x), out, (b, 1, h, w), arg,
                          sig().CTFOptions(fftfreq_range=(0.0, stop, True), ctf_abs=ctf_abs, ctf_squared=squared))
    got = cpu(out)
    tol = 1e-12 if f64 else 1e-5
    assert rel_l2(got, want) <= tol
    assert float(np.abs(got.astype(np.complex128) - want).max()) <= 10 * tol * max(1.0, float(np.abs(want).max()))


@pytest.mark.parametrize("remap", ["h2h", "hc2hc", "hc2h"])
def test_fused_weighted_equals_the_two_calls(remap):
    """The fused prologue == ctf_anisotropic(input) and ctf_anisotropic(None, ctf_squared) on the GPU, exactly."""
    print(f"case ctf weighted {remap}", flush=True)
    rng = np.random.default_rng(62)
    b, h, w = 7, 64, 64
    ctfs = _ctfs(b)
    x = to_dev(random_spectrum((b, 1, h, w // 2 + 1), rng))
    v1 = torch.empty_like(x)
    w1 = torch.empty((b, 1, h, w // 2 + 1), dtype=torch.float32, device=DEV)
    sig().ctf_anisotropic_weighted(remap, x, v1, w1, (b, 1, h, w), ctfs)
    v2 = torch.empty_like(x)
    w2 = torch.empty_like(w1)
    sig().ctf_anisotropic(remap, x, v2, (b, 1, h, w), ctfs)
    sig().ctf_anisotropic(remap, None, w2, (b, 1, h, w), ctfs, sig().CTFOptions(ctf_squared=True))
    torch.cuda.synchronize()
    assert torch.equal(v1, v2) and torch.equal(w1, w2)
    # And against the oracle's two calls.
    raw = [c.to_bytes() for c in ctfs]
    want_v = np.zeros((b, 1, h, w // 2 + 1), np.complex64)
    want_w = np.zeros((b, 1, h, w // 2 + 1), np.float32)
    O.ctf_anisotropic_2d(remap, cpu(x), want_v, (b, 1, h, w), raw)
    O.ctf_anisotropic_2d(remap, None, want_w, (b, 1, h, w), raw, ctf_squared=True)
    assert rel_l2(cpu(v1), want_v) <= 1e-5 and rel_l2(cpu(w1), want_w) <= 1e-5


def test_ctf_in_place_and_errors():
    b, h, w = 2, 32, 32
    ctfs = _ctfs(b)
    x = to_dev(random_spectrum((b, 1, h, w // 2 + 1), np.random.default_rng(63)))
    ref = torch.empty_like(x)
    sig().ctf_anisotropic("hc2hc", x, ref, (b, 1, h, w), ctfs)
    sig().ctf_anisotropic("hc2hc", x, x, (b, 1, h, w), ctfs)
    torch.cuda.synchronize()
    assert torch.equal(x, ref)
    with pytest.raises(sig().NoaError, match="in-place multiplication and a remapping"):
        sig().ctf_anisotropic("hc2h", x, x, (b, 1, h, w), ctfs)
    with pytest.raises(sig().NoaError, match="contiguous vector with 2 elements"):
        sig().ctf_anisotropic("h2h", x, ref, (b, 1, h, w), ctfs[:1] * 3)
    with pytest.raises(sig().NoaError, match="starting fftfreq should be 0"):
        sig().ctf_anisotropic("h2h", x, ref, (b, 1, h, w), ctfs, sig().CTFOptions(fftfreq_range=(0.1, 0.5, True)))


# ---------------------------------------------------------------------------------------------------------------
# ctf_isotropic over 1, 2 and 3 dimensions and the full layouts of both CTFs (the general operator).
def _iso(b, f64=False):
    rng = np.random.default_rng(64)
    out = []
    for i in range(b):
        extra = dict(phase_shift=0.3, bfactor=-20.0, scale=1.2) if i == 0 else {}
        out.append(sig().CTFIsotropic(rng.uniform(0.9, 1.4), rng.uniform(0.5, 3.0), 300, 0.07, 2.7,
                                      dtype=np.float64 if f64 else np.float32, **extra))
    return out


def test_python_isotropic_ctf_objects_equal_the_oracles():
    for f64 in (False, True):
        c = sig().CTFIsotropic(1.1, 1.7, 200, 0.1, 2.0, 0.4, -30.0, 0.8, dtype=np.float64 if f64 else np.float32)
        ref = O.make_ctf_isotropic(1.1, 1.7, 200, 0.1, 2.0, 0.4, -30.0, 0.8, f64=f64)
        dt = np.float64 if f64 else np.float32
        a, b = np.frombuffer(c.to_bytes(), dt), np.frombuffer(ref, dt)
        np.testing.assert_array_equal(a[:-1], b[:-1])
        assert a[-1] == pytest.approx(b[-1], rel=2e-7)
    # noa's conversion from an anisotropic CTF: mean pixel size, no astigmatism.
    an = sig().CTFAnisotropic((1.0, 1.2), (1.7, 0.05, 0.3), 300, 0.07, 2.7, 0.4, -30.0, 0.8)
    iso = sig().CTFIsotropic.from_anisotropic(an)
    assert iso.pixel_size == np.float32(1.1) and iso.defocus == pytest.approx(an.defocus_value, rel=1e-6)


def _out_shape(shape, remap):
    return tuple(shape[:3]) + ((shape[3] // 2 + 1) if remap.startswith("h") else shape[3],)


ISO_CASES = [
    # shape, remap, input dtype, output dtype, batched, abs, squared, f64, (start, stop)
    ((3, 1, 1, 64), "h2h", None, np.float32, True, False, False, False, (0.0, -1.0)),
    ((3, 1, 1, 64), "h2h", None, np.float64, True, False, False, True, (0.1, 0.4)),
    ((2, 1, 1, 35), "fc2fc", None, np.float32, False, True, False, False, (0.0, -1.0)),
    ((2, 1, 40, 1), "h2h", np.float32, np.float32, True, False, False, False, (0.0, -1.0)),
    ((4, 1, 48, 40), "hc2hc", None, np.float32, True, False, True, False, (0.0, -1.0)),
    ((4, 1, 48, 40), "hc2h", np.complex64, np.complex64, True, False, False, False, (0.0, 0.3)),
    ((2, 1, 33, 36), "fc2f", np.complex64, np.float32, False, False, False, False, (0.0, -1.0)),
    ((2, 1, 32, 30), "f2fc", None, np.complex64, True, False, False, False, (0.0, -1.0)),
    ((2, 24, 20, 18), "h2h", None, np.float32, True, False, False, False, (0.0, -1.0)),
    ((2, 15, 16, 17), "fc2fc", np.complex128, np.complex128, True, False, False, True, (0.0, -1.0)),
    ((1, 9, 10, 11), "hc2h", np.complex64, np.float32, False, False, False, False, (0.0, -1.0)),
]


@pytest.mark.parametrize("shape,remap,tin,tout,batched,ctf_abs,squared,f64,rng_", ISO_CASES)
def test_ctf_isotropic_matches_the_oracle(shape, remap, tin, tout, batched, ctf_abs, squared, f64, rng_):
    print(f"case ctf_isotropic {shape} {remap} {tin}->{np.dtype(tout)} batched={batched} f64={f64}", flush=True)
    rng = np.random.default_rng(65)
    b = shape[0]
    ctfs = _iso(b, f64)
    arg = ctfs if batched else ctfs[0]
    raw = [c.to_bytes() for c in ctfs] if batched else ctfs[0].to_bytes()
    oshape = _out_shape(shape, remap)
    x = None if tin is None else random_spectrum(oshape, rng, tin)
    want = np.zeros(oshape, tout)
    O.ctf(remap, x, want, shape, raw, fftfreq_range=(rng_[0], rng_[1], True), ctf_abs=ctf_abs, ctf_squared=squared)
    out = torch.zeros(oshape, dtype=getattr(torch, np.dtype(tout).name), device=DEV)
    sig().ctf_isotropic(remap, None if x is None else to_dev(x), out, shape, arg,
                        sig().CTFOptions(fftfreq_range=(rng_[0], rng_[1], True), ctf_abs=ctf_abs, ctf_squared=squared))
    got = cpu(out)
    tol = 1e-12 if f64 else 1e-5
    assert rel_l2(got, want) <= tol
    assert float(np.abs(got.astype(np.complex128) - want).max()) <= 10 * tol * max(1.0, float(np.abs(want).max()))


def test_ctf_isotropic_strided_broadcast_and_in_place():
    """A column-major full output (the dimensions reordered as noa does), a batch-broadcast input, a transposed
    input, and the product in place, against the oracle on the same views."""
    print("case ctf_isotropic strides", flush=True)
    rng = np.random.default_rng(66)
    b, d, h, w = 2, 12, 14, 16
    ctfs = _iso(b, True)
    raw = [c.to_bytes() for c in ctfs]
    want = np.zeros((w, h, d, b), np.float64).transpose(3, 2, 1, 0)
    O.ctf("f2f", None, want, (b, d, h, w), raw)
    out = torch.zeros((w, h, d, b), dtype=torch.float64, device=DEV).permute(3, 2, 1, 0)
    sig().ctf_isotropic("f2f", None, out, (b, d, h, w), ctfs)
    assert rel_l2(cpu(out), want) <= 1e-12
    x = random_spectrum((1, d, h, w // 2 + 1), rng, np.complex128)
    want2 = np.zeros((b, d, h, w // 2 + 1), np.complex128)
    O.ctf("hc2hc", x, want2, (b, d, h, w), raw)
    out2 = torch.zeros(want2.shape, dtype=torch.complex128, device=DEV)
    sig().ctf_isotropic("hc2hc", to_dev(x), out2, (b, d, h, w), ctfs)
    assert rel_l2(cpu(out2), want2) <= 1e-12
    xt = to_dev(np.ascontiguousarray(np.transpose(random_spectrum((b, w // 2 + 1, h, d), rng, np.complex128), (0, 3, 2, 1))))
    xt = xt.permute(0, 3, 2, 1).contiguous().permute(0, 3, 2, 1)   # strides of a (B, W', H, D) array
    want3 = np.zeros((b, d, h, w // 2 + 1), np.complex128)
    O.ctf("h2h", cpu(xt), want3, (b, d, h, w), raw)
    out3 = torch.zeros(want3.shape, dtype=torch.complex128, device=DEV)
    sig().ctf_isotropic("h2h", xt, out3, (b, d, h, w), ctfs)
    assert rel_l2(cpu(out3), want3) <= 1e-12
    y = to_dev(want3.copy())
    sig().ctf_isotropic("h2h", y, y, (b, d, h, w), ctfs)
    O.ctf("h2h", want3.copy(), want3, (b, d, h, w), raw)
    assert rel_l2(cpu(y), want3) <= 1e-12
    with pytest.raises(sig().NoaError, match="innermost"):
        bad = torch.zeros((w // 2 + 1, h, d, b), dtype=torch.float32, device=DEV).permute(3, 2, 1, 0)
        sig().ctf_isotropic("h2h", None, bad, (b, d, h, w), _iso(b))


@pytest.mark.parametrize("remap,tin,tout,colmajor", [
    ("f2f", None, np.float32, False), ("fc2fc", None, np.float32, True), ("fc2f", np.complex64, np.complex64, False),
    ("f2fc", np.complex64, np.float32, False),
])
def test_ctf_anisotropic_full_layouts_match_the_oracle(remap, tin, tout, colmajor):
    print(f"case ctf_anisotropic full {remap} {tin}->{np.dtype(tout)} colmajor={colmajor}", flush=True)
    rng = np.random.default_rng(67)
    b, h, w = 3, 40, 36
    ctfs = _ctfs(b)
    raw = [c.to_bytes() for c in ctfs]
    x = None if tin is None else random_spectrum((b, 1, h, w), rng, tin)
    if colmajor:
        want = np.zeros((w, h, 1, b), tout).transpose(3, 2, 1, 0)
        out = torch.zeros((w, h, 1, b), dtype=getattr(torch, np.dtype(tout).name), device=DEV).permute(3, 2, 1, 0)
    else:
        want = np.zeros((b, 1, h, w), tout)
        out = torch.zeros(want.shape, dtype=getattr(torch, np.dtype(tout).name), device=DEV)
    O.ctf(remap, x, want, (b, 1, h, w), raw, isotropic=False)
    sig().ctf_anisotropic(remap, None if x is None else to_dev(x), out, (b, 1, h, w), ctfs)
    got = cpu(out)
    assert rel_l2(got, want) <= 1e-5
    # The corners of a full spectrum reach |f| = 0.71 (0.5 in a half one): twice the phase, twice the ulp spread.
    assert float(np.abs(got.astype(np.complex128) - want).max()) <= 3e-4 * max(1.0, float(np.abs(want).max()))
