"""noa::fft r2c / c2r / c2c on the GPU (rocFFT) against the oracle (the DFT with noa's normalisation), and noa's own
FFT tests (tests/noa/fft/TestFFTTransform.cpp): round trips on padded subregions, in place, and the plan cache.

Bar: FFT libraries round differently, so the comparison is to tolerance: rel-L2 <= 2e-6 (f32) and 1e-12 (f64) of the
spectrum's norm, and round trips within noa's own test bounds (1e-5 / 1e-9 absolute for values in [-5, 5]).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"
TOL = {torch.float32: 2e-6, torch.float64: 1e-12}
CX = {torch.float32: torch.complex64, torch.float64: torch.complex128}


def F():
    import noa_amd.fft as m
    return m


def _cpu(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _rand(shape, dtype, seed, lo=-5.0, hi=5.0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return (torch.rand(shape, generator=g, dtype=torch.float64) * (hi - lo) + lo).to(dtype)


SHAPES = [(3, 1, 1, 64), (2, 1, 1, 81), (2, 1, 64, 48), (1, 1, 45, 30), (1, 32, 24, 20), (2, 9, 10, 11)]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_c2r_match_the_oracle(shape, dtype):
    print(f"fft r2c/c2r {shape} {dtype}", flush=True)
    x = _rand(shape, dtype, sum(shape))
    xd = x.to(DEV)
    X = F().r2c(xd)
    Xn = _cpu(X)
    assert np.array_equal(_cpu(xd), x.numpy()), "r2c changed its input"
    ref = O.fft_r2c(x.numpy())
    assert rel_l2(Xn, ref) <= TOL[dtype]
    # c2r of the oracle's spectrum (its input is overwritten for 2D/3D, so give it a copy).
    spec = torch.from_numpy(ref.astype(np.complex64 if dtype == torch.float32 else np.complex128)).to(DEV)
    y = F().c2r(spec, shape)
    assert rel_l2(_cpu(y), O.fft_c2r(ref, shape)) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("shape", [(2, 1, 1, 81), (1, 1, 45, 30), (2, 9, 10, 11), (1, 1, 24, 1)])
def test_c2c_matches_the_oracle_both_signs(shape, dtype):
    print(f"fft c2c {shape} {dtype}", flush=True)
    x = torch.complex(_rand(shape, dtype, 3), _rand(shape, dtype, 4))
    xd = x.to(DEV)
    for sign in (F().Sign.FORWARD, F().Sign.BACKWARD):
        y = F().c2c(xd, sign=sign)
        assert np.array_equal(_cpu(xd), x.numpy()), "c2c changed its input"
        assert rel_l2(_cpu(y), O.fft_c2c(x.numpy(), int(sign))) <= TOL[dtype]


@pytest.mark.parametrize("norm", ["forward", "ortho", "backward", "none"])
def test_norms(norm):
    """Every norm on the forward and the backward transform, against the oracle's scaling."""
    print(f"fft norm {norm}", flush=True)
    shape = (2, 1, 32, 40)
    x = _rand(shape, torch.float64, 9)
    X = F().r2c(x.to(DEV), norm=norm)
    ref = O.fft_r2c(x.numpy(), norm)
    assert rel_l2(_cpu(X), ref) <= 1e-12
    y = F().c2r(torch.from_numpy(ref).to(DEV), shape, norm=norm)
    assert rel_l2(_cpu(y), O.fft_c2r(ref, shape, norm)) <= 1e-12
    xc = torch.complex(x, x.flip(-1)).to(DEV)
    for sign in (-1, 1):
        yc = F().c2c(xc, sign=sign, norm=norm)
        assert rel_l2(_cpu(yc), O.fft_c2c(xc.cpu().numpy(), sign, norm)) <= 1e-12


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_round_trips_like_noas_tests(ndim, dtype):
    """TestFFTTransform.cpp:14-67 and :131-177: out of place on padded subregions (strided views), in place on
    nf::empty's aliased arrays, c2c in place; default norm; |error| <= 1e-5 (f32) / 1e-9 (f64)."""
    print(f"fft round trips ndim={ndim} {dtype}", flush=True)
    eps = 1e-5 if dtype == torch.float32 else 1e-9
    sub = {1: (3, 1, 1, 60), 2: (2, 1, 48, 50), 3: (2, 20, 24, 30)}[ndim]
    pad = (0, 11 if ndim == 3 else 0, 12 if ndim >= 2 else 0, 13)
    full = tuple(s + p for s, p in zip(sub, pad))
    # Out of place, padded subregions.
    buf = _rand(full, dtype, 1).to(DEV)
    expected = buf[:, :sub[1], :sub[2], :sub[3]]
    fbuf = torch.empty((*full[:3], full[3] // 2 + 1), dtype=CX[dtype], device=DEV)
    fft = fbuf[:, :sub[1], :sub[2], :sub[3] // 2 + 1]
    F().r2c(expected, fft)
    result = F().c2r(fft, expected.shape)
    assert (result - expected).abs().max().item() <= eps
    # In place.
    real, cplx = F().empty(full, dtype=dtype, device=DEV)
    real.copy_(_rand(full, dtype, 2).to(DEV))
    before = real.clone()
    F().r2c(real, cplx)
    F().c2r(cplx, real)
    assert (real - before).abs().max().item() <= eps
    # c2c in place.
    z = torch.complex(_rand(full, dtype, 5), _rand(full, dtype, 6)).to(DEV)
    z0 = z.clone()
    F().c2c(z, z, sign=F().Sign.FORWARD)
    F().c2c(z, z, sign=F().Sign.BACKWARD)
    assert (z - z0).abs().max().item() <= eps


def test_plan_cache_like_noas_tests():
    """TestFFTTransform.cpp:179-210: the cache grows with new transforms, not with repeated ones, and clears;
    plus the limit evicting the oldest plans and plan_only computing nothing."""
    print("fft plan cache", flush=True)
    f = F()
    f.clear_cache()
    assert f.set_cache_limit(5) == 0 and f.cache_limit() == 5
    a = torch.zeros(64, device=DEV)
    f.r2c(a)
    a_s = f.cache_size()
    assert a_s > 0
    b = torch.zeros(128, device=DEV)
    f.r2c(b)
    assert f.cache_size() >= a_s
    assert f.clear_cache() >= 2 and f.cache_size() == 0
    c = torch.zeros((2, 1, 1024, 1024), device=DEV)
    f.r2c(c)
    c_s = f.cache_size()
    d = torch.zeros((2, 1, 1024, 1024), device=DEV)
    f.r2c(d)
    assert f.cache_size() == c_s
    for n in (16, 32, 48, 80, 96, 112):
        f.r2c(torch.zeros(n, device=DEV))
    assert f.cache_size() == 5
    assert f.set_cache_limit(2) == 3 and f.cache_size() == 2
    # plan_only: the plan is cached, the output untouched.
    f.clear_cache()
    out = torch.full((1, 1, 1, 33), 7 + 0j, dtype=torch.complex64, device=DEV)
    f.r2c(torch.ones(64, device=DEV), out.view(33), plan_only=True)
    torch.cuda.synchronize()
    assert f.cache_size() == 1 and bool((out == 7).all())
    f.set_cache_limit(4)
    f.clear_cache()


def test_shared_workspace():
    """TestFFTTransform.cpp:212-270: plans recorded with record_and_share_workspace get one shared work buffer, either
    from the next ordinary transform or from set_workspace; the results are the ordinary ones."""
    print("fft shared workspace", flush=True)
    f = F()
    f.clear_cache()
    f.set_cache_limit(50)
    shape = (6, 1, 384, 384)
    a = _rand(shape, torch.float32, 11).to(DEV)
    ref = f.r2c(a, cache_plan=False)
    f.clear_cache()
    outs = torch.empty((*shape[:3], shape[3] // 2 + 1), dtype=torch.complex64, device=DEV)
    for i in range(shape[0]):
        f.r2c(a[i:i + 1], outs[i:i + 1], record_and_share_workspace=True)
    left = f.workspace_left_to_allocate()
    assert left >= 0
    if left > 0:
        ws = torch.empty(left, dtype=torch.uint8, device=DEV)
        assert f.set_workspace(torch.empty(max(1, left // 2), dtype=torch.uint8, device=DEV)) == 0
        assert f.set_workspace(ws) >= 1
        assert f.workspace_left_to_allocate() == 0
    for i in range(shape[0]):
        f.r2c(a[i:i + 1], outs[i:i + 1])
    torch.cuda.synchronize()
    assert rel_l2(_cpu(outs), _cpu(ref)) <= 1e-7
    f.clear_cache()
    f.set_cache_limit(4)


def test_errors_are_noas():
    print("fft errors", flush=True)
    f = F()
    x = torch.zeros((1, 1, 8, 8), device=DEV)
    with pytest.raises(f.NoaError, match="non-redundant shape of the complex output"):
        f.r2c(x, torch.zeros((1, 1, 8, 4), dtype=torch.complex64, device=DEV))
    with pytest.raises(f.NoaError, match="non-redundant shape of the complex input"):
        f.c2r(torch.zeros((1, 1, 8, 4), dtype=torch.complex64, device=DEV), (1, 1, 8, 8))
    with pytest.raises(f.NoaError, match="no broadcasting allowed"):
        f.c2c(torch.zeros((1, 1, 8, 8), dtype=torch.complex64, device=DEV),
              torch.zeros((1, 1, 8, 4), dtype=torch.complex64, device=DEV))
    with pytest.raises(f.NoaError, match="same precision"):
        f.r2c(x, torch.zeros((1, 1, 8, 5), dtype=torch.complex128, device=DEV))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("case", ["odd_distance", "odd_pitch", "x_stride_2", "odd_length"])
def test_real_operands_with_any_strides(case, dtype):
    """Real arrays whose layout rocFFT's even-length real kernels do not handle (odd row pitch or batch distance,
    x stride 2) are staged through a contiguous buffer; the results equal the oracle's either way."""
    print(f"fft strided real {case} {dtype}", flush=True)
    if case == "odd_distance":
        buf, view = (3, 1, 1, 73), (slice(None), slice(None), slice(None), slice(0, 60))
    elif case == "odd_pitch":
        buf, view = (2, 1, 40, 63), (slice(None), slice(None), slice(0, 36), slice(0, 50))
    elif case == "x_stride_2":
        buf, view = (2, 1, 20, 96), (slice(None), slice(None), slice(None), slice(0, 96, 2))
    else:
        buf, view = (2, 1, 20, 61), (slice(None), slice(None), slice(None), slice(0, 45))
    base = _rand(buf, dtype, 17).to(DEV)
    x = base[view]
    X = F().r2c(x)
    ref = O.fft_r2c(x.cpu().numpy())
    assert rel_l2(_cpu(X), ref) <= TOL[dtype]
    out_base = torch.zeros(buf, dtype=dtype, device=DEV)
    out = out_base[view]
    spec = torch.from_numpy(ref).to(CX[dtype]).to(DEV)
    F().c2r(spec, out)
    assert rel_l2(_cpu(out), O.fft_c2r(ref, tuple(x.shape))) <= TOL[dtype]
    mask = torch.ones(buf, dtype=torch.bool, device=DEV)
    mask[view] = False
    assert bool((out_base[mask] == 0).all()), "c2r wrote outside its output view"
