"""cubic_bspline_prefilter on the GPU against the oracle.

Bar: bit-exact. Each line runs noa's recursion with noa's expressions in noa's order (real * value products, one
add, no contraction), in the value's precision, along the same axes in the same order, so the device and the
oracle produce the same bits for f32, f64, c32 and c64. At 256^3 the result is checked through what the filter is
for: the cubic B-spline of the coefficients passes through the samples.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda"
DTYPES = {"f32": (np.float32, torch.float32), "f64": (np.float64, torch.float64),
          "c32": (np.complex64, torch.complex64), "c64": (np.complex128, torch.complex128)}


def nx():
    import noa_amd.xform as m
    return m


def _random(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(dtype)


def _expected(x_np, out_like_np=None):
    out = np.empty_like(x_np) if out_like_np is None else out_like_np
    O.cubic_bspline_prefilter(x_np, out)
    return out


def _cpu(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("dt", list(DTYPES))
@pytest.mark.parametrize("shape", [(3, 1, 1, 1), (3, 1, 1, 5), (2, 1, 1, 12), (2, 1, 1, 13), (2, 1, 1, 33),
                                   (70, 1, 1, 64), (2, 1, 1, 100), (2, 1, 37, 70), (1, 1, 64, 33),
                                   (2, 17, 33, 65), (1, 40, 31, 32), (1, 5, 1, 1), (1, 3, 1, 9)])
def test_prefilter_matches_the_oracle(dt, shape):
    npt, tt = DTYPES[dt]
    x = _random(shape, npt, 100 + sum(shape))
    print(f"prefilter {dt} {shape}", flush=True)
    xin = torch.from_numpy(x).to(DEV)
    out = torch.empty_like(xin)
    nx().cubic_bspline_prefilter(xin, out)
    assert np.array_equal(_cpu(out), _expected(x))
    # In place gives the same bits.
    nx().cubic_bspline_prefilter(xin, xin)
    assert np.array_equal(_cpu(xin), _expected(x))


@pytest.mark.parametrize("dt", ["f32", "c64"])
def test_prefilter_strided_views_and_broadcast(dt):
    """A column-major output (filtered along its contiguous axis first), a sliced non-contiguous input, and an
    input broadcast over the batch."""
    npt, tt = DTYPES[dt]
    x = _random((1, 12, 20, 36), npt, 7)
    print(f"prefilter views {dt}", flush=True)
    xin = torch.from_numpy(x).to(DEV)
    # Column-major output.
    out_f = torch.empty((1, 36, 20, 12), dtype=tt, device=DEV).permute(0, 3, 2, 1)
    nx().cubic_bspline_prefilter(xin, out_f)
    exp_f = np.empty((1, 36, 20, 12), dtype=npt).transpose(0, 3, 2, 1)
    O.cubic_bspline_prefilter(x, exp_f)
    assert np.array_equal(_cpu(out_f), exp_f)
    # Sliced input (every other x) into a contiguous output.
    big = _random((1, 12, 20, 72), npt, 8)
    xs = torch.from_numpy(big).to(DEV)[..., ::2]
    out = torch.empty((1, 12, 20, 36), dtype=tt, device=DEV)
    nx().cubic_bspline_prefilter(xs, out)
    assert np.array_equal(_cpu(out), _expected(np.ascontiguousarray(big[..., ::2])))
    # Batch-broadcast input.
    outb = torch.empty((3, 12, 20, 36), dtype=tt, device=DEV)
    nx().cubic_bspline_prefilter(xin, outb)
    expb = np.empty((3, 12, 20, 36), dtype=npt)
    O.cubic_bspline_prefilter(np.broadcast_to(x, (3, 12, 20, 36)), expb)
    assert np.array_equal(_cpu(outb), expb)


def test_prefilter_errors_are_noas():
    print("prefilter errors", flush=True)
    x = torch.zeros((1, 4, 4, 4), device=DEV)
    with pytest.raises(nx().NoaError, match="Cannot broadcast"):
        nx().cubic_bspline_prefilter(torch.zeros((1, 4, 3, 4), device=DEV), x)
    with pytest.raises(nx().NoaError, match="types should match"):
        nx().cubic_bspline_prefilter(x.double(), x)
    h = torch.zeros((1, 4, 4, 4), dtype=torch.complex32, device=DEV)
    with pytest.raises(nx().NoaError, match="f32, f64, c32 or c64"):
        nx().cubic_bspline_prefilter(h, h)


def _spline_at_samples(c):
    """Cubic B-spline of coefficients c evaluated at the integer samples of a 3D volume: the separable
    [1, 4, 1] / 6 kernel, the coefficients extended by half-sample symmetry (edge replication)."""
    out = c
    for axis in (1, 2, 3):
        pad = torch.cat([out.narrow(axis, 0, 1), out, out.narrow(axis, out.shape[axis] - 1, 1)], dim=axis)
        n = out.shape[axis]
        out = (pad.narrow(axis, 0, n) + 4 * pad.narrow(axis, 1, n) + pad.narrow(axis, 2, n)) / 6
    return out


def test_prefilter_256_cube_interpolates_its_samples():
    """At 256^3 (f32, in place): B-spline evaluation of the coefficients at the integers returns the volume, to
    float rounding plus the 12-term truncation of the causal start (|z|^13 of the signal)."""
    print("prefilter 256^3 property", flush=True)
    g = torch.Generator(device=DEV).manual_seed(5)
    v = torch.randn((1, 256, 256, 256), device=DEV, generator=g, dtype=torch.float32)
    c = v.clone()
    nx().cubic_bspline_prefilter(c, c)
    rec = _spline_at_samples(c.double())
    err = (rec - v.double()).abs().max().item()
    assert err < 2e-5 * v.abs().max().item(), err
