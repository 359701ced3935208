"""MRC files between disk and device memory: the library's writer and reader against the oracle's MRC encoder,
and noa's own MRC tests (tests/noa/io/TestIOMRCFile.cpp). Bar: bit-exact bytes on disk and bit-exact values.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mrc as M

pytestmark = pytest.mark.gpu

DEV = "cuda"


def io():
    import noa_amd.io as m
    return m


def _values(shape, dtype, seed, complex_=False):
    """Values every dtype stores exactly when clamping is off (noa's tests draw integers in [0, 127]), plus
    out-of-range and fractional values for the clamping types."""
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 128, size=shape).astype(np.float32)
    if dtype in ("I8", "I16", "F16", "F32", "CI16", "C32"):
        v -= 64
    if dtype == "U4":
        v = np.mod(v, 16)
    if complex_:
        v = v + 1j * rng.integers(-50, 50, size=shape).astype(np.float32)
    return v


@pytest.mark.parametrize("dtype", ["I8", "U8", "I16", "U16", "F16", "F32", "U4", "CI16", "C32"])
@pytest.mark.parametrize("shape", [(1, 12, 16, 20), (7, 1, 16, 20), (3, 4, 8, 10)])
def test_write_matches_the_oracle_bytes(tmp_path, dtype, shape):
    print(f"mrc write {dtype} {shape}", flush=True)
    cx = dtype in ("CI16", "C32")
    v = _values(shape, dtype, 3, cx)
    x = torch.from_numpy(v.astype(np.complex64 if cx else np.float32)).to(DEV)
    path = tmp_path / "w.mrc"
    io().write_image(x, path, spacing=(1.0, 1.5, 2.0), dtype=dtype, stats=(-1.0, 1.0, 0.0, 1.0))
    data = open(path, "rb").read()
    want_hdr = M.make_header(shape, (1.0, 1.5, 2.0), dtype, (-1.0, 1.0, 0.0, 1.0))
    assert data[:1024] == want_hdr
    assert data[1024:] == M.encode(v, dtype)


@pytest.mark.parametrize("dtype", ["I8", "U8", "I16", "U16", "F16", "F32", "U4", "CI16", "C32"])
@pytest.mark.parametrize("big_endian", [False, True])
def test_read_matches_the_oracle(tmp_path, dtype, big_endian):
    print(f"mrc read {dtype} big_endian={big_endian}", flush=True)
    cx = dtype in ("CI16", "C32")
    shape = (2, 5, 12, 14)
    v = _values(shape, dtype, 4, cx)
    path = tmp_path / "r.mrc"
    M.write_file(path, v, dtype, spacing=(2.0, 2.0, 2.0), big_endian=big_endian, extended=64 if big_endian else 0,
                 imod_u8=(dtype == "U8"))
    h, want = M.read_file(path)
    for out_dt in ((torch.complex64, torch.complex128) if cx else (torch.float32, torch.float64, torch.complex64)):
        got, hdr = io().read_image(path, dtype=out_dt)
        assert hdr.shape == h["shape"] and hdr.dtype == h["dtype"]
        g = got.cpu().numpy()
        assert np.array_equal(g, want.astype(g.dtype)), out_dt


def test_read_into_strided_views_and_many_chunks(tmp_path):
    """A 64 MiB F32 volume (three 32 MiB chunks with the header) into a column-major device view; and a written
    file from a strided (transposed) tensor."""
    print("mrc chunks and views", flush=True)
    shape = (1, 256, 256, 257)
    v = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
    path = tmp_path / "big.mrc"
    M.write_file(path, v, "F32")
    out = torch.empty((1, 257, 256, 256), dtype=torch.float32, device=DEV).permute(0, 3, 2, 1)
    io().read_image(path, output=out)
    assert np.array_equal(out.cpu().numpy(), v)
    path2 = tmp_path / "big2.mrc"
    io().write_image(out, path2)
    assert open(path2, "rb").read()[1024:] == v.tobytes()


@pytest.mark.parametrize("dtype", ["I16", "U16", "U8", "I8", "F16", "F32"])
def test_noas_mrc_round_trips(tmp_path, dtype):
    """TestIOMRCFile.cpp:11-130: a 64^3 volume, a stack of five 64^3 volumes and a stack of 41 64^2 images of
    integers in [0, 127], written with stats (-1, 1, 0, 1) and spacing 1.23, read back identical with the same
    shape, spacing and stats."""
    print(f"mrc round trips {dtype}", flush=True)
    rng = np.random.default_rng(6)
    for shape, spacing, stats in (((1, 64, 64, 64), (1.23, 1.23, 1.23), (-1.0, 1.0, 0.0, 1.0)),
                                  ((5, 64, 64, 64), (1.23, 1.23, 1.23), (-1.0, 1.0, 0.0, 1.0)),
                                  ((41, 1, 64, 64), (1.0, 1.23, 1.23), None)):
        v = rng.integers(0, 128, size=shape).astype(np.float32)
        x = torch.from_numpy(v).to(DEV)
        path = tmp_path / "rt.mrc"
        io().write_image(x, path, spacing=spacing, dtype=dtype, stats=stats)
        y, h = io().read_image(path)
        assert h.shape == shape
        assert np.allclose(np.float32(h.spacing), np.float32(spacing))
        if stats is not None:
            assert h.stats[0] == -1.0 and h.stats[1] == 1.0
        assert torch.equal(x, y)


def test_errors_are_noas(tmp_path):
    print("mrc errors", flush=True)
    from noa_amd.xform import NoaError
    path = tmp_path / "e.mrc"
    M.write_file(path, np.zeros((1, 2, 4, 4), np.complex64), "C32")
    with pytest.raises(NoaError, match="complex data into a real array"):
        io().read_image(path, dtype=torch.float32)
    with pytest.raises(NoaError, match="does not match the file shape"):
        io().read_image(path, output=torch.empty((1, 2, 4, 5), dtype=torch.complex64, device=DEV))
    with pytest.raises(NoaError, match="complex value to a non-complex"):
        io().write_image(torch.zeros((1, 1, 4, 4), dtype=torch.complex64, device=DEV), tmp_path / "x.mrc", dtype="F32")
