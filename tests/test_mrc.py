"""MRC files without a GPU: the oracle pinned to the MRC2014 layout and to noa's clamp_cast rules, and the library's
host-side header parser (noa_hip_mrc_read_header) against the oracle, including noa's error cases.
"""
import struct

import numpy as np
import pytest

import mrc as M


def test_oracle_header_follows_the_mrc2014_layout():
    """Word offsets of the published format: nx, ny, nz (0), mode (12), mx, my, mz (28), cell lengths (40), cell
    angles (52), mapc, mapr, maps (64), dmin, dmax, dmean (76), ispg (88), nsymbt (92), exttyp (104), map (208),
    machst (212), rms (216). A stack of volumes is ispg 401 with mz sections per volume."""
    h = M.make_header((3, 5, 6, 7), spacing=(1.5, 2.0, 2.5), dtype="I16", stats=(-1.0, 4.0, 0.5, 0.25))
    assert len(h) == 1024
    assert struct.unpack_from("<3i", h, 0) == (7, 6, 15)
    assert struct.unpack_from("<i", h, 12) == (1,)
    assert struct.unpack_from("<3i", h, 28) == (7, 6, 5)
    assert struct.unpack_from("<3f", h, 40) == (17.5, 12.0, 7.5)
    assert struct.unpack_from("<3f", h, 52) == (90.0, 90.0, 90.0)
    assert struct.unpack_from("<3i", h, 64) == (1, 2, 3)
    assert struct.unpack_from("<3f", h, 76) == (-1.0, 4.0, 0.5)
    assert struct.unpack_from("<2i", h, 88) == (401, 0)
    assert h[104:108] == b"SERI" and h[208:212] == b"MAP " and tuple(h[212:216]) == (68, 68, 0, 0)
    assert struct.unpack_from("<f", h, 216) == (0.25,)
    p = M.parse_header(h)
    assert p["shape"] == (3, 5, 6, 7) and p["dtype"] == "I16" and p["spacing"] == (1.5, 2.0, 2.5)
    # Unset statistics are written as max = -1, mean = 2 and read back as unset.
    assert M.parse_header(M.make_header((1, 1, 8, 8)))["stats"] is None


def test_oracle_values_follow_clamp_cast():
    """clamp_cast to integers: NaN -> 0, saturation (inf included), truncation toward zero; 4-bit values rounded
    half away from zero, clamped to [0, 15], packed low nibble first; f16 is IEEE round-to-nearest-even."""
    x = np.array([np.nan, np.inf, -np.inf, 300.7, -2.9, 2.9, -128.5, 127.4], np.float32)
    assert np.frombuffer(M.encode(x, "I8"), np.int8).tolist() == [0, 127, -128, 127, -2, 2, -128, 127]
    assert np.frombuffer(M.encode(x, "U8"), np.uint8).tolist() == [0, 255, 0, 255, 0, 2, 0, 127]
    assert np.frombuffer(M.encode(np.array([70000.0, -5.5]), "U16"), np.uint16).tolist() == [65535, 0]
    assert M.encode(np.array([1.0, 2.0, 2.5, -1.0, 20.0, 0.49]), "U4") == bytes([0x21, 0x03, 0x0F])
    assert np.array_equal(M.decode(bytes([0x21, 0x03, 0x0F]), "U4", (6,)), [1, 2, 3, 0, 15, 0])
    h = np.array([1.0009765625, 1.00048828125, 70000.0, -1e6], np.float64)
    assert np.frombuffer(M.encode(h, "F16"), np.float16).tolist() == [1.0009765625, 1.0, 65504.0, -65504.0]
    c = np.array([1.6 - 2.4j, 40000 + 0j], np.complex64)
    assert np.frombuffer(M.encode(c, "CI16"), np.int16).tolist() == [1, -2, 32767, 0]


def _lib():
    import noa_amd.io as io
    return io


@pytest.mark.parametrize("shape, dtype, kw", [
    ((1, 1, 32, 48), "F32", {}),
    ((1, 20, 16, 12), "I16", {"spacing": (1.1, 1.2, 1.3), "stats": (0.0, 1.0, 0.5, 0.1)}),
    ((9, 1, 16, 16), "U16", {}),
    ((4, 6, 8, 10), "F16", {}),
    ((1, 4, 8, 8), "U8", {"imod_u8": True}),
    ((1, 4, 8, 8), "U8", {}),
    ((1, 4, 8, 8), "I8", {}),
    ((2, 1, 6, 8), "U4", {}),
    ((1, 1, 8, 8), "CI16", {}),
    ((1, 3, 8, 8), "C32", {"big_endian": True, "extended": 96}),
    ((1, 1, 1, 64), "F32", {}),
])
def test_library_header_parser_matches_the_oracle(tmp_path, shape, dtype, kw):
    rng = np.random.default_rng(1)
    vals = rng.integers(0, 15, size=shape).astype(np.float32)
    if dtype in ("C32", "CI16"):
        vals = vals + 1j * vals[..., ::-1]
    path = tmp_path / "f.mrc"
    M.write_file(path, vals, dtype, **kw)
    want = M.parse_header(open(path, "rb").read())
    got = _lib().read_header(path)
    assert got.shape == want["shape"] and got.dtype == want["dtype"]
    assert np.allclose(got.spacing, want["spacing"], rtol=1e-7)
    assert got.data_offset == want["data_offset"] and got.endian_swapped == want["big_endian"]
    assert (got.stats is None) == (want["stats"] is None)
    if got.stats is not None:
        assert np.allclose(got.stats, want["stats"])


@pytest.mark.parametrize("mutate, message", [
    (lambda b: b.__setitem__(slice(212, 216), bytes([1, 2, 3, 4])), "Endianness was not recognized"),
    (lambda b: struct.pack_into("<3i", b, 64, 2, 1, 3), "Map order"),
    (lambda b: struct.pack_into("<i", b, 12, 16), "mode 16"),
    (lambda b: struct.pack_into("<i", b, 12, 7), "mode not recognized"),
    (lambda b: struct.pack_into("<i", b, 88, 7), "not recognized"),
    (lambda b: struct.pack_into("<i", b, 0, 0), "greater than zero"),
    (lambda b: b.__delitem__(slice(2000, None)), "truncated"),
])
def test_library_header_errors_are_noas(tmp_path, mutate, message):
    path = tmp_path / "bad.mrc"
    M.write_file(path, np.zeros((1, 4, 16, 16), np.float32), "F32")
    b = bytearray(open(path, "rb").read())
    mutate(b)
    open(path, "wb").write(bytes(b))
    from noa_amd.xform import NoaError
    with pytest.raises(NoaError, match=message):
        _lib().read_header(path)
