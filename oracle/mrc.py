"""TEST INFRASTRUCTURE -- CPU restatement of noa's MRC encoder and of its value encoding (numpy).

Loaded only by tests/. It follows noa's code, not this library's:
  - header parse  src/noa/io/ImageFile.cpp:14-166 (ImageFileEncoderMrc::read_header)
  - header write  src/noa/io/ImageFile.cpp:169-296 (ImageFileEncoderMrc::write_header)
  - dtypes        src/noa/io/ImageFile.hpp:606-628 (closest_supported_dtype), Encoding.hpp:85-113 (n_bytes)
  - values        src/noa/io/Encoding.cpp:33-160 (Encoder/Decoder, 4-bit packing), base/ClampCast.hpp (clamp_cast)
The header fields sit at the offsets of the MRC2014 format (https://www.ccpem.ac.uk/mrc_format/mrc2014.php).
Big-endian headers are read word by word (noa reverses the first 24 bytes first; see the library's documentation).
"""
from __future__ import annotations

import struct

import numpy as np

IO = dict(UNKNOWN=0, I8=1, U8=2, I16=3, U16=4, I32=5, U32=6, I64=7, U64=8, F16=9, F32=10, F64=11, C16=12, C32=13,
          C64=14, U4=15, CI16=16)
NAME = {v: k for k, v in IO.items()}
MODE_OF = {"I8": 0, "U8": 0, "I16": 1, "F32": 2, "CI16": 3, "C32": 4, "U16": 6, "F16": 12, "U4": 101}
NP = {"I8": np.int8, "U8": np.uint8, "I16": np.int16, "U16": np.uint16, "F16": np.float16, "F32": np.float32,
      "C32": np.complex64}
IMOD_STAMP = 1146047817


class MrcError(RuntimeError):
    pass


def n_bytes(dtype: str, n: int) -> int:
    per = {"U4": 0.5, "I8": 1, "U8": 1, "I16": 2, "U16": 2, "F16": 2, "I32": 4, "U32": 4, "F32": 4, "CI16": 4, "C16": 4,
           "I64": 8, "U64": 8, "F64": 8, "C32": 8, "C64": 16}[dtype]
    return int(n * per)


def closest_supported(dtype: str) -> str:
    if dtype in ("I8", "U8", "I16", "U16", "F16", "U4", "CI16"):
        return dtype
    if dtype in ("I32", "U32", "I64", "U64", "F32", "F64"):
        return "F32"
    if dtype in ("C16", "C32", "C64"):
        return "C32"
    return "UNKNOWN"


def make_header(shape, spacing=(0.0, 0.0, 0.0), dtype="F32", stats=None, *, big_endian=False, extended=0,
                imod_u8=False, stamp=None, map_order=(1, 2, 3)) -> bytes:
    """write_header (ImageFile.cpp:169-296) for a BDHW shape, optionally big-endian (stamp 17,17), with an extended
    header size, and (mode 0) the IMOD flag that marks unsigned bytes (which noa itself never sets)."""
    b, d, h, w = shape
    if b == 1:
        whd, grid = (w, h, d), (w, h, d)
        ndim = 3 if d > 1 else (2 if h > 1 and w > 1 else 1)
        group = 1 if ndim == 3 else 0
    elif d == 1:
        whd, grid, group = (w, h, b), (w, h, 1), 0
    else:
        whd, grid, group = (w, h, d * b), (w, h, d), 401
    sp = (spacing[2], spacing[1], spacing[0])
    cell = [np.float32(np.float32(g) * np.float32(s)) for g, s in zip(grid, sp)]
    st = [0.0, -1.0, 2.0, -1.0] if stats is None else list(stats)
    e = ">" if big_endian else "<"
    hdr = bytearray(1024)
    struct.pack_into(e + "3i", hdr, 0, *whd)
    struct.pack_into(e + "i", hdr, 12, MODE_OF[dtype])
    struct.pack_into(e + "3i", hdr, 28, *grid)
    struct.pack_into(e + "3f", hdr, 40, *cell)
    struct.pack_into(e + "3f", hdr, 52, 90.0, 90.0, 90.0)
    struct.pack_into(e + "3i", hdr, 64, *map_order)
    struct.pack_into(e + "3f", hdr, 76, *st[:3])
    struct.pack_into(e + "i", hdr, 88, group)
    struct.pack_into(e + "i", hdr, 92, extended)
    hdr[104:108] = b"SERI"
    if dtype == "U8":
        struct.pack_into(e + "i", hdr, 152, IMOD_STAMP)
        struct.pack_into(e + "i", hdr, 156, 1 if imod_u8 else 0)
    hdr[208:212] = b"MAP "
    hdr[212:216] = bytes(stamp) if stamp is not None else (bytes([17, 17, 0, 0]) if big_endian else bytes([68, 68, 0, 0]))
    struct.pack_into(e + "f", hdr, 216, st[3])
    return bytes(hdr)


def parse_header(buf: bytes) -> dict:
    """read_header (ImageFile.cpp:14-166)."""
    if len(buf) < 1024:
        raise MrcError("Failed to read the header (1024 bytes)")
    st = tuple(buf[212:216])
    if st in ((68, 65, 0, 0), (68, 68, 0, 0)):
        e = "<"
    elif st == (17, 17, 0, 0):
        e = ">"
    else:
        raise MrcError("Endianness was not recognized")
    nx, ny, nz, mode = struct.unpack_from(e + "4i", buf, 0)
    mx, my, mz = struct.unpack_from(e + "3i", buf, 28)
    cx, cy, cz = struct.unpack_from(e + "3f", buf, 40)
    order = struct.unpack_from(e + "3i", buf, 64)
    dmin, dmax, dmean = struct.unpack_from(e + "3f", buf, 76)
    group, next_ = struct.unpack_from(e + "2i", buf, 88)
    imod_stamp, imod_flags = struct.unpack_from(e + "2i", buf, 152)
    (rms,) = struct.unpack_from(e + "f", buf, 216)
    if min(nx, ny, nz) <= 0:
        raise MrcError("Logical shape should be greater than zero")
    ndim = 3 if nz > 1 else (2 if ny > 1 and nx > 1 else 1)
    if ndim <= 2:
        shape = (1, 1, ny, nx)
    elif group == 0:
        shape = (nz, 1, ny, nx)
    elif group == 1:
        shape = (1, nz, ny, nx)
    elif 401 <= group <= 630:
        shape = (nz // mz, mz, ny, nx)
    else:
        raise MrcError("Data shape is not recognized")
    spacing = tuple(float(np.float32(c) / np.float32(m)) for c, m in ((cz, mz), (cy, my), (cx, mx)))
    modes = {1: "I16", 2: "F32", 3: "CI16", 4: "C32", 6: "U16", 12: "F16", 101: "U4"}
    if mode == 0:
        dtype = "U8" if (imod_stamp == IMOD_STAMP and imod_flags & 1) else "I8"
    elif mode in modes:
        dtype = modes[mode]
    else:
        raise MrcError(f"MRC mode not recognized, got {mode}")
    if tuple(order) != (1, 2, 3):
        raise MrcError("Map order is not supported")
    stats = None if (np.float32(dmax) == np.float32(-1) and np.float32(dmean) == np.float32(2)) else (dmin, dmax, dmean, rms)
    return dict(shape=shape, spacing=spacing, dtype=dtype, stats=stats, data_offset=1024 + next_,
                big_endian=(e == ">"))


def _clamp_to_int(x, lo, hi):
    """clamp_cast<int type>(float): NaN -> 0, saturate, truncate toward zero (ClampCast.hpp:56-75)."""
    x = np.asarray(x, np.float64)
    out = np.where(np.isnan(x), 0.0, np.clip(np.trunc(np.nan_to_num(x, nan=0.0, posinf=hi, neginf=lo)), lo, hi))
    return out


def encode(values: np.ndarray, dtype: str, clamp: bool = True) -> bytes:
    """encode (Encoding.cpp:33-160) of a BDHW array of real or complex values into little-endian bytes."""
    v = np.asarray(values)
    flat = v.reshape(-1)
    if dtype == "U4":
        r = np.asarray(flat.real if np.iscomplexobj(flat) else flat, np.float64)
        rr = np.where(np.isnan(r), 0.0, np.sign(r) * np.floor(np.abs(r) + 0.5))   # round half away from zero
        u = np.clip(np.where(rr <= 0, 0, np.minimum(rr, 4294967295.0)), 0, 15).astype(np.uint8)
        return (u[0::2] | (u[1::2] << 4)).astype(np.uint8).tobytes()
    if dtype in ("I8", "U8", "I16", "U16"):
        lo, hi = {"I8": (-128, 127), "U8": (0, 255), "I16": (-32768, 32767), "U16": (0, 65535)}[dtype]
        return _clamp_to_int(np.real(flat), lo, hi).astype(NP[dtype]).tobytes()
    if dtype == "F16":
        x = np.real(flat).astype(np.float64)
        if clamp:
            x = np.where(np.isnan(x), x, np.clip(x, -65504.0, 65504.0))
        return x.astype(np.float16).tobytes()
    if dtype == "F32":
        x = np.real(flat).astype(np.float64)
        if clamp and np.real(flat).dtype == np.float64:
            x = np.where(np.isnan(x), x, np.clip(x, -3.4028234663852886e38, 3.4028234663852886e38))
        return x.astype(np.float32).tobytes()
    if dtype == "C32":
        return flat.astype(np.complex64).tobytes()
    if dtype == "CI16":
        c = flat.astype(np.complex128)
        out = np.empty(2 * c.size, np.int16)
        out[0::2] = _clamp_to_int(c.real, -32768, 32767)
        out[1::2] = _clamp_to_int(c.imag, -32768, 32767)
        return out.tobytes()
    raise MrcError(f"unsupported dtype {dtype}")


def decode(raw: bytes, dtype: str, shape, big_endian: bool = False, complex_out: bool = False) -> np.ndarray:
    """decode (Encoding.cpp:85-160): byte swap of big-endian data, then the value as float64 (complex128)."""
    n = int(np.prod(shape))
    e = ">" if big_endian else "<"
    if dtype == "U4":
        b = np.frombuffer(raw, np.uint8, n // 2)
        out = np.empty(n, np.float64)
        out[0::2] = b & 0x0F
        out[1::2] = (b >> 4) & 0x0F
    elif dtype == "CI16":
        a = np.frombuffer(raw, np.dtype(np.int16).newbyteorder(e), 2 * n).astype(np.float64)
        out = a[0::2] + 1j * a[1::2]
    elif dtype == "C32":
        out = np.frombuffer(raw, np.dtype(np.complex64).newbyteorder(e), n).astype(np.complex128)
    else:
        out = np.frombuffer(raw, np.dtype(NP[dtype]).newbyteorder(e), n).astype(np.float64)
    if complex_out and not np.iscomplexobj(out):
        out = out.astype(np.complex128)
    return out.reshape(shape)


def write_file(path, values: np.ndarray, dtype: str, spacing=(0.0, 0.0, 0.0), stats=None, **header_kw) -> None:
    """A whole MRC file as noa writes it (header then data); header_kw as make_header (big_endian also swaps the
    data; extended adds that many zero bytes after the header)."""
    shape = values.shape
    hdr = make_header(shape, spacing, dtype, stats, **header_kw)
    raw = encode(values, dtype)
    if header_kw.get("big_endian"):
        item = {"I16": 2, "U16": 2, "F16": 2, "F32": 4, "C32": 4, "CI16": 2}.get(dtype, 1)
        if item > 1:
            raw = np.frombuffer(raw, np.dtype(f"u{item}")).byteswap().tobytes()
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(bytes(header_kw.get("extended", 0)))
        f.write(raw)


def read_file(path) -> tuple[dict, np.ndarray]:
    with open(path, "rb") as f:
        buf = f.read()
    h = parse_header(buf)
    n = int(np.prod(h["shape"]))
    raw = buf[h["data_offset"]:h["data_offset"] + n_bytes(h["dtype"], n)]
    return h, decode(raw, h["dtype"], h["shape"], h["big_endian"])
