"""noa::io for MRC files, between disk and device memory (the C ABI's MRC reader and writer).

Mirrors src/noa/io/IO.hpp (read_image, write_image with ReadOption/WriteOption) and the MRC encoder of
src/noa/io/ImageFile.hpp: the header is parsed on the host, the values are decoded and encoded on the GPU.
"""
from __future__ import annotations

import ctypes
import dataclasses
from pathlib import Path
from typing import Sequence

import torch

from . import abi
from ._lib import last_error, lib
from .xform import NoaError, _view

DTYPE_NAMES = {v: k for k, v in abi.IO_DTYPES.items()}


@dataclasses.dataclass
class Header:
    """What noa's ImageFile reports for an MRC file: BDHW shape, DHW spacing (Angstrom/pixel, 0 = unset), the
    stored data type (noa::io::DataType name), the statistics if set (min, max, mean, stddev)."""
    shape: tuple[int, int, int, int]
    spacing: tuple[float, float, float]
    dtype: str
    stats: tuple[float, float, float, float] | None
    data_offset: int
    endian_swapped: bool


def read_header(path: str | Path) -> Header:
    h = abi.MrcHeader()
    if lib().noa_hip_mrc_read_header(str(path).encode(), ctypes.byref(h)) != abi.OK:
        raise NoaError(last_error())
    return Header(tuple(h.shape), tuple(h.spacing), DTYPE_NAMES[h.dtype],
                  tuple(h.stats) if h.has_stats else None, int(h.data_offset), bool(h.endian_swapped))


_TORCH_OUT = {torch.float32, torch.float64, torch.complex64, torch.complex128}


def read_image(path: str | Path, dtype: torch.dtype = torch.float32, device="cuda", clamp: bool = True,
               enforce_2d_stack: bool = False, output: torch.Tensor | None = None) -> tuple[torch.Tensor, Header]:
    """noa::io::read_image (IO.hpp:34-57): the whole file into a new BDHW device tensor (or into `output`, a
    device tensor of the file's shape), decoded on the GPU; and the header. enforce_2d_stack turns a single volume
    into a stack of 2D images."""
    header = read_header(path)
    if output is None:
        if dtype not in _TORCH_OUT:
            raise NoaError("read_image: the output should be f32, f64, c64 or c128")
        output = torch.empty(header.shape, dtype=dtype, device=device)
    keep: list = []
    v = _view(output, keep, abi.F32, abi.C32, False, "output")
    stream = torch.cuda.current_stream(output.device).cuda_stream
    if lib().noa_hip_mrc_read(str(path).encode(), ctypes.byref(v), int(bool(clamp)), ctypes.c_void_p(stream)) != abi.OK:
        raise NoaError(last_error())
    if enforce_2d_stack and output.dim() == 4 and output.shape[0] == 1 and output.shape[1] > 1:
        output = output.permute(1, 0, 2, 3)
    return output, header


def write_image(input: torch.Tensor, path: str | Path, spacing: Sequence[float] | None = None,
                dtype: str = "UNKNOWN", stats: Sequence[float] | None = None, clamp: bool = True) -> None:
    """noa::io::write_image (IO.hpp:119-163) into a new MRC file: `dtype` is a noa::io::DataType name ("F32",
    "I16", "U4", ...; UNKNOWN: F32 or C32), replaced by the closest type MRC stores; spacing is DHW; stats is
    (min, max, mean, stddev) or None."""
    if not isinstance(input, torch.Tensor) or input.device.type != "cuda":
        raise NoaError("write_image: the input should be a device tensor")
    keep: list = []
    v = _view(input, keep, abi.F32, abi.C32, False, "input")
    sp = (ctypes.c_double * 3)(*spacing) if spacing is not None else None
    st = (ctypes.c_double * 4)(*stats) if stats is not None else None
    code = abi.IO_DTYPES[dtype.upper()] if isinstance(dtype, str) else int(dtype)
    stream = torch.cuda.current_stream(input.device).cuda_stream
    status = lib().noa_hip_mrc_write(str(path).encode(), ctypes.byref(v),
                                     ctypes.cast(sp, ctypes.c_void_p) if sp is not None else None, code,
                                     ctypes.cast(st, ctypes.c_void_p) if st is not None else None, int(bool(clamp)),
                                     ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())
