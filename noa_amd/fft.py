"""noa::fft transforms on the MI355X (rocFFT through the C ABI): r2c, c2r, c2c and the plan-cache controls.

Mirrors src/noa/fft/Transform.hpp: the same argument meaning, defaults (norm FORWARD: the forward transform is
scaled by 1/N, N = D*H*W), shape rules and error messages. Arrays are torch device tensors of up to 4 dimensions
(BDHW, leading dimensions of size 1 implied). The normalisation is applied by the transform itself (rocFFT's scale
factor), not by a second pass.
"""
from __future__ import annotations

import ctypes
import enum
from typing import Sequence

import torch

from . import abi
from ._lib import last_error, lib
from .xform import NoaError, _pad4, _view


class Norm(enum.IntEnum):
    """noa::fft::Norm (src/noa/fft/core/Transform.hpp:15-20)."""
    FORWARD = abi.NORM_FORWARD
    ORTHO = abi.NORM_ORTHO
    BACKWARD = abi.NORM_BACKWARD
    NONE = abi.NORM_NONE


class Sign(enum.IntEnum):
    """noa::fft::Sign (src/noa/fft/core/Transform.hpp:8-11)."""
    FORWARD = abi.SIGN_FORWARD
    BACKWARD = abi.SIGN_BACKWARD


def _options(norm, cache_plan: bool, plan_only: bool, record_and_share_workspace: bool) -> abi.FftOptions:
    o = abi.FftOptions()
    if isinstance(norm, str):
        norm = Norm[norm.upper()]
    o.norm = int(norm)
    o.cache_plan = int(bool(cache_plan))
    o.plan_only = int(bool(plan_only))
    o.record_and_share_workspace = int(bool(record_and_share_workspace))
    return o


def _check_device(t, name):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise NoaError(f"{name}: expected a device tensor")


def _call(fn: str, input, output, opts: abi.FftOptions, *extra) -> None:
    keep: list = []
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input")
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    stream = torch.cuda.current_stream(output.device).cuda_stream
    f = getattr(lib(), fn)
    if extra:
        status = f(ctypes.byref(in_v), ctypes.byref(out_v), *extra, ctypes.byref(opts), ctypes.c_void_p(stream))
    else:
        status = f(ctypes.byref(in_v), ctypes.byref(out_v), ctypes.byref(opts), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def r2c(input: torch.Tensor, output: torch.Tensor | None = None, norm=Norm.FORWARD, cache_plan: bool = True,
        plan_only: bool = False, record_and_share_workspace: bool = False) -> torch.Tensor:
    """noa::fft::r2c (Transform.hpp:160-223): the forward transform of a real array into its non-redundant,
    non-centred ("h") spectrum, shape (..., W//2+1). Allocates the output when none is given. In place when the
    output shares the input's memory in noa's padded layout."""
    _check_device(input, "input")
    if output is None:
        shape, _ = _pad4(input.shape, input.stride())
        cdt = torch.complex64 if input.dtype == torch.float32 else torch.complex128
        output = torch.empty((*shape[:3], shape[3] // 2 + 1), dtype=cdt, device=input.device)
    _call("noa_hip_fft_r2c", input, output, _options(norm, cache_plan, plan_only, record_and_share_workspace))
    return output


def c2r(input: torch.Tensor, output: torch.Tensor | Sequence[int], norm=Norm.FORWARD, cache_plan: bool = True,
        plan_only: bool = False, record_and_share_workspace: bool = False) -> torch.Tensor:
    """noa::fft::c2r (Transform.hpp:225-288): the backward transform of a non-redundant spectrum into a real array.
    `output` is the real array or its logical BDHW shape (then allocated). For 2D and 3D transforms the input is
    overwritten, as in noa."""
    _check_device(input, "input")
    if not isinstance(output, torch.Tensor):
        rdt = torch.float32 if input.dtype == torch.complex64 else torch.float64
        output = torch.empty(tuple(int(s) for s in output), dtype=rdt, device=input.device)
    _call("noa_hip_fft_c2r", input, output, _options(norm, cache_plan, plan_only, record_and_share_workspace))
    return output


def c2c(input: torch.Tensor, output: torch.Tensor | None = None, sign=Sign.FORWARD, norm=Norm.FORWARD,
        cache_plan: bool = True, plan_only: bool = False, record_and_share_workspace: bool = False) -> torch.Tensor:
    """noa::fft::c2c (Transform.hpp:290-349): the full transform with the exponent's sign (FORWARD -1,
    BACKWARD +1) into a non-centred ("f") layout. In place when output is input."""
    _check_device(input, "input")
    if output is None:
        output = torch.empty_like(input)
    _call("noa_hip_fft_c2c", input, output, _options(norm, cache_plan, plan_only, record_and_share_workspace),
          ctypes.c_int32(int(sign)))
    return output


def _dev_index(device) -> int:
    if device is None:
        return torch.cuda.current_device()
    d = torch.device(device)
    return d.index if d.index is not None else torch.cuda.current_device()


def clear_cache(device=None) -> int:
    """noa::fft::clear_cache: releases the calling thread's plans on the device; returns how many."""
    return int(lib().noa_hip_fft_clear_cache(_dev_index(device)))


def set_cache_limit(count: int, device=None) -> int:
    """noa::fft::set_cache_limit: returns how many plans were evicted by the new limit."""
    return int(lib().noa_hip_fft_set_cache_limit(int(count), _dev_index(device)))


def cache_limit(device=None) -> int:
    return int(lib().noa_hip_fft_cache_limit(_dev_index(device)))


def cache_size(device=None) -> int:
    return int(lib().noa_hip_fft_cache_size(_dev_index(device)))


def workspace_left_to_allocate(device=None) -> int:
    """Bytes still to allocate for the plans created with record_and_share_workspace."""
    return int(lib().noa_hip_fft_workspace_left_to_allocate(_dev_index(device)))


def set_workspace(buffer: torch.Tensor, device=None) -> int:
    """noa::fft::set_workspace: the cached plans still without a work buffer use `buffer` (a contiguous device
    tensor the caller keeps alive while they run). Returns how many plans took it (0 if it is too small)."""
    _check_device(buffer, "buffer")
    if not buffer.is_contiguous():
        raise NoaError("The workspace should be a contiguous array")
    n = lib().noa_hip_fft_set_workspace(_dev_index(device if device is not None else buffer.device),
                                        ctypes.c_void_p(buffer.data_ptr()),
                                        ctypes.c_int64(buffer.numel() * buffer.element_size()))
    if n < 0:
        raise NoaError(last_error())
    return int(n)


def empty(shape: Sequence[int], dtype=torch.float32, device="cuda") -> tuple[torch.Tensor, torch.Tensor]:
    """noa::fft::empty (src/noa/fft/Factory.hpp:75-85): a real array of logical BDHW `shape` and the complex array
    of its rfft sharing the same memory (rows padded to 2 * (W//2+1) reals), for in-place r2c/c2r."""
    b, d, h, w = [1] * (4 - len(shape)) + [int(s) for s in shape]
    wr = w // 2 + 1
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    buf = torch.empty((b, d, h, wr), dtype=cdt, device=device)
    real = torch.view_as_real(buf).reshape(b, d, h, 2 * wr)[..., :w]
    return real, buf
