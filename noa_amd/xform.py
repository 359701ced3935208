"""Host-side mirror of noa's central-slice API (``noa::xform``, src/noa/xform/FourierProject.hpp:1601-1914).

The three operators keep noa's argument order and meaning. The template argument ``REMAP`` becomes the first
argument (``"h2h"``, ``"hc2hc"``/``"hc"``, ``"hc2h"``, ``"h2hc"``). Arrays are torch tensors on a HIP device,
with rfft BDHW shapes (``shape.rfft()``) and element strides; shapes are logical BDHW tuples. Work is
enqueued on torch's current stream of the arrays' device and is asynchronous, like noa's GPU path.

Everything is computed by the HIP library through its C ABI (include/noa_hip.h); nothing here computes a
result on the host. Errors raise :class:`NoaError` with the library's message.
"""
from __future__ import annotations

import ctypes
import enum
import math
from dataclasses import dataclass, field
from typing import Any, Sequence

import numpy as np
import torch

from . import abi
from ._lib import last_error, lib


class NoaError(RuntimeError):
    """Raised for invalid arguments or runtime failures (noa::Exception, src/noa/base/Error.hpp:15-172)."""


# ---------------------------------------------------------------------------------------------------------------
# Enums and options (Interp.hpp:17-41, FourierProject.hpp:1602-1627, :1703-1739, :1817-1838).
# ---------------------------------------------------------------------------------------------------------------
class Interp(enum.IntEnum):
    NEAREST = 0
    LINEAR = 1
    CUBIC = 2
    CUBIC_BSPLINE = 3
    LANCZOS4 = 4
    LANCZOS6 = 5
    LANCZOS8 = 6
    NEAREST_FAST = 100
    LINEAR_FAST = 101
    CUBIC_FAST = 102
    CUBIC_BSPLINE_FAST = 103
    LANCZOS4_FAST = 104
    LANCZOS6_FAST = 105
    LANCZOS8_FAST = 106


_LAYOUTS = {
    "h2h": abi.H2H, "h": abi.H2H,
    "hc2hc": abi.HC2HC, "hc": abi.HC2HC,
    "hc2h": abi.HC2H,
    "h2hc": abi.H2HC,
}
_FULL_LAYOUTS = {"h2f", "h2fc", "hc2f", "hc2fc", "f2h", "f2hc", "f2f", "f", "f2fc", "fc2h", "fc2hc", "fc2f",
                 "fc2fc", "fc"}


def layout(remap: str | int) -> int:
    """noa::fft::Layout::from_string restricted to half-to-half layouts (Layout.hpp:76-120, is_hx2hx)."""
    if isinstance(remap, int):
        if remap not in (abi.H2H, abi.HC2H, abi.H2HC, abi.HC2HC):
            raise NoaError(f"invalid remap value {remap}")
        return remap
    key = remap.strip().lower()
    if key in _LAYOUTS:
        return _LAYOUTS[key]
    if key in _FULL_LAYOUTS:
        raise NoaError(f"the remap should be HX2HX, but got {remap!r}")
    raise NoaError(f"invalid layout {remap!r}")


@dataclass
class WindowedSinc:
    fftfreq_sinc: float = -1.0
    fftfreq_blackman: float = -1.0


@dataclass
class RasterizeCentralSlicesOptions:
    fftfreq_cutoff: float = 0.5
    target_shape: Sequence[int] = (0, 0, 0, 0)
    ews_radius: Sequence[float] = (0.0, 0.0)


@dataclass
class InsertCentralSlicesOptions:
    interp: Interp = Interp.LINEAR
    windowed_sinc: WindowedSinc = field(default_factory=WindowedSinc)
    fftfreq_cutoff: float = 0.5
    target_shape: Sequence[int] = (0, 0, 0, 0)
    ews_radius: Sequence[float] = (0.0, 0.0)


@dataclass
class ExtractCentralSlicesOptions:
    interp: Interp = Interp.LINEAR
    w_windowed_sinc: WindowedSinc = field(default_factory=WindowedSinc)
    fftfreq_cutoff: float = 0.5
    target_shape: Sequence[int] = (0, 0, 0, 0)
    ews_radius: Sequence[float] = (0.0, 0.0)


@dataclass
class InsertAndExtractCentralSlicesOptions:
    """InsertAndExtractCentralSlicesOptions (FourierProject.hpp:1916-1946)."""
    interp: Interp = Interp.LINEAR
    input_windowed_sinc: WindowedSinc = field(default_factory=WindowedSinc)
    w_windowed_sinc: WindowedSinc = field(default_factory=WindowedSinc)
    add_to_output: bool = False
    correct_weights: bool = False
    fftfreq_cutoff: float = 0.5
    ews_radius: Sequence[float] = (0.0, 0.0)


# ---------------------------------------------------------------------------------------------------------------
# Input forms.
# ---------------------------------------------------------------------------------------------------------------
class Texture:
    """noa::xform::Texture (src/noa/xform/Texture.hpp:62-300).

    On MI355X a texture is a device buffer read with the library's FP32 software interpolation (no hardware
    texture units are used), so results equal the array path. As in noa, the texture's interpolation method
    overrides the options' (FourierProject.hpp:1203-1217), and its border is ZERO.
    """

    def __init__(self, array: torch.Tensor, interp: Interp = Interp.LINEAR):
        if not isinstance(array, torch.Tensor) or array.device.type != "cuda":
            raise NoaError("Texture: a device array is expected")
        self.array = array.contiguous()
        self.interp = Interp(interp)

    @property
    def shape(self):
        return self.array.shape


class VolumeBlocks:
    """A c32 rfft volume (1, D, H, W/2+1) copied once into the 4x2x4-block layout that the extraction gathers from
    (noa_hip_volume_blocks). The single-call extraction makes this copy on every call; a caller that extracts many
    times from the same resident volume (a serving loop) makes it once and passes the copy to the batched extraction
    (noa_amd.batch.extract_call). Like noa's Texture it is a copy: `update` copies a new volume in."""

    def __init__(self, volume: torch.Tensor):
        self._check(volume)
        self.shape = tuple(volume.shape)
        self.device = volume.device
        shape = (ctypes.c_int64 * 4)(*self.shape)
        nbytes = int(lib().noa_hip_volume_blocks_bytes(shape))
        self.blocks = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=volume.device)
        self.update(volume)

    @staticmethod
    def _check(volume):
        if (not isinstance(volume, torch.Tensor) or volume.device.type != "cuda" or volume.dtype != torch.complex64
                or volume.dim() != 4 or volume.shape[0] != 1 or volume.stride(-1) != 1):
            raise NoaError("VolumeBlocks: expected a complex64 device volume (1, D, H, W/2+1) with a unit x stride")

    def update(self, volume: torch.Tensor) -> None:
        self._check(volume)
        if tuple(volume.shape) != self.shape or volume.device != self.device:
            raise NoaError(f"VolumeBlocks.update: expected shape {self.shape} on {self.device}")
        keep: list = []
        v = _view(volume, keep, abi.F32, abi.C32, False, "volume")
        stream = torch.cuda.current_stream(volume.device).cuda_stream
        if lib().noa_hip_volume_blocks(ctypes.byref(v), self.blocks.data_ptr(), self.blocks.numel(), stream) != abi.OK:
            raise NoaError(last_error())


class ComplexBF16:
    """A complex array stored as {bf16 real, bf16 imag} pairs: a bfloat16 tensor whose last dimension is 2."""

    def __init__(self, pairs: torch.Tensor):
        if pairs.dtype != torch.bfloat16 or pairs.shape[-1] != 2 or pairs.stride(-1) != 1:
            raise NoaError("ComplexBF16: expected a bfloat16 tensor with a trailing dimension of 2")
        self.pairs = pairs

    @staticmethod
    def from_complex(x: torch.Tensor) -> "ComplexBF16":
        return ComplexBF16(torch.view_as_real(x.to(torch.complex64)).to(torch.bfloat16).contiguous())

    @property
    def shape(self):
        return self.pairs.shape[:-1]


class Quaternions:
    """Quaternion<f32|f64> values in {z, y, x, w} order (xform/core/Quaternion.hpp:15-23): a (B, 4) or (4,) array."""

    def __init__(self, zyxw):
        self.zyxw = zyxw


class Mat33Array:
    """A batch of Mat33 in noa's memory layout: rows of 3 elements, 48-byte (f32) or 80-byte (f64) stride
    (alignas(16) Mat, src/noa/base/Mat.hpp:177). Build it once to avoid a per-call repack."""

    def __init__(self, matrices: torch.Tensor | np.ndarray, device: torch.device | str | None = None):
        m = torch.as_tensor(matrices)
        if m.ndim != 3 or tuple(m.shape[1:]) != (3, 3) or m.dtype not in (torch.float32, torch.float64):
            raise NoaError("Mat33Array: expected a (B, 3, 3) float32 or float64 array")
        dev = torch.device(device) if device is not None else (m.device if m.device.type == "cuda" else torch.device("cuda"))
        pad = 12 if m.dtype == torch.float32 else 10
        packed = torch.zeros((m.shape[0], pad), dtype=m.dtype, device=dev)
        packed[:, :9] = m.reshape(m.shape[0], 9).to(dev)
        self.packed = packed
        self.count = m.shape[0]
        self.kind = abi.MAT33_F32 if m.dtype == torch.float32 else abi.MAT33_F64


# ---------------------------------------------------------------------------------------------------------------
# Poses (xform/core/Euler.cpp:218-274, xform/core/Transform.hpp:82-129, xform/core/Quaternion.hpp:29-94).
# ---------------------------------------------------------------------------------------------------------------
def rotate_z(angle, dtype=np.float64):
    s, c = np.sin(angle), np.cos(angle)
    return np.array([[1, 0, 0], [0, c, s], [0, -s, c]], dtype=dtype)


def rotate_y(angle, dtype=np.float64):
    s, c = np.sin(angle), np.cos(angle)
    return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]], dtype=dtype)


def rotate_x(angle, dtype=np.float64):
    s, c = np.sin(angle), np.cos(angle)
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]], dtype=dtype)


_AXIS = {"Z": rotate_z, "Y": rotate_y, "X": rotate_x}


def euler2matrix(angles, axes: str = "zyx", intrinsic: bool = True, right_handed: bool = True,
                 dtype=np.float64) -> np.ndarray:
    """DHW rotation matrices from Euler angles in radians, for an array of (..., 3) angles.

    Same conventions as noa::xform::euler2matrix: 12 axis orders, `intrinsic ? r1*r2*r3 : r3*r2*r1`, and
    `right_handed=False` negates the angles. The trig is evaluated in `dtype` (f32 like noa's f32 overload).
    """
    ax = axes.strip().upper()
    if len(ax) != 3 or any(a not in "ZYX" for a in ax) or ax[0] == ax[1] or ax[1] == ax[2]:
        raise NoaError(f'Axes "{ax}" are not valid')
    a = np.asarray(angles, dtype=dtype)
    if not right_handed:
        a = a * dtype(-1)
    flat = a.reshape(-1, 3)
    out = np.empty((flat.shape[0], 3, 3), dtype=dtype)
    for i, (a0, a1, a2) in enumerate(flat):
        r1 = _AXIS[ax[0]](a0, dtype)
        r2 = _AXIS[ax[1]](a1, dtype)
        r3 = _AXIS[ax[2]](a2, dtype)
        out[i] = (r1 @ r2 @ r3) if intrinsic else (r3 @ r2 @ r1)
    return out.reshape(a.shape[:-1] + (3, 3))


def quaternion_from_matrix(m) -> np.ndarray:
    """Quaternion::from_matrix (Quaternion.hpp:29-45) for (..., 3, 3) matrices; returns (..., 4) zyxw."""
    m = np.asarray(m)
    w = np.sqrt(np.maximum(0, 1 + m[..., 2, 2] + m[..., 1, 1] + m[..., 0, 0])) / 2
    z = np.sqrt(np.maximum(0, 1 - m[..., 2, 2] - m[..., 1, 1] + m[..., 0, 0])) / 2
    y = np.sqrt(np.maximum(0, 1 - m[..., 2, 2] + m[..., 1, 1] - m[..., 0, 0])) / 2
    x = np.sqrt(np.maximum(0, 1 + m[..., 2, 2] - m[..., 1, 1] - m[..., 0, 0])) / 2
    z = np.copysign(z, m[..., 1, 2] - m[..., 2, 1])
    y = np.copysign(y, m[..., 2, 0] - m[..., 0, 2])
    x = np.copysign(x, m[..., 0, 1] - m[..., 1, 0])
    return np.stack([z, y, x, w], axis=-1).astype(m.dtype)


def quaternion_to_matrix(q) -> np.ndarray:
    """Quaternion::to_matrix (Quaternion.hpp:69-94) for (..., 4) zyxw quaternions."""
    q = np.asarray(q)
    z, y, x, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    tx, ty, tz = 2 * x, 2 * y, 2 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    m = np.empty(q.shape[:-1] + (3, 3), dtype=q.dtype)
    m[..., 0, 0] = 1 - (txx + tyy)
    m[..., 0, 1] = tyz + twx
    m[..., 0, 2] = txz - twy
    m[..., 1, 0] = tyz - twx
    m[..., 1, 1] = 1 - (txx + tzz)
    m[..., 1, 2] = txy + twz
    m[..., 2, 0] = txz + twy
    m[..., 2, 1] = txy - twz
    m[..., 2, 2] = 1 - (tyy + tzz)
    return m


# ---------------------------------------------------------------------------------------------------------------
# Argument conversion.
# ---------------------------------------------------------------------------------------------------------------
_TORCH_DTYPES = {
    torch.float32: abi.F32,
    torch.float64: abi.F64,
    torch.complex64: abi.C32,
    torch.complex128: abi.C64,
    torch.complex32: abi.C16,
}


def _pad4(shape, strides):
    shape = list(shape)
    strides = list(strides)
    if len(shape) > 4:
        raise NoaError(f"arrays are at most 4D (BDHW), got shape={tuple(shape)}")
    while len(shape) < 4:
        strides.insert(0, (shape[0] * strides[0]) if shape else 1)
        shape.insert(0, 1)
    return shape, strides


def _view(obj: Any, keep: list, real_dtype: int, cx_dtype: int, allow_value: bool, name: str) -> abi.View:
    v = abi.View()
    if obj is None:
        v.dtype = abi.DTYPE_NONE
        return v
    if isinstance(obj, Texture):
        obj = obj.array
    if isinstance(obj, (int, float, complex, np.number)) and not isinstance(obj, bool):
        if not allow_value:
            raise NoaError(f"{name}: a constant value is not accepted here")
        c = complex(obj)
        v.is_value = 1
        v.dtype = cx_dtype if isinstance(obj, (complex, np.complexfloating)) else real_dtype
        v.value[0] = c.real
        v.value[1] = c.imag
        for i in range(4):
            v.shape[i] = 1
        return v
    if isinstance(obj, VolumeBlocks):
        # The block copy: the volume's rfft shape with contiguous strides (the layout fixes the offsets).
        d, h, w = obj.shape[1], obj.shape[2], obj.shape[3]
        v.data = obj.blocks.data_ptr()
        v.dtype = abi.C32_BLOCKS
        for i, (n_, s_) in enumerate(zip(obj.shape, (d * h * w, h * w, w, 1))):
            v.shape[i] = n_
            v.strides[i] = s_
        keep.append(obj)
        return v
    if isinstance(obj, ComplexBF16):
        t = obj.pairs
        if t.device.type != "cuda":
            raise NoaError(f"{name}: the array should be on a HIP device")
        shape, strides = _pad4(t.shape[:-1], [s // 2 for s in t.stride()[:-1]])
        if any(s % 2 for s in t.stride()[:-1]):
            raise NoaError(f"{name}: bf16 complex pairs must be 4-byte aligned")
        v.data = t.data_ptr()
        v.dtype = abi.CBF16
    elif isinstance(obj, torch.Tensor):
        if obj.device.type != "cuda":
            raise NoaError(f"The arrays should be on the same device, but got {name}:device={obj.device}")
        if obj.dtype not in _TORCH_DTYPES:
            raise NoaError(f"{name}: unsupported dtype {obj.dtype}")
        shape, strides = _pad4(obj.shape, obj.stride())
        v.data = obj.data_ptr() if obj.numel() else None
        v.dtype = _TORCH_DTYPES[obj.dtype]
    else:
        raise NoaError(f"{name}: expected a device tensor, a Texture, a value or None, got {type(obj).__name__}")
    for i in range(4):
        v.shape[i] = shape[i]
        v.strides[i] = strides[i]
    keep.append(obj)
    return v


def _transform(obj: Any, keep: list, count: int, device: torch.device, is_rotation: bool, name: str) -> abi.Transform:
    t = abi.Transform()
    if obj is None:
        if is_rotation:
            raise NoaError(f"{name} should not be empty")
        t.kind = abi.XFORM_NONE
        return t
    if isinstance(obj, Mat33Array):
        if not is_rotation:
            raise NoaError(f"{name}: expected Mat22")
        t.data = obj.packed.data_ptr()
        t.kind = obj.kind
        t.batched = 1
        t.count = obj.count
        keep.append(obj)
        return t
    quat = isinstance(obj, Quaternions)
    arr = obj.zyxw if quat else obj
    if isinstance(arr, torch.Tensor) and arr.device.type == "cuda":
        tensor = arr
    else:
        tensor = None
        host = np.asarray(arr.cpu() if isinstance(arr, torch.Tensor) else arr)
        if host.dtype not in (np.float32, np.float64):
            host = host.astype(np.float64)
    dtype = tensor.dtype if tensor is not None else (torch.float32 if host.dtype == np.float32 else torch.float64)
    f64 = dtype == torch.float64
    if quat:
        elem = (4,)
        kind = abi.QUAT_F64 if f64 else abi.QUAT_F32
    elif is_rotation:
        elem = (3, 3)
        kind = abi.MAT33_F64 if f64 else abi.MAT33_F32
    else:
        elem = (2, 2)
        kind = abi.MAT22_F64 if f64 else abi.MAT22_F32
    shape = tuple(tensor.shape if tensor is not None else host.shape)
    if shape == elem:
        # One value broadcast over the batch: passed by value (host memory).
        vals = (tensor.detach().cpu().numpy() if tensor is not None else host).reshape(-1)
        buf = (ctypes.c_double * len(vals))(*vals.astype(np.float64)) if f64 else \
            (ctypes.c_float * len(vals))(*vals.astype(np.float32))
        keep.append(buf)
        t.data = ctypes.cast(buf, ctypes.c_void_p).value
        t.kind = kind
        t.batched = 0
        t.count = 1
        return t
    if len(shape) != len(elem) + 1 or shape[1:] != elem:
        raise NoaError(f"{name}: expected {elem} or (B,)+{elem}, got {shape}")
    n = shape[0]
    if tensor is None:
        tensor = torch.from_numpy(np.ascontiguousarray(host)).to(device)
    elif tensor.device != device:
        raise NoaError(f"{name} should be on the compute device")
    if kind in (abi.MAT33_F32, abi.MAT33_F64):
        pad = 12 if kind == abi.MAT33_F32 else 10
        packed = torch.zeros((n, pad), dtype=dtype, device=device)
        packed[:, :9] = tensor.reshape(n, 9)
    else:
        packed = tensor.reshape(n, -1).contiguous()
    keep.append(packed)
    t.data = packed.data_ptr()
    t.kind = kind
    t.batched = 1
    t.count = n
    return t


def _device_of(*objs) -> torch.device:
    for o in objs:
        if isinstance(o, Texture):
            o = o.array
        if isinstance(o, ComplexBF16):
            o = o.pairs
        if isinstance(o, torch.Tensor):
            return o.device
    raise NoaError("Empty array detected")


def _value_dtypes(output: Any) -> tuple[int, int]:
    if isinstance(output, torch.Tensor) and output.dtype in (torch.float64, torch.complex128):
        return abi.F64, abi.C64
    return abi.F32, abi.C32


def _shape4(s, name):
    s = tuple(int(v) for v in s)
    if len(s) != 4:
        raise NoaError(f"{name} should be a BDHW shape, got {s}")
    return s


def _call(fn_name: str, args: abi.ProjectArgs, device: torch.device) -> None:
    stream = torch.cuda.current_stream(device).cuda_stream
    status = getattr(lib(), fn_name)(ctypes.byref(args), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def _fill(args, *, remap, inp, inp_w, out, out_w, in_shape, out_shape, rotation, scaling, count, opts, keep,
          allow_value, interp):
    device = _device_of(out)
    real_t, cx_t = _value_dtypes(out)
    args.input = _view(inp, keep, real_t, cx_t, allow_value, "input")
    args.input_weight = _view(inp_w, keep, real_t, real_t, allow_value, "input_weight")
    args.output = _view(out, keep, real_t, cx_t, False, "output")
    args.output_weight = _view(out_w, keep, real_t, real_t, False, "output_weight")
    for i in range(4):
        args.input_shape[i] = in_shape[i]
        args.output_shape[i] = out_shape[i]
        args.target_shape[i] = int(tuple(opts.target_shape)[i]) if len(tuple(opts.target_shape)) == 4 else 0
    args.remap = layout(remap)
    args.interp = int(interp)
    args.rotation = _transform(rotation, keep, count, device, True, "rotation")
    args.scaling = _transform(scaling, keep, count, device, False, "scaling")
    args.fftfreq_cutoff = float(opts.fftfreq_cutoff)
    ews = tuple(opts.ews_radius) if isinstance(opts.ews_radius, (tuple, list)) else (opts.ews_radius,) * 2
    args.ews_radius[0] = float(ews[0])
    args.ews_radius[1] = float(ews[1])
    return device


def _interp_of(inp, inp_w, interp):
    """fourier_insert_extract_interp_mode (FourierProject.hpp:1203-1217)."""
    if isinstance(inp, Texture):
        interp = inp.interp
        if isinstance(inp_w, Texture) and inp_w.interp != interp:
            raise NoaError("Input textures should have the same interpolation method, "
                           f"but got input:interp={interp.name} and input_weight:interp={inp_w.interp.name}")
    elif isinstance(inp_w, Texture):
        interp = inp_w.interp
    return interp


# ---------------------------------------------------------------------------------------------------------------
# Public operators.
# ---------------------------------------------------------------------------------------------------------------
def rasterize_central_slices_3d(remap, slice, slice_weight, slice_shape, volume, volume_weight, volume_shape,
                                inv_scaling=None, fwd_rotation=None,
                                options: RasterizeCentralSlicesOptions | None = None) -> None:
    """Inserts 2D rfft central slices into a 3D rfft volume by trilinear rasterisation (accumulates).

    noa::xform::rasterize_central_slices_3d (FourierProject.hpp:1653-1692). `slice` may be a constant value.
    """
    opts = options or RasterizeCentralSlicesOptions()
    slice_shape = _shape4(slice_shape, "slice_shape")
    volume_shape = _shape4(volume_shape, "volume_shape")
    keep: list = []
    args = abi.default_args()
    device = _fill(args, remap=remap, inp=slice, inp_w=slice_weight, out=volume, out_w=volume_weight,
                   in_shape=slice_shape, out_shape=volume_shape, rotation=fwd_rotation, scaling=inv_scaling,
                   count=slice_shape[0], opts=opts, keep=keep, allow_value=True, interp=Interp.LINEAR)
    _call("noa_hip_rasterize_central_slices_3d", args, device)


def insert_central_slices_3d(remap, slice, slice_weight, slice_shape, volume, volume_weight, volume_shape,
                             fwd_scaling=None, inv_rotation=None,
                             options: InsertCentralSlicesOptions | None = None) -> None:
    """Inserts 2D rfft central slices by 2D interpolation and a windowed sinc along their normal (accumulates).

    noa::xform::insert_central_slices_3d (FourierProject.hpp:1770-1815). Takes INVERSE rotations.
    """
    opts = options or InsertCentralSlicesOptions()
    slice_shape = _shape4(slice_shape, "slice_shape")
    volume_shape = _shape4(volume_shape, "volume_shape")
    keep: list = []
    args = abi.default_args()
    interp = _interp_of(slice, slice_weight, opts.interp)
    device = _fill(args, remap=remap, inp=slice, inp_w=slice_weight, out=volume, out_w=volume_weight,
                   in_shape=slice_shape, out_shape=volume_shape, rotation=inv_rotation, scaling=fwd_scaling,
                   count=slice_shape[0], opts=opts, keep=keep, allow_value=True, interp=interp)
    args.fftfreq_sinc = float(opts.windowed_sinc.fftfreq_sinc)
    args.fftfreq_blackman = float(opts.windowed_sinc.fftfreq_blackman)
    _call("noa_hip_insert_central_slices_3d", args, device)


def extract_central_slices_3d(remap, volume, volume_weight, volume_shape, slice, slice_weight, slice_shape,
                              inv_scaling=None, fwd_rotation=None,
                              options: ExtractCentralSlicesOptions | None = None) -> None:
    """Extracts 2D rfft central slices from a 3D rfft volume by 3D interpolation (overwrites the slices).

    noa::xform::extract_central_slices_3d (FourierProject.hpp:1876-1914). A w-windowed sinc wider than one
    sample convolves the volume along w (:661-694).
    """
    opts = options or ExtractCentralSlicesOptions()
    slice_shape = _shape4(slice_shape, "slice_shape")
    volume_shape = _shape4(volume_shape, "volume_shape")
    keep: list = []
    args = abi.default_args()
    interp = _interp_of(volume, volume_weight, opts.interp)
    device = _fill(args, remap=remap, inp=volume, inp_w=volume_weight, out=slice, out_w=slice_weight,
                   in_shape=volume_shape, out_shape=slice_shape, rotation=fwd_rotation, scaling=inv_scaling,
                   count=slice_shape[0], opts=opts, keep=keep, allow_value=False, interp=interp)
    args.fftfreq_sinc = float(opts.w_windowed_sinc.fftfreq_sinc)
    args.fftfreq_blackman = float(opts.w_windowed_sinc.fftfreq_blackman)
    _call("noa_hip_extract_central_slices_3d", args, device)


def insert_and_extract_central_slices_3d(remap, input_slice, input_weight, input_slice_shape, output_slice,
                                         output_weight, output_slice_shape, input_fwd_scaling=None,
                                         input_inv_rotation=None, output_inv_scaling=None, output_fwd_rotation=None,
                                         options: InsertAndExtractCentralSlicesOptions | None = None) -> None:
    """Extracts 2D rfft central slices from a virtual volume made of other 2D rfft central slices.

    noa::xform::insert_and_extract_central_slices_3d (FourierProject.hpp:1916-2049): the input slices (INVERSE
    rotations, forward scaling) are inserted, with a windowed sinc along their normal, into a volume that is never
    stored, and sampled at the output slices (FORWARD rotations, inverse scaling). The outputs are overwritten, or
    added to with `add_to_output`; a w window wider than one sample sums the virtual volume along w (:890-924).
    `input_slice` and `input_weight` may be constant values or Textures.
    """
    opts = options or InsertAndExtractCentralSlicesOptions()
    input_slice_shape = _shape4(input_slice_shape, "input_slice_shape")
    output_slice_shape = _shape4(output_slice_shape, "output_slice_shape")
    keep: list = []
    args = abi.default_insert_extract_args()
    device = _device_of(output_slice)
    real_t, cx_t = _value_dtypes(output_slice)
    args.input = _view(input_slice, keep, real_t, cx_t, True, "input")
    args.input_weight = _view(input_weight, keep, real_t, real_t, True, "input_weight")
    args.output = _view(output_slice, keep, real_t, cx_t, False, "output")
    args.output_weight = _view(output_weight, keep, real_t, real_t, False, "output_weight")
    for i in range(4):
        args.input_shape[i] = input_slice_shape[i]
        args.output_shape[i] = output_slice_shape[i]
    args.remap = layout(remap)
    args.interp = int(_interp_of(input_slice, input_weight, opts.interp))
    args.input_scaling = _transform(input_fwd_scaling, keep, input_slice_shape[0], device, False, "input_scaling")
    args.input_rotation = _transform(input_inv_rotation, keep, input_slice_shape[0], device, True, "input_rotation")
    args.output_scaling = _transform(output_inv_scaling, keep, output_slice_shape[0], device, False,
                                     "output_scaling")
    args.output_rotation = _transform(output_fwd_rotation, keep, output_slice_shape[0], device, True,
                                      "output_rotation")
    args.input_fftfreq_sinc = float(opts.input_windowed_sinc.fftfreq_sinc)
    args.input_fftfreq_blackman = float(opts.input_windowed_sinc.fftfreq_blackman)
    args.w_fftfreq_sinc = float(opts.w_windowed_sinc.fftfreq_sinc)
    args.w_fftfreq_blackman = float(opts.w_windowed_sinc.fftfreq_blackman)
    args.fftfreq_cutoff = float(opts.fftfreq_cutoff)
    ews = tuple(opts.ews_radius) if isinstance(opts.ews_radius, (tuple, list)) else (opts.ews_radius,) * 2
    args.ews_radius[0] = float(ews[0])
    args.ews_radius[1] = float(ews[1])
    args.add_to_output = int(bool(opts.add_to_output))
    args.correct_weights = int(bool(opts.correct_weights))
    stream = torch.cuda.current_stream(device).cuda_stream
    status = lib().noa_hip_insert_and_extract_central_slices_3d(ctypes.byref(args), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def _shift2(shifts, count: int, device: torch.device, keep: list, cutoff: float) -> abi.Shift2:
    """One (y, x) shift (host value) or (B, 2) shifts (device array), float32 or float64."""
    sh = abi.Shift2()
    sh.fftfreq_cutoff = float(cutoff)
    if isinstance(shifts, torch.Tensor) and shifts.ndim == 2:
        if shifts.device != device:
            raise NoaError("The shift and output arrays must be on the same device")
        if shifts.dtype not in (torch.float32, torch.float64):
            raise NoaError("shifts should be float32 or float64")
        t = shifts.contiguous()
        keep.append(t)
        sh.data = t.data_ptr()
        sh.dtype = abi.F64 if t.dtype == torch.float64 else abi.F32
        sh.batched = 1
        sh.count = t.shape[0]
        return sh
    a = np.asarray(shifts.cpu() if isinstance(shifts, torch.Tensor) else shifts)
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float64)
    if a.ndim == 2:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
        return _shift2(t, count, device, keep, cutoff)
    if a.shape != (2,):
        raise NoaError(f"shifts: expected one (y, x) shift or (B, 2) shifts, got {a.shape}")
    buf = np.ascontiguousarray(a)
    keep.append(buf)
    sh.data = buf.ctypes.data
    sh.dtype = abi.F64 if buf.dtype == np.float64 else abi.F32
    sh.batched = 0
    sh.count = 1
    return sh


def extract_central_slices_3d_shifted(remap, volume, volume_weight, volume_shape, slice, slice_weight, slice_shape,
                                      inv_scaling=None, fwd_rotation=None, shifts=None, shift_fftfreq_cutoff=1.0,
                                      options: ExtractCentralSlicesOptions | None = None) -> None:
    """extract_central_slices_3d followed by noa::signal::phase_shift_2d of the slices in their own layout, fused
    into the extraction (the slices are written once). `shifts`: (y, x) in pixels, one or one per slice."""
    opts = options or ExtractCentralSlicesOptions()
    slice_shape = _shape4(slice_shape, "slice_shape")
    volume_shape = _shape4(volume_shape, "volume_shape")
    keep: list = []
    args = abi.default_args()
    interp = _interp_of(volume, volume_weight, opts.interp)
    device = _fill(args, remap=remap, inp=volume, inp_w=volume_weight, out=slice, out_w=slice_weight,
                   in_shape=volume_shape, out_shape=slice_shape, rotation=fwd_rotation, scaling=inv_scaling,
                   count=slice_shape[0], opts=opts, keep=keep, allow_value=False, interp=interp)
    args.fftfreq_sinc = float(opts.w_windowed_sinc.fftfreq_sinc)
    args.fftfreq_blackman = float(opts.w_windowed_sinc.fftfreq_blackman)
    sh = _shift2(shifts, slice_shape[0], device, keep, shift_fftfreq_cutoff)
    stream = torch.cuda.current_stream(device).cuda_stream
    status = lib().noa_hip_extract_central_slices_3d_shifted(ctypes.byref(args), ctypes.byref(sh),
                                                             ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def phase_shift_2d(remap, input, output, shape, shifts, fftfreq_cutoff: float = 1.0) -> None:
    """noa::signal::phase_shift_2d<REMAP> (signal/PhaseShift.hpp:215-289) for the rfft layouts h2h, h2hc, hc2h,
    hc2hc: output = input * exp(-2 pi i shift . f). `input` may be None (the phase shifts are written) or `output`
    itself when the layout does not change; shifts: one (y, x) pair or one per batch."""
    shape = _shape4(shape, "shape")
    if not isinstance(output, torch.Tensor) or output.device.type != "cuda":
        raise NoaError("phase_shift_2d: the output should be a device tensor")
    keep: list = []
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input") if input is not None else abi.View()
    sh = _shift2(shifts, shape[0], output.device, keep, fftfreq_cutoff)
    c_shape = (ctypes.c_int64 * 4)(*shape)
    stream = torch.cuda.current_stream(output.device).cuda_stream
    status = lib().noa_hip_phase_shift_2d(ctypes.byref(in_v), ctypes.byref(out_v), c_shape, layout(remap),
                                          ctypes.byref(sh), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def remap(remap, input, output, shape) -> None:
    """noa::fft::remap (fft/Remap.hpp:205-340) between the half-to-half layouts h2h, h2hc, hc2h, hc2hc of rfft
    arrays (B, D, H, W//2+1): the D and H axes are fftshifted or ifftshifted."""
    shape = _shape4(shape, "shape")
    if not isinstance(output, torch.Tensor) or output.device.type != "cuda":
        raise NoaError("remap: the output should be a device tensor")
    keep: list = []
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input")
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    c_shape = (ctypes.c_int64 * 4)(*shape)
    stream = torch.cuda.current_stream(output.device).cuda_stream
    status = lib().noa_hip_remap(layout(remap), ctypes.byref(in_v), ctypes.byref(out_v), c_shape,
                                 ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def fourier_interpolation_correction(input, output, interp: Interp = Interp.LINEAR,
                                     post_correction: bool = True) -> None:
    """noa::xform::fourier_interpolation_correction (FourierProject.hpp:2062-2087): divides (post) or multiplies
    (pre) a real-space image or volume by linear interpolation's transfer function sinc^2(pi r)."""
    if not isinstance(output, torch.Tensor) or output.device.type != "cuda":
        raise NoaError("fourier_interpolation_correction: the output should be a device tensor")
    keep: list = []
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input")
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    stream = torch.cuda.current_stream(output.device).cuda_stream
    status = lib().noa_hip_fourier_interpolation_correction(ctypes.byref(in_v), ctypes.byref(out_v), int(interp),
                                                            int(bool(post_correction)), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def cubic_bspline_prefilter(input, output) -> None:
    """noa::xform::cubic_bspline_prefilter (src/noa/xform/CubicBSplinePrefilter.hpp:24): prefilters `input` into
    `output` (f32, f64, c64 or c128 device tensors of up to 4 dimensions, BDHW) so that CUBIC_BSPLINE
    interpolation of `output` passes through the samples of `input`. The input may be broadcast into the output's
    shape and may be the output itself (in place)."""
    if not isinstance(output, torch.Tensor) or output.device.type != "cuda":
        raise NoaError("cubic_bspline_prefilter: the output should be a device tensor")
    keep: list = []
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input")
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    stream = torch.cuda.current_stream(output.device).cuda_stream
    status = lib().noa_hip_cubic_bspline_prefilter(ctypes.byref(in_v), ctypes.byref(out_v), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def rfft_shape(shape: Sequence[int]) -> tuple[int, int, int, int]:
    """Shape4::rfft (src/noa/runtime/core/Shape.hpp:380): (B, D, H, W//2+1)."""
    b, d, h, w = _shape4(shape, "shape")
    return (b, d, h, w // 2 + 1)
