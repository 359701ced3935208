"""Host-side mirror of the parts of ``noa::signal`` on either side of the central-slice path.

- :class:`CTFAnisotropic` builds noa's ``CTFAnisotropic<T>`` objects (src/noa/signal/core/CTF.hpp:291-575) with the
  constructor's arithmetic in T; :func:`ctf_anisotropic` is ``noa::signal::ctf_anisotropic<REMAP>``
  (src/noa/signal/CTF.hpp:430-500) for 2D rfft layouts; :func:`ctf_anisotropic_weighted` writes the CTF-multiplied
  slices and the squared CTF in one pass (the two calls that precede a CTF-weighted insertion).
- :class:`CTFIsotropic` builds noa's ``CTFIsotropic<T>`` (signal/core/CTF.hpp:30-272); :func:`ctf_isotropic` is
  ``noa::signal::ctf_isotropic<REMAP>`` (signal/CTF.hpp:216-421) for 1D, 2D and 3D spectra, half and full layouts.
- :func:`phase_shift_2d` is ``noa::signal::phase_shift_2d`` (re-exported from :mod:`noa_amd.xform`).

Everything is computed by the HIP library through its C ABI; nothing here computes a result on the host.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import abi
from ._lib import last_error, lib
from .xform import NoaError, _shape4, _view, layout, phase_shift_2d  # noqa: F401


@dataclass
class CTFOptions:
    """CTFOptions (signal/CTF.hpp:196-209): fftfreq_range = (start, stop, endpoint); stop <= 0 is the highest
    frequency of each dimension."""
    fftfreq_range: tuple = (0.0, -1.0, True)
    ctf_abs: bool = False
    ctf_squared: bool = False


def relativistic_electron_wavelength(energy: float) -> float:
    """Electron wavelength in meters for an acceleration potential in volts (signal/core/CTF.hpp:12-25)."""
    h, c, m0, e = 6.62607015e-34, 299792458.0, 9.1093837015e-31, 1.602176634e-19
    ev = e * energy
    return (h * c) / math.sqrt(ev * (2 * m0 * c * c + ev))


class CTFAnisotropic:
    """noa::signal::CTFAnisotropic<T> (T = float32 or float64).

    pixel_size: (h, w) in A/px; defocus: (value, astigmatism, angle) in um, um, rad (positive underfocus);
    voltage in kV; amplitude fraction; cs in mm; phase_shift in rad; bfactor in A^2; scale. The stored members
    follow noa's constructor (:318-345): defocus and astigmatism in angstroms with the sign flipped, cs in
    angstroms, voltage in volts, bfactor / 4, k1 = pi lambda, k2 = pi / 2 cs lambda^3, k3 = atan(a / sqrt(1 - a^2)).
    """

    def __init__(self, pixel_size, defocus, voltage, amplitude, cs, phase_shift=0.0, bfactor=0.0, scale=1.0,
                 dtype=np.float32):
        t = np.dtype(dtype).type
        if t not in (np.float32, np.float64):
            raise NoaError("CTFAnisotropic: float32 or float64")
        self.dtype = np.dtype(t)
        ps = (t(pixel_size[0]), t(pixel_size[1]))
        d_value, d_astig, d_angle = (t(v) for v in defocus)
        self.pixel_size = ps
        self.phase_shift = t(phase_shift)
        self.cs = t(cs) * t(1e7)
        self.voltage = t(voltage) * t(1e3)
        self.amplitude = t(amplitude)
        self.scale = t(scale)
        self.defocus_value = -d_value * t(1e4)
        self.defocus_astigmatism = -d_astig * t(1e4)
        self.defocus_angle = d_angle
        self.bfactor_forth = t(bfactor) / t(4)
        lam = relativistic_electron_wavelength(float(self.voltage)) * 1e10
        self.k1 = t(math.pi * lam)
        self.k2 = t(math.pi * 0.5 * float(self.cs) * lam * lam * lam)
        a = self.amplitude
        self.k3 = t(np.arctan(a / np.sqrt(t(1) - a * a)))

    def to_struct(self):
        cls = abi.CtfAnisotropicF64 if self.dtype == np.float64 else abi.CtfAnisotropicF32
        s = cls()
        s.pixel_size[0], s.pixel_size[1] = (float(v) for v in self.pixel_size)
        for name in ("defocus_value", "defocus_astigmatism", "defocus_angle", "phase_shift", "cs", "voltage",
                     "amplitude", "bfactor_forth", "scale", "k1", "k2", "k3"):
            setattr(s, name, float(getattr(self, name)))
        return s

    def to_bytes(self) -> bytes:
        return bytes(self.to_struct())


class CTFIsotropic:
    """noa::signal::CTFIsotropic<T> (T = float32 or float64).

    pixel_size in A/px; defocus in um (positive underfocus); voltage in kV; amplitude fraction; cs in mm;
    phase_shift in rad; bfactor in A^2; scale. The stored members follow noa's constructor (:61-82): defocus in
    angstroms with the sign flipped, cs in angstroms, voltage in volts, bfactor / 4, k1, k2, k3 as for
    CTFAnisotropic. ``CTFIsotropic.from_anisotropic(ctf)`` is noa's conversion (:95-106): mean pixel size, the
    defocus value without astigmatism.
    """

    def __init__(self, pixel_size, defocus, voltage, amplitude, cs, phase_shift=0.0, bfactor=0.0, scale=1.0,
                 dtype=np.float32):
        t = np.dtype(dtype).type
        if t not in (np.float32, np.float64):
            raise NoaError("CTFIsotropic: float32 or float64")
        self.dtype = np.dtype(t)
        self.pixel_size = t(pixel_size)
        self.defocus = -t(defocus) * t(1e4)
        self.phase_shift = t(phase_shift)
        self.cs = t(cs) * t(1e7)
        self.voltage = t(voltage) * t(1e3)
        self.amplitude = t(amplitude)
        self.scale = t(scale)
        self.bfactor_forth = t(bfactor) / t(4)
        lam = relativistic_electron_wavelength(float(self.voltage)) * 1e10
        self.k1 = t(math.pi * lam)
        self.k2 = t(math.pi * 0.5 * float(self.cs) * lam * lam * lam)
        a = self.amplitude
        self.k3 = t(np.arctan(a / np.sqrt(t(1) - a * a)))

    @classmethod
    def from_anisotropic(cls, ctf: "CTFAnisotropic") -> "CTFIsotropic":
        t = ctf.dtype.type   # the getters of core/CTF.hpp:384-397 (mean pixel size, defocus value, cs, voltage)
        return cls((ctf.pixel_size[0] + ctf.pixel_size[1]) / t(2), -ctf.defocus_value * t(1e-4),
                   ctf.voltage * t(1e-3), ctf.amplitude, ctf.cs * t(1e-7), ctf.phase_shift, ctf.bfactor_forth * t(4),
                   ctf.scale, dtype=ctf.dtype)

    def to_struct(self):
        cls = abi.CtfIsotropicF64 if self.dtype == np.float64 else abi.CtfIsotropicF32
        s = cls()
        for name, _ in cls._fields_:
            setattr(s, name, float(getattr(self, name)))
        return s

    def to_bytes(self) -> bytes:
        return bytes(self.to_struct())


_CTF_LAYOUTS = {"f2f": abi.F2F, "f": abi.F2F, "fc2fc": abi.FC2FC, "fc": abi.FC2FC, "fc2f": abi.FC2F,
                "f2fc": abi.F2FC}


def ctf_layout(remap) -> int:
    """noa::fft::Layout::from_string restricted to the layouts of the CTF functions (hx2hx or fx2fx)."""
    if isinstance(remap, int):
        if remap not in (abi.H2H, abi.HC2H, abi.H2HC, abi.HC2HC, abi.F2F, abi.FC2F, abi.F2FC, abi.FC2FC):
            raise NoaError(f"invalid remap value {remap}")
        return remap
    key = remap.strip().lower()
    if key in _CTF_LAYOUTS:
        return _CTF_LAYOUTS[key]
    return layout(remap)


def ctf_array(ctfs: Sequence[CTFAnisotropic], device) -> torch.Tensor:
    """A device array of noa CTF objects (one per batch), as a uint8 tensor of the objects' bytes."""
    if not ctfs:
        raise NoaError("Empty array detected")
    dt = ctfs[0].dtype
    if any(c.dtype != dt for c in ctfs):
        raise NoaError("ctf: all CTFs should have the same precision")
    raw = b"".join(c.to_bytes() for c in ctfs)
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


def _ctf_arg(ctf, count: int, device: torch.device, keep: list, kind=None) -> abi.Ctf:
    kind = kind or CTFAnisotropic
    c = abi.Ctf()
    sizes = (44, 88) if kind is CTFIsotropic else (56, 112)
    if isinstance(ctf, kind):
        s = ctf.to_struct()
        keep.append(s)
        c.data = ctypes.addressof(s)
        c.dtype = abi.F64 if ctf.dtype == np.float64 else abi.F32
        c.batched = 0
        c.count = 1
        return c
    if isinstance(ctf, (list, tuple)):
        if not all(isinstance(v, kind) for v in ctf):
            raise NoaError(f"ctf: expected {kind.__name__} objects")
        t = ctf_array(ctf, device)
        keep.append(t)
        c.data = t.data_ptr()
        c.dtype = abi.F64 if ctf[0].dtype == np.float64 else abi.F32
        c.batched = 1
        c.count = len(ctf)
        return c
    if isinstance(ctf, torch.Tensor):
        # A device array of objects built elsewhere: (count, bytes per object) uint8.
        if ctf.dtype != torch.uint8 or ctf.ndim != 2 or ctf.shape[1] not in sizes:
            raise NoaError(f"ctf: expected a (count, {sizes[0]}|{sizes[1]}) uint8 device tensor of {kind.__name__} "
                           "objects")
        if ctf.device != device:
            raise NoaError("The ctf and output arrays must be on the same device")
        t = ctf.contiguous()
        keep.append(t)
        c.data = t.data_ptr()
        c.dtype = abi.F64 if ctf.shape[1] == sizes[1] else abi.F32
        c.batched = 1
        c.count = ctf.shape[0]
        return c
    raise NoaError(f"ctf: unsupported type {type(ctf).__name__}")


def _options(options: CTFOptions | None) -> abi.CtfOptions:
    o = options or CTFOptions()
    r = tuple(o.fftfreq_range)
    start, stop = float(r[0]), float(r[1])
    endpoint = bool(r[2]) if len(r) > 2 else True
    return abi.CtfOptions(start, stop, int(endpoint), int(bool(o.ctf_abs)), int(bool(o.ctf_squared)))


def ctf_anisotropic(remap, input, output, shape, ctf, options: CTFOptions | None = None) -> None:
    """noa::signal::ctf_anisotropic<REMAP> (signal/CTF.hpp:430-500) on (B, 1, H, W) spectra, half ((B, 1, H, W/2+1)
    arrays) or full layouts: output = input * CTF (|.|^2 for complex -> real), or the CTF when `input` is None.
    `ctf`: one CTFAnisotropic (broadcast), a list of them (one per batch) or a (B, 56|112) uint8 device tensor of
    noa CTF objects."""
    shape = _shape4(shape, "shape")
    if not isinstance(output, torch.Tensor) or output.device.type != "cuda":
        raise NoaError("ctf_anisotropic: the output should be a device tensor")
    keep: list = []
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input") if input is not None else abi.View()
    c = _ctf_arg(ctf, shape[0], output.device, keep)
    opts = _options(options)
    c_shape = (ctypes.c_int64 * 4)(*shape)
    stream = torch.cuda.current_stream(output.device).cuda_stream
    status = lib().noa_hip_ctf_anisotropic_2d(ctypes.byref(in_v), ctypes.byref(out_v), c_shape, ctf_layout(remap),
                                              ctypes.byref(c), ctypes.byref(opts), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def ctf_anisotropic_weighted(remap, input, output, output_weight, shape, ctf,
                             options: CTFOptions | None = None) -> None:
    """output = input * CTF and output_weight = CTF^2 in one pass: the same results as
    ``ctf_anisotropic(remap, input, output, ...)`` followed by ``ctf_anisotropic(remap, None, output_weight, ...,
    ctf_squared=True)``."""
    shape = _shape4(shape, "shape")
    if not isinstance(output, torch.Tensor) or output.device.type != "cuda":
        raise NoaError("ctf_anisotropic_weighted: the output should be a device tensor")
    keep: list = []
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input")
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    w_v = _view(output_weight, keep, abi.F32, abi.C32, False, "output_weight")
    c = _ctf_arg(ctf, shape[0], output.device, keep)
    opts = _options(options)
    c_shape = (ctypes.c_int64 * 4)(*shape)
    stream = torch.cuda.current_stream(output.device).cuda_stream
    status = lib().noa_hip_ctf_anisotropic_2d_weighted(ctypes.byref(in_v), ctypes.byref(out_v), ctypes.byref(w_v),
                                                       c_shape, layout(remap), ctypes.byref(c), ctypes.byref(opts),
                                                       ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())


def ctf_isotropic(remap, input, output, shape, ctf, options: CTFOptions | None = None) -> None:
    """noa::signal::ctf_isotropic<REMAP> (signal/CTF.hpp:216-421) on 1D, 2D or 3D spectra (the dimensionality of the
    logical BDHW `shape`), half (last dimension W/2+1) or full layouts: output = input * CTF(|f|) (|.|^2 for
    complex -> real), or the CTF when `input` is None. 1D spectra may start the fftfreq range above 0. `ctf`: one
    CTFIsotropic (broadcast), a list of them (one per batch) or a (B, 44|88) uint8 device tensor of noa objects."""
    shape = _shape4(shape, "shape")
    if not isinstance(output, torch.Tensor) or output.device.type != "cuda":
        raise NoaError("ctf_isotropic: the output should be a device tensor")
    keep: list = []
    out_v = _view(output, keep, abi.F32, abi.C32, False, "output")
    in_v = _view(input, keep, abi.F32, abi.C32, False, "input") if input is not None else abi.View()
    c = _ctf_arg(ctf, shape[0], output.device, keep, CTFIsotropic)
    opts = _options(options)
    c_shape = (ctypes.c_int64 * 4)(*shape)
    stream = torch.cuda.current_stream(output.device).cuda_stream
    status = lib().noa_hip_ctf_isotropic(ctypes.byref(in_v), ctypes.byref(out_v), c_shape, ctf_layout(remap),
                                         ctypes.byref(c), ctypes.byref(opts), ctypes.c_void_p(stream))
    if status != abi.OK:
        raise NoaError(last_error())
