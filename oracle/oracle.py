"""TEST INFRASTRUCTURE -- Python front-end of the CPU oracle (oracle/_build/liboracle.so).

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. The functions take numpy
arrays (host memory) with the same argument order as noa_amd.xform, build the same C descriptor
(include/noa_hip.h) and run noa's CPU algorithm restated in oracle/noa_oracle.cpp.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
LIB_PATH = ROOT / "_build" / "liboracle.so"

# ctypes layout of include/noa_hip.h, restated here so that the oracle does not import the product.
from ctypes import c_double, c_int32, c_int64, c_void_p  # noqa: E402


class View(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("shape", c_int64 * 4), ("strides", c_int64 * 4), ("dtype", c_int32),
                ("is_value", c_int32), ("value", c_double * 2)]


class Transform(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("kind", c_int32), ("batched", c_int32), ("count", c_int64)]


class ProjectArgs(ctypes.Structure):
    _fields_ = [("input", View), ("input_weight", View), ("output", View), ("output_weight", View),
                ("input_shape", c_int64 * 4), ("output_shape", c_int64 * 4), ("target_shape", c_int64 * 4),
                ("remap", c_int32), ("interp", c_int32), ("rotation", Transform), ("scaling", Transform),
                ("fftfreq_cutoff", c_double), ("fftfreq_sinc", c_double), ("fftfreq_blackman", c_double),
                ("ews_radius", c_double * 2)]


class InsertExtractArgs(ctypes.Structure):
    _fields_ = [("input", View), ("input_weight", View), ("output", View), ("output_weight", View),
                ("input_shape", c_int64 * 4), ("output_shape", c_int64 * 4), ("remap", c_int32),
                ("interp", c_int32), ("input_scaling", Transform), ("input_rotation", Transform),
                ("output_scaling", Transform), ("output_rotation", Transform),
                ("input_fftfreq_sinc", c_double), ("input_fftfreq_blackman", c_double),
                ("w_fftfreq_sinc", c_double), ("w_fftfreq_blackman", c_double), ("fftfreq_cutoff", c_double),
                ("ews_radius", c_double * 2), ("add_to_output", c_int32), ("correct_weights", c_int32)]


F32, F64, C32, C64 = 1, 2, 3, 4
MAT33_F32, MAT33_F64, QUAT_F32, QUAT_F64, MAT22_F32, MAT22_F64 = 1, 2, 3, 4, 5, 6
LAYOUTS = {"h2h": 0, "h": 0, "hc2h": 4, "h2hc": 8, "hc2hc": 12, "hc": 12}

_NP_DTYPES = {np.dtype(np.float32): F32, np.dtype(np.float64): F64, np.dtype(np.complex64): C32,
              np.dtype(np.complex128): C64}

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(ROOT)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.POINTER(ProjectArgs)
        for n in ("oracle_rasterize_central_slices_3d", "oracle_insert_central_slices_3d",
                  "oracle_extract_central_slices_3d"):
            getattr(L, n).argtypes = [P, ctypes.c_int, ctypes.c_int]
            getattr(L, n).restype = ctypes.c_int
        L.oracle_insert_and_extract_central_slices_3d.argtypes = [ctypes.POINTER(InsertExtractArgs), ctypes.c_int,
                                                                  ctypes.c_int]
        L.oracle_insert_and_extract_central_slices_3d.restype = ctypes.c_int
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_index2frequency.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong]
        L.oracle_index2frequency.restype = ctypes.c_longlong
        L.oracle_frequency2index.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong]
        L.oracle_frequency2index.restype = ctypes.c_longlong
        L.oracle_windowed_sinc_f64.argtypes = [c_double] * 3
        L.oracle_windowed_sinc_f64.restype = c_double
        L.oracle_windowed_sinc_f32.argtypes = [ctypes.c_float] * 3
        L.oracle_windowed_sinc_f32.restype = ctypes.c_float
        L.oracle_blackman_window_size.argtypes = [c_double, c_double]
        L.oracle_blackman_window_size.restype = ctypes.c_longlong
        L.oracle_iwise_threads.argtypes = [ctypes.c_longlong, ctypes.c_int]
        L.oracle_iwise_threads.restype = ctypes.c_int
        L.oracle_interpolate_spectrum_3d_c32.argtypes = [c_void_p, c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                                         ctypes.c_longlong, ctypes.c_int, ctypes.c_int, c_void_p,
                                                         ctypes.c_longlong, c_void_p]
        L.oracle_interpolate_spectrum_2d_c32.argtypes = [c_void_p, c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                                         ctypes.c_longlong, ctypes.c_int, ctypes.c_int, c_void_p,
                                                         ctypes.c_longlong, c_void_p]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _pad4(a: np.ndarray):
    shape = list(a.shape)
    strides = [s // a.itemsize for s in a.strides]
    while len(shape) < 4:
        strides.insert(0, shape[0] * strides[0] if shape else 1)
        shape.insert(0, 1)
    return shape, strides


def _view(obj, keep, real_t, cx_t):
    v = View()
    if obj is None:
        return v
    if isinstance(obj, (int, float, complex, np.number)):
        c = complex(obj)
        v.is_value = 1
        v.dtype = cx_t if isinstance(obj, (complex, np.complexfloating)) else real_t
        v.value[0], v.value[1] = c.real, c.imag
        for i in range(4):
            v.shape[i] = 1
        return v
    a = np.asarray(obj)
    shape, strides = _pad4(a)
    v.data = a.ctypes.data
    v.dtype = _NP_DTYPES[a.dtype]
    for i in range(4):
        v.shape[i] = shape[i]
        v.strides[i] = strides[i]
    keep.append(a)
    return v


def _transform(obj, keep, is_rotation, quaternion=False):
    t = Transform()
    if obj is None:
        return t
    a = np.asarray(obj)
    f64 = a.dtype == np.float64
    if quaternion:
        elem, kind = (4,), (QUAT_F64 if f64 else QUAT_F32)
    elif is_rotation:
        elem, kind = (3, 3), (MAT33_F64 if f64 else MAT33_F32)
    else:
        elem, kind = (2, 2), (MAT22_F64 if f64 else MAT22_F32)
    if a.shape == elem:
        packed = np.ascontiguousarray(a.reshape(-1))
        t.batched = 0
        t.count = 1
    else:
        n = a.shape[0]
        if kind in (MAT33_F32, MAT33_F64):
            packed = np.zeros((n, 12 if kind == MAT33_F32 else 10), dtype=a.dtype)
            packed[:, :9] = a.reshape(n, 9)
        else:
            packed = np.ascontiguousarray(a.reshape(n, -1))
        t.batched = 1
        t.count = n
    keep.append(packed)
    t.data = packed.ctypes.data
    t.kind = kind
    return t


def _args(remap, inp, inp_w, in_shape, out, out_w, out_shape, rotation, scaling, cutoff, target_shape, ews,
          sinc, blackman, interp, keep, quaternion):
    a = ProjectArgs()
    real_t, cx_t = (F64, C64) if out.dtype in (np.float64, np.complex128) else (F32, C32)
    a.input = _view(inp, keep, real_t, cx_t)
    a.input_weight = _view(inp_w, keep, real_t, real_t)
    a.output = _view(out, keep, real_t, cx_t)
    a.output_weight = _view(out_w, keep, real_t, real_t)
    for i in range(4):
        a.input_shape[i] = in_shape[i]
        a.output_shape[i] = out_shape[i]
        a.target_shape[i] = target_shape[i]
    a.remap = LAYOUTS[remap.lower()] if isinstance(remap, str) else remap
    a.interp = int(interp)
    a.rotation = _transform(rotation, keep, True, quaternion)
    a.scaling = _transform(scaling, keep, False)
    a.fftfreq_cutoff = cutoff
    a.fftfreq_sinc = sinc
    a.fftfreq_blackman = blackman
    a.ews_radius[0], a.ews_radius[1] = ews
    return a


def _run(name, a, n_threads, force):
    status = getattr(lib(), name)(ctypes.byref(a), n_threads, 1 if force else 0)
    if status != 0:
        raise OracleError(lib().oracle_last_error().decode())


def rasterize_central_slices_3d(remap, slice, slice_weight, slice_shape, volume, volume_weight, volume_shape,
                                inv_scaling=None, fwd_rotation=None, *, fftfreq_cutoff=0.5,
                                target_shape=(0, 0, 0, 0), ews_radius=(0.0, 0.0), quaternion=False,
                                n_threads=1, force_threads=False):
    keep = []
    a = _args(remap, slice, slice_weight, slice_shape, volume, volume_weight, volume_shape, fwd_rotation,
              inv_scaling, fftfreq_cutoff, target_shape, ews_radius, -1.0, -1.0, 1, keep, quaternion)
    _run("oracle_rasterize_central_slices_3d", a, n_threads, force_threads)


def insert_central_slices_3d(remap, slice, slice_weight, slice_shape, volume, volume_weight, volume_shape,
                             fwd_scaling=None, inv_rotation=None, *, interp=1, fftfreq_sinc=-1.0,
                             fftfreq_blackman=-1.0, fftfreq_cutoff=0.5, target_shape=(0, 0, 0, 0),
                             ews_radius=(0.0, 0.0), quaternion=False, n_threads=1, force_threads=False):
    keep = []
    a = _args(remap, slice, slice_weight, slice_shape, volume, volume_weight, volume_shape, inv_rotation,
              fwd_scaling, fftfreq_cutoff, target_shape, ews_radius, fftfreq_sinc, fftfreq_blackman, interp, keep,
              quaternion)
    _run("oracle_insert_central_slices_3d", a, n_threads, force_threads)


def insert_central_slices_3d_voxels(remap, slice, slice_weight, slice_shape, volume_shape, inv_rotation, voxels, *,
                                    fwd_scaling=None, interp=1, fftfreq_sinc=-1.0, fftfreq_blackman=-1.0,
                                    fftfreq_cutoff=0.5, target_shape=(0, 0, 0, 0), ews_radius=(0.0, 0.0),
                                    quaternion=False, n_threads=1):
    """noa's per-voxel insertion sums (FourierProject.hpp:478-515) at the (z, y, x) voxels of `voxels` ((n, 3)
    integers), before they are added to a volume: returns (values, weights, inside), `inside` False (and both sums
    0) for voxels beyond the cutoff, which noa leaves untouched. Runs in parallel over the voxels."""
    vox = np.ascontiguousarray(np.asarray(voxels, dtype=np.int64).reshape(-1, 3))
    n = vox.shape[0]
    sl = np.asarray(slice)
    vdt = sl.dtype
    rdt = np.float64 if vdt in (np.float64, np.complex128) else np.float32
    dummy_v = np.zeros((1, 1, 1, 1), vdt)
    dummy_w = np.zeros((1, 1, 1, 1), rdt)
    keep = []
    a = _args(remap, slice, slice_weight, slice_shape, dummy_v, dummy_w, volume_shape, inv_rotation, fwd_scaling,
              fftfreq_cutoff, target_shape, ews_radius, fftfreq_sinc, fftfreq_blackman, interp, keep, quaternion)
    values = np.zeros(n, vdt)
    weights = np.zeros(n, rdt)
    inside = np.zeros(n, np.int8)
    L = lib()
    fn = L.oracle_insert_central_slices_3d_voxels
    fn.argtypes = [ctypes.POINTER(ProjectArgs), ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    fn.restype = ctypes.c_int
    status = fn(ctypes.byref(a), vox.ctypes.data, n, values.ctypes.data, weights.ctypes.data, inside.ctypes.data,
                n_threads)
    if status != 0:
        raise OracleError(L.oracle_last_error().decode())
    return values, weights, inside.astype(bool)


def extract_central_slices_3d(remap, volume, volume_weight, volume_shape, slice, slice_weight, slice_shape,
                              inv_scaling=None, fwd_rotation=None, *, interp=1, fftfreq_sinc=-1.0,
                              fftfreq_blackman=-1.0, fftfreq_cutoff=0.5, target_shape=(0, 0, 0, 0),
                              ews_radius=(0.0, 0.0), quaternion=False, n_threads=1, force_threads=False):
    keep = []
    a = _args(remap, volume, volume_weight, volume_shape, slice, slice_weight, slice_shape, fwd_rotation,
              inv_scaling, fftfreq_cutoff, target_shape, ews_radius, fftfreq_sinc, fftfreq_blackman, interp, keep,
              quaternion)
    _run("oracle_extract_central_slices_3d", a, n_threads, force_threads)


def insert_and_extract_central_slices_3d(remap, input_slice, input_weight, input_slice_shape, output_slice,
                                         output_weight, output_slice_shape, input_fwd_scaling=None,
                                         input_inv_rotation=None, output_inv_scaling=None,
                                         output_fwd_rotation=None, *, interp=1, input_fftfreq_sinc=-1.0,
                                         input_fftfreq_blackman=-1.0, w_fftfreq_sinc=-1.0, w_fftfreq_blackman=-1.0,
                                         add_to_output=False, correct_weights=False, fftfreq_cutoff=0.5,
                                         ews_radius=(0.0, 0.0), input_quaternion=False, output_quaternion=False,
                                         n_threads=1, force_threads=False):
    """FourierInsertExtract (FourierProject.hpp:727-1006) on host arrays; the output slices are updated in place."""
    keep = []
    a = InsertExtractArgs()
    out = np.asarray(output_slice)
    real_t, cx_t = (F64, C64) if out.dtype in (np.float64, np.complex128) else (F32, C32)
    a.input = _view(input_slice, keep, real_t, cx_t)
    a.input_weight = _view(input_weight, keep, real_t, real_t)
    a.output = _view(output_slice, keep, real_t, cx_t)
    a.output_weight = _view(output_weight, keep, real_t, real_t)
    for i in range(4):
        a.input_shape[i] = input_slice_shape[i]
        a.output_shape[i] = output_slice_shape[i]
    a.remap = LAYOUTS[remap.lower()] if isinstance(remap, str) else remap
    a.interp = int(interp)
    a.input_scaling = _transform(input_fwd_scaling, keep, False)
    a.input_rotation = _transform(input_inv_rotation, keep, True, input_quaternion)
    a.output_scaling = _transform(output_inv_scaling, keep, False)
    a.output_rotation = _transform(output_fwd_rotation, keep, True, output_quaternion)
    a.input_fftfreq_sinc = input_fftfreq_sinc
    a.input_fftfreq_blackman = input_fftfreq_blackman
    a.w_fftfreq_sinc = w_fftfreq_sinc
    a.w_fftfreq_blackman = w_fftfreq_blackman
    a.fftfreq_cutoff = fftfreq_cutoff
    a.ews_radius[0], a.ews_radius[1] = ews_radius
    a.add_to_output = int(bool(add_to_output))
    a.correct_weights = int(bool(correct_weights))
    _run("oracle_insert_and_extract_central_slices_3d", a, n_threads, force_threads)


def index2frequency(centered, index, size):
    return lib().oracle_index2frequency(int(centered), index, size)


def frequency2index(centered, frequency, size):
    return lib().oracle_frequency2index(int(centered), frequency, size)


def windowed_sinc(f, fs, fb, f32=False):
    return (lib().oracle_windowed_sinc_f32 if f32 else lib().oracle_windowed_sinc_f64)(f, fs, fb)


def blackman_window_size(fb, n):
    return lib().oracle_blackman_window_size(fb, n)


def iwise_threads(n, n_threads):
    return lib().oracle_iwise_threads(n, n_threads)


def interpolate_spectrum_3d(volume: np.ndarray, freqs: np.ndarray, shape, centered: bool, interp=1) -> np.ndarray:
    """volume: (D, H, W//2+1) complex64; freqs: (N, 3) float32 centred frequencies in samples."""
    v = np.ascontiguousarray(volume.reshape((1,) + volume.shape[-3:]), dtype=np.complex64)
    st = np.array([s // 8 for s in v.strides], dtype=np.int64)
    f = np.ascontiguousarray(freqs, dtype=np.float32)
    out = np.zeros(len(f), dtype=np.complex64)
    lib().oracle_interpolate_spectrum_3d_c32(v.ctypes.data, st.ctypes.data, shape[0], shape[1], shape[2],
                                             int(centered), interp, f.ctypes.data, len(f), out.ctypes.data)
    return out


def interpolate_spectrum_2d(slices: np.ndarray, batch: int, freqs: np.ndarray, shape, centered: bool,
                            interp=1) -> np.ndarray:
    """slices: (B, H, W//2+1) complex64; freqs: (N, 2) float32 centred frequencies in samples."""
    s = np.ascontiguousarray(slices, dtype=np.complex64)
    st = np.array([x // 8 for x in s.strides], dtype=np.int64)
    f = np.ascontiguousarray(freqs, dtype=np.float32)
    out = np.zeros(len(f), dtype=np.complex64)
    lib().oracle_interpolate_spectrum_2d_c32(s.ctypes.data, st.ctypes.data, batch, shape[0], shape[1],
                                             int(centered), interp, f.ctypes.data, len(f), out.ctypes.data)
    return out


def phase_shift_2d(remap, input, output, shape, shifts, fftfreq_cutoff=1.0):
    """noa::signal::phase_shift_2d<REMAP> (signal/PhaseShift.hpp:215-289) on contiguous (B, 1, H, W/2+1) arrays.

    input: complex64/complex128 array or None (the phase shifts are written); shifts: one (y, x) pair or one per
    batch (B, 2), float32 or float64 (the coordinate precision)."""
    out = np.asarray(output)
    if out.dtype not in (np.complex64, np.complex128) or not out.flags.c_contiguous:
        raise OracleError("phase_shift_2d: the output should be a contiguous complex array")
    b, d, h, wr = out.shape
    w = shape[3]
    if d != 1 or shape[0] != b or shape[1] != 1 or shape[2] != h or w // 2 + 1 != wr:
        raise OracleError("phase_shift_2d: shape mismatch")
    inp = None
    if input is not None:
        inp = np.ascontiguousarray(input, dtype=out.dtype)
        if inp.shape != out.shape:
            raise OracleError("phase_shift_2d: the input should have the output's shape")
    sh = np.asarray(shifts)
    if sh.dtype not in (np.float32, np.float64):
        sh = sh.astype(np.float64)
    batched = sh.ndim == 2
    if batched and sh.shape != (b, 2) or not batched and sh.shape != (2,):
        raise OracleError("phase_shift_2d: expected one (y, x) shift or one per batch")
    sh = np.ascontiguousarray(sh)
    fn = lib().oracle_phase_shift_2d
    fn.restype = ctypes.c_int
    status = fn(ctypes.c_void_p(inp.ctypes.data if inp is not None else 0), ctypes.c_void_p(out.ctypes.data),
                int(out.dtype == np.complex128), ctypes.c_longlong(b), ctypes.c_longlong(h), ctypes.c_longlong(w),
                LAYOUTS[remap.lower()] if isinstance(remap, str) else int(remap), ctypes.c_void_p(sh.ctypes.data),
                int(sh.dtype == np.float64), int(batched), ctypes.c_double(fftfreq_cutoff))
    if status != 0:
        raise OracleError(lib().oracle_last_error().decode())


def fourier_interpolation_correction(input, output, post_correction=True):
    """noa::xform::fourier_interpolation_correction (FourierProject.hpp:2062-2087), LINEAR, on contiguous real
    (B, D, H, W) arrays; the input may have a batch of 1 (broadcast)."""
    out = np.asarray(output)
    inp = np.ascontiguousarray(input)
    if out.dtype not in (np.float32, np.float64) or inp.dtype not in (np.float32, np.float64) or out.ndim != 4:
        raise OracleError("fourier_interpolation_correction: real (B, D, H, W) arrays expected")
    shape = (ctypes.c_longlong * 4)(*out.shape)
    lib().oracle_fourier_interpolation_correction(
        ctypes.c_void_p(inp.ctypes.data), int(inp.dtype == np.float64), ctypes.c_longlong(inp.shape[0]),
        ctypes.c_void_p(out.ctypes.data), int(out.dtype == np.float64), shape, int(bool(post_correction)))


def remap_hx2hx(remap, input, output, shape):
    """nf::remap between h2h, h2hc, hc2h, hc2hc on contiguous rfft arrays (B, D, H, W//2+1)."""
    out = np.asarray(output)
    inp = np.ascontiguousarray(input)
    sh = (ctypes.c_longlong * 4)(*shape)
    status = lib().oracle_remap_hx2hx(ctypes.c_void_p(inp.ctypes.data), ctypes.c_void_p(out.ctypes.data),
                                      inp.itemsize, sh, LAYOUTS[remap.lower()] if isinstance(remap, str) else remap)
    if status != 0:
        raise OracleError("remap_hx2hx: unsupported element size")


def cubic_bspline_prefilter(input, output):
    """noa::xform::cubic_bspline_prefilter (xform/CubicBSplinePrefilter.hpp:24, cpu/CubicBSplinePrefilter.hpp:129-158)
    on (B, D, H, W) numpy arrays of f32, f64, c64 or c128 with any strides (a view); `input is output` filters in
    place. The input may be broadcast (stride 0)."""
    out = output
    if out.ndim != 4 or input.ndim != 4 or out.dtype != input.dtype or np.dtype(out.dtype) not in _CTF_DT:
        raise OracleError("cubic_bspline_prefilter: (B, D, H, W) arrays of one of f32, f64, c64, c128 expected")
    item = out.itemsize
    ist = (ctypes.c_longlong * 4)(*[s // item for s in input.strides])
    ost = (ctypes.c_longlong * 4)(*[s // item for s in out.strides])
    sh = (ctypes.c_longlong * 4)(*out.shape)
    L = lib()
    L.oracle_cubic_bspline_prefilter.restype = ctypes.c_int
    status = L.oracle_cubic_bspline_prefilter(ctypes.c_void_p(input.ctypes.data), ist, ctypes.c_void_p(out.ctypes.data),
                                              ost, sh, _CTF_DT[np.dtype(out.dtype)])
    if status != 0:
        raise OracleError("cubic_bspline_prefilter: unsupported type")


_CTF_DT = {np.dtype(np.float32): 1, np.dtype(np.float64): 2, np.dtype(np.complex64): 3, np.dtype(np.complex128): 4}


def make_ctf(pixel_size, defocus, voltage, amplitude, cs, phase_shift=0.0, bfactor=0.0, scale=1.0, f64=False):
    """noa::signal::CTFAnisotropic<T> from physical parameters (signal/core/CTF.hpp:318-345), as raw bytes.
    defocus = (value um, astigmatism um, angle rad)."""
    L = lib()
    L.oracle_ctf_anisotropic_make.argtypes = [ctypes.c_int, c_void_p] + [c_double] * 9 + [c_void_p]
    L.oracle_ctf_anisotropic_make.restype = None
    ps = (c_double * 2)(*pixel_size)
    buf = ctypes.create_string_buffer(112 if f64 else 56)
    L.oracle_ctf_anisotropic_make(int(f64), ctypes.addressof(ps), defocus[0], defocus[1], defocus[2], voltage,
                                  amplitude, cs, phase_shift, bfactor, scale, ctypes.addressof(buf))
    return bytes(buf.raw)


def relativistic_electron_wavelength(volts):
    L = lib()
    L.oracle_relativistic_electron_wavelength.argtypes = [c_double]
    L.oracle_relativistic_electron_wavelength.restype = c_double
    return L.oracle_relativistic_electron_wavelength(volts)


def ctf_anisotropic_2d(remap, input, output, shape, ctfs, *, fftfreq_range=(0.0, -1.0, True), ctf_abs=False,
                       ctf_squared=False):
    """noa::signal::ctf_anisotropic<REMAP> (signal/CTF.hpp:14-125) on contiguous (B, 1, H, W/2+1) arrays.
    ctfs: one CTF (bytes from make_ctf) or a list of B of them; input None writes the CTF."""
    out = np.asarray(output)
    if not out.flags.c_contiguous:
        raise OracleError("ctf_anisotropic_2d: contiguous output expected")
    b, h, w = shape[0], shape[2], shape[3]
    items = ctfs if isinstance(ctfs, (list, tuple)) else [ctfs]
    f64 = len(items[0]) == 112
    raw = ctypes.create_string_buffer(b"".join(items), len(items) * len(items[0]))
    inp = None if input is None else np.ascontiguousarray(input)
    L = lib()
    L.oracle_ctf_anisotropic_2d.argtypes = [c_void_p, ctypes.c_int, ctypes.c_longlong, c_void_p, ctypes.c_int,
                                            ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int,
                                            c_void_p, ctypes.c_int, ctypes.c_int, c_double, c_double, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int]
    L.oracle_ctf_anisotropic_2d.restype = ctypes.c_int
    status = L.oracle_ctf_anisotropic_2d(
        None if inp is None else inp.ctypes.data, 0 if inp is None else _CTF_DT[inp.dtype],
        1 if inp is None else inp.shape[0], out.ctypes.data, _CTF_DT[out.dtype], b, h, w,
        LAYOUTS[remap.lower()] if isinstance(remap, str) else remap, ctypes.addressof(raw), int(f64),
        int(len(items) > 1), float(fftfreq_range[0]), float(fftfreq_range[1]),
        int(bool(fftfreq_range[2]) if len(fftfreq_range) > 2 else 1), int(bool(ctf_abs)), int(bool(ctf_squared)))
    if status != 0:
        raise OracleError(lib().oracle_last_error().decode())


CTF_LAYOUTS = {"h2h": 0, "hc2h": 4, "h2hc": 8, "hc2hc": 12, "f2f": 3, "fc2f": 7, "f2fc": 11, "fc2fc": 15}


def make_ctf_isotropic(pixel_size, defocus, voltage, amplitude, cs, phase_shift=0.0, bfactor=0.0, scale=1.0,
                       f64=False):
    """noa::signal::CTFIsotropic<T> from physical parameters (signal/core/CTF.hpp:61-82), as raw bytes."""
    L = lib()
    L.oracle_ctf_isotropic_make.argtypes = [ctypes.c_int] + [c_double] * 8 + [c_void_p]
    L.oracle_ctf_isotropic_make.restype = None
    buf = ctypes.create_string_buffer(88 if f64 else 44)
    L.oracle_ctf_isotropic_make(int(f64), pixel_size, defocus, voltage, amplitude, cs, phase_shift, bfactor, scale,
                                ctypes.addressof(buf))
    return bytes(buf.raw)


def ctf(remap, input, output, shape, ctfs, *, isotropic=True, fftfreq_range=(0.0, -1.0, True), ctf_abs=False,
        ctf_squared=False):
    """noa::signal::ctf_isotropic / ctf_anisotropic<REMAP> (signal/CTF.hpp:14-125, :216-500) on strided numpy
    arrays, every half and full layout: output = input * CTF (|.|^2 for complex -> real), or the CTF when `input` is
    None. `input` is broadcast to the output's shape. ctfs: one CTF (bytes) or a list of B of them."""
    out = output
    if not isinstance(out, np.ndarray):
        raise OracleError("ctf: numpy output expected")
    b = shape[0]
    items = ctfs if isinstance(ctfs, (list, tuple)) else [ctfs]
    size = len(items[0])
    f64 = size in (88, 112)
    raw = ctypes.create_string_buffer(b"".join(items), len(items) * size)
    inp = None
    if input is not None:
        inp = np.broadcast_to(input, out.shape)
    el = lambda a: (ctypes.c_longlong * 4)(*[st // a.itemsize for st in a.strides])
    L = lib()
    L.oracle_ctf.argtypes = [ctypes.c_int, c_void_p, ctypes.c_int, c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p,
                             c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, ctypes.c_int, c_double, c_double,
                             ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.oracle_ctf.restype = ctypes.c_int
    in_strides = el(inp) if inp is not None else (ctypes.c_longlong * 4)()
    out_strides = el(out)
    out_shape = (ctypes.c_longlong * 4)(*out.shape)
    c_shape = (ctypes.c_longlong * 4)(*shape)
    status = L.oracle_ctf(
        int(bool(isotropic)), None if inp is None else inp.ctypes.data, 0 if inp is None else _CTF_DT[inp.dtype],
        ctypes.addressof(in_strides), out.ctypes.data, _CTF_DT[out.dtype], ctypes.addressof(out_strides),
        ctypes.addressof(out_shape), ctypes.addressof(c_shape), CTF_LAYOUTS[remap.lower()], ctypes.addressof(raw), int(f64), int(len(items) > 1), float(fftfreq_range[0]),
        float(fftfreq_range[1]), int(bool(fftfreq_range[2]) if len(fftfreq_range) > 2 else 1), int(bool(ctf_abs)),
        int(bool(ctf_squared)))
    if status == 1:
        raise OracleError("ctf: the reordered iteration would leave the output")
    if status != 0:
        raise OracleError(lib().oracle_last_error().decode())


def window_weights(interp, f, f32=False):
    """interpolation_weights of CUBIC, CUBIC_BSPLINE, LANCZOS4/6/8 at fraction f (Interpolation.hpp:81-141)."""
    L = lib()
    if f32:
        out = (ctypes.c_float * 8)()
        L.oracle_window_weights_f32.argtypes = [ctypes.c_int, ctypes.c_float, c_void_p]
        n = L.oracle_window_weights_f32(int(interp), f, ctypes.addressof(out))
        return np.array(out[:n], np.float32)
    out = (c_double * 8)()
    L.oracle_window_weights_f64.argtypes = [ctypes.c_int, c_double, c_void_p]
    n = L.oracle_window_weights_f64(int(interp), f, ctypes.addressof(out))
    return np.array(out[:n], np.float64)


# ---- FFTs (noa::fft, src/noa/fft/Transform.hpp:145-349) -------------------------------------------------------
# noa's CPU path calls FFTW3 (cmake/ext/fftw.cmake, tag `main`; not vendored in the reference). Its algorithm is the
# discrete Fourier transform, X[k] = sum_n x[n] exp(sign 2 pi i k.n / N), which numpy.fft (pocketfft) evaluates
# here in float64; noa's own part is the rank selection (fft/cuda/Plan.cpp:411-470: D, H, W of the logical shape,
# 1D along W, or along H for a column vector) and the normalisation (Transform.hpp:146-158), restated below.

FFT_NORMS = {"forward": 0, "ortho": 1, "backward": 2, "none": 3}


def _fft_axes(shape):
    b, d, h, w = shape
    if d > 1:
        return (1, 2, 3)
    if h > 1 and w > 1:
        return (2, 3)
    return (2,) if (w == 1 and h > 1) else (3,)


def _fft_scale(shape, sign, norm):
    """Transform.hpp:146-158: the forward transform is scaled for FORWARD and ORTHO, the backward one for BACKWARD
    and ORTHO; by 1/N or 1/sqrt(N), N = D * H * W of the logical shape."""
    norm = FFT_NORMS[norm] if isinstance(norm, str) else int(norm)
    n = float(shape[1] * shape[2] * shape[3])
    scale = np.sqrt(n) if norm == 1 else n
    if (sign < 0 and norm in (0, 1)) or (sign > 0 and norm in (2, 1)):
        return 1.0 / scale
    return 1.0


def fft_r2c(x, norm="forward"):
    """noa::fft::r2c of a real (B, D, H, W) array: the non-redundant half along W, non-centred."""
    x = np.asarray(x, np.float64)
    axes = _fft_axes(x.shape)
    if axes == (2,):
        raise OracleError("r2c of a column vector is not part of this oracle")
    return np.fft.rfftn(x, axes=axes, norm="backward") * _fft_scale(x.shape, -1, norm)


def fft_c2r(X, shape, norm="forward"):
    """noa::fft::c2r into a real array of logical (B, D, H, W) shape."""
    X = np.asarray(X, np.complex128)
    axes = _fft_axes(shape)
    if axes == (2,):
        raise OracleError("c2r of a column vector is not part of this oracle")
    s = [shape[a] for a in axes]
    # numpy's irfftn includes 1/N; noa's unnormalised backward transform does not.
    out = np.fft.irfftn(X, s=s, axes=axes, norm="forward")
    return out * _fft_scale(shape, +1, norm)


def fft_c2c(x, sign, norm="forward"):
    """noa::fft::c2c with the exponent's sign (-1 forward, +1 backward)."""
    x = np.asarray(x, np.complex128)
    axes = _fft_axes(x.shape)
    if sign < 0:
        out = np.fft.fftn(x, axes=axes, norm="backward")
    else:
        out = np.fft.ifftn(x, axes=axes, norm="forward")
    return out * _fft_scale(x.shape, sign, norm)
