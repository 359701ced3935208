"""ctypes mirror of include/noa_hip.h (the C ABI of the central-slice library).

Only plain data definitions live here; loading the library is done by :mod:`noa_amd._lib`.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int32, c_int64, c_void_p

ABI_VERSION = 1

# Status codes.
OK = 0
ERROR_INVALID_ARGUMENT = 1
ERROR_UNSUPPORTED = 2
ERROR_RUNTIME = 3

# Element types.
DTYPE_NONE = 0
F32 = 1
F64 = 2
C32 = 3
C64 = 4
C16 = 5
CBF16 = 6
C32_BLOCKS = 7      # c32 volume in the 4x2x4-block layout of noa_hip_volume_blocks (batched extraction only)

# Remaps (noa::fft::Layout bits, src/noa/fft/core/Layout.hpp:39-68).
H2H = 0
HC2H = 4
H2HC = 8
HC2HC = 12
F2F = 3
FC2F = 7
F2FC = 11
FC2FC = 15

# Transform kinds.
XFORM_NONE = 0
MAT33_F32 = 1
MAT33_F64 = 2
QUAT_F32 = 3
QUAT_F64 = 4
MAT22_F32 = 5
MAT22_F64 = 6


class View(ctypes.Structure):
    _fields_ = [
        ("data", c_void_p),
        ("shape", c_int64 * 4),
        ("strides", c_int64 * 4),
        ("dtype", c_int32),
        ("is_value", c_int32),
        ("value", c_double * 2),
    ]


class Transform(ctypes.Structure):
    _fields_ = [
        ("data", c_void_p),
        ("kind", c_int32),
        ("batched", c_int32),
        ("count", c_int64),
    ]


class ProjectArgs(ctypes.Structure):
    _fields_ = [
        ("input", View),
        ("input_weight", View),
        ("output", View),
        ("output_weight", View),
        ("input_shape", c_int64 * 4),
        ("output_shape", c_int64 * 4),
        ("target_shape", c_int64 * 4),
        ("remap", c_int32),
        ("interp", c_int32),
        ("rotation", Transform),
        ("scaling", Transform),
        ("fftfreq_cutoff", c_double),
        ("fftfreq_sinc", c_double),
        ("fftfreq_blackman", c_double),
        ("ews_radius", c_double * 2),
    ]


def default_args() -> ProjectArgs:
    """noa's defaults, as noa_hip_project_args_init."""
    a = ProjectArgs()
    a.remap = H2H
    a.interp = 1  # LINEAR
    a.fftfreq_cutoff = 0.5
    a.fftfreq_sinc = -1.0
    a.fftfreq_blackman = -1.0
    return a


class InsertExtractArgs(ctypes.Structure):
    _fields_ = [
        ("input", View),
        ("input_weight", View),
        ("output", View),
        ("output_weight", View),
        ("input_shape", c_int64 * 4),
        ("output_shape", c_int64 * 4),
        ("remap", c_int32),
        ("interp", c_int32),
        ("input_scaling", Transform),
        ("input_rotation", Transform),
        ("output_scaling", Transform),
        ("output_rotation", Transform),
        ("input_fftfreq_sinc", c_double),
        ("input_fftfreq_blackman", c_double),
        ("w_fftfreq_sinc", c_double),
        ("w_fftfreq_blackman", c_double),
        ("fftfreq_cutoff", c_double),
        ("ews_radius", c_double * 2),
        ("add_to_output", c_int32),
        ("correct_weights", c_int32),
    ]


def default_insert_extract_args() -> InsertExtractArgs:
    """InsertAndExtractCentralSlicesOptions' defaults, as noa_hip_insert_extract_args_init."""
    a = InsertExtractArgs()
    a.remap = H2H
    a.interp = 1  # LINEAR
    a.input_fftfreq_sinc = -1.0
    a.input_fftfreq_blackman = -1.0
    a.w_fftfreq_sinc = -1.0
    a.w_fftfreq_blackman = -1.0
    a.fftfreq_cutoff = 0.5
    return a


class CtfAnisotropicF32(ctypes.Structure):
    """noa::signal::CTFAnisotropic<f32> members (signal/core/CTF.hpp:564-575): 56 bytes, 8-byte aligned."""
    _fields_ = [("pixel_size", ctypes.c_float * 2), ("defocus_value", ctypes.c_float),
                ("defocus_astigmatism", ctypes.c_float), ("defocus_angle", ctypes.c_float),
                ("phase_shift", ctypes.c_float), ("cs", ctypes.c_float), ("voltage", ctypes.c_float),
                ("amplitude", ctypes.c_float), ("bfactor_forth", ctypes.c_float), ("scale", ctypes.c_float),
                ("k1", ctypes.c_float), ("k2", ctypes.c_float), ("k3", ctypes.c_float)]


class CtfAnisotropicF64(ctypes.Structure):
    """noa::signal::CTFAnisotropic<f64>: 112 bytes, 16-byte aligned."""
    _fields_ = [(name, ctypes.c_double * 2 if name == "pixel_size" else ctypes.c_double)
                for name, _ in CtfAnisotropicF32._fields_]


class CtfIsotropicF32(ctypes.Structure):
    """noa::signal::CTFIsotropic<f32> members (signal/core/CTF.hpp:259-271): 44 bytes."""
    _fields_ = [(name, ctypes.c_float) for name in ("pixel_size", "defocus", "phase_shift", "cs", "voltage",
                                                     "amplitude", "bfactor_forth", "scale", "k1", "k2", "k3")]


class CtfIsotropicF64(ctypes.Structure):
    """noa::signal::CTFIsotropic<f64>: 88 bytes."""
    _fields_ = [(name, ctypes.c_double) for name, _ in CtfIsotropicF32._fields_]


class Ctf(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("dtype", c_int32), ("batched", c_int32), ("count", c_int64)]


class CtfOptions(ctypes.Structure):
    _fields_ = [("fftfreq_start", c_double), ("fftfreq_stop", c_double), ("endpoint", c_int32),
                ("ctf_abs", c_int32), ("ctf_squared", c_int32)]


class FftOptions(ctypes.Structure):
    _fields_ = [("norm", c_int32), ("cache_plan", c_int32), ("plan_only", c_int32),
                ("record_and_share_workspace", c_int32)]


NORM_FORWARD, NORM_ORTHO, NORM_BACKWARD, NORM_NONE = 0, 1, 2, 3
SIGN_FORWARD, SIGN_BACKWARD = -1, 1


class MrcHeader(ctypes.Structure):
    _fields_ = [("shape", c_int64 * 4), ("spacing", c_double * 3), ("dtype", c_int32), ("has_stats", c_int32),
                ("stats", c_double * 4), ("data_offset", c_int64), ("endian_swapped", c_int32),
                ("reserved", c_int32)]


IO_DTYPES = dict(UNKNOWN=0, I8=1, U8=2, I16=3, U16=4, I32=5, U32=6, I64=7, U64=8, F16=9, F32=10, F64=11, C16=12,
                 C32=13, C64=14, U4=15, CI16=16)


class Shift2(ctypes.Structure):
    _fields_ = [
        ("data", c_void_p),
        ("dtype", c_int32),
        ("batched", c_int32),
        ("count", c_int64),
        ("fftfreq_cutoff", c_double),
    ]


# Launch tuning (noa_hip_tuning): operators and the table entry.
OP_EXTRACT = 1
OP_INSERT = 2
OP_EXTRACT_BATCH = 3
OP_INSERT_BATCH = 4


class Tuning(ctypes.Structure):
    _fields_ = [("op", c_int32), ("box", c_int32), ("remap", c_int32), ("interp", c_int32), ("dtype", c_int32),
                ("values", c_int32 * 4)]


def declare(lib: ctypes.CDLL) -> ctypes.CDLL:
    """Set the argument/return types of the exported functions."""
    P = ctypes.POINTER(ProjectArgs)
    for name in ("noa_hip_rasterize_central_slices_3d", "noa_hip_insert_central_slices_3d",
                 "noa_hip_insert_central_slices_3d_noalloc", "noa_hip_extract_central_slices_3d"):
        f = getattr(lib, name)
        f.argtypes = [P, c_void_p]
        f.restype = ctypes.c_int
    lib.noa_hip_volume_blocks_bytes.argtypes = [ctypes.POINTER(c_int64)]
    lib.noa_hip_volume_blocks_bytes.restype = c_int64
    lib.noa_hip_volume_blocks.argtypes = [ctypes.POINTER(View), c_void_p, c_int64, c_void_p]
    lib.noa_hip_volume_blocks.restype = ctypes.c_int
    for name in ("noa_hip_extract_batch_prepare", "noa_hip_insert_batch_prepare"):
        f = getattr(lib, name)
        f.argtypes = [P, c_int32, c_void_p, c_int64, ctypes.POINTER(c_int64)]
        f.restype = ctypes.c_int
    for name in ("noa_hip_extract_batch_launch", "noa_hip_insert_batch_launch"):
        f = getattr(lib, name)
        f.argtypes = [c_void_p, c_void_p]
        f.restype = ctypes.c_int
    lib.noa_hip_set_tuning.argtypes = [ctypes.POINTER(Tuning), c_int32]
    lib.noa_hip_set_tuning.restype = ctypes.c_int
    lib.noa_hip_get_tuning.argtypes = [c_int32, c_int32, c_int32, c_int32, c_int32, ctypes.POINTER(c_int32)]
    lib.noa_hip_get_tuning.restype = ctypes.c_int
    lib.noa_hip_last_error.argtypes = []
    lib.noa_hip_last_error.restype = ctypes.c_char_p
    lib.noa_hip_abi_version.argtypes = []
    lib.noa_hip_abi_version.restype = ctypes.c_int
    lib.noa_hip_project_args_init.argtypes = [P]
    lib.noa_hip_project_args_init.restype = None
    S = ctypes.POINTER(Shift2)
    lib.noa_hip_extract_central_slices_3d_shifted.argtypes = [P, S, c_void_p]
    lib.noa_hip_extract_central_slices_3d_shifted.restype = ctypes.c_int
    lib.noa_hip_phase_shift_2d.argtypes = [ctypes.POINTER(View), ctypes.POINTER(View), ctypes.POINTER(c_int64),
                                           c_int32, S, c_void_p]
    lib.noa_hip_phase_shift_2d.restype = ctypes.c_int
    V = ctypes.POINTER(View)
    IE = ctypes.POINTER(InsertExtractArgs)
    lib.noa_hip_insert_and_extract_central_slices_3d.argtypes = [IE, c_void_p]
    lib.noa_hip_insert_and_extract_central_slices_3d.restype = ctypes.c_int
    lib.noa_hip_insert_extract_args_init.argtypes = [IE]
    lib.noa_hip_insert_extract_args_init.restype = None
    C = ctypes.POINTER(Ctf)
    CO = ctypes.POINTER(CtfOptions)
    lib.noa_hip_ctf_anisotropic_2d.argtypes = [V, V, ctypes.POINTER(c_int64), c_int32, C, CO, c_void_p]
    lib.noa_hip_ctf_anisotropic_2d.restype = ctypes.c_int
    lib.noa_hip_ctf_anisotropic_2d_weighted.argtypes = [V, V, V, ctypes.POINTER(c_int64), c_int32, C, CO, c_void_p]
    lib.noa_hip_ctf_anisotropic_2d_weighted.restype = ctypes.c_int
    lib.noa_hip_ctf_isotropic.argtypes = [V, V, ctypes.POINTER(c_int64), c_int32, C, CO, c_void_p]
    lib.noa_hip_ctf_isotropic.restype = ctypes.c_int
    lib.noa_hip_remap.argtypes = [c_int32, V, V, ctypes.POINTER(c_int64), c_void_p]
    lib.noa_hip_remap.restype = ctypes.c_int
    lib.noa_hip_fourier_interpolation_correction.argtypes = [V, V, c_int32, c_int32, c_void_p]
    lib.noa_hip_fourier_interpolation_correction.restype = ctypes.c_int
    lib.noa_hip_cubic_bspline_prefilter.argtypes = [V, V, c_void_p]
    lib.noa_hip_cubic_bspline_prefilter.restype = ctypes.c_int
    lib.noa_hip_mrc_read_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(MrcHeader)]
    lib.noa_hip_mrc_read_header.restype = ctypes.c_int
    lib.noa_hip_mrc_read.argtypes = [ctypes.c_char_p, V, c_int32, c_void_p]
    lib.noa_hip_mrc_read.restype = ctypes.c_int
    lib.noa_hip_mrc_write.argtypes = [ctypes.c_char_p, V, c_void_p, c_int32, c_void_p, c_int32, c_void_p]
    lib.noa_hip_mrc_write.restype = ctypes.c_int
    FO = ctypes.POINTER(FftOptions)
    lib.noa_hip_fft_options_init.argtypes = [FO]
    lib.noa_hip_fft_options_init.restype = None
    for name in ("noa_hip_fft_r2c", "noa_hip_fft_c2r"):
        getattr(lib, name).argtypes = [V, V, FO, c_void_p]
        getattr(lib, name).restype = ctypes.c_int
    lib.noa_hip_fft_c2c.argtypes = [V, V, c_int32, FO, c_void_p]
    lib.noa_hip_fft_c2c.restype = ctypes.c_int
    for name in ("noa_hip_fft_clear_cache", "noa_hip_fft_cache_limit", "noa_hip_fft_cache_size"):
        getattr(lib, name).argtypes = [c_int32]
        getattr(lib, name).restype = c_int32
    lib.noa_hip_fft_set_cache_limit.argtypes = [c_int32, c_int32]
    lib.noa_hip_fft_set_cache_limit.restype = c_int32
    lib.noa_hip_fft_workspace_left_to_allocate.argtypes = [c_int32]
    lib.noa_hip_fft_workspace_left_to_allocate.restype = c_int64
    lib.noa_hip_fft_set_workspace.argtypes = [c_int32, c_void_p, c_int64]
    lib.noa_hip_fft_set_workspace.restype = c_int32
    return lib
