"""The C ABI without a GPU: exported symbols, struct layouts, enum values, defaults and validation errors.

Every check runs on the host; calls that reach the device checks fail there with a runtime error, after the
argument rules (which run first) have passed.
"""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "noa_hip.h"
LIB = ROOT / "noa_amd" / "_build" / "libnoa_hip.so"
FACADE_TEST = ROOT / "tests" / "_build" / "test_facade"


def header_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(noa_hip_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        pytest.fail(f"{LIB} is not built (run __graft_entry__.build())")
    from noa_amd import _lib
    return _lib.lib()


def test_library_exports_every_header_function():
    names = header_functions()
    assert {"noa_hip_rasterize_central_slices_3d", "noa_hip_insert_central_slices_3d",
            "noa_hip_extract_central_slices_3d", "noa_hip_last_error"} <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared in noa_hip.h but not exported: {missing}"


def test_struct_layout_matches_the_c_compiler(tmp_path):
    """offsetof/sizeof of every ABI struct from gcc equals the ctypes mirror."""
    from noa_amd import abi
    fields = {
        "noa_hip_view": ("View", ["data", "shape", "strides", "dtype", "is_value", "value"]),
        "noa_hip_transform": ("Transform", ["data", "kind", "batched", "count"]),
        "noa_hip_project_args": ("ProjectArgs", [f[0] for f in abi.ProjectArgs._fields_]),
        "noa_hip_insert_extract_args": ("InsertExtractArgs", [f[0] for f in abi.InsertExtractArgs._fields_]),
        "noa_hip_shift2": ("Shift2", [f[0] for f in abi.Shift2._fields_]),
        "noa_hip_ctf_anisotropic_f32": ("CtfAnisotropicF32", [f[0] for f in abi.CtfAnisotropicF32._fields_]),
        "noa_hip_ctf_anisotropic_f64": ("CtfAnisotropicF64", [f[0] for f in abi.CtfAnisotropicF64._fields_]),
        "noa_hip_ctf_isotropic_f32": ("CtfIsotropicF32", [f[0] for f in abi.CtfIsotropicF32._fields_]),
        "noa_hip_ctf_isotropic_f64": ("CtfIsotropicF64", [f[0] for f in abi.CtfIsotropicF64._fields_]),
        "noa_hip_ctf": ("Ctf", [f[0] for f in abi.Ctf._fields_]),
        "noa_hip_ctf_options": ("CtfOptions", [f[0] for f in abi.CtfOptions._fields_]),
        "noa_hip_tuning": ("Tuning", [f[0] for f in abi.Tuning._fields_]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "noa_hip.h"', "int main(void) {"]
    for c_name, (_, names) in fields.items():
        lines.append(f'printf("{c_name} size %zu\\n", sizeof({c_name}));')
        for n in names:
            lines.append(f'printf("{c_name} {n} %zu\\n", offsetof({c_name}, {n}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        struct, field, value = line.split()
        got[(struct, field)] = int(value)
    for c_name, (py_name, names) in fields.items():
        cls = getattr(abi, py_name)
        assert got[(c_name, "size")] == ctypes.sizeof(cls), c_name
        for n in names:
            assert got[(c_name, n)] == getattr(cls, n).offset, (c_name, n)


def test_enum_values_match_noa(lib):
    from noa_amd import abi
    text = HEADER.read_text()
    for name, value in [("NOA_HIP_H2H", abi.H2H), ("NOA_HIP_HC2H", abi.HC2H), ("NOA_HIP_H2HC", abi.H2HC),
                        ("NOA_HIP_HC2HC", abi.HC2HC), ("NOA_HIP_C32", abi.C32), ("NOA_HIP_MAT33_F32", abi.MAT33_F32),
                        ("NOA_HIP_QUAT_F64", abi.QUAT_F64)]:
        m = re.search(rf"\b{name}\s*=\s*(\d+)", text)
        assert m and int(m.group(1)) == value, name
    # noa's Layout bits (fft/core/Layout.hpp:39-46): SRC_FULL=1, DST_FULL=2, SRC_CENTERED=4, DST_CENTERED=8.
    assert (abi.H2H, abi.HC2H, abi.H2HC, abi.HC2HC) == (0, 4, 8, 12)
    assert (abi.F2F, abi.FC2F, abi.F2FC, abi.FC2FC) == (3, 7, 11, 15)
    for name, value in [("NOA_HIP_F2F", 3), ("NOA_HIP_FC2F", 7), ("NOA_HIP_F2FC", 11), ("NOA_HIP_FC2FC", 15)]:
        m = re.search(rf"\b{name}\s*=\s*(\d+)", text)
        assert m and int(m.group(1)) == value, name
    assert lib.noa_hip_abi_version() == abi.ABI_VERSION


def test_args_init_sets_noa_defaults(lib):
    from noa_amd import abi
    a = abi.ProjectArgs()
    lib.noa_hip_project_args_init(ctypes.byref(a))
    assert a.fftfreq_cutoff == 0.5 and a.fftfreq_sinc == -1 and a.fftfreq_blackman == -1
    assert a.interp == 1 and a.remap == abi.H2H and a.scaling.kind == abi.XFORM_NONE
    assert list(a.ews_radius) == [0, 0] and list(a.target_shape) == [0, 0, 0, 0]


def test_insert_extract_args_init_sets_noa_defaults(lib):
    """InsertAndExtractCentralSlicesOptions (FourierProject.hpp:1916-1946)."""
    from noa_amd import abi
    a = abi.InsertExtractArgs()
    lib.noa_hip_insert_extract_args_init(ctypes.byref(a))
    assert a.fftfreq_cutoff == 0.5 and a.interp == 1 and a.remap == abi.H2H
    assert (a.input_fftfreq_sinc, a.input_fftfreq_blackman, a.w_fftfreq_sinc, a.w_fftfreq_blackman) == (-1, -1, -1, -1)
    assert a.add_to_output == 0 and a.correct_weights == 0
    assert a.input_scaling.kind == abi.XFORM_NONE and a.output_scaling.kind == abi.XFORM_NONE


def _ie_args(lib, n=16, b_in=3, b_out=2):
    """A virtual-volume extraction with host buffers (never dereferenced by the checks)."""
    from noa_amd import abi
    inp = np.zeros((b_in, 1, n, n // 2 + 1), np.complex64)
    out = np.zeros((b_out, 1, n, n // 2 + 1), np.complex64)
    rin = np.zeros((b_in, 12), np.float32)
    rout = np.zeros((b_out, 12), np.float32)
    a = abi.InsertExtractArgs()
    lib.noa_hip_insert_extract_args_init(ctypes.byref(a))
    for v, arr in ((a.input, inp), (a.output, out)):
        v.data = arr.ctypes.data
        for i in range(4):
            v.shape[i] = arr.shape[i]
            v.strides[i] = arr.strides[i] // arr.itemsize
        v.dtype = abi.C32
    for i in range(4):
        a.input_shape[i] = (b_in, 1, n, n)[i]
        a.output_shape[i] = (b_out, 1, n, n)[i]
    for t, arr, cnt in ((a.input_rotation, rin, b_in), (a.output_rotation, rout, b_out)):
        t.data = arr.ctypes.data
        t.kind = abi.MAT33_F32
        t.batched = 1
        t.count = cnt
    return a, (inp, out, rin, rout)


@pytest.mark.parametrize("mutate, message", [
    (lambda a: a.output_shape.__setitem__(1, 4), "shape"),
    (lambda a: setattr(a.input_rotation, "count", 2), "input_rotation"),
    (lambda a: setattr(a.output_rotation, "count", 3), "output_rotation"),
    (lambda a: setattr(a, "correct_weights", 1) or setattr(a, "add_to_output", 1), "correct_weights"),
    (lambda a: setattr(a, "correct_weights", 1) or setattr(a, "w_fftfreq_blackman", 0.3), "correct_weights"),
    (lambda a: setattr(a, "interp", 7), "nterp"),
    (lambda a: setattr(a.output_rotation, "kind", 2), "same precision"),
])
def test_insert_extract_validation_errors(lib, mutate, message):
    from noa_amd import abi
    a, keep = _ie_args(lib)
    mutate(a)
    status = lib.noa_hip_insert_and_extract_central_slices_3d(ctypes.byref(a), None)
    err = lib.noa_hip_last_error().decode()
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_UNSUPPORTED), err
    assert message in err, err
    del keep


def test_insert_extract_valid_arguments_reach_the_device_check(lib):
    from noa_amd import abi
    a, keep = _ie_args(lib)
    status = lib.noa_hip_insert_and_extract_central_slices_3d(ctypes.byref(a), None)
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_RUNTIME)
    assert "device" in lib.noa_hip_last_error().decode().lower()
    del keep


def _args(lib, n=16, b=2, remap=0):
    """An extraction of b slices from an n^3 volume, with host buffers (never dereferenced by the checks)."""
    from noa_amd import abi
    vol = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    out = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
    rot = np.zeros((b, 12), np.float32)
    a = abi.ProjectArgs()
    lib.noa_hip_project_args_init(ctypes.byref(a))

    def view(v, arr, dtype):
        v.data = arr.ctypes.data
        for i in range(4):
            v.shape[i] = arr.shape[i]
            v.strides[i] = arr.strides[i] // arr.itemsize
        v.dtype = dtype

    view(a.input, vol, abi.C32)
    view(a.output, out, abi.C32)
    for i, s in enumerate((1, n, n, n)):
        a.input_shape[i] = s
    for i, s in enumerate((b, 1, n, n)):
        a.output_shape[i] = s
    a.remap = remap
    a.rotation.data = rot.ctypes.data
    a.rotation.kind = abi.MAT33_F32
    a.rotation.batched = 1
    a.rotation.count = b
    return a, (vol, out, rot)


@pytest.mark.parametrize("mutate, message", [
    (lambda a: a.input_shape.__setitem__(1, 8), "shape"),
    (lambda a: setattr(a, "remap", 1), "emap"),
    (lambda a: setattr(a.rotation, "count", 3), "rotation"),
    (lambda a: setattr(a, "interp", 9), "nterp"),
    (lambda a: (a.output_shape.__setitem__(0, 0), a.output.shape.__setitem__(0, 0)), "Empty array detected"),
    # (the volume placed 1 MiB below the output, so that the output's huge extent cannot overlap it; host addresses
    # are never dereferenced by the checks)
    (lambda a: (a.output.strides.__setitem__(2, 1 << 31), setattr(a.input, "data", a.output.data - (1 << 20))),
     "exceed 32-bit indexing"),
])
def test_validation_errors_before_any_device_access(lib, mutate, message):
    from noa_amd import abi
    a, keep = _args(lib)
    mutate(a)
    status = lib.noa_hip_extract_central_slices_3d(ctypes.byref(a), None)
    err = lib.noa_hip_last_error().decode()
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_UNSUPPORTED), err
    assert message in err, err
    del keep


def test_valid_arguments_reach_the_device_check(lib):
    """Host buffers pass every argument rule and are rejected only by the device-memory check."""
    from noa_amd import abi
    a, keep = _args(lib)
    status = lib.noa_hip_extract_central_slices_3d(ctypes.byref(a), None)
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_RUNTIME)
    assert "device" in lib.noa_hip_last_error().decode().lower()
    del keep


def test_cpp_facade_descriptors_and_errors():
    """include/noa_hip.hpp: noa-style arguments -> descriptor fields, and errors as noa_hip::Exception."""
    if not FACADE_TEST.exists():
        pytest.fail(f"{FACADE_TEST} is not built (run __graft_entry__.build())")
    res = subprocess.run([str(FACADE_TEST)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all facade checks passed" in res.stdout


@pytest.mark.gpu
def test_cpp_facade_on_the_gpu():
    """The facade end to end on device 0: identity-pose extraction (exact) and insertion (closed form)."""
    res = subprocess.run([str(FACADE_TEST), "--gpu"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ok gpu extraction" in res.stdout and "ok gpu insertion" in res.stdout


def test_ctf_struct_sizes_are_noas():
    """CTFAnisotropic<f32> is 56 bytes and CTFAnisotropic<f64> 112 bytes (Vec<f64,2> is 16-byte aligned)."""
    from noa_amd import abi
    assert ctypes.sizeof(abi.CtfAnisotropicF32) == 56
    assert ctypes.sizeof(abi.CtfAnisotropicF64) == 112


@pytest.mark.parametrize("mutate, message", [
    (lambda k: k["shape"].__setitem__(1, 2), "2d arrays"),
    (lambda k: k.__setitem__("remap", 1), "rfft layouts"),
    (lambda k: setattr(k["ctf"], "count", 3), "contiguous vector"),
    (lambda k: setattr(k["opt"], "fftfreq_start", 0.1), "starting fftfreq"),
    (lambda k: setattr(k["ctf"], "dtype", 3), "CTFAnisotropic"),
])
def test_ctf_validation_errors(lib, mutate, message):
    from noa_amd import abi
    b, h, w = 2, 16, 16
    out = np.zeros((b, 1, h, w // 2 + 1), np.float32)
    ctfs = np.zeros((b, 14), np.float32)
    v = abi.View()
    v.data = out.ctypes.data
    for i in range(4):
        v.shape[i] = out.shape[i]
        v.strides[i] = out.strides[i] // 4
    v.dtype = abi.F32
    c = abi.Ctf(ctfs.ctypes.data, abi.F32, 1, b)
    k = {"shape": (ctypes.c_int64 * 4)(b, 1, h, w), "remap": abi.H2H, "ctf": c, "opt": abi.CtfOptions(0, -1, 1, 0, 0)}
    mutate(k)
    status = lib.noa_hip_ctf_anisotropic_2d(None, ctypes.byref(v), k["shape"], k["remap"], ctypes.byref(k["ctf"]),
                                            ctypes.byref(k["opt"]), None)
    err = lib.noa_hip_last_error().decode()
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_UNSUPPORTED), err
    assert message in err, err


@pytest.mark.parametrize("mutate, message", [
    (lambda k: k.__setitem__("remap", 1), "should be h2h, hc2h, h2hc, hc2hc, f2f"),
    (lambda k: k["shape"].__setitem__(3, 18), "doesn't match the expected shape"),
    (lambda k: setattr(k["ctf"], "count", 3), "contiguous vector"),
    (lambda k: setattr(k["opt"], "fftfreq_start", 0.1), "starting fftfreq"),
    (lambda k: setattr(k["ctf"], "dtype", 3), "CTFIsotropic"),
    (lambda k: k.__setitem__("input", np.zeros((2, 3, 16, 9), np.float32)), "Cannot broadcast"),
    (lambda k: k.__setitem__("out", np.zeros((9, 16, 4, 2), np.float32).transpose(3, 2, 1, 0)), "innermost"),
    (lambda k: None, "not device memory"),
])
def test_ctf_isotropic_validation_errors(lib, mutate, message):
    """ctf_check_parameters (signal/CTF.hpp:143-189) and the layout rules, on host buffers (rejected before any
    device access; a valid call reaches the device check)."""
    from noa_amd import abi
    b, d, h, w = 2, 4, 16, 16
    ctfs = np.zeros((b, 11), np.float32)
    c = abi.Ctf(ctfs.ctypes.data, abi.F32, 1, b)
    k = {"shape": (ctypes.c_int64 * 4)(b, d, h, w), "remap": abi.HC2HC, "ctf": c, "opt": abi.CtfOptions(0, -1, 1, 0, 0),
         "out": np.zeros((b, d, h, w // 2 + 1), np.float32), "input": None}
    mutate(k)
    out_v = _host_view(k["out"], abi.F32)
    in_v = _host_view(k["input"], abi.F32) if k["input"] is not None else None
    status = lib.noa_hip_ctf_isotropic(None if in_v is None else ctypes.byref(in_v), ctypes.byref(out_v), k["shape"],
                                       k["remap"], ctypes.byref(k["ctf"]), ctypes.byref(k["opt"]), None)
    err = lib.noa_hip_last_error().decode()
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_UNSUPPORTED), err
    assert message in err, err


def test_ctf_isotropic_1d_allows_a_start_and_anisotropic_full_layouts_are_2d(lib):
    from noa_amd import abi
    line = np.zeros((3, 1, 1, 33), np.float32)
    ctfs = np.zeros((3, 11), np.float32)
    c = abi.Ctf(ctfs.ctypes.data, abi.F32, 1, 3)
    v = _host_view(line, abi.F32)
    status = lib.noa_hip_ctf_isotropic(None, ctypes.byref(v), (ctypes.c_int64 * 4)(3, 1, 1, 64), abi.H2H,
                                       ctypes.byref(c), ctypes.byref(abi.CtfOptions(0.1, 0.4, 1, 0, 0)), None)
    assert "not device memory" in lib.noa_hip_last_error().decode() and status == abi.ERROR_INVALID_ARGUMENT
    vol = np.zeros((1, 4, 8, 8), np.float32)
    a = np.zeros((1, 14), np.float32)
    ca = abi.Ctf(a.ctypes.data, abi.F32, 1, 1)
    status = lib.noa_hip_ctf_anisotropic_2d(None, ctypes.byref(_host_view(vol, abi.F32)),
                                            (ctypes.c_int64 * 4)(1, 4, 8, 8), abi.F2F, ctypes.byref(ca), None, None)
    assert "2d arrays" in lib.noa_hip_last_error().decode() and status == abi.ERROR_INVALID_ARGUMENT


def _host_view(arr, dtype):
    from noa_amd import abi
    v = abi.View()
    v.data = arr.ctypes.data
    a4 = arr.reshape((1,) * (4 - arr.ndim) + arr.shape)
    for i in range(4):
        v.shape[i] = a4.shape[i]
        v.strides[i] = a4.strides[i] // arr.itemsize
    v.dtype = dtype
    return v


@pytest.mark.parametrize("case, message", [
    ("f16", "f32, f64, c32 or c64"),
    ("types", "types should match"),
    ("broadcast", "Cannot broadcast"),
    ("overlap", "should not overlap"),
    ("empty", "Empty array"),
    ("outbcast", "should not be broadcast"),
    ("host", "not device memory"),
])
def test_bspline_prefilter_validation_errors(lib, case, message):
    """cubic_bspline_prefilter's argument checks (CubicBSplinePrefilter.hpp:24-35, runtime/Utils.hpp:106-117) run
    before anything touches a device; host pointers stop at the device check."""
    from noa_amd import abi
    buf = np.zeros((2, 4, 8, 16), np.float32)
    inp = _host_view(buf[0:1], abi.F32)
    out = _host_view(buf[1:2], abi.F32)
    if case == "f16":
        inp.dtype = out.dtype = abi.C16
    elif case == "types":
        inp.dtype = abi.F64
    elif case == "broadcast":
        inp.shape[2] = 3
    elif case == "overlap":
        inp = _host_view(buf.reshape(-1)[1:].reshape(1, 1, 1, -1)[..., :512], abi.F32)
        out = _host_view(buf.reshape(-1)[:512].reshape(1, 1, 1, 512), abi.F32)
    elif case == "empty":
        out.shape[3] = 0
    elif case == "outbcast":
        out.strides[1] = 0
    status = lib.noa_hip_cubic_bspline_prefilter(ctypes.byref(inp), ctypes.byref(out), None)
    err = lib.noa_hip_last_error().decode()
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_UNSUPPORTED), err
    assert message in err, err


def test_fft_options_init_sets_noas_defaults(lib):
    """FFTOptions (fft/Transform.hpp:39-88): norm FORWARD, cache_plan true, plan_only and
    record_and_share_workspace false."""
    from noa_amd import abi
    o = abi.FftOptions(9, 9, 9, 9)
    lib.noa_hip_fft_options_init(ctypes.byref(o))
    assert (o.norm, o.cache_plan, o.plan_only, o.record_and_share_workspace) == (abi.NORM_FORWARD, 1, 0, 0)


@pytest.mark.parametrize("kind, case, message", [
    ("r2c", "shape", "non-redundant shape of the complex output"),
    ("r2c", "types", "same precision"),
    ("c2r", "shape", "non-redundant shape of the complex input"),
    ("c2c", "shape", "no broadcasting allowed"),
    ("c2c", "sign", "the sign should be"),
    ("r2c", "norm", "invalid norm"),
    ("r2c", "column", "column vectors"),
    ("r2c", "depth", "vector along the depth"),
    ("r2c", "negative", "negative strides"),
    ("r2c", "host", "not device memory"),
])
def test_fft_validation_errors(lib, kind, case, message):
    """noa's checks of r2c/c2r/c2c (Transform.hpp:160-349) and this library's own (strides, vectors), all before any
    device access."""
    from noa_amd import abi
    shape = (2, 1, 8, 8)
    if case == "column":
        shape = (2, 1, 8, 1)
    elif case == "depth":
        shape = (2, 8, 1, 1)
    real = np.zeros(shape, np.float32)
    cplx = np.zeros(shape[:3] + (shape[3] // 2 + 1,), np.complex64)
    rv, cv = _host_view(real, abi.F32), _host_view(cplx, abi.C32)
    o = abi.FftOptions()
    lib.noa_hip_fft_options_init(ctypes.byref(o))
    if case == "shape":
        cv.shape[3] = 4
    elif case == "types":
        cv.dtype = abi.C64
    elif case == "norm":
        o.norm = 7
    elif case == "negative":
        rv.strides[2] = -8
    if kind == "r2c":
        status = lib.noa_hip_fft_r2c(ctypes.byref(rv), ctypes.byref(cv), ctypes.byref(o), None)
    elif kind == "c2r":
        status = lib.noa_hip_fft_c2r(ctypes.byref(cv), ctypes.byref(rv), ctypes.byref(o), None)
    else:
        c2 = np.zeros(shape, np.complex64)
        a, b = _host_view(c2, abi.C32), _host_view(c2.copy(), abi.C32)
        if case == "shape":
            b.shape[3] = 4
        status = lib.noa_hip_fft_c2c(ctypes.byref(a), ctypes.byref(b), 3 if case == "sign" else -1, ctypes.byref(o),
                                     None)
    err = lib.noa_hip_last_error().decode()
    assert status in (abi.ERROR_INVALID_ARGUMENT, abi.ERROR_UNSUPPORTED), err
    assert message in err, err


def test_tuning_table_most_specific_entry_wins(lib):
    """The launch tuning table: exact key fields beat wildcards, the later entry wins a tie, count 0 clears."""
    T = abi_tuning()
    entries = (T * 4)()
    rows = [(2, 0, -1, -1, -1, (1, 100, 0, 0)),      # insert, any key
            (2, 256, -1, -1, -1, (2, 200, 0, 0)),    # insert, box 256
            (2, 256, 0, -1, -1, (2, 300, 1, 0)),     # insert, box 256, h2h
            (2, 256, 0, -1, -1, (1, 400, 0, 0))]     # same key, later: wins the tie
    for e, (op, box, remap, interp, dtype, vals) in zip(entries, rows):
        e.op, e.box, e.remap, e.interp, e.dtype = op, box, remap, interp, dtype
        for i in range(4):
            e.values[i] = vals[i]
    assert lib.noa_hip_set_tuning(entries, 4) == 0
    v = (ctypes.c_int32 * 4)()
    assert lib.noa_hip_get_tuning(2, 256, 0, 1, 3, v) == 1 and tuple(v) == (1, 400, 0, 0)
    assert lib.noa_hip_get_tuning(2, 256, 12, 1, 3, v) == 1 and tuple(v) == (2, 200, 0, 0)
    assert lib.noa_hip_get_tuning(2, 128, 12, 1, 3, v) == 1 and tuple(v) == (1, 100, 0, 0)
    assert lib.noa_hip_get_tuning(1, 128, 12, 1, 3, v) == 0 and tuple(v) == (0, 0, 0, 0)
    bad = (T * 1)()
    bad[0].op = 9
    assert lib.noa_hip_set_tuning(bad, 1) != 0
    assert "unknown operator" in lib.noa_hip_last_error().decode()
    assert lib.noa_hip_set_tuning(entries, 0) == 0
    assert lib.noa_hip_get_tuning(2, 256, 0, 1, 3, v) == 0


def abi_tuning():
    from noa_amd import abi
    return abi.Tuning


def test_autotune_table_round_trip(lib, tmp_path):
    from noa_amd import autotune
    try:
        autotune.apply([autotune.Entry(op="extract", box=256, values=(3, 96, 0, 0)),
                        autotune.Entry(op="insert_batch", values=(1536, 0, 0, 0), ms=0.5)])
        assert autotune.lookup("extract", 256) == (3, 96, 0, 0)
        assert autotune.lookup("extract", 128) == (0, 0, 0, 0)
        assert autotune.lookup("insert_batch") == (1536, 0, 0, 0)
        autotune.save(tmp_path / "t.json")
        autotune.apply([])
        assert autotune.lookup("insert_batch") == (0, 0, 0, 0)
        autotune.load(tmp_path / "t.json")
        assert autotune.lookup("extract", 256) == (3, 96, 0, 0)
    finally:
        autotune.apply([])


def test_batch_prepare_argument_errors(lib):
    """Argument errors of the batch calls are reported before any device access."""
    from noa_amd import abi
    size = ctypes.c_int64(-1)
    calls = (abi.ProjectArgs * 1)()
    assert lib.noa_hip_extract_batch_prepare(calls, -1, None, 0, ctypes.byref(size)) == abi.ERROR_INVALID_ARGUMENT
    assert "negative count" in lib.noa_hip_last_error().decode()
    assert lib.noa_hip_insert_batch_prepare(calls, 1, None, 0, None) == abi.ERROR_INVALID_ARGUMENT
    assert "null size" in lib.noa_hip_last_error().decode()
    # Zero calls: a valid empty batch; a too-small image reports the size needed.
    assert lib.noa_hip_extract_batch_prepare(calls, 0, None, 0, ctypes.byref(size)) == abi.ERROR_INVALID_ARGUMENT
    assert size.value == 64
    img = (ctypes.c_uint8 * 64)()
    assert lib.noa_hip_extract_batch_prepare(calls, 0, img, 64, ctypes.byref(size)) == 0
    assert lib.noa_hip_extract_batch_launch(None, None) == abi.ERROR_INVALID_ARGUMENT
    # A first call whose own rules fail names the call.
    lib.noa_hip_project_args_init(calls)
    assert lib.noa_hip_insert_batch_prepare(calls, 1, None, 0, ctypes.byref(size)) != 0
    assert lib.noa_hip_last_error().decode().startswith("batch call 0: ")


def test_volume_blocks_sizes_and_argument_errors(lib):
    """The block copy's size (4x2x4 blocks of 32 c32 values, each extent rounded up to the block's) and the argument
    checks of noa_hip_volume_blocks, all made before any device access; a block-copy view is refused by the single
    extraction call."""
    from noa_amd import abi
    shape = (ctypes.c_int64 * 4)(1, 256, 256, 129)
    assert lib.noa_hip_volume_blocks_bytes(shape) == 256 * 256 * 132 * 8
    odd = (ctypes.c_int64 * 4)(1, 33, 33, 17)
    assert lib.noa_hip_volume_blocks_bytes(odd) == 36 * 34 * 20 * 8
    vol = np.zeros((1, 16, 16, 9), np.complex64)
    v = abi.View()
    v.data = vol.ctypes.data
    for i in range(4):
        v.shape[i] = vol.shape[i]
        v.strides[i] = vol.strides[i] // vol.itemsize
    v.dtype = abi.C32
    need = lib.noa_hip_volume_blocks_bytes((ctypes.c_int64 * 4)(*vol.shape))
    dst = np.zeros(need // 8, np.complex64)
    assert lib.noa_hip_volume_blocks(None, dst.ctypes.data, need, None) == abi.ERROR_INVALID_ARGUMENT
    assert "null" in lib.noa_hip_last_error().decode()
    assert lib.noa_hip_volume_blocks(ctypes.byref(v), dst.ctypes.data, need - 1, None) == abi.ERROR_INVALID_ARGUMENT
    assert f"needs {need} bytes" in lib.noa_hip_last_error().decode()
    v.dtype = abi.F32
    assert lib.noa_hip_volume_blocks(ctypes.byref(v), dst.ctypes.data, need, None) == abi.ERROR_INVALID_ARGUMENT
    assert "c32" in lib.noa_hip_last_error().decode()
    v.dtype = abi.C32
    v.shape[0] = 2
    assert lib.noa_hip_volume_blocks(ctypes.byref(v), dst.ctypes.data, need, None) == abi.ERROR_INVALID_ARGUMENT
    a, keep = _args(lib)
    a.input.dtype = abi.C32_BLOCKS
    assert lib.noa_hip_extract_central_slices_3d(ctypes.byref(a), None) == abi.ERROR_UNSUPPORTED
    assert "batched extraction only" in lib.noa_hip_last_error().decode()
    del keep
