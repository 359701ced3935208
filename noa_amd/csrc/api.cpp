// C ABI of the central-slice library: validation, descriptor flattening and dispatch.
//
// The checks restate noa's fourier_projection_check_parameters (src/noa/xform/FourierProject.hpp:1095-1185)
// with the same rules and close to the same messages, plus what a C boundary must check itself (null
// pointers, element types, memory kinds, 32-bit in-volume indexing). Nothing is launched unless every check
// passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>
#include <exception>
#include <functional>
#include <limits>
#include <string>

#include "../../include/noa_hip.h"
#include "kernels.hpp"

namespace noa_hip {
namespace {

thread_local std::string g_last_error;

enum class Kind { RASTERIZE, INSERT, EXTRACT };

std::string shape_str(const int64_t* s) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "(%lld,%lld,%lld,%lld)", (long long)s[0], (long long)s[1], (long long)s[2],
                  (long long)s[3]);
    return buf;
}

bool is_real(int dtype) { return dtype == NOA_HIP_F32 || dtype == NOA_HIP_F64; }
bool is_complex(int dtype) {
    return dtype == NOA_HIP_C32 || dtype == NOA_HIP_C64 || dtype == NOA_HIP_C16 || dtype == NOA_HIP_CBF16 ||
           dtype == NOA_HIP_C32_BLOCKS;
}
size_t dtype_size(int dtype) {
    switch (dtype) {
        case NOA_HIP_F32: return 4;
        case NOA_HIP_F64: return 8;
        case NOA_HIP_C32: return 8;
        case NOA_HIP_C32_BLOCKS: return 8;
        case NOA_HIP_C64: return 16;
        case NOA_HIP_C16: return 4;
        case NOA_HIP_CBF16: return 4;
        default: return 0;
    }
}
bool is_empty_view(const noa_hip_view& v) { return v.data == nullptr && !v.is_value; }

void rfft_shape(const int64_t* logical, int64_t* out) {
    out[0] = logical[0]; out[1] = logical[1]; out[2] = logical[2]; out[3] = logical[3] / 2 + 1;
}
bool same_shape(const int64_t* a, const int64_t* b) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && a[3] == b[3];
}

// Byte range [lo, hi) covered by a view.
void extent(const noa_hip_view& v, uintptr_t& lo, uintptr_t& hi) {
    int64_t min_off = 0, max_off = 0;
    for (int i = 0; i < 4; ++i) {
        const int64_t span = (v.shape[i] - 1) * v.strides[i];
        if (span < 0) min_off += span; else max_off += span;
    }
    const size_t es = dtype_size(v.dtype);
    lo = reinterpret_cast<uintptr_t>(v.data) + static_cast<uintptr_t>(min_off * static_cast<int64_t>(es));
    hi = reinterpret_cast<uintptr_t>(v.data) + static_cast<uintptr_t>((max_off + 1) * static_cast<int64_t>(es));
}
bool overlap(const noa_hip_view& a, const noa_hip_view& b) {
    uintptr_t alo, ahi, blo, bhi;
    extent(a, alo, ahi);
    extent(b, blo, bhi);
    return alo < bhi && blo < ahi;
}

// The device owning a pointer, or -1 when it is not device-accessible memory of this process.
int device_of(const void* p, bool* is_managed) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    if (is_managed) *is_managed = attr.isManaged != 0;
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged || attr.isManaged)
        return attr.device;
    return -1;
}

bool fits_i32(const noa_hip_view& v, int first_dim) {
    int64_t max_off = 0;
    for (int i = first_dim; i < 4; ++i) max_off += (v.shape[i] - 1) * (v.strides[i] < 0 ? -v.strides[i] : v.strides[i]);
    return max_off <= 0x7fffffffLL;
}

struct Fail {
    int code;
    std::string msg;
};

#define NOA_CHECK(cond, code, ...)                                                   \
    do {                                                                             \
        if (!(cond)) {                                                               \
            char buf_[512];                                                          \
            std::snprintf(buf_, sizeof(buf_), __VA_ARGS__);                          \
            throw Fail{code, buf_};                                                  \
        }                                                                            \
    } while (0)

constexpr int INVALID = NOA_HIP_ERROR_INVALID_ARGUMENT;
constexpr int UNSUPPORTED = NOA_HIP_ERROR_UNSUPPORTED;

Arr to_arr(const noa_hip_view& v, bool volume_dims) {
    Arr a;
    a.data = v.data;
    a.dtype = v.dtype;
    a.is_value = v.is_value != 0;
    a.value[0] = v.value[0];
    a.value[1] = v.value[1];
    if (volume_dims) { a.s0 = v.strides[1]; a.s1 = v.strides[2]; a.s2 = v.strides[3]; }   // filter {1,2,3}
    else { a.s0 = v.strides[0]; a.s1 = v.strides[2]; a.s2 = v.strides[3]; }               // filter {0,2,3}
    if (a.is_value) a.data = nullptr;
    return a;
}

Xform to_xform(const noa_hip_transform& t) {
    Xform x;
    x.data = t.data;
    x.kind = t.kind;
    x.batched = t.batched != 0;
    if (t.kind != NOA_HIP_XFORM_NONE && !t.batched && t.data) {
        // Single (broadcast) value: copied from host memory, keeping the stored element order.
        const unsigned char* p = static_cast<const unsigned char*>(t.data);
        switch (t.kind) {
            case NOA_HIP_MAT33_F32: for (int i = 0; i < 9; ++i) { float e; std::memcpy(&e, p + 4 * i, 4); x.single[i] = e; } break;
            case NOA_HIP_MAT33_F64: for (int i = 0; i < 9; ++i) { double e; std::memcpy(&e, p + 8 * i, 8); x.single[i] = e; } break;
            case NOA_HIP_QUAT_F32: for (int i = 0; i < 4; ++i) { float e; std::memcpy(&e, p + 4 * i, 4); x.single[i] = e; } break;
            case NOA_HIP_QUAT_F64: for (int i = 0; i < 4; ++i) { double e; std::memcpy(&e, p + 8 * i, 8); x.single[i] = e; } break;
            case NOA_HIP_MAT22_F32: for (int i = 0; i < 4; ++i) { float e; std::memcpy(&e, p + 4 * i, 4); x.single[i] = e; } break;
            case NOA_HIP_MAT22_F64: for (int i = 0; i < 4; ++i) { double e; std::memcpy(&e, p + 8 * i, 8); x.single[i] = e; } break;
            default: break;
        }
    }
    return x;
}

void check_view_common(const noa_hip_view& v, const char* name) {
    NOA_CHECK(dtype_size(v.dtype) != 0, INVALID, "%s: invalid element type %d", name, v.dtype);
    for (int i = 0; i < 4; ++i)
        NOA_CHECK(v.shape[i] > 0, INVALID, "Empty array detected (%s:shape=%s)", name, shape_str(v.shape).c_str());
}

ProjectDesc validate(Kind kind, const noa_hip_project_args& a, int& device) {
    ProjectDesc d;
    const bool extract = kind == Kind::EXTRACT;

    // Output (:1109-1113).
    NOA_CHECK(a.output.data != nullptr && !a.output.is_value, INVALID, "Empty array detected");
    check_view_common(a.output, "output");
    int64_t expected[4];
    rfft_shape(a.output_shape, expected);
    NOA_CHECK(same_shape(a.output.shape, expected), INVALID,
              "The shape of the rfft output does not match the expected shape. Got output:shape=%s and expected:shape=%s",
              shape_str(a.output.shape).c_str(), shape_str(expected).c_str());
    // Inputs (:1115-1136). Values are allowed for rasterize/insert inputs, not for extraction.
    // Device checks need the HIP runtime and run last, after every host-only check.
    auto check_input = [&](const noa_hip_view& v, const char* name, bool is_weight) {
        if (is_empty_view(v)) return;
        if (v.is_value) {
            NOA_CHECK(!extract, INVALID, "%s: a constant value is not a valid input of the extraction", name);
            if (is_weight) NOA_CHECK(is_real(v.dtype), INVALID, "%s: weights should be real", name);
            return;
        }
        check_view_common(v, name);
        NOA_CHECK(!overlap(v, a.output), INVALID, "Input and output arrays should not overlap");
        int64_t exp[4];
        rfft_shape(a.input_shape, exp);
        NOA_CHECK(same_shape(v.shape, exp), INVALID,
                  "The shape of the rfft does not match the expected shape. Got %s:shape=%s and shape=%s", name,
                  shape_str(v.shape).c_str(), shape_str(exp).c_str());
        if (is_weight) NOA_CHECK(is_real(v.dtype), INVALID, "%s: weights should be real", name);
    };
    check_input(a.input, "input", false);
    check_input(a.input_weight, "input_weight", true);
    NOA_CHECK(!is_empty_view(a.input), INVALID, "input: Empty array detected");

    // Output weight (:1138-1147).
    if (!is_empty_view(a.output_weight)) {
        NOA_CHECK(!a.output_weight.is_value && a.output_weight.data, INVALID, "Empty array detected");
        check_view_common(a.output_weight, "output_weight");
        NOA_CHECK(!overlap(a.output_weight, a.output), INVALID, "Output arrays should not overlap");
        NOA_CHECK(same_shape(a.output_weight.shape, expected), INVALID,
                  "The shape of the rfft does not match the expected shape. Got output_weight:shape=%s and shape=%s",
                  shape_str(a.output_weight.shape).c_str(), shape_str(expected).c_str());
        NOA_CHECK(is_real(a.output_weight.dtype), INVALID, "output_weight: weights should be real");
        if (!is_empty_view(a.input_weight) && !a.input_weight.is_value)
            NOA_CHECK(!overlap(a.input_weight, a.output_weight), INVALID, "Input and output arrays should not overlap");
        if (!a.input.is_value) NOA_CHECK(!overlap(a.input, a.output_weight), INVALID, "Input and output arrays should not overlap");
    }

    // Dimensionality (:1149-1164).
    if (!extract) {
        NOA_CHECK(a.input_shape[1] == 1, INVALID, "2d input slices are expected, but got shape=%s", shape_str(a.input_shape).c_str());
        NOA_CHECK(a.output_shape[0] == 1 && a.target_shape[0] <= 1, INVALID,
                  "A single 3d volume is expected, but got output_shape=%s and target_shape=%s (optional)",
                  shape_str(a.output_shape).c_str(), shape_str(a.target_shape).c_str());
    } else {
        NOA_CHECK(a.input_shape[0] == 1 && a.target_shape[0] <= 1, INVALID,
                  "A single 3d volume is expected, but got input_shape=%s and target_shape=%s (optional)",
                  shape_str(a.input_shape).c_str(), shape_str(a.target_shape).c_str());
        NOA_CHECK(a.output_shape[1] == 1, INVALID, "2d input slices are expected but got shape %s", shape_str(a.output_shape).c_str());
    }
    for (int i = 0; i < 4; ++i) {
        NOA_CHECK(a.input_shape[i] > 0 && a.output_shape[i] > 0, INVALID, "Empty array detected");
        NOA_CHECK(a.target_shape[i] >= 0, INVALID, "target_shape should be positive or zero");
    }

    // Remap (requires REMAP.is_hx2hx(), :1659).
    NOA_CHECK(a.remap == NOA_HIP_H2H || a.remap == NOA_HIP_HC2H || a.remap == NOA_HIP_H2HC || a.remap == NOA_HIP_HC2HC,
              INVALID, "The remap should be H2H, HC2H, H2HC or HC2HC, but got %d", a.remap);

    // Interpolation.
    int interp = a.interp >= 100 ? a.interp - 100 : a.interp;
    NOA_CHECK(interp >= 0 && interp <= 6, INVALID, "Invalid interpolation method %d", a.interp);


    // Value types: spectrum types (Traits.hpp:87) = c->c, r->r, c->r.
    const int in = a.input.dtype, out = a.output.dtype;
    NOA_CHECK((is_complex(in) && is_complex(out)) || (is_real(in) && is_real(out)) || (is_complex(in) && is_real(out)),
              INVALID, "Invalid input/output types: %d -> %d", in, out);
    if (in == NOA_HIP_C16 || in == NOA_HIP_CBF16)
        NOA_CHECK(extract && !a.input.is_value, UNSUPPORTED, "FP16/BF16 storage is only supported for the extracted volume");
    if (in == NOA_HIP_C32_BLOCKS)
        NOA_CHECK(extract && !a.input.is_value && out == NOA_HIP_C32, UNSUPPORTED,
                  "A block-copy volume is only accepted as the c32 volume of an extraction");
    NOA_CHECK(out != NOA_HIP_C32_BLOCKS && a.input_weight.dtype != NOA_HIP_C32_BLOCKS &&
              a.output_weight.dtype != NOA_HIP_C32_BLOCKS, UNSUPPORTED, "A block-copy volume is only accepted as the input of an extraction");

    // Transforms (:1166-1178).
    const int64_t required = extract ? a.output_shape[0] : a.input_shape[0];
    auto check_transform = [&](const noa_hip_transform& t, const char* name, bool rotation) {
        if (rotation) {
            NOA_CHECK(t.kind >= NOA_HIP_MAT33_F32 && t.kind <= NOA_HIP_QUAT_F64, INVALID,
                      "%s should be Mat33 or Quaternion (f32 or f64), got kind=%d", name, t.kind);
        } else if (t.kind == NOA_HIP_XFORM_NONE) {
            return;
        } else {
            NOA_CHECK(t.kind == NOA_HIP_MAT22_F32 || t.kind == NOA_HIP_MAT22_F64, INVALID,
                      "%s should be Mat22 (f32 or f64), got kind=%d", name, t.kind);
        }
        NOA_CHECK(t.data != nullptr, INVALID, "%s should not be empty", name);
        if (t.batched) {
            NOA_CHECK(t.count == required, INVALID,
                      "%s should be a contiguous vector with n_slices=%lld elements, but got %lld elements", name,
                      (long long)required, (long long)t.count);
            NOA_CHECK(reinterpret_cast<uintptr_t>(t.data) % 16 == 0 || t.kind == NOA_HIP_MAT33_F64, INVALID,
                      "%s should be 16-byte aligned", name);
        }
    };
    check_transform(a.rotation, "rotation", true);
    check_transform(a.scaling, "scaling", false);
    const bool rot64 = a.rotation.kind == NOA_HIP_MAT33_F64 || a.rotation.kind == NOA_HIP_QUAT_F64;
    if (a.scaling.kind != NOA_HIP_XFORM_NONE) {
        const bool s64 = a.scaling.kind == NOA_HIP_MAT22_F64;
        NOA_CHECK(s64 == rot64, INVALID, "The scaling and rotation should have the same precision");
    }

    // 32-bit in-volume / in-slice indexing (the batch offset of slice stacks is 64-bit).
    auto check_i32 = [&](const noa_hip_view& v, bool is_volume, const char* name) {
        if (is_empty_view(v) || v.is_value) return;
        NOA_CHECK(fits_i32(v, is_volume ? 1 : 2), UNSUPPORTED, "%s: offsets within a %s exceed 32-bit indexing", name,
                  is_volume ? "volume" : "slice");
    };
    check_i32(a.input, extract, "input");
    check_i32(a.input_weight, extract, "input_weight");
    check_i32(a.output, !extract, "output");
    check_i32(a.output_weight, !extract, "output_weight");

    // Weights: real, same precision as the output values.
    const bool out64 = out == NOA_HIP_F64 || out == NOA_HIP_C64;
    const int wtype = out64 ? NOA_HIP_F64 : NOA_HIP_F32;
    if (!is_empty_view(a.input_weight)) NOA_CHECK(a.input_weight.dtype == wtype, UNSUPPORTED, "input_weight: expected %s weights", out64 ? "f64" : "f32");
    if (!is_empty_view(a.output_weight)) NOA_CHECK(a.output_weight.dtype == wtype, UNSUPPORTED, "output_weight: expected %s weights", out64 ? "f64" : "f32");
    if (!a.input.is_value) {
        const bool in64 = in == NOA_HIP_F64 || in == NOA_HIP_C64;
        NOA_CHECK(in64 == out64 || in == NOA_HIP_C16 || in == NOA_HIP_CBF16, UNSUPPORTED,
                  "input and output should have the same precision");
    }

    // Device checks (HIP runtime): every array and batched transform on the output's device.
    bool managed = false;
    device = device_of(a.output.data, &managed);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    auto check_device = [&](const noa_hip_view& v, const char* name) {
        if (is_empty_view(v) || v.is_value) return;
        const int dev = device_of(v.data, nullptr);
        NOA_CHECK(dev == device, INVALID, "The arrays should be on the same device, but got %s:device=%d and output:device=%d",
                  name, dev, device);
    };
    check_device(a.input, "input");
    check_device(a.input_weight, "input_weight");
    check_device(a.output_weight, "output_weight");
    if (a.rotation.batched)
        NOA_CHECK(device_of(a.rotation.data, nullptr) == device, INVALID, "rotation should be on the compute device");
    if (a.scaling.kind != NOA_HIP_XFORM_NONE && a.scaling.batched)
        NOA_CHECK(device_of(a.scaling.data, nullptr) == device, INVALID, "scaling should be on the compute device");
    // Float atomics on managed memory are not device-wide safe without XNACK: refuse for accumulating ops.
    if (kind == Kind::RASTERIZE)
        NOA_CHECK(!managed, UNSUPPORTED, "rasterize: the volume should be device (hipMalloc) memory, not managed memory");

    // Flatten.
    d.input = to_arr(a.input, extract);
    d.input_weight = to_arr(a.input_weight, extract);
    d.output = to_arr(a.output, !extract);
    d.output_weight = to_arr(a.output_weight, !extract);
    if (is_empty_view(a.input_weight)) { d.input_weight.data = nullptr; d.input_weight.is_value = false; }
    if (is_empty_view(a.output_weight)) { d.output_weight.data = nullptr; }
    for (int i = 0; i < 4; ++i) { d.in_shape[i] = a.input_shape[i]; d.out_shape[i] = a.output_shape[i]; }
    const int64_t* grid = extract ? a.input_shape : a.output_shape;
    const bool any0 = a.target_shape[0] == 0 || a.target_shape[1] == 0 || a.target_shape[2] == 0 || a.target_shape[3] == 0;
    for (int i = 0; i < 3; ++i) d.target[i] = any0 ? grid[i + 1] : a.target_shape[i + 1];
    d.src_centered = (a.remap & 4) != 0;
    d.dst_centered = (a.remap & 8) != 0;
    d.interp = interp;
    d.rot = to_xform(a.rotation);
    d.scale = to_xform(a.scaling);
    d.coord_f64 = rot64;
    d.cutoff = a.fftfreq_cutoff;
    d.fftfreq_sinc = a.fftfreq_sinc;
    d.fftfreq_blackman = a.fftfreq_blackman;
    d.ews_radius[0] = a.ews_radius[0];
    d.ews_radius[1] = a.ews_radius[1];
    const bool has_ews = a.ews_radius[0] != 0 || a.ews_radius[1] != 0;
    bool has_scale = a.scaling.kind != NOA_HIP_XFORM_NONE;
    if (has_scale && !a.scaling.batched) {   // a single identity Mat22 is no scaling (:1234-1244)
        const double* s = d.scale.single;
        has_scale = !(s[0] == 1 && s[1] == 0 && s[2] == 0 && s[3] == 1);
    }
    d.general = has_ews || has_scale;
    if (!has_scale) d.scale.kind = NOA_HIP_XFORM_NONE;
    return d;
}

// Shifts of phase_shift_2d (PhaseShift.hpp:190-208 checks, :250-267 single-shift special cases).
bool allclose_f64(double x, double y) {   // noa::allclose with noa's defaults (base/Math.hpp:296-305)
    const double diff = std::abs(x - y);
    if (!std::isfinite(diff)) return false;
    return diff <= 1e-6 || diff <= std::max(std::abs(x), std::abs(y)) * 2.220446049250313e-16 * 2;
}
bool allclose_f32(float x, float y) {
    const float diff = std::abs(x - y);
    if (!std::isfinite(diff)) return false;
    return diff <= 1e-6f || diff <= std::max(std::abs(x), std::abs(y)) * 1.1920929e-07f * 2;
}

ShiftDesc resolve_shift(const noa_hip_shift2* sh, int64_t batch, int64_t h, int64_t w) {
    ShiftDesc d;
    NOA_CHECK(sh != nullptr, INVALID, "shifts: null");
    NOA_CHECK(sh->dtype == NOA_HIP_F32 || sh->dtype == NOA_HIP_F64, INVALID,
              "shifts should be Vec<f32,2> or Vec<f64,2>, got dtype=%d", sh->dtype);
    NOA_CHECK(sh->data != nullptr, INVALID, "shifts: Empty array detected");
    d.present = true;
    d.f64 = sh->dtype == NOA_HIP_F64;
    d.batched = sh->batched != 0;
    d.cutoff = sh->fftfreq_cutoff;
    d.data = sh->data;
    if (d.batched) {
        NOA_CHECK(sh->count == batch, INVALID,
                  "The input shift(s) should be entered as a 1d contiguous vector, with one shift per output batch, "
                  "but got %lld shifts and %lld batches", (long long)sh->count, (long long)batch);
        NOA_CHECK(reinterpret_cast<uintptr_t>(sh->data) % (d.f64 ? 16 : 8) == 0, INVALID,
                  "shifts should be aligned to their pair size");
        return d;
    }
    const bool zero = d.f64 ? (allclose_f64(static_cast<const double*>(sh->data)[0], 0) &&
                               allclose_f64(static_cast<const double*>(sh->data)[1], 0))
                            : (allclose_f32(static_cast<const float*>(sh->data)[0], 0) &&
                               allclose_f32(static_cast<const float*>(sh->data)[1], 0));
    for (int i = 0; i < 2; ++i)
        d.single[i] = d.f64 ? static_cast<const double*>(sh->data)[i] : static_cast<const float*>(sh->data)[i];
    const bool half = d.f64 ? (allclose_f64(std::abs(d.single[0]), static_cast<double>(h) / 2) &&
                               allclose_f64(std::abs(d.single[1]), static_cast<double>(w) / 2))
                            : (allclose_f32(std::abs(static_cast<float>(d.single[0])), static_cast<float>(h) / 2) &&
                               allclose_f32(std::abs(static_cast<float>(d.single[1])), static_cast<float>(w) / 2));
    d.mode = zero ? 1 : (half && sh->fftfreq_cutoff >= std::sqrt(0.5) ? 2 : 0);
    d.data = nullptr;
    return d;
}

PhaseShiftDesc validate_phase_shift(const noa_hip_view* in, const noa_hip_view* out, const int64_t* shape,
                                    int32_t remap, const noa_hip_shift2* shifts, int& device) {
    NOA_CHECK(out != nullptr && shape != nullptr, INVALID, "null arguments");
    NOA_CHECK(out->data != nullptr && !out->is_value, INVALID, "Empty array detected");
    NOA_CHECK(out->dtype == NOA_HIP_C32 || out->dtype == NOA_HIP_C64, INVALID,
              "phase_shift_2d: the output should be c32 or c64, got dtype=%d", out->dtype);
    NOA_CHECK(remap == NOA_HIP_H2H || remap == NOA_HIP_HC2H || remap == NOA_HIP_H2HC || remap == NOA_HIP_HC2HC, INVALID,
              "phase_shift_2d: the remap should be H2H, HC2H, H2HC or HC2HC (rfft layouts), but got %d", remap);
    for (int i = 0; i < 4; ++i) NOA_CHECK(shape[i] > 0, INVALID, "Empty array detected");
    NOA_CHECK(shape[1] == 1, UNSUPPORTED, "phase_shift_2d: only 2d arrays (depth 1) are supported, got shape=%s",
              shape_str(shape).c_str());
    int64_t expected[4];
    rfft_shape(shape, expected);
    NOA_CHECK(same_shape(out->shape, expected), INVALID,
              "Given the logical shape %s and FFT layout %d, the expected physical shape should be %s, but got %s",
              shape_str(shape).c_str(), remap, shape_str(expected).c_str(), shape_str(out->shape).c_str());
    const bool has_input = in != nullptr && in->data != nullptr;
    const bool layout_change = ((remap & 4) != 0) != ((remap & 8) != 0);
    if (has_input) {
        NOA_CHECK(!in->is_value, INVALID, "phase_shift_2d: the input should be an array");
        NOA_CHECK(in->dtype == out->dtype, UNSUPPORTED, "phase_shift_2d: input and output should have the same type");
        NOA_CHECK(same_shape(in->shape, expected), INVALID, "phase_shift_2d: the input should have the shape %s, got %s",
                  shape_str(expected).c_str(), shape_str(in->shape).c_str());
        NOA_CHECK(!layout_change || in->data != out->data, INVALID, "In-place remapping is not allowed");
        NOA_CHECK(in->data == out->data || !overlap(*in, *out), INVALID, "Input and output arrays should not overlap");
    }
    auto i32_ok = [](const noa_hip_view& v) { return fits_i32(v, 2); };
    NOA_CHECK(i32_ok(*out) && (!has_input || i32_ok(*in)), UNSUPPORTED, "offsets within a slice exceed 32-bit indexing");
    PhaseShiftDesc d;
    d.shift = resolve_shift(shifts, shape[0], shape[2], shape[3]);
    device = device_of(out->data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    if (has_input)
        NOA_CHECK(device_of(in->data, nullptr) == device, INVALID,
                  "The input and output arrays must be on the same device");
    if (d.shift.batched)
        NOA_CHECK(device_of(d.shift.data, nullptr) == device, INVALID,
                  "The shift and output arrays must be on the same device");
    if (has_input) d.input = to_arr(*in, false);
    d.output = to_arr(*out, false);
    d.b = shape[0];
    d.h = shape[2];
    d.w = shape[3];
    d.src_centered = (remap & 4) != 0;
    d.dst_centered = (remap & 8) != 0;
    d.c64 = out->dtype == NOA_HIP_C64;
    return d;
}

int guarded(int device, const std::function<int(std::string&)>& launch) {
    int current = 0;
    (void)hipGetDevice(&current);
    if (current != device) (void)hipSetDevice(device);
    std::string err;
    const int status = launch(err);
    if (current != device) (void)hipSetDevice(current);
    if (status != NOA_HIP_OK) g_last_error = err;
    return status;
}

Arr4 to_arr4(const noa_hip_view& v) {
    Arr4 a;
    a.data = v.data;
    a.dtype = v.dtype;
    for (int i = 0; i < 4; ++i) a.s[i] = v.strides[i];
    return a;
}

RemapDesc validate_remap(int32_t remap, const noa_hip_view* in, const noa_hip_view* out, const int64_t* shape,
                         int& device) {
    NOA_CHECK(in != nullptr && out != nullptr && shape != nullptr, INVALID, "null arguments");
    NOA_CHECK(in->data != nullptr && out->data != nullptr && !in->is_value && !out->is_value, INVALID,
              "Empty array detected");
    NOA_CHECK(remap == NOA_HIP_H2H || remap == NOA_HIP_HC2H || remap == NOA_HIP_H2HC || remap == NOA_HIP_HC2HC,
              UNSUPPORTED, "remap: only the half-to-half layouts are implemented by this library, got %d", remap);
    NOA_CHECK(dtype_size(in->dtype) != 0 && in->dtype == out->dtype, UNSUPPORTED,
              "remap: input and output should have the same element type");
    for (int i = 0; i < 4; ++i) NOA_CHECK(shape[i] > 0, INVALID, "Empty array detected");
    int64_t expected[4];
    rfft_shape(shape, expected);
    NOA_CHECK(same_shape(in->shape, expected), INVALID,
              "Given the remap, the input fft is expected to have a physical shape of %s, but got %s",
              shape_str(expected).c_str(), shape_str(in->shape).c_str());
    NOA_CHECK(same_shape(out->shape, expected), INVALID,
              "Given the remap, the output fft is expected to have a physical shape of %s, but got %s",
              shape_str(expected).c_str(), shape_str(out->shape).c_str());
    const bool change = ((remap & 4) != 0) != ((remap & 8) != 0);
    NOA_CHECK(!change || !overlap(*in, *out), UNSUPPORTED,
              "remap: in-place remapping with a layout change is not supported");
    device = device_of(out->data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    NOA_CHECK(device_of(in->data, nullptr) == device, INVALID, "The arrays should be on the same device");
    RemapDesc d;
    d.input = to_arr4(*in);
    d.output = to_arr4(*out);
    for (int i = 0; i < 4; ++i) d.shape[i] = shape[i];
    d.src_centered = (remap & 4) != 0;
    d.dst_centered = (remap & 8) != 0;
    return d;
}

GriddingDesc validate_gridding(const noa_hip_view* in, const noa_hip_view* out, int32_t interp, int32_t post,
                               int& device) {
    NOA_CHECK(in != nullptr && out != nullptr, INVALID, "null arguments");
    NOA_CHECK(in->data != nullptr && out->data != nullptr && !in->is_value && !out->is_value, INVALID,
              "Empty array detected");
    NOA_CHECK(interp == NOA_HIP_INTERP_LINEAR || interp == NOA_HIP_INTERP_LINEAR_FAST, INVALID,
              "Interpolation method %d is currently not supported", interp);
    NOA_CHECK(is_real(in->dtype) && is_real(out->dtype), INVALID,
              "fourier_interpolation_correction: real f32 or f64 arrays are expected");
    for (int i = 0; i < 4; ++i) NOA_CHECK(out->shape[i] > 0, INVALID, "Empty array detected");
    const bool broadcast = in->shape[0] == 1 && out->shape[0] > 1;
    NOA_CHECK((in->shape[0] == out->shape[0] || broadcast) && in->shape[1] == out->shape[1] &&
              in->shape[2] == out->shape[2] && in->shape[3] == out->shape[3], INVALID,
              "Cannot broadcast an array of shape %s into an array of shape %s", shape_str(in->shape).c_str(),
              shape_str(out->shape).c_str());
    NOA_CHECK(in->data == out->data || !overlap(*in, *out), INVALID, "Input and output arrays should not overlap");
    device = device_of(out->data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    NOA_CHECK(device_of(in->data, nullptr) == device, INVALID, "The arrays should be on the same device");
    GriddingDesc d;
    d.input = to_arr4(*in);
    if (broadcast) d.input.s[0] = 0;
    d.output = to_arr4(*out);
    for (int i = 0; i < 4; ++i) d.shape[i] = out->shape[i];
    d.post = post != 0;
    return d;
}

// fft::r2c / c2r / c2c (fft/Transform.hpp:160-349): non-empty arrays, the type pairs noa's templates accept, the
// shape rules of each transform, one device. Plus what rocFFT needs: non-negative strides, and for an in-place
// real transform noa's padded layout (nf::empty, fft/Factory.hpp): the complex strides are half the real ones.
FftDesc validate_fft(int type, const noa_hip_view* in, const noa_hip_view* out, int32_t sign,
                     const noa_hip_fft_options* opt, int& device) {
    NOA_CHECK(in != nullptr && out != nullptr, INVALID, "null arguments");
    NOA_CHECK(in->data != nullptr && out->data != nullptr && !in->is_value && !out->is_value, INVALID,
              "Empty array detected");
    for (int i = 0; i < 4; ++i)
        NOA_CHECK(in->shape[i] > 0 && out->shape[i] > 0, INVALID, "Empty array detected");
    FftDesc d;
    d.type = type;
    int64_t expected[4];
    if (type == FFT_R2C) {
        NOA_CHECK((in->dtype == NOA_HIP_F32 && out->dtype == NOA_HIP_C32) ||
                  (in->dtype == NOA_HIP_F64 && out->dtype == NOA_HIP_C64), INVALID,
                  "r2c: the input should be f32 or f64 and the output the complex type of the same precision");
        rfft_shape(in->shape, expected);
        NOA_CHECK(same_shape(out->shape, expected), INVALID,
                  "Given the real input with a shape of %s, the non-redundant shape of the complex output should be %s, "
                  "but got %s", shape_str(in->shape).c_str(), shape_str(expected).c_str(), shape_str(out->shape).c_str());
        for (int i = 0; i < 4; ++i) d.shape[i] = in->shape[i];
        d.sign = NOA_HIP_SIGN_FORWARD;
    } else if (type == FFT_C2R) {
        NOA_CHECK((in->dtype == NOA_HIP_C32 && out->dtype == NOA_HIP_F32) ||
                  (in->dtype == NOA_HIP_C64 && out->dtype == NOA_HIP_F64), INVALID,
                  "c2r: the input should be c32 or c64 and the output the real type of the same precision");
        rfft_shape(out->shape, expected);
        NOA_CHECK(same_shape(in->shape, expected), INVALID,
                  "Given the real output with a shape of %s, the non-redundant shape of the complex input should be %s, "
                  "but got %s", shape_str(out->shape).c_str(), shape_str(expected).c_str(), shape_str(in->shape).c_str());
        for (int i = 0; i < 4; ++i) d.shape[i] = out->shape[i];
        d.sign = NOA_HIP_SIGN_BACKWARD;
    } else {
        NOA_CHECK((in->dtype == NOA_HIP_C32 || in->dtype == NOA_HIP_C64) && in->dtype == out->dtype, INVALID,
                  "c2c: the input and output should both be c32 or both be c64");
        NOA_CHECK(same_shape(in->shape, out->shape), INVALID,
                  "The input and output shape should match (no broadcasting allowed), but got input %s and output %s",
                  shape_str(in->shape).c_str(), shape_str(out->shape).c_str());
        NOA_CHECK(sign == NOA_HIP_SIGN_FORWARD || sign == NOA_HIP_SIGN_BACKWARD, INVALID,
                  "c2c: the sign should be -1 (forward) or +1 (backward), got %d", sign);
        for (int i = 0; i < 4; ++i) d.shape[i] = in->shape[i];
        d.sign = sign;
    }
    d.f64 = in->dtype == NOA_HIP_F64 || in->dtype == NOA_HIP_C64;
    for (int i = 0; i < 4; ++i) {
        NOA_CHECK(in->strides[i] >= 0 && out->strides[i] >= 0, UNSUPPORTED, "fft: negative strides are not supported");
        d.in_s[i] = in->strides[i];
        d.out_s[i] = out->strides[i];
    }
    // Rank (cuda/Plan.cpp:420-424): ndim of the DHW shape; a vector along D alone is refused, as in noa.
    const int64_t D = d.shape[1], H = d.shape[2], W = d.shape[3];
    d.rank = D > 1 ? 3 : (H > 1 && W > 1) ? 2 : 1;
    NOA_CHECK(d.rank != 3 || H > 1 || W > 1, UNSUPPORTED,
              "fft: a vector along the depth dimension is not supported (shape %s)", shape_str(d.shape).c_str());
    // noa turns a column vector into a row vector for the plan, so its r2c/c2r would store W/2+1 values along a
    // dimension of the rfft shape that holds W = 1: not supported here.
    NOA_CHECK(type == FFT_C2C || !(d.rank == 1 && W == 1 && H > 1), UNSUPPORTED,
              "fft: r2c and c2r of column vectors are not supported (shape %s)", shape_str(d.shape).c_str());
    d.in = in->data;
    d.out = out->data;
    if (in->data == out->data) {
        if (type == FFT_C2C) {
            for (int i = 0; i < 4; ++i)
                NOA_CHECK(in->strides[i] == out->strides[i], INVALID,
                          "c2c: an in-place transform needs equal input and output strides");
        } else {
            const noa_hip_view& re = type == FFT_R2C ? *in : *out;
            const noa_hip_view& cx = type == FFT_R2C ? *out : *in;
            bool padded = re.strides[3] == 1 && cx.strides[3] == 1;
            for (int i = 0; i < 3; ++i) padded = padded && re.strides[i] == 2 * cx.strides[i];
            NOA_CHECK(padded, INVALID,
                      "In-place r2c/c2r transforms need the padded layout: the real array's strides should be twice "
                      "the complex array's, with unit x strides");
        }
    } else {
        NOA_CHECK(!overlap(*in, *out), INVALID, "Input and output arrays should not overlap");
    }
    const noa_hip_fft_options defaults{NOA_HIP_NORM_FORWARD, 1, 0, 0};
    const noa_hip_fft_options& o = opt ? *opt : defaults;
    NOA_CHECK(o.norm >= NOA_HIP_NORM_FORWARD && o.norm <= NOA_HIP_NORM_NONE, INVALID, "fft: invalid norm %d", o.norm);
    d.norm = o.norm;
    d.cache_plan = o.cache_plan != 0;
    d.plan_only = o.plan_only != 0;
    d.record_workspace = o.record_and_share_workspace != 0;
    device = device_of(out->data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    const int in_dev = device_of(in->data, nullptr);
    NOA_CHECK(in_dev == device, INVALID,
              "The input and output arrays must be on the same device, but got input:device=%d, output:device=%d",
              in_dev, device);
    NOA_CHECK(device < 64, UNSUPPORTED, "fft: the plan cache is limited to 64 visible devices");
    d.device = device;
    return d;
}

// cubic_bspline_prefilter (xform/CubicBSplinePrefilter.hpp:24-63): non-empty arrays of one of f32, f64, c32, c64 on
// one device, the input broadcast into the output (runtime/Utils.hpp:106-117). The dimensions are then reordered to
// the output's rightmost order (cpu/CubicBSplinePrefilter.hpp:136-141, Shape.hpp:871-886), and ndim() of the
// reordered shape (Shape.hpp:315-335) picks the passes. In place when the data pointers are equal; other overlaps
// are rejected, where noa would race.
BSplineDesc validate_bspline(const noa_hip_view* in, const noa_hip_view* out, int& device) {
    NOA_CHECK(in != nullptr && out != nullptr, INVALID, "null arguments");
    NOA_CHECK(in->data != nullptr && out->data != nullptr && !in->is_value && !out->is_value, INVALID,
              "Empty array detected");
    for (int i = 0; i < 4; ++i)
        NOA_CHECK(in->shape[i] > 0 && out->shape[i] > 0, INVALID, "Empty array detected");
    NOA_CHECK(out->dtype == NOA_HIP_F32 || out->dtype == NOA_HIP_F64 || out->dtype == NOA_HIP_C32 ||
              out->dtype == NOA_HIP_C64, INVALID, "cubic_bspline_prefilter: f32, f64, c32 or c64 arrays are expected");
    NOA_CHECK(in->dtype == out->dtype, INVALID, "cubic_bspline_prefilter: the input and output types should match");
    BSplineDesc d;
    d.input = to_arr4(*in);
    d.output = to_arr4(*out);
    for (int i = 0; i < 4; ++i) {
        d.shape[i] = out->shape[i];
        if (in->shape[i] == 1 && out->shape[i] != 1) d.input.s[i] = 0;
        else NOA_CHECK(in->shape[i] == out->shape[i], INVALID, "Cannot broadcast an array of shape %s into an array of shape %s",
                       shape_str(in->shape).c_str(), shape_str(out->shape).c_str());
    }
    for (int i = 0; i < 4; ++i)
        NOA_CHECK(out->shape[i] == 1 || out->strides[i] != 0, INVALID,
                  "cubic_bspline_prefilter: the output should not be broadcast (stride 0 along dimension %d)", i);
    const bool inplace = in->data == out->data;
    NOA_CHECK(inplace || !overlap(*in, *out), INVALID, "Input and output arrays should not overlap");
    if (inplace) d.input = d.output;   // filter_inplace reads the output (cpu/CubicBSplinePrefilter.hpp:70-90)
    device = device_of(out->data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    const int in_dev = device_of(in->data, nullptr);
    NOA_CHECK(in_dev == device, INVALID,
              "The input and output arrays must be on the same device, but got input:device=%d, output:device=%d",
              in_dev, device);
    // Rightmost order: a stable sort of D, H, W by decreasing output stride, size-1 dimensions first.
    int order[3] = {0, 1, 2};
    int64_t key[3];
    for (int i = 0; i < 3; ++i) key[i] = d.shape[i + 1] <= 1 ? std::numeric_limits<int64_t>::max() : d.output.s[i + 1];
    std::stable_sort(order, order + 3, [&](int x, int y) { return key[x] > key[y]; });
    BSplineDesc r = d;
    for (int i = 0; i < 3; ++i) {
        r.shape[i + 1] = d.shape[order[i] + 1];
        r.input.s[i + 1] = d.input.s[order[i] + 1];
        r.output.s[i + 1] = d.output.s[order[i] + 1];
    }
    r.ndim = r.shape[1] > 1 ? 3 : (r.shape[2] > 1 && r.shape[3] > 1) ? 2 : 1;
    // The kernels index lines and samples with 32-bit sizes.
    NOA_CHECK(r.shape[1] <= 0x7fffffffLL && r.shape[2] <= 0x7fffffffLL && r.shape[3] <= 0x7fffffffLL, UNSUPPORTED,
              "cubic_bspline_prefilter: dimensions above 2^31 are not supported");
    const int64_t rows = r.shape[0] * r.shape[1] * r.shape[2], cols = r.shape[0] * r.shape[1] * r.shape[3],
                  pages = r.shape[0] * r.shape[2] * r.shape[3];
    NOA_CHECK(rows / 64 < 0xffffffffLL && cols / 256 < 0xffffffffLL && pages / 256 < 0xffffffffLL, UNSUPPORTED,
              "cubic_bspline_prefilter: too many lines for one launch");
    return r;
}

// insert_and_extract_central_slices_3d: fourier_projection_check_parameters<INSERT_EXTRACT> (:1095-1185) and the
// public function's own check of correct_weights (:2011-2018).
InsertExtractDesc validate_insert_extract(const noa_hip_insert_extract_args& a, int& device) {
    // Output slices (:1109-1113).
    NOA_CHECK(a.output.data != nullptr && !a.output.is_value, INVALID, "Empty array detected");
    check_view_common(a.output, "output");
    int64_t expected[4];
    rfft_shape(a.output_shape, expected);
    NOA_CHECK(same_shape(a.output.shape, expected), INVALID,
              "The shape of the rfft output does not match the expected shape. Got output:shape=%s and expected:shape=%s",
              shape_str(a.output.shape).c_str(), shape_str(expected).c_str());
    // Inputs (:1115-1136): arrays, textures (as their buffer) or values.
    auto check_input = [&](const noa_hip_view& v, const char* name, bool is_weight) {
        if (is_empty_view(v)) return;
        if (v.is_value) {
            if (is_weight) NOA_CHECK(is_real(v.dtype), INVALID, "%s: weights should be real", name);
            return;
        }
        check_view_common(v, name);
        NOA_CHECK(!overlap(v, a.output), INVALID, "Input and output arrays should not overlap");
        if (!is_empty_view(a.output_weight) && a.output_weight.data)
            NOA_CHECK(!overlap(v, a.output_weight), INVALID, "Input and output arrays should not overlap");
        int64_t exp[4];
        rfft_shape(a.input_shape, exp);
        NOA_CHECK(same_shape(v.shape, exp), INVALID,
                  "The shape of the rfft does not match the expected shape. Got %s:shape=%s and shape=%s", name,
                  shape_str(v.shape).c_str(), shape_str(exp).c_str());
        if (is_weight) NOA_CHECK(is_real(v.dtype), INVALID, "%s: weights should be real", name);
    };
    NOA_CHECK(!is_empty_view(a.input), INVALID, "input: Empty array detected");
    check_input(a.input, "input", false);
    check_input(a.input_weight, "input_weight", true);
    // Output weights (:1138-1147).
    if (!is_empty_view(a.output_weight)) {
        NOA_CHECK(!a.output_weight.is_value && a.output_weight.data, INVALID, "Empty array detected");
        check_view_common(a.output_weight, "output_weight");
        NOA_CHECK(!overlap(a.output_weight, a.output), INVALID, "Output arrays should not overlap");
        NOA_CHECK(same_shape(a.output_weight.shape, expected), INVALID,
                  "The shape of the rfft does not match the expected shape. Got output_weight:shape=%s and shape=%s",
                  shape_str(a.output_weight.shape).c_str(), shape_str(expected).c_str());
        NOA_CHECK(is_real(a.output_weight.dtype), INVALID, "output_weight: weights should be real");
    }
    // 2D slices on both sides (:1160-1164).
    for (int i = 0; i < 4; ++i)
        NOA_CHECK(a.input_shape[i] > 0 && a.output_shape[i] > 0, INVALID, "Empty array detected");
    NOA_CHECK(a.input_shape[1] == 1 && a.output_shape[1] == 1, INVALID,
              "2d slices are expected but got shape input:shape=%s and output:shape=%s",
              shape_str(a.input_shape).c_str(), shape_str(a.output_shape).c_str());
    NOA_CHECK(a.remap == NOA_HIP_H2H || a.remap == NOA_HIP_HC2H || a.remap == NOA_HIP_H2HC || a.remap == NOA_HIP_HC2HC,
              INVALID, "The remap should be H2H, HC2H, H2HC or HC2HC, but got %d", a.remap);
    const int interp = a.interp >= 100 ? a.interp - 100 : a.interp;
    NOA_CHECK(interp >= 0 && interp <= 6, INVALID, "Invalid interpolation method %d", a.interp);

    const int in = a.input.dtype, out = a.output.dtype;
    NOA_CHECK((is_complex(in) && is_complex(out)) || (is_real(in) && is_real(out)) || (is_complex(in) && is_real(out)),
              INVALID, "Invalid input/output types: %d -> %d", in, out);
    NOA_CHECK(in != NOA_HIP_C16 && in != NOA_HIP_CBF16, UNSUPPORTED,
              "FP16/BF16 storage is only supported for the extracted volume");

    // Transforms (:1166-1183): the input ones have one element per input slice, the output ones per output slice.
    auto check_transform = [&](const noa_hip_transform& t, const char* name, bool rotation, int64_t required) {
        if (rotation) {
            NOA_CHECK(t.kind >= NOA_HIP_MAT33_F32 && t.kind <= NOA_HIP_QUAT_F64, INVALID,
                      "%s should be Mat33 or Quaternion (f32 or f64), got kind=%d", name, t.kind);
        } else if (t.kind == NOA_HIP_XFORM_NONE) {
            return;
        } else {
            NOA_CHECK(t.kind == NOA_HIP_MAT22_F32 || t.kind == NOA_HIP_MAT22_F64, INVALID,
                      "%s should be Mat22 (f32 or f64), got kind=%d", name, t.kind);
        }
        NOA_CHECK(t.data != nullptr, INVALID, "%s should not be empty", name);
        if (t.batched) {
            NOA_CHECK(t.count == required, INVALID,
                      "%s should be a contiguous vector with n_slices=%lld elements, but got %lld elements", name,
                      (long long)required, (long long)t.count);
            NOA_CHECK(reinterpret_cast<uintptr_t>(t.data) % 16 == 0 || t.kind == NOA_HIP_MAT33_F64, INVALID,
                      "%s should be 16-byte aligned", name);
        }
    };
    check_transform(a.input_scaling, "input_scaling", false, a.input_shape[0]);
    check_transform(a.input_rotation, "input_rotation", true, a.input_shape[0]);
    check_transform(a.output_scaling, "output_scaling", false, a.output_shape[0]);
    check_transform(a.output_rotation, "output_rotation", true, a.output_shape[0]);
    auto rot64 = [](const noa_hip_transform& t) { return t.kind == NOA_HIP_MAT33_F64 || t.kind == NOA_HIP_QUAT_F64; };
    const bool coord64 = rot64(a.input_rotation);
    NOA_CHECK(rot64(a.output_rotation) == coord64, UNSUPPORTED,
              "input_rotation and output_rotation should have the same precision");
    for (const noa_hip_transform* t : {&a.input_scaling, &a.output_scaling})
        if (t->kind != NOA_HIP_XFORM_NONE)
            NOA_CHECK((t->kind == NOA_HIP_MAT22_F64) == coord64, INVALID,
                      "The scaling and rotation should have the same precision");

    // correct_weights (:2011-2018): the w window size in the output rotation's precision, from the option as given.
    const double volume_z = static_cast<double>(std::min(a.output_shape[2], a.output_shape[3]));
    const double fb_raw = coord64 ? a.w_fftfreq_blackman : static_cast<double>(static_cast<float>(a.w_fftfreq_blackman));
    const double z_c = coord64 ? volume_z : static_cast<double>(static_cast<float>(volume_z));
    const int64_t w_raw = blackman_window_size(fb_raw, z_c);
    NOA_CHECK(!a.correct_weights || (!a.add_to_output && w_raw == 1), INVALID,
              "options.correct_weights=true is not compatible with options.add_to_output=true and "
              "options.w_windowed_sinc.fftfreq_blackman=%g (=%lld pixels)", a.w_fftfreq_blackman, (long long)w_raw);

    // 32-bit offsets within a slice, weights of the values' precision.
    auto check_i32 = [&](const noa_hip_view& v, const char* name) {
        if (is_empty_view(v) || v.is_value) return;
        NOA_CHECK(fits_i32(v, 2), UNSUPPORTED, "%s: offsets within a slice exceed 32-bit indexing", name);
    };
    check_i32(a.input, "input");
    check_i32(a.input_weight, "input_weight");
    check_i32(a.output, "output");
    check_i32(a.output_weight, "output_weight");
    const bool out64 = out == NOA_HIP_F64 || out == NOA_HIP_C64;
    const bool in64 = in == NOA_HIP_F64 || in == NOA_HIP_C64;
    NOA_CHECK(in64 == out64, UNSUPPORTED, "input and output should have the same precision");
    const int wtype = out64 ? NOA_HIP_F64 : NOA_HIP_F32;
    if (!is_empty_view(a.input_weight)) NOA_CHECK(a.input_weight.dtype == wtype, UNSUPPORTED, "input_weight: expected %s weights", out64 ? "f64" : "f32");
    if (!is_empty_view(a.output_weight)) NOA_CHECK(a.output_weight.dtype == wtype, UNSUPPORTED, "output_weight: expected %s weights", out64 ? "f64" : "f32");

    // Device checks last.
    device = device_of(a.output.data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    auto check_device = [&](const noa_hip_view& v, const char* name) {
        if (is_empty_view(v) || v.is_value) return;
        const int dev = device_of(v.data, nullptr);
        NOA_CHECK(dev == device, INVALID, "The arrays should be on the same device, but got %s:device=%d and output:device=%d",
                  name, dev, device);
    };
    check_device(a.input, "input");
    check_device(a.input_weight, "input_weight");
    check_device(a.output_weight, "output_weight");
    for (const auto& [t, name] : {std::pair{&a.input_scaling, "input_scaling"}, std::pair{&a.input_rotation, "input_rotation"},
                                  std::pair{&a.output_scaling, "output_scaling"}, std::pair{&a.output_rotation, "output_rotation"}})
        if (t->kind != NOA_HIP_XFORM_NONE && t->batched)
            NOA_CHECK(device_of(t->data, nullptr) == device, INVALID, "%s should be on the compute device", name);

    // Flatten.
    InsertExtractDesc d;
    d.input = to_arr(a.input, false);
    d.input_weight = to_arr(a.input_weight, false);
    d.output = to_arr(a.output, false);
    d.output_weight = to_arr(a.output_weight, false);
    if (is_empty_view(a.input_weight)) { d.input_weight.data = nullptr; d.input_weight.is_value = false; }
    if (is_empty_view(a.output_weight)) d.output_weight.data = nullptr;
    for (int i = 0; i < 4; ++i) { d.in_shape[i] = a.input_shape[i]; d.out_shape[i] = a.output_shape[i]; }
    d.src_centered = (a.remap & 4) != 0;
    d.dst_centered = (a.remap & 8) != 0;
    d.interp = interp;
    d.rot_in = to_xform(a.input_rotation);
    d.scale_in = to_xform(a.input_scaling);
    d.rot_out = to_xform(a.output_rotation);
    d.scale_out = to_xform(a.output_scaling);
    auto scaled = [](Xform& x) {   // a single identity Mat22 is no scaling (:1234-1244)
        bool has = x.kind != NOA_HIP_XFORM_NONE;
        if (has && !x.batched) {
            const double* s = x.single;
            has = !(s[0] == 1 && s[1] == 0 && s[2] == 0 && s[3] == 1);
        }
        if (!has) x.kind = NOA_HIP_XFORM_NONE;
        return has;
    };
    const bool has_scale_in = scaled(d.scale_in), has_scale_out = scaled(d.scale_out);
    d.ews_radius[0] = a.ews_radius[0];
    d.ews_radius[1] = a.ews_radius[1];
    d.general = a.ews_radius[0] != 0 || a.ews_radius[1] != 0 || has_scale_in || has_scale_out;
    d.cutoff = a.fftfreq_cutoff;
    d.in_sinc = a.input_fftfreq_sinc;
    d.in_blackman = a.input_fftfreq_blackman;
    d.w_sinc = a.w_fftfreq_sinc;
    d.w_blackman = a.w_fftfreq_blackman;
    // The operator's window (:842-844): the option clamped to 1 / volume_z, in the coordinate precision.
    if (coord64) {
        const double fb = std::max(a.w_fftfreq_blackman, 1 / volume_z);
        d.w_size = blackman_window_size(fb, volume_z);
    } else {
        const float zf = static_cast<float>(volume_z);
        const float fb = std::max(static_cast<float>(a.w_fftfreq_blackman), 1 / zf);
        d.w_size = blackman_window_size(static_cast<double>(fb), static_cast<double>(zf));
    }
    d.add = a.add_to_output != 0;
    d.correct = a.correct_weights != 0;
    NOA_CHECK(d.w_size == 1 || !d.correct, INVALID, "correct_weights is not compatible with a w window");
    d.coord_f64 = coord64;
    return d;
}

// ctf_anisotropic (signal/CTF.hpp:143-189 checks): 2D rfft arrays, one CTF or one per batch.
CtfDesc validate_ctf(const noa_hip_view* in, const noa_hip_view* out, const noa_hip_view* wout, const int64_t* shape,
                     int32_t remap, const noa_hip_ctf* ctf, const noa_hip_ctf_options* opt, bool fused, int& device) {
    NOA_CHECK(out != nullptr && shape != nullptr && ctf != nullptr, INVALID, "null arguments");
    NOA_CHECK(out->data != nullptr && !out->is_value, INVALID, "Empty array detected");
    NOA_CHECK(remap == NOA_HIP_H2H || remap == NOA_HIP_HC2H || remap == NOA_HIP_H2HC || remap == NOA_HIP_HC2HC,
              UNSUPPORTED, "ctf_anisotropic: only the rfft layouts (h2h, hc2h, h2hc, hc2hc) are implemented, got %d",
              remap);
    for (int i = 0; i < 4; ++i) NOA_CHECK(shape[i] > 0, INVALID, "Empty array detected");
    NOA_CHECK(shape[1] == 1 && shape[2] > 1 && shape[3] > 1, INVALID,
              "Only (batched) 2d arrays are supported with anisotropic CTFs, but got ctf:shape=%s",
              shape_str(shape).c_str());
    int64_t expected[4];
    rfft_shape(shape, expected);
    NOA_CHECK(same_shape(out->shape, expected), INVALID,
              "The output shape doesn't match the expected shape. Got output:shape=%s and expected:shape=%s",
              shape_str(out->shape).c_str(), shape_str(expected).c_str());
    NOA_CHECK(out->strides[3] <= out->strides[2] || out->shape[2] == 1, UNSUPPORTED,
              "ctf_anisotropic: the output should be row-major along (H, W)");
    NOA_CHECK(is_real(out->dtype) || out->dtype == NOA_HIP_C32 || out->dtype == NOA_HIP_C64, INVALID,
              "ctf_anisotropic: the output should be f32, f64, c32 or c64, got dtype=%d", out->dtype);
    const bool has_input = in != nullptr && in->data != nullptr;
    NOA_CHECK(!fused || has_input, INVALID, "ctf_anisotropic (weighted): the input should not be empty");
    const bool layout_change = ((remap & 4) != 0) != ((remap & 8) != 0);
    if (has_input) {
        NOA_CHECK(!in->is_value, INVALID, "ctf_anisotropic: the input should be an array");
        NOA_CHECK((in->shape[0] == shape[0] || in->shape[0] == 1) && in->shape[1] == 1 &&
                  in->shape[2] == expected[2] && in->shape[3] == expected[3], INVALID,
                  "Cannot broadcast an array of shape %s into an array of shape %s", shape_str(in->shape).c_str(),
                  shape_str(expected).c_str());
        NOA_CHECK(!layout_change || !overlap(*in, *out), INVALID,
                  "This function cannot execute an in-place multiplication and a remapping");
        NOA_CHECK(in->data == out->data || !overlap(*in, *out), INVALID, "Input and output arrays should not overlap");
        const int i = in->dtype, o = out->dtype;
        NOA_CHECK((is_complex(i) && is_complex(o)) || (is_real(i) && is_real(o)) || (is_complex(i) && is_real(o)),
                  INVALID, "Invalid input/output types: %d -> %d", i, o);
        NOA_CHECK(i != NOA_HIP_C16 && i != NOA_HIP_CBF16, UNSUPPORTED, "ctf_anisotropic: f16/bf16 inputs are not supported");
        const bool i64p = i == NOA_HIP_F64 || i == NOA_HIP_C64, o64p = o == NOA_HIP_F64 || o == NOA_HIP_C64;
        NOA_CHECK(i64p == o64p, UNSUPPORTED, "input and output should have the same precision");
    }
    if (fused) {
        NOA_CHECK(wout != nullptr && wout->data != nullptr && !wout->is_value, INVALID, "output_weight: Empty array detected");
        NOA_CHECK(same_shape(wout->shape, expected), INVALID,
                  "The output_weight shape doesn't match the expected shape. Got %s and expected %s",
                  shape_str(wout->shape).c_str(), shape_str(expected).c_str());
        const bool o64p = out->dtype == NOA_HIP_F64 || out->dtype == NOA_HIP_C64;
        NOA_CHECK(wout->dtype == (o64p ? NOA_HIP_F64 : NOA_HIP_F32), UNSUPPORTED,
                  "output_weight: expected %s weights", o64p ? "f64" : "f32");
        NOA_CHECK(!overlap(*wout, *out) && !overlap(*wout, *in), INVALID, "Output arrays should not overlap");
    }
    // CTFs.
    NOA_CHECK(ctf->dtype == NOA_HIP_F32 || ctf->dtype == NOA_HIP_F64, INVALID,
              "ctf: expected CTFAnisotropic<f32> or CTFAnisotropic<f64>, got dtype=%d", ctf->dtype);
    NOA_CHECK(ctf->data != nullptr, INVALID, "ctf: Empty array detected");
    if (ctf->batched) {
        NOA_CHECK(ctf->count == shape[0], INVALID,
                  "The CTFs should be specified as a contiguous vector with %lld elements, but got %lld elements",
                  (long long)shape[0], (long long)ctf->count);
        NOA_CHECK(reinterpret_cast<uintptr_t>(ctf->data) % (ctf->dtype == NOA_HIP_F64 ? 16 : 8) == 0, INVALID,
                  "ctf: the array should be aligned as CTFAnisotropic");
    }
    // Options (:196-209; the start must be 0 for multidimensional spectra).
    noa_hip_ctf_options o{0, -1, 1, 0, 0};
    if (opt) o = *opt;
    NOA_CHECK(allclose_f64(o.fftfreq_start, 0.), INVALID,
              "For multidimensional cases, the starting fftfreq should be 0, but got %g", o.fftfreq_start);
    auto i32_ok = [](const noa_hip_view& v) { return fits_i32(v, 2); };
    NOA_CHECK(i32_ok(*out) && (!has_input || i32_ok(*in)) && (!fused || i32_ok(*wout)), UNSUPPORTED,
              "offsets within a slice exceed 32-bit indexing");

    device = device_of(out->data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    if (has_input)
        NOA_CHECK(device_of(in->data, nullptr) == device, INVALID,
                  "The input and output arrays must be on the same device");
    if (fused) NOA_CHECK(device_of(wout->data, nullptr) == device, INVALID, "The arrays should be on the same device");
    if (ctf->batched)
        NOA_CHECK(device_of(ctf->data, nullptr) == device, INVALID, "The ctf and output arrays must be on the same device");

    CtfDesc d;
    if (has_input) {
        d.input = to_arr(*in, false);
        if (in->shape[0] == 1) d.input.s0 = 0;   // broadcast over the batch
    }
    d.output = to_arr(*out, false);
    if (fused) d.output_weight = to_arr(*wout, false);
    for (int i = 0; i < 4; ++i) d.shape[i] = shape[i];
    d.src_centered = (remap & 4) != 0;
    d.dst_centered = (remap & 8) != 0;
    d.ctf_abs = !fused && o.ctf_abs != 0;
    d.ctf_squared = !fused && o.ctf_squared != 0;
    d.range_start = o.fftfreq_start;
    d.range_stop = o.fftfreq_stop;
    d.endpoint = o.endpoint != 0;
    d.ctf_f64 = ctf->dtype == NOA_HIP_F64;
    d.ctf_batched = ctf->batched != 0;
    if (d.ctf_batched) d.ctf_data = ctf->data;
    else std::memcpy(d.ctf_single, ctf->data, d.ctf_f64 ? 112 : 56);
    return d;
}

// Shape::ndim() (runtime/core/Shape.hpp:315-335) of a BDHW shape.
int ndim4(const int64_t* s) {
    if (s[1] > 1) return 3;
    if (s[2] > 1 && s[3] > 1) return 2;
    return 1;
}

// ctf_isotropic (signal/CTF.hpp:216-421) and ctf_anisotropic (:430-500) through the general operator: the checks of
// ctf_check_parameters (:143-189), the input broadcast (runtime/Utils.hpp:119-131), then noa's reordering of the
// dimensions (1D: the first non-empty one; 2D: H and W swapped unless W is innermost; 3D: the rightmost order of
// the output's strides, Shape.hpp:871-886).
CtfNdDesc validate_ctf_nd(const noa_hip_view* in, const noa_hip_view* out, const int64_t* shape, int32_t remap,
                          const noa_hip_ctf* ctf, const noa_hip_ctf_options* opt, bool isotropic, int& device) {
    const char* fn = isotropic ? "ctf_isotropic" : "ctf_anisotropic";
    NOA_CHECK(out != nullptr && shape != nullptr && ctf != nullptr, INVALID, "null arguments");
    NOA_CHECK(out->data != nullptr && !out->is_value, INVALID, "Empty array detected");
    const bool half = remap == NOA_HIP_H2H || remap == NOA_HIP_HC2H || remap == NOA_HIP_H2HC || remap == NOA_HIP_HC2HC;
    const bool full = remap == NOA_HIP_F2F || remap == NOA_HIP_FC2F || remap == NOA_HIP_F2FC || remap == NOA_HIP_FC2FC;
    NOA_CHECK(half || full, INVALID, "%s: the layout should be h2h, hc2h, h2hc, hc2hc, f2f, fc2f, f2fc or fc2fc, got %d",
              fn, remap);
    for (int i = 0; i < 4; ++i) NOA_CHECK(shape[i] > 0, INVALID, "Empty array detected");
    int64_t expected[4];
    if (half) rfft_shape(shape, expected);
    else for (int i = 0; i < 4; ++i) expected[i] = shape[i];
    NOA_CHECK(same_shape(out->shape, expected), INVALID,
              "The output shape doesn't match the expected shape. Got output:shape=%s and expected:shape=%s",
              shape_str(out->shape).c_str(), shape_str(expected).c_str());
    NOA_CHECK(is_real(out->dtype) || out->dtype == NOA_HIP_C32 || out->dtype == NOA_HIP_C64, INVALID,
              "%s: the output should be f32, f64, c32 or c64, got dtype=%d", fn, out->dtype);
    const bool has_input = in != nullptr && in->data != nullptr;
    const bool src_c = (remap & 4) != 0, dst_c = (remap & 8) != 0;
    int64_t is[4]{};
    if (has_input) {
        NOA_CHECK(!in->is_value, INVALID, "%s: the input should be an array", fn);
        for (int i = 0; i < 4; ++i) {
            NOA_CHECK(in->shape[i] == expected[i] || in->shape[i] == 1, INVALID,
                      "Cannot broadcast an array of shape %s into an array of shape %s", shape_str(in->shape).c_str(),
                      shape_str(expected).c_str());
            is[i] = in->shape[i] == expected[i] ? in->strides[i] : 0;
        }
        NOA_CHECK(src_c == dst_c || !overlap(*in, *out), INVALID,
                  "This function cannot execute an in-place multiplication and a remapping");
        NOA_CHECK(in->data == out->data || !overlap(*in, *out), INVALID, "Input and output arrays should not overlap");
        const int i = in->dtype, o = out->dtype;
        NOA_CHECK((is_complex(i) && is_complex(o)) || (is_real(i) && is_real(o)) || (is_complex(i) && is_real(o)),
                  INVALID, "Invalid input/output types: %d -> %d", i, o);
        NOA_CHECK(i != NOA_HIP_C16 && i != NOA_HIP_CBF16, UNSUPPORTED, "%s: f16/bf16 inputs are not supported", fn);
        const bool i64p = i == NOA_HIP_F64 || i == NOA_HIP_C64, o64p = o == NOA_HIP_F64 || o == NOA_HIP_C64;
        NOA_CHECK(i64p == o64p, UNSUPPORTED, "input and output should have the same precision");
    }
    NOA_CHECK(ctf->dtype == NOA_HIP_F32 || ctf->dtype == NOA_HIP_F64, INVALID,
              "ctf: expected %s<f32> or %s<f64>, got dtype=%d", isotropic ? "CTFIsotropic" : "CTFAnisotropic",
              isotropic ? "CTFIsotropic" : "CTFAnisotropic", ctf->dtype);
    NOA_CHECK(ctf->data != nullptr, INVALID, "ctf: Empty array detected");
    const bool f64 = ctf->dtype == NOA_HIP_F64;
    const size_t ctf_bytes = isotropic ? (f64 ? 88 : 44) : (f64 ? 112 : 56);
    const size_t ctf_align = isotropic ? (f64 ? 8 : 4) : (f64 ? 16 : 8);
    if (ctf->batched) {
        NOA_CHECK(ctf->count == shape[0], INVALID,
                  "The CTFs should be specified as a contiguous vector with %lld elements, but got %lld elements",
                  (long long)shape[0], (long long)ctf->count);
        NOA_CHECK(reinterpret_cast<uintptr_t>(ctf->data) % ctf_align == 0, INVALID,
                  "ctf: the array should be aligned as %s", isotropic ? "CTFIsotropic" : "CTFAnisotropic");
    }
    if (!isotropic)
        NOA_CHECK(ndim4(shape) == 2, INVALID,
                  "Only (batched) 2d arrays are supported with anisotropic CTFs, but got ctf:shape=%s",
                  shape_str(shape).c_str());
    noa_hip_ctf_options o{0, -1, 1, 0, 0};
    if (opt) o = *opt;
    NOA_CHECK(ndim4(expected) == 1 || allclose_f64(o.fftfreq_start, 0.), INVALID,
              "For multidimensional cases, the starting fftfreq should be 0, but got %g", o.fftfreq_start);


    // noa's reordering: sel[i] is the output's axis of the i-th iterated dimension (outermost first).
    int64_t sh[4], os[4];
    for (int i = 0; i < 4; ++i) { sh[i] = shape[i]; os[i] = out->strides[i]; }
    const int nd = ndim4(shape);
    int sel[3] = {0, 0, 0};
    int n = nd;
    if (nd == 1) {
        int index = 3;   // non_empty_dhw_dimension (runtime/core/Utils.hpp:30-35)
        for (int i = 1; i < 4; ++i) if (sh[i] > 1) { index = i; break; }
        sel[0] = index;
    } else if (nd == 2) {
        const int64_t k2 = sh[2] <= 1 ? std::numeric_limits<int64_t>::max() : os[2];
        const int64_t k3 = sh[3] <= 1 ? std::numeric_limits<int64_t>::max() : os[3];
        sel[0] = 2; sel[1] = 3;
        if (!(k2 >= k3)) { sel[0] = 3; sel[1] = 2; }
    } else {
        int order[3] = {0, 1, 2};
        int64_t key[3];
        for (int i = 0; i < 3; ++i) key[i] = sh[i + 1] <= 1 ? std::numeric_limits<int64_t>::max() : os[i + 1];
        std::stable_sort(order, order + 3, [&](int a, int b) { return key[a] > key[b]; });
        for (int i = 0; i < 3; ++i) sel[i] = order[i] + 1;
    }
    CtfNdDesc d;
    d.n_dims = n;
    d.batch = shape[0];
    d.output = to_arr4(*out);
    if (has_input) { d.input = to_arr4(*in); d.input.s[0] = is[0]; }
    for (int i = 0; i < n; ++i) {
        const int a = sel[i];
        d.dims[i] = sh[a];
        const int64_t iter = (half && i == n - 1) ? sh[a] / 2 + 1 : sh[a];
        NOA_CHECK(iter <= expected[a], UNSUPPORTED,
                  "%s: with a half layout, the output's width should be its innermost dimension (got strides %s)", fn,
                  shape_str(out->strides).c_str());
        d.output.s[i + 1] = os[a];
        if (has_input) d.input.s[i + 1] = is[a];
    }
    for (int i = n; i < 3; ++i) { d.output.s[i + 1] = 0; d.input.s[i + 1] = 0; }
    d.half = half;
    d.src_centered = src_c;
    d.dst_centered = dst_c;
    d.isotropic = isotropic;
    d.ctf_abs = o.ctf_abs != 0;
    d.ctf_squared = o.ctf_squared != 0;
    d.range_start = o.fftfreq_start;
    d.range_stop = o.fftfreq_stop;
    d.endpoint = o.endpoint != 0;
    d.ctf_f64 = f64;
    d.ctf_batched = ctf->batched != 0;
    if (d.ctf_batched) d.ctf_data = ctf->data;
    else std::memcpy(d.ctf_single, ctf->data, ctf_bytes);

    device = device_of(out->data, nullptr);
    NOA_CHECK(device >= 0, INVALID, "output: the pointer is not device memory");
    if (has_input)
        NOA_CHECK(device_of(in->data, nullptr) == device, INVALID,
                  "The input and output arrays must be on the same device");
    if (ctf->batched)
        NOA_CHECK(device_of(ctf->data, nullptr) == device, INVALID, "The ctf and output arrays must be on the same device");
    return d;
}

int call(Kind kind, const noa_hip_project_args* args, void* stream, const noa_hip_shift2* shifts = nullptr,
         bool no_alloc = false) {
    g_last_error.clear();
    if (!args) {
        g_last_error = "null arguments";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    int device = -1;
    ProjectDesc d;
    if (args->input.dtype == NOA_HIP_C32_BLOCKS) {
        g_last_error = "A block-copy volume (NOA_HIP_C32_BLOCKS) is accepted by the batched extraction only";
        return NOA_HIP_ERROR_UNSUPPORTED;
    }
    try {
        if (shifts) {
            // Host-only checks of the shifts first, then the extraction's own checks (device checks last).
            NOA_CHECK(args->output.dtype == NOA_HIP_C32 || args->output.dtype == NOA_HIP_C64, INVALID,
                      "phase shift: the extracted slices should be complex");
            const ShiftDesc sd = resolve_shift(shifts, args->output_shape[0], args->output_shape[2],
                                               args->output_shape[3]);
            d = validate(kind, *args, device);
            if (sd.batched)
                NOA_CHECK(device_of(sd.data, nullptr) == device, INVALID,
                          "The shift and output arrays must be on the same device");
            d.shift = sd;
        } else {
            d = validate(kind, *args, device);
        }
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    int current = 0;
    (void)hipGetDevice(&current);
    if (current != device) (void)hipSetDevice(device);
    std::string err;
    int status = NOA_HIP_OK;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    switch (kind) {
        case Kind::RASTERIZE: status = launch_rasterize(d, s, err); break;
        case Kind::INSERT: status = launch_insert(d, s, err, no_alloc); break;
        case Kind::EXTRACT: status = launch_extract(d, s, err); break;
    }
    if (current != device) (void)hipSetDevice(current);
    if (status != NOA_HIP_OK) g_last_error = err;
    return status;
}

// Validates every call of a batch with the single-call rules; all calls on one device.
int validate_batch(Kind kind, const noa_hip_project_args* calls, int32_t n_calls, std::vector<ProjectDesc>& descs,
                   int& device) {
    if (n_calls < 0 || (n_calls > 0 && !calls)) {
        g_last_error = "batch: null calls or a negative count";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    descs.resize(static_cast<size_t>(n_calls));
    device = -1;
    for (int32_t k = 0; k < n_calls; ++k) {
        try {
            int dev = -1;
            descs[static_cast<size_t>(k)] = validate(kind, calls[k], dev);
            NOA_CHECK(device < 0 || dev == device, INVALID, "the calls of a batch must be on one device");
            device = dev;
        } catch (const Fail& f) {
            g_last_error = "batch call " + std::to_string(k) + ": " + f.msg;
            return f.code;
        }
    }
    if (device < 0) (void)hipGetDevice(&device);
    return NOA_HIP_OK;
}

} // namespace
namespace {
std::mutex g_tuning_mutex;
std::vector<noa_hip_tuning> g_tuning;
} // namespace

bool tuning_lookup(int op, i64 box, int remap, int interp, int dtype, i32 values[4]) {
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    int best = -1;
    const noa_hip_tuning* found = nullptr;
    for (const noa_hip_tuning& t : g_tuning) {
        if (t.op != op) continue;
        int score = 0;
        if (t.box != 0) { if (t.box != box) continue; ++score; }
        if (t.remap != -1) { if (t.remap != remap) continue; ++score; }
        if (t.interp != -1) { if (t.interp != interp) continue; ++score; }
        if (t.dtype != -1) { if (t.dtype != dtype) continue; ++score; }
        if (score >= best) { best = score; found = &t; }
    }
    for (int i = 0; i < 4; ++i) values[i] = found ? found->values[i] : 0;
    return found != nullptr;
}
} // namespace noa_hip

extern "C" {

void noa_hip_project_args_init(noa_hip_project_args* args) {
    if (!args) return;
    std::memset(args, 0, sizeof(*args));
    args->remap = NOA_HIP_H2H;
    args->interp = NOA_HIP_INTERP_LINEAR;
    args->fftfreq_cutoff = 0.5;
    args->fftfreq_sinc = -1;
    args->fftfreq_blackman = -1;
}

void noa_hip_insert_extract_args_init(noa_hip_insert_extract_args* args) {
    if (!args) return;
    std::memset(args, 0, sizeof(*args));
    args->remap = NOA_HIP_H2H;
    args->interp = NOA_HIP_INTERP_LINEAR;
    args->input_fftfreq_sinc = -1;
    args->input_fftfreq_blackman = -1;
    args->w_fftfreq_sinc = -1;
    args->w_fftfreq_blackman = -1;
    args->fftfreq_cutoff = 0.5;
}

int noa_hip_abi_version(void) { return NOA_HIP_ABI_VERSION; }

const char* noa_hip_last_error(void) { return noa_hip::g_last_error.c_str(); }

int noa_hip_rasterize_central_slices_3d(const noa_hip_project_args* args, void* stream) {
    return noa_hip::call(noa_hip::Kind::RASTERIZE, args, stream);
}
int noa_hip_insert_central_slices_3d(const noa_hip_project_args* args, void* stream) {
    return noa_hip::call(noa_hip::Kind::INSERT, args, stream);
}
int noa_hip_insert_central_slices_3d_noalloc(const noa_hip_project_args* args, void* stream) {
    return noa_hip::call(noa_hip::Kind::INSERT, args, stream, nullptr, true);
}
int noa_hip_extract_batch_prepare(const noa_hip_project_args* calls, int32_t n_calls, void* image, int64_t capacity,
                                  int64_t* bytes) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!bytes) { g_last_error = "batch: null size pointer"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    *bytes = extract_batch_bytes(std::max<int32_t>(n_calls, 0));
    std::vector<ProjectDesc> descs;
    int device = -1;
    const int st = validate_batch(Kind::EXTRACT, calls, n_calls, descs, device);
    if (st != NOA_HIP_OK) return st;
    i32 v[4];
    tuning_lookup(NOA_HIP_OP_EXTRACT_BATCH, 0, -1, -1, -1, v);
    const i32 ppt = v[0] > 0 ? v[0] : 3;
    std::string err;
    i64 b = 0;
    const int status = extract_batch_prepare(descs.data(), n_calls, image, capacity, ppt, b, err);
    *bytes = b;
    if (status != NOA_HIP_OK) g_last_error = err;
    return status;
}
int64_t noa_hip_volume_blocks_bytes(const int64_t rfft_shape[4]) {
    if (!rfft_shape) return -1;
    return noa_hip::volume_blocks_bytes(rfft_shape);
}
int noa_hip_volume_blocks(const noa_hip_view* volume, void* blocks, int64_t capacity, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!volume || !blocks) { g_last_error = "volume blocks: null argument"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    const noa_hip_view& v = *volume;
    if (v.dtype != NOA_HIP_C32 || v.is_value || !v.data || v.shape[0] != 1 || v.strides[3] != 1) {
        g_last_error = "volume blocks: expected a c32 rfft volume (1, D, H, W/2+1) with a unit x stride";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    const int64_t need = volume_blocks_bytes(v.shape);
    if (capacity < need) {
        g_last_error = "volume blocks: the copy needs " + std::to_string(need) + " bytes";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    std::string err;
    const int st = volume_blocks(v.data, v.strides[1], v.strides[2], v.shape, blocks, static_cast<hipStream_t>(stream), err);
    if (st != NOA_HIP_OK) g_last_error = err;
    return st;
}
int noa_hip_extract_batch_launch(const void* device_image, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!device_image) { g_last_error = "batch: null image"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    i32 v[4];
    tuning_lookup(NOA_HIP_OP_EXTRACT_BATCH, 0, -1, -1, -1, v);
    std::string err;
    const int status = extract_batch_launch(device_image, v[1] > 0 ? v[1] : 2048, static_cast<hipStream_t>(stream), err);
    if (status != NOA_HIP_OK) g_last_error = err;
    return status;
}
int noa_hip_insert_batch_prepare(const noa_hip_project_args* calls, int32_t n_calls, void* image, int64_t capacity,
                                 int64_t* bytes) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!bytes) { g_last_error = "batch: null size pointer"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    *bytes = 0;
    std::vector<ProjectDesc> descs;
    int device = -1;
    const int st = validate_batch(Kind::INSERT, calls, n_calls, descs, device);
    if (st != NOA_HIP_OK) return st;
    return guarded(device, [&](std::string& err) {
        i64 b = 0;
        const int status = insert_batch_prepare(descs.data(), n_calls, image, capacity, b, err);
        *bytes = b;
        return status;
    });
}
int noa_hip_insert_batch_launch(const void* device_image, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!device_image) { g_last_error = "batch: null image"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    i32 v[4];
    tuning_lookup(NOA_HIP_OP_INSERT_BATCH, 0, -1, -1, -1, v);
    std::string err;
    const int status = insert_batch_launch(device_image, v[0] > 0 ? v[0] : 3072, static_cast<hipStream_t>(stream), err);
    if (status != NOA_HIP_OK) g_last_error = err;
    return status;
}
int noa_hip_set_tuning(const noa_hip_tuning* entries, int32_t count) {
    using namespace noa_hip;
    g_last_error.clear();
    if (count < 0 || (count > 0 && !entries)) { g_last_error = "tuning: null entries or a negative count"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    for (int32_t k = 0; k < count; ++k) {
        if (entries[k].op < NOA_HIP_OP_EXTRACT || entries[k].op > NOA_HIP_OP_INSERT_BATCH) {
            g_last_error = "tuning: unknown operator in entry " + std::to_string(k);
            return NOA_HIP_ERROR_INVALID_ARGUMENT;
        }
    }
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    g_tuning.assign(entries, entries + count);
    return NOA_HIP_OK;
}
int noa_hip_get_tuning(int32_t op, int32_t box, int32_t remap, int32_t interp, int32_t dtype, int32_t values[4]) {
    if (!values) return 0;
    return noa_hip::tuning_lookup(op, box, remap, interp, dtype, values) ? 1 : 0;
}
int noa_hip_extract_central_slices_3d(const noa_hip_project_args* args, void* stream) {
    return noa_hip::call(noa_hip::Kind::EXTRACT, args, stream);
}
int noa_hip_extract_central_slices_3d_shifted(const noa_hip_project_args* args, const noa_hip_shift2* shifts,
                                              void* stream) {
    if (!shifts) {
        noa_hip::g_last_error = "shifts: null";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    return noa_hip::call(noa_hip::Kind::EXTRACT, args, stream, shifts);
}
int noa_hip_insert_and_extract_central_slices_3d(const noa_hip_insert_extract_args* args, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!args) {
        g_last_error = "null arguments";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    int device = -1;
    InsertExtractDesc d;
    try {
        d = validate_insert_extract(*args, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) {
        return launch_insert_extract(d, static_cast<hipStream_t>(stream), err);
    });
}
int noa_hip_ctf_anisotropic_2d(const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                               int32_t remap, const noa_hip_ctf* ctf, const noa_hip_ctf_options* options,
                               void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    const bool full = remap == NOA_HIP_F2F || remap == NOA_HIP_FC2F || remap == NOA_HIP_F2FC || remap == NOA_HIP_FC2FC;
    if (full) {   // full layouts: the general operator
        CtfNdDesc d;
        try {
            d = validate_ctf_nd(input, output, shape, remap, ctf, options, false, device);
        } catch (const Fail& f) {
            g_last_error = f.msg;
            return f.code;
        }
        return guarded(device, [&](std::string& err) { return launch_ctf_nd(d, static_cast<hipStream_t>(stream), err); });
    }
    CtfDesc d;
    try {
        d = validate_ctf(input, output, nullptr, shape, remap, ctf, options, false, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) { return launch_ctf(d, static_cast<hipStream_t>(stream), err); });
}
int noa_hip_ctf_isotropic(const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                          int32_t remap, const noa_hip_ctf* ctf, const noa_hip_ctf_options* options, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    CtfNdDesc d;
    try {
        d = validate_ctf_nd(input, output, shape, remap, ctf, options, true, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) { return launch_ctf_nd(d, static_cast<hipStream_t>(stream), err); });
}
int noa_hip_ctf_anisotropic_2d_weighted(const noa_hip_view* input, const noa_hip_view* output,
                                        const noa_hip_view* output_weight, const int64_t shape[4], int32_t remap,
                                        const noa_hip_ctf* ctf, const noa_hip_ctf_options* options, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    CtfDesc d;
    try {
        d = validate_ctf(input, output, output_weight, shape, remap, ctf, options, true, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) { return launch_ctf(d, static_cast<hipStream_t>(stream), err); });
}
int noa_hip_mrc_read_header(const char* path, noa_hip_mrc_header* header) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!header) { g_last_error = "null arguments"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    std::string err;
    const int st = mrc_read_header(path, *header, err);
    if (st != NOA_HIP_OK) g_last_error = err;
    return st;
}

int noa_hip_mrc_read(const char* path, const noa_hip_view* output, int32_t clamp, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!output || !output->data) { g_last_error = "Empty array detected"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    const int device = device_of(output->data, nullptr);
    if (device < 0) { g_last_error = "output: the pointer is not device memory"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    return guarded(device, [&](std::string& err) {
        return mrc_read(path, *output, clamp != 0, static_cast<hipStream_t>(stream), err);
    });
}

int noa_hip_mrc_write(const char* path, const noa_hip_view* input, const double* spacing, int32_t dtype,
                      const double* stats, int32_t clamp, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    if (!input || !input->data) { g_last_error = "Empty array detected"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    const int device = device_of(input->data, nullptr);
    if (device < 0) { g_last_error = "input: the pointer is not device memory"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    return guarded(device, [&](std::string& err) {
        return mrc_write(path, *input, spacing, dtype, stats, clamp != 0, static_cast<hipStream_t>(stream), err);
    });
}

void noa_hip_fft_options_init(noa_hip_fft_options* options) {
    if (!options) return;
    options->norm = NOA_HIP_NORM_FORWARD;
    options->cache_plan = 1;
    options->plan_only = 0;
    options->record_and_share_workspace = 0;
}

namespace {
int fft_call(int type, const noa_hip_view* input, const noa_hip_view* output, int32_t sign,
             const noa_hip_fft_options* options, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    FftDesc d;
    try {
        d = validate_fft(type, input, output, sign, options, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) {
        try {
            return launch_fft(d, static_cast<hipStream_t>(stream), err);
        } catch (const std::exception& e) {
            err = e.what();
            return static_cast<int>(NOA_HIP_ERROR_RUNTIME);
        }
    });
}
} // namespace

int noa_hip_fft_r2c(const noa_hip_view* input, const noa_hip_view* output, const noa_hip_fft_options* options,
                    void* stream) {
    return fft_call(noa_hip::FFT_R2C, input, output, NOA_HIP_SIGN_FORWARD, options, stream);
}
int noa_hip_fft_c2r(const noa_hip_view* input, const noa_hip_view* output, const noa_hip_fft_options* options,
                    void* stream) {
    return fft_call(noa_hip::FFT_C2R, input, output, NOA_HIP_SIGN_BACKWARD, options, stream);
}
int noa_hip_fft_c2c(const noa_hip_view* input, const noa_hip_view* output, int32_t sign,
                    const noa_hip_fft_options* options, void* stream) {
    return fft_call(noa_hip::FFT_C2C, input, output, sign, options, stream);
}
int32_t noa_hip_fft_clear_cache(int32_t device) { return noa_hip::fft_clear_cache(device); }
int32_t noa_hip_fft_set_cache_limit(int32_t count, int32_t device) { return noa_hip::fft_set_cache_limit(count, device); }
int32_t noa_hip_fft_cache_limit(int32_t device) { return noa_hip::fft_cache_limit(device); }
int32_t noa_hip_fft_cache_size(int32_t device) { return noa_hip::fft_cache_size(device); }
int64_t noa_hip_fft_workspace_left_to_allocate(int32_t device) {
    return noa_hip::fft_workspace_left_to_allocate(device);
}
int32_t noa_hip_fft_set_workspace(int32_t device, void* buffer, int64_t bytes) {
    using namespace noa_hip;
    g_last_error.clear();
    if (buffer == nullptr || device_of(buffer, nullptr) != device) {
        g_last_error = "The workspace should be device memory of the given device";
        return -1;
    }
    int current = 0;
    (void)hipGetDevice(&current);
    if (current != device) (void)hipSetDevice(device);
    std::string err;
    const int n = fft_set_workspace(device, buffer, bytes, err);
    if (current != device) (void)hipSetDevice(current);
    if (n < 0) g_last_error = err;
    return n;
}
int noa_hip_cubic_bspline_prefilter(const noa_hip_view* input, const noa_hip_view* output, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    BSplineDesc d;
    try {
        d = validate_bspline(input, output, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) {
        return launch_bspline_prefilter(d, static_cast<hipStream_t>(stream), err);
    });
}
int noa_hip_remap(int32_t remap, const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                  void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    RemapDesc d;
    try {
        d = validate_remap(remap, input, output, shape, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) { return launch_remap(d, static_cast<hipStream_t>(stream), err); });
}
int noa_hip_fourier_interpolation_correction(const noa_hip_view* input, const noa_hip_view* output, int32_t interp,
                                             int32_t post_correction, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    GriddingDesc d;
    try {
        d = validate_gridding(input, output, interp, post_correction, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) {
        return launch_gridding_correction(d, static_cast<hipStream_t>(stream), err);
    });
}
int noa_hip_phase_shift_2d(const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                           int32_t remap, const noa_hip_shift2* shifts, void* stream) {
    using namespace noa_hip;
    g_last_error.clear();
    int device = -1;
    PhaseShiftDesc d;
    try {
        d = validate_phase_shift(input, output, shape, remap, shifts, device);
    } catch (const Fail& f) {
        g_last_error = f.msg;
        return f.code;
    }
    return guarded(device, [&](std::string& err) { return launch_phase_shift(d, static_cast<hipStream_t>(stream), err); });
}

} // extern "C"
