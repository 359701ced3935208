/*
 * noa_hip.hpp -- C++20 facade of the MI355X central-slice library with noa's spelling.
 *
 * Header-only, over the C ABI of noa_hip.h. It mirrors the public functions of noa::xform
 * (src/noa/xform/FourierProject.hpp:1601-1914) for C++ code that does not link noa itself:
 *
 *   noa_hip::xform::rasterize_central_slices_3d<"h2h">(slice, slice_weight, slice_shape,
 *                                                      volume, volume_weight, volume_shape,
 *                                                      inv_scaling, fwd_rotation, options, stream);
 *   noa_hip::xform::insert_central_slices_3d<"h2h">(..., fwd_scaling, inv_rotation, options, stream);
 *   noa_hip::xform::extract_central_slices_3d<"hc2hc">(volume, volume_weight, volume_shape,
 *                                                      slice, slice_weight, slice_shape,
 *                                                      inv_scaling, fwd_rotation, options, stream);
 *
 * Arguments follow noa: BDHW logical shapes (Shape4), rfft-shaped strided arrays (View, element strides),
 * an Empty for absent weights or transforms, a plain value for a constant input (rasterisation, insertion),
 * one transform (Mat33, Quaternion, Mat22) or one per slice (a View of transforms in device memory).
 * The layout template argument is a string literal as in noa ("h2h", "h", "hc2hc", "hc", "hc2h", "h2hc").
 * Errors throw noa_hip::Exception with the library's message. Calls are asynchronous on `stream`
 * (a hipStream_t, nullptr = the null stream), and borrow the buffers.
 */
#ifndef NOA_HIP_HPP
#define NOA_HIP_HPP

#include "noa_hip.h"

#include <array>
#include <cmath>
#include <complex>
#include <concepts>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>

namespace noa_hip {

using i64 = std::int64_t;
using f32 = float;
using f64 = double;
using c32 = std::complex<float>;    // same layout as noa::Complex<f32>: {real, imag}
using c64 = std::complex<double>;

/* noa::Exception stand-in. */
class Exception : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

/* noa::Empty: an absent argument. */
struct Empty {};

/* BDHW shape (noa::Shape4 = Shape<isize, 4>). */
struct Shape4 {
    std::array<i64, 4> v{};
    constexpr Shape4() = default;
    constexpr Shape4(i64 b, i64 d, i64 h, i64 w) : v{b, d, h, w} {}
    constexpr i64 operator[](int i) const { return v[static_cast<std::size_t>(i)]; }
    constexpr i64& operator[](int i) { return v[static_cast<std::size_t>(i)]; }
    constexpr Shape4 rfft() const { return {v[0], v[1], v[2], v[3] / 2 + 1}; }
    constexpr bool is_empty() const { return v[0] == 0 && v[1] == 0 && v[2] == 0 && v[3] == 0; }
    constexpr i64 elements() const { return v[0] * v[1] * v[2] * v[3]; }
    constexpr bool operator==(const Shape4&) const = default;
};

/* Strided 4D array of device memory (noa::View): rfft shape, strides in elements. */
template<class T>
struct View {
    T* ptr{};
    Shape4 shape{};
    Shape4 strides{};
    View() = default;
    View(T* p, Shape4 s, Shape4 st) : ptr(p), shape(s), strides(st) {}
    /* C-contiguous view. */
    View(T* p, Shape4 s)
        : ptr(p), shape(s), strides(s[1] * s[2] * s[3], s[2] * s[3], s[3], 1) {}
    T* get() const { return ptr; }
    i64 ssize() const { return shape.elements(); }
};

/* noa::Mat33<f32> (48-byte stride: alignas(16)), Mat33<f64> (80-byte stride), row-major. */
template<class T>
struct alignas(16) Mat33 {
    T m[3][3];
};
/* noa::Quaternion: {z, y, x, w}. */
template<class T>
struct alignas(sizeof(T) * 4) Quaternion {
    T z, y, x, w;
};
/* noa::Mat22, row-major. */
template<class T>
struct alignas(sizeof(T) * 4) Mat22 {
    T m[2][2];
};
static_assert(sizeof(Mat33<f32>) == 48 && sizeof(Mat33<f64>) == 80);
static_assert(sizeof(Quaternion<f32>) == 16 && sizeof(Quaternion<f64>) == 32);
static_assert(sizeof(Mat22<f32>) == 16 && sizeof(Mat22<f64>) == 32);

namespace xform {

/* noa::fft::Layout restricted to the half-to-half remaps, constructible from a string literal. */
struct Layout {
    int32_t bits{};
    static constexpr int32_t from_string(std::string_view name) {
        if (name == "h2h" || name == "H2H" || name == "h" || name == "H") return NOA_HIP_H2H;
        if (name == "h2hc" || name == "H2HC") return NOA_HIP_H2HC;
        if (name == "hc2h" || name == "HC2H") return NOA_HIP_HC2H;
        if (name == "hc2hc" || name == "HC2HC" || name == "hc" || name == "HC") return NOA_HIP_HC2HC;
        // Full-to-full layouts: accepted by the CTF functions only (the other operators reject them).
        if (name == "f2f" || name == "F2F" || name == "f" || name == "F") return NOA_HIP_F2F;
        if (name == "f2fc" || name == "F2FC") return NOA_HIP_F2FC;
        if (name == "fc2f" || name == "FC2F") return NOA_HIP_FC2F;
        if (name == "fc2fc" || name == "FC2FC" || name == "fc" || name == "FC") return NOA_HIP_FC2FC;
        throw Exception("invalid layout: the half-to-half (h2h, h2hc, hc2h, hc2hc) or full-to-full (f2f, f2fc, "
                        "fc2f, fc2fc) remaps are supported");
    }
    template<std::size_t N>
    constexpr Layout(const char (&name)[N]) : bits(from_string(std::string_view(name, N - 1))) {}
};

/* noa::xform::Interp (Interp.hpp:17-41). */
enum class Interp : int32_t {
    NEAREST = NOA_HIP_INTERP_NEAREST,
    LINEAR = NOA_HIP_INTERP_LINEAR,
    CUBIC = NOA_HIP_INTERP_CUBIC,
    CUBIC_BSPLINE = NOA_HIP_INTERP_CUBIC_BSPLINE,
    LANCZOS4 = NOA_HIP_INTERP_LANCZOS4,
    LANCZOS6 = NOA_HIP_INTERP_LANCZOS6,
    LANCZOS8 = NOA_HIP_INTERP_LANCZOS8,
    NEAREST_FAST = NOA_HIP_INTERP_NEAREST_FAST,
    LINEAR_FAST = NOA_HIP_INTERP_LINEAR_FAST,
    CUBIC_FAST = NOA_HIP_INTERP_CUBIC_FAST,
    CUBIC_BSPLINE_FAST = NOA_HIP_INTERP_CUBIC_BSPLINE_FAST,
    LANCZOS4_FAST = NOA_HIP_INTERP_LANCZOS4_FAST,
    LANCZOS6_FAST = NOA_HIP_INTERP_LANCZOS6_FAST,
    LANCZOS8_FAST = NOA_HIP_INTERP_LANCZOS8_FAST,
};

/* Options (FourierProject.hpp:1602-1627, :1703-1739, :1817-1838), same fields and defaults. */
struct RasterizeCentralSlicesOptions {
    f64 fftfreq_cutoff{0.5};
    Shape4 target_shape{};
    std::array<f64, 2> ews_radius{};
};
struct WindowedSinc {
    f64 fftfreq_sinc{-1};
    f64 fftfreq_blackman{-1};
};
struct InsertCentralSlicesOptions {
    Interp interp{Interp::LINEAR};
    WindowedSinc windowed_sinc{};
    f64 fftfreq_cutoff{0.5};
    Shape4 target_shape{};
    std::array<f64, 2> ews_radius{};
};
struct ExtractCentralSlicesOptions {
    Interp interp{Interp::LINEAR};
    WindowedSinc w_windowed_sinc{};
    f64 fftfreq_cutoff{0.5};
    Shape4 target_shape{};
    std::array<f64, 2> ews_radius{};
};

/* InsertAndExtractCentralSlicesOptions (FourierProject.hpp:1916-1946). */
struct InsertAndExtractCentralSlicesOptions {
    Interp interp{Interp::LINEAR};
    WindowedSinc input_windowed_sinc{};
    WindowedSinc w_windowed_sinc{};
    bool add_to_output{false};
    bool correct_weights{false};
    f64 fftfreq_cutoff{0.5};
    std::array<f64, 2> ews_radius{};
};

namespace details {
    template<class T> struct is_view : std::false_type {};
    template<class T> struct is_view<View<T>> : std::true_type {};
    template<class T> constexpr bool is_view_v = is_view<std::remove_cvref_t<T>>::value;

    template<class T>
    constexpr int32_t dtype() {
        using U = std::remove_cv_t<T>;
        if constexpr (std::is_same_v<U, f32>) return NOA_HIP_F32;
        else if constexpr (std::is_same_v<U, f64>) return NOA_HIP_F64;
        else if constexpr (std::is_same_v<U, c32>) return NOA_HIP_C32;
        else if constexpr (std::is_same_v<U, c64>) return NOA_HIP_C64;
        else static_assert(sizeof(U) == 0, "unsupported value type");
    }

    template<class T>
    constexpr int32_t xform_kind() {
        using U = std::remove_cv_t<T>;
        if constexpr (std::is_same_v<U, Mat33<f32>>) return NOA_HIP_MAT33_F32;
        else if constexpr (std::is_same_v<U, Mat33<f64>>) return NOA_HIP_MAT33_F64;
        else if constexpr (std::is_same_v<U, Quaternion<f32>>) return NOA_HIP_QUAT_F32;
        else if constexpr (std::is_same_v<U, Quaternion<f64>>) return NOA_HIP_QUAT_F64;
        else if constexpr (std::is_same_v<U, Mat22<f32>>) return NOA_HIP_MAT22_F32;
        else if constexpr (std::is_same_v<U, Mat22<f64>>) return NOA_HIP_MAT22_F64;
        else static_assert(sizeof(U) == 0, "unsupported transform type");
    }

    /* A View -> strided view; a value -> constant (noa's AccessorValue); Empty -> absent. */
    template<class T>
    noa_hip_view view(const T& a) {
        noa_hip_view v{};
        using U = std::remove_cvref_t<T>;
        if constexpr (std::is_same_v<U, Empty>) {
            return v;
        } else if constexpr (is_view_v<U>) {
            v.data = const_cast<void*>(static_cast<const void*>(a.ptr));
            for (int i = 0; i < 4; ++i) {
                v.shape[i] = a.shape[i];
                v.strides[i] = a.strides[i];
            }
            v.dtype = dtype<std::remove_pointer_t<decltype(a.ptr)>>();
        } else {
            v.is_value = 1;
            v.dtype = dtype<U>();
            if constexpr (std::is_same_v<U, c32> || std::is_same_v<U, c64>) {
                v.value[0] = static_cast<f64>(a.real());
                v.value[1] = static_cast<f64>(a.imag());
            } else {
                v.value[0] = static_cast<f64>(a);
            }
        }
        return v;
    }

    /* Empty -> none; one transform -> broadcast (host value, copied at the call); a View -> one per slice. */
    template<class T>
    noa_hip_transform transform(const T& t) {
        noa_hip_transform x{};
        using U = std::remove_cvref_t<T>;
        if constexpr (std::is_same_v<U, Empty>) {
            x.kind = NOA_HIP_XFORM_NONE;
        } else if constexpr (is_view_v<U>) {
            x.data = t.ptr;
            x.kind = xform_kind<std::remove_pointer_t<decltype(t.ptr)>>();
            x.batched = 1;
            x.count = t.ssize();
        } else {
            x.data = &t;
            x.kind = xform_kind<U>();
            x.batched = 0;
            x.count = 1;
        }
        return x;
    }

    template<class Options>
    void options(noa_hip_project_args& a, const Options& o) {
        a.fftfreq_cutoff = o.fftfreq_cutoff;
        for (int i = 0; i < 4; ++i) a.target_shape[i] = o.target_shape[i];
        a.ews_radius[0] = o.ews_radius[0];
        a.ews_radius[1] = o.ews_radius[1];
        if constexpr (requires { o.interp; }) a.interp = static_cast<int32_t>(o.interp);
        if constexpr (requires { o.windowed_sinc; }) {
            a.fftfreq_sinc = o.windowed_sinc.fftfreq_sinc;
            a.fftfreq_blackman = o.windowed_sinc.fftfreq_blackman;
        }
        if constexpr (requires { o.w_windowed_sinc; }) {
            a.fftfreq_sinc = o.w_windowed_sinc.fftfreq_sinc;
            a.fftfreq_blackman = o.w_windowed_sinc.fftfreq_blackman;
        }
    }

    /* The descriptor of one call (exposed so that a caller can inspect or reuse it). */
    template<class Input, class InputWeight, class Output, class OutputWeight, class Scale, class Rotate,
             class Options>
    noa_hip_project_args make_args(int32_t remap, const Input& input, const InputWeight& input_weight,
                                   const Shape4& input_shape, const Output& output,
                                   const OutputWeight& output_weight, const Shape4& output_shape,
                                   const Scale& scaling, const Rotate& rotation, const Options& opt) {
        noa_hip_project_args a;
        noa_hip_project_args_init(&a);
        a.input = view(input);
        a.input_weight = view(input_weight);
        a.output = view(output);
        a.output_weight = view(output_weight);
        for (int i = 0; i < 4; ++i) {
            a.input_shape[i] = input_shape[i];
            a.output_shape[i] = output_shape[i];
        }
        a.remap = remap;
        a.rotation = transform(rotation);
        a.scaling = transform(scaling);
        options(a, opt);
        return a;
    }

    inline void check(int status) {
        if (status != NOA_HIP_OK) throw Exception(noa_hip_last_error());
    }
} // namespace details

/* noa::xform::rasterize_central_slices_3d (FourierProject.hpp:1653-1692). Accumulates into volume. */
template<Layout REMAP, class Input, class InputWeight = Empty, class Output, class OutputWeight = Empty,
         class Scale = Empty, class Rotate>
void rasterize_central_slices_3d(const Input& slice, const InputWeight& slice_weight, const Shape4& slice_shape,
                                 const Output& volume, const OutputWeight& volume_weight,
                                 const Shape4& volume_shape, const Scale& inv_scaling, const Rotate& fwd_rotation,
                                 const RasterizeCentralSlicesOptions& options = {}, void* stream = nullptr) {
    const noa_hip_project_args a = details::make_args(REMAP.bits, slice, slice_weight, slice_shape, volume,
                                                      volume_weight, volume_shape, inv_scaling, fwd_rotation,
                                                      options);
    details::check(noa_hip_rasterize_central_slices_3d(&a, stream));
}

/* noa::xform::insert_central_slices_3d (FourierProject.hpp:1770-1815). Accumulates into volume. */
template<Layout REMAP, class Input, class InputWeight = Empty, class Output, class OutputWeight = Empty,
         class Scale = Empty, class Rotate>
void insert_central_slices_3d(const Input& slice, const InputWeight& slice_weight, const Shape4& slice_shape,
                              const Output& volume, const OutputWeight& volume_weight, const Shape4& volume_shape,
                              const Scale& fwd_scaling, const Rotate& inv_rotation,
                              const InsertCentralSlicesOptions& options = {}, void* stream = nullptr) {
    const noa_hip_project_args a = details::make_args(REMAP.bits, slice, slice_weight, slice_shape, volume,
                                                      volume_weight, volume_shape, fwd_scaling, inv_rotation,
                                                      options);
    details::check(noa_hip_insert_central_slices_3d(&a, stream));
}

/* noa::xform::extract_central_slices_3d (FourierProject.hpp:1869-1914). Overwrites slice. */
template<Layout REMAP, class Input, class InputWeight = Empty, class Output, class OutputWeight = Empty,
         class Scale = Empty, class Rotate>
void extract_central_slices_3d(const Input& volume, const InputWeight& volume_weight, const Shape4& volume_shape,
                               const Output& slice, const OutputWeight& slice_weight, const Shape4& slice_shape,
                               const Scale& inv_scaling, const Rotate& fwd_rotation,
                               const ExtractCentralSlicesOptions& options = {}, void* stream = nullptr) {
    const noa_hip_project_args a = details::make_args(REMAP.bits, volume, volume_weight, volume_shape, slice,
                                                      slice_weight, slice_shape, inv_scaling, fwd_rotation,
                                                      options);
    details::check(noa_hip_extract_central_slices_3d(&a, stream));
}

namespace details {
    template<class InputScale, class InputRotate, class OutputScale, class OutputRotate, class Input,
             class InputWeight, class Output, class OutputWeight>
    noa_hip_insert_extract_args make_insert_extract_args(
        int32_t remap, const Input& input, const InputWeight& input_weight, const Shape4& input_shape,
        const Output& output, const OutputWeight& output_weight, const Shape4& output_shape,
        const InputScale& input_scaling, const InputRotate& input_rotation, const OutputScale& output_scaling,
        const OutputRotate& output_rotation, const InsertAndExtractCentralSlicesOptions& o) {
        noa_hip_insert_extract_args a;
        noa_hip_insert_extract_args_init(&a);
        a.input = view(input);
        a.input_weight = view(input_weight);
        a.output = view(output);
        a.output_weight = view(output_weight);
        for (int i = 0; i < 4; ++i) {
            a.input_shape[i] = input_shape[i];
            a.output_shape[i] = output_shape[i];
        }
        a.remap = remap;
        a.interp = static_cast<int32_t>(o.interp);
        a.input_scaling = transform(input_scaling);
        a.input_rotation = transform(input_rotation);
        a.output_scaling = transform(output_scaling);
        a.output_rotation = transform(output_rotation);
        a.input_fftfreq_sinc = o.input_windowed_sinc.fftfreq_sinc;
        a.input_fftfreq_blackman = o.input_windowed_sinc.fftfreq_blackman;
        a.w_fftfreq_sinc = o.w_windowed_sinc.fftfreq_sinc;
        a.w_fftfreq_blackman = o.w_windowed_sinc.fftfreq_blackman;
        a.fftfreq_cutoff = o.fftfreq_cutoff;
        a.ews_radius[0] = o.ews_radius[0];
        a.ews_radius[1] = o.ews_radius[1];
        a.add_to_output = o.add_to_output ? 1 : 0;
        a.correct_weights = o.correct_weights ? 1 : 0;
        return a;
    }
} // namespace details

/* noa::xform::insert_and_extract_central_slices_3d (FourierProject.hpp:1995-2049): extraction from a virtual volume
 * made of the input slices. Overwrites the outputs (or adds to them with add_to_output). */
template<Layout REMAP, class Input, class InputWeight = Empty, class Output, class OutputWeight = Empty,
         class InputScale = Empty, class InputRotate, class OutputScale = Empty, class OutputRotate>
void insert_and_extract_central_slices_3d(const Input& input_slice, const InputWeight& input_weight,
                                          const Shape4& input_slice_shape, const Output& output_slice,
                                          const OutputWeight& output_weight, const Shape4& output_slice_shape,
                                          const InputScale& input_fwd_scaling, const InputRotate& input_inv_rotation,
                                          const OutputScale& output_inv_scaling,
                                          const OutputRotate& output_fwd_rotation,
                                          const InsertAndExtractCentralSlicesOptions& options = {},
                                          void* stream = nullptr) {
    const noa_hip_insert_extract_args a = details::make_insert_extract_args(
        REMAP.bits, input_slice, input_weight, input_slice_shape, output_slice, output_weight, output_slice_shape,
        input_fwd_scaling, input_inv_rotation, output_inv_scaling, output_fwd_rotation, options);
    details::check(noa_hip_insert_and_extract_central_slices_3d(&a, stream));
}

/* noa::xform::fourier_interpolation_correction (FourierProject.hpp:2062-2087) on real (B, D, H, W) arrays. */
template<class Input, class Output>
void fourier_interpolation_correction(const Input& input, const Output& output, Interp interp = Interp::LINEAR,
                                      bool post_correction = true, void* stream = nullptr) {
    const noa_hip_view in = details::view(input), out = details::view(output);
    details::check(noa_hip_fourier_interpolation_correction(&in, &out, static_cast<int32_t>(interp),
                                                            post_correction ? 1 : 0, stream));
}

/* noa::xform::cubic_bspline_prefilter (CubicBSplinePrefilter.hpp:24): f32, f64, c32 or c64 (B, D, H, W) arrays;
 * the input may be broadcast into the output and may be the output. */
template<class Input, class Output>
void cubic_bspline_prefilter(const Input& input, const Output& output, void* stream = nullptr) {
    const noa_hip_view in = details::view(input), out = details::view(output);
    details::check(noa_hip_cubic_bspline_prefilter(&in, &out, stream));
}

namespace details {
    /* One (y, x) shift on the host, or a View of (y, x) pairs (Vec<f32,2> / Vec<f64,2>) on the device. */
    template<class T> struct Vec2 { T y, x; };
    template<class Shift>
    noa_hip_shift2 shifts(const Shift& s, f64 cutoff) {
        noa_hip_shift2 r{};
        r.fftfreq_cutoff = cutoff;
        using U = std::remove_cvref_t<Shift>;
        if constexpr (is_view_v<U>) {
            using E = std::remove_cv_t<std::remove_pointer_t<decltype(s.ptr)>>;
            r.data = s.ptr;
            r.dtype = std::is_same_v<E, Vec2<f64>> ? NOA_HIP_F64 : NOA_HIP_F32;
            r.batched = 1;
            r.count = s.ssize();
        } else {
            r.data = &s;
            r.dtype = std::is_same_v<U, Vec2<f64>> ? NOA_HIP_F64 : NOA_HIP_F32;
            r.batched = 0;
            r.count = 1;
        }
        return r;
    }
} // namespace details

template<class T> using Vec2 = details::Vec2<T>;

/* extract_central_slices_3d followed by noa::signal::phase_shift_2d of the slices in their own layout, fused
 * (the slices are written once). `shifts`: one Vec2 (host) or a View of one Vec2 per slice (device). */
template<Layout REMAP, class Input, class InputWeight = Empty, class Output, class OutputWeight = Empty,
         class Scale = Empty, class Rotate, class Shift>
void extract_central_slices_3d_shifted(const Input& volume, const InputWeight& volume_weight,
                                       const Shape4& volume_shape, const Output& slice,
                                       const OutputWeight& slice_weight, const Shape4& slice_shape,
                                       const Scale& inv_scaling, const Rotate& fwd_rotation, const Shift& shifts,
                                       const ExtractCentralSlicesOptions& options = {}, f64 shift_cutoff = 1,
                                       void* stream = nullptr) {
    const noa_hip_project_args a = details::make_args(REMAP.bits, volume, volume_weight, volume_shape, slice,
                                                      slice_weight, slice_shape, inv_scaling, fwd_rotation,
                                                      options);
    const noa_hip_shift2 sh = details::shifts(shifts, shift_cutoff);
    details::check(noa_hip_extract_central_slices_3d_shifted(&a, &sh, stream));
}

} // namespace xform

namespace signal {
/* noa::signal::phase_shift_2d<REMAP> (src/noa/signal/PhaseShift.hpp:215-289) for the rfft layouts. `input` may be
 * Empty (the phase shifts are written). */
template<xform::Layout REMAP, class Input, class Output, class Shift>
void phase_shift_2d(const Input& input, const Output& output, const Shape4& shape, const Shift& shifts,
                    f64 fftfreq_cutoff = 1, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input);
    const noa_hip_view out = xform::details::view(output);
    const noa_hip_shift2 sh = xform::details::shifts(shifts, fftfreq_cutoff);
    xform::details::check(noa_hip_phase_shift_2d(&in, &out, shape.v.data(), REMAP.bits, &sh, stream));
}

/* noa::signal::CTFAnisotropic<T> (src/noa/signal/core/CTF.hpp:291-575): the same members in the same order (the
 * memory layout of noa's objects), built by noa's constructor from physical parameters. */
template<class T>
struct alignas(2 * sizeof(T)) CTFAnisotropic {
    static_assert(std::is_same_v<T, f32> || std::is_same_v<T, f64>);
    struct Defocus { T value, astigmatism, angle; };   // um (positive underfocus), um, radians
    T pixel_size[2]{};
    T defocus_value{}, defocus_astigmatism{}, defocus_angle{};   // angstroms (noa's sign), radians
    T phase_shift{}, cs{}, voltage{}, amplitude{}, bfactor_forth{}, scale{}, k1{}, k2{}, k3{};

    CTFAnisotropic() = default;
    CTFAnisotropic(const std::array<T, 2>& pixel_size_, Defocus defocus, T voltage_kv, T amplitude_, T cs_mm,
                   T phase_shift_ = 0, T bfactor = 0, T scale_ = 1) {
        pixel_size[0] = pixel_size_[0];
        pixel_size[1] = pixel_size_[1];
        phase_shift = phase_shift_;
        cs = cs_mm * static_cast<T>(1e7);
        voltage = voltage_kv * static_cast<T>(1e3);
        amplitude = amplitude_;
        scale = scale_;
        defocus_value = -defocus.value * static_cast<T>(1e4);
        defocus_astigmatism = -defocus.astigmatism * static_cast<T>(1e4);
        defocus_angle = defocus.angle;
        bfactor_forth = bfactor / 4;
        const f64 h = 6.62607015e-34, c = 299792458.0, m0 = 9.1093837015e-31, e = 1.602176634e-19;
        const f64 ev = e * static_cast<f64>(voltage);
        const f64 lambda = (h * c) / std::sqrt(ev * (2 * m0 * c * c + ev)) * 1e10;
        const f64 pi = 3.1415926535897932384626433832795;
        k1 = static_cast<T>(pi * lambda);
        k2 = static_cast<T>(pi * 0.5 * static_cast<f64>(cs) * lambda * lambda * lambda);
        k3 = std::atan(amplitude / std::sqrt(1 - amplitude * amplitude));
    }
};
static_assert(sizeof(CTFAnisotropic<f32>) == sizeof(noa_hip_ctf_anisotropic_f32));
static_assert(sizeof(CTFAnisotropic<f64>) == sizeof(noa_hip_ctf_anisotropic_f64));

/* noa::signal::CTFIsotropic<T> (src/noa/signal/core/CTF.hpp:30-272): the same members in the same order, built by
 * noa's constructor from physical parameters (:61-82). */
template<class T>
struct CTFIsotropic {
    static_assert(std::is_same_v<T, f32> || std::is_same_v<T, f64>);
    T pixel_size{}, defocus{}, phase_shift{}, cs{}, voltage{}, amplitude{}, bfactor_forth{}, scale{}, k1{}, k2{}, k3{};

    CTFIsotropic() = default;
    CTFIsotropic(T pixel_size_, T defocus_um, T voltage_kv, T amplitude_, T cs_mm, T phase_shift_ = 0, T bfactor = 0,
                 T scale_ = 1) {
        pixel_size = pixel_size_;
        defocus = -defocus_um * static_cast<T>(1e4);
        phase_shift = phase_shift_;
        cs = cs_mm * static_cast<T>(1e7);
        voltage = voltage_kv * static_cast<T>(1e3);
        amplitude = amplitude_;
        scale = scale_;
        bfactor_forth = bfactor / 4;
        const f64 h = 6.62607015e-34, c = 299792458.0, m0 = 9.1093837015e-31, e = 1.602176634e-19;
        const f64 ev = e * static_cast<f64>(voltage);
        const f64 lambda = (h * c) / std::sqrt(ev * (2 * m0 * c * c + ev)) * 1e10;
        const f64 pi = 3.1415926535897932384626433832795;
        k1 = static_cast<T>(pi * lambda);
        k2 = static_cast<T>(pi * 0.5 * static_cast<f64>(cs) * lambda * lambda * lambda);
        k3 = std::atan(amplitude / std::sqrt(1 - amplitude * amplitude));
    }
};
static_assert(sizeof(CTFIsotropic<f32>) == sizeof(noa_hip_ctf_isotropic_f32));
static_assert(sizeof(CTFIsotropic<f64>) == sizeof(noa_hip_ctf_isotropic_f64));

/* CTFOptions (src/noa/signal/CTF.hpp:196-209). */
struct CTFOptions {
    struct Range { f64 start{0}, stop{-1}; bool endpoint{true}; } fftfreq_range{};
    bool ctf_abs{false};
    bool ctf_squared{false};
};

namespace details {
    template<class CTF>
    noa_hip_ctf ctf(const CTF& c) {
        noa_hip_ctf r{};
        using U = std::remove_cvref_t<CTF>;
        if constexpr (xform::details::is_view_v<U>) {
            using E = std::remove_cv_t<std::remove_pointer_t<decltype(c.ptr)>>;
            r.data = c.ptr;
            r.dtype = std::is_same_v<E, CTFAnisotropic<f64>> || std::is_same_v<E, CTFIsotropic<f64>> ? NOA_HIP_F64
                                                                                                     : NOA_HIP_F32;
            r.batched = 1;
            r.count = c.ssize();
        } else {
            r.data = &c;
            r.dtype = std::is_same_v<U, CTFAnisotropic<f64>> || std::is_same_v<U, CTFIsotropic<f64>> ? NOA_HIP_F64
                                                                                                     : NOA_HIP_F32;
            r.batched = 0;
            r.count = 1;
        }
        return r;
    }
    inline noa_hip_ctf_options ctf_options(const CTFOptions& o) {
        return {o.fftfreq_range.start, o.fftfreq_range.stop, o.fftfreq_range.endpoint ? 1 : 0, o.ctf_abs ? 1 : 0,
                o.ctf_squared ? 1 : 0};
    }
} // namespace details

/* noa::signal::ctf_isotropic<REMAP> (src/noa/signal/CTF.hpp:216-421) for 1D, 2D or 3D spectra, half or full
 * layouts. `input` may be Empty (the CTF is written). `ctf`: one CTFIsotropic (host) or a View of one per batch
 * (device). */
template<xform::Layout REMAP, class Input, class Output, class CTF>
void ctf_isotropic(const Input& input, const Output& output, const Shape4& shape, const CTF& ctf,
                   const CTFOptions& options = {}, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input), out = xform::details::view(output);
    const noa_hip_ctf c = details::ctf(ctf);
    const noa_hip_ctf_options o = details::ctf_options(options);
    xform::details::check(noa_hip_ctf_isotropic(&in, &out, shape.v.data(), REMAP.bits, &c, &o, stream));
}

/* noa::signal::ctf_anisotropic<REMAP> (src/noa/signal/CTF.hpp:430-500) for 2D spectra, half or full layouts.
 * `input` may be Empty (the CTF is written). `ctf`: one CTFAnisotropic (host) or a View of one per batch (device). */
template<xform::Layout REMAP, class Input, class Output, class CTF>
void ctf_anisotropic(const Input& input, const Output& output, const Shape4& shape, const CTF& ctf,
                     const CTFOptions& options = {}, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input), out = xform::details::view(output);
    const noa_hip_ctf c = details::ctf(ctf);
    const noa_hip_ctf_options o = details::ctf_options(options);
    xform::details::check(noa_hip_ctf_anisotropic_2d(&in, &out, shape.v.data(), REMAP.bits, &c, &o, stream));
}

/* output = input * CTF and output_weight = CTF^2 in one pass (the CTF-weighted insertion's prologue). */
template<xform::Layout REMAP, class Input, class Output, class OutputWeight, class CTF>
void ctf_anisotropic_weighted(const Input& input, const Output& output, const OutputWeight& output_weight,
                              const Shape4& shape, const CTF& ctf, const CTFOptions& options = {},
                              void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input), out = xform::details::view(output);
    const noa_hip_view w = xform::details::view(output_weight);
    const noa_hip_ctf c = details::ctf(ctf);
    const noa_hip_ctf_options o = details::ctf_options(options);
    xform::details::check(noa_hip_ctf_anisotropic_2d_weighted(&in, &out, &w, shape.v.data(), REMAP.bits, &c, &o,
                                                              stream));
}
} // namespace signal

namespace fft {
/* noa::fft::remap (src/noa/fft/Remap.hpp:205-340) between the half-to-half layouts. */
template<xform::Layout REMAP, class Input, class Output>
void remap(const Input& input, const Output& output, const Shape4& shape, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input), out = xform::details::view(output);
    xform::details::check(noa_hip_remap(REMAP.bits, &in, &out, shape.v.data(), stream));
}

/* noa::fft::Norm, Sign and FFTOptions (src/noa/fft/core/Transform.hpp:6-20, fft/Transform.hpp:39-88). */
enum class Norm : int32_t { FORWARD = NOA_HIP_NORM_FORWARD, ORTHO = NOA_HIP_NORM_ORTHO,
                            BACKWARD = NOA_HIP_NORM_BACKWARD, NONE = NOA_HIP_NORM_NONE };
enum class Sign : int32_t { FORWARD = NOA_HIP_SIGN_FORWARD, BACKWARD = NOA_HIP_SIGN_BACKWARD };
struct FFTOptions {
    Norm norm = Norm::FORWARD;
    bool cache_plan = true;
    bool plan_only = false;
    bool record_and_share_workspace = false;
};

namespace details {
    inline noa_hip_fft_options options(const FFTOptions& o) {
        return {static_cast<int32_t>(o.norm), o.cache_plan ? 1 : 0, o.plan_only ? 1 : 0,
                o.record_and_share_workspace ? 1 : 0};
    }
}

/* noa::fft::r2c, c2r and c2c (fft/Transform.hpp:160-349): real (f32/f64) <-> "h" rfft, and the full complex
 * transform; in place when input and output share their data pointer (noa's padded layout for r2c/c2r). */
template<class Input, class Output>
void r2c(const Input& input, const Output& output, FFTOptions options = {}, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input), out = xform::details::view(output);
    const noa_hip_fft_options o = details::options(options);
    xform::details::check(noa_hip_fft_r2c(&in, &out, &o, stream));
}
template<class Input, class Output>
void c2r(const Input& input, const Output& output, FFTOptions options = {}, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input), out = xform::details::view(output);
    const noa_hip_fft_options o = details::options(options);
    xform::details::check(noa_hip_fft_c2r(&in, &out, &o, stream));
}
template<class Input, class Output>
void c2c(const Input& input, const Output& output, Sign sign, FFTOptions options = {}, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input), out = xform::details::view(output);
    const noa_hip_fft_options o = details::options(options);
    xform::details::check(noa_hip_fft_c2c(&in, &out, static_cast<int32_t>(sign), &o, stream));
}

/* The plan cache of the calling thread (fft/Transform.hpp:90-142). */
inline int32_t clear_cache(int32_t device) { return noa_hip_fft_clear_cache(device); }
inline int32_t set_cache_limit(int32_t count, int32_t device) { return noa_hip_fft_set_cache_limit(count, device); }
inline int32_t cache_limit(int32_t device) { return noa_hip_fft_cache_limit(device); }
inline int32_t cache_size(int32_t device) { return noa_hip_fft_cache_size(device); }
inline int64_t workspace_left_to_allocate(int32_t device) { return noa_hip_fft_workspace_left_to_allocate(device); }
inline int32_t set_workspace(int32_t device, void* buffer, int64_t bytes) {
    const int32_t n = noa_hip_fft_set_workspace(device, buffer, bytes);
    if (n < 0) throw Exception(noa_hip_last_error());
    return n;
}
} // namespace fft

namespace io {
/* noa::io::DataType (src/noa/io/Encoding.hpp:19-26). */
enum class DataType : int32_t {
    UNKNOWN = NOA_HIP_IO_UNKNOWN, I8 = NOA_HIP_IO_I8, U8 = NOA_HIP_IO_U8, I16 = NOA_HIP_IO_I16, U16 = NOA_HIP_IO_U16,
    I32 = NOA_HIP_IO_I32, U32 = NOA_HIP_IO_U32, I64 = NOA_HIP_IO_I64, U64 = NOA_HIP_IO_U64, F16 = NOA_HIP_IO_F16,
    F32 = NOA_HIP_IO_F32, F64 = NOA_HIP_IO_F64, C16 = NOA_HIP_IO_C16, C32 = NOA_HIP_IO_C32, C64 = NOA_HIP_IO_C64,
    U4 = NOA_HIP_IO_U4, CI16 = NOA_HIP_IO_CI16
};

/* The MRC header as noa's ImageFile reports it. */
inline noa_hip_mrc_header read_header(const char* path) {
    noa_hip_mrc_header h{};
    xform::details::check(noa_hip_mrc_read_header(path, &h));
    return h;
}

/* noa::io::read_image into an existing device View of the file's shape (IO.hpp:34-57); synchronous. */
template<class Output>
void read_image(const char* path, const Output& output, bool clamp = true, void* stream = nullptr) {
    const noa_hip_view out = xform::details::view(output);
    xform::details::check(noa_hip_mrc_read(path, &out, clamp ? 1 : 0, stream));
}

/* noa::io::write_image of a device View into a new MRC file (IO.hpp:119-163); synchronous. spacing is DHW. */
struct WriteOption {
    std::array<f64, 3> spacing{};
    DataType dtype{DataType::UNKNOWN};
    bool has_stats{false};
    std::array<f64, 4> stats{};   // min, max, mean, stddev
    bool clamp{true};
};
template<class Input>
void write_image(const Input& input, const char* path, const WriteOption& option = {}, void* stream = nullptr) {
    const noa_hip_view in = xform::details::view(input);
    xform::details::check(noa_hip_mrc_write(path, &in, option.spacing.data(), static_cast<int32_t>(option.dtype),
                                            option.has_stats ? option.stats.data() : nullptr, option.clamp ? 1 : 0,
                                            stream));
}
} // namespace io
} // namespace noa_hip

#endif // NOA_HIP_HPP
