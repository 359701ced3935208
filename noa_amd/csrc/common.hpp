// Device-side building blocks of the central-slice kernels (gfx950).
//
// Every expression that changes a result is written in the same evaluation order as noa's operators, so the
// kernels agree with the CPU path to the last bit wherever no transcendental function is involved (the
// library is compiled with -ffp-contract=off):
//   index <-> frequency           src/noa/fft/core/Frequency.hpp:80-145, bounds :258-288
//   slice <-> grid transforms     src/noa/xform/FourierProject.hpp:41-95
//   windowed sinc                 src/noa/xform/FourierProject.hpp:97-110, noa::sinc base/Math.hpp:28-29
//   interpolate_spectrum          src/noa/xform/core/Interpolation.hpp:435-526
//   Mat*Vec (row dot, left fold)  src/noa/base/Mat.hpp:419-424, base/Vec.hpp:846-855
//   Quaternion::rotate            src/noa/xform/core/Quaternion.hpp:167-181
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "../../include/noa_hip.h"

namespace noa_hip {

using i32 = int32_t;
using i64 = int64_t;
using u32 = uint32_t;
using u64 = uint64_t;

// ---------------------------------------------------------------------------------------------------------
// Value types: real (float, double) and complex (float2, double2 as {re, im}).
// ---------------------------------------------------------------------------------------------------------
template<class T> struct vt;
template<> struct vt<float>   { using real = float;  static constexpr bool cx = false; };
template<> struct vt<double>  { using real = double; static constexpr bool cx = false; };
template<> struct vt<float2>  { using real = float;  static constexpr bool cx = true; };
template<> struct vt<double2> { using real = double; static constexpr bool cx = true; };

template<class T> __device__ __forceinline__ T vzero() { return T{}; }

__device__ __forceinline__ float   vadd(float a, float b) { return a + b; }
__device__ __forceinline__ double  vadd(double a, double b) { return a + b; }
__device__ __forceinline__ float2  vadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 vadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// Complex * real = {s*re, s*im} (base/Complex.hpp:203-216).
__device__ __forceinline__ float   vmul(float a, float s) { return a * s; }
__device__ __forceinline__ double  vmul(double a, double s) { return a * s; }
__device__ __forceinline__ float2  vmul(float2 a, float s) { return make_float2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 vmul(double2 a, double s) { return make_double2(s * a.x, s * a.y); }

template<class T, class R> __device__ __forceinline__ void vimag_mul(T& v, R s) {
    if constexpr (vt<T>::cx) v.y *= s;
}

// static_cast between spectrum types; complex -> real is |z|^2 (cast_or_abs_squared, base/Math.hpp:220-225).
template<class To, class From> __device__ __forceinline__ To vcast(const From& v) {
    if constexpr (vt<To>::cx && vt<From>::cx) {
        using R = typename vt<To>::real;
        To out; out.x = static_cast<R>(v.x); out.y = static_cast<R>(v.y); return out;
    } else if constexpr (!vt<To>::cx && !vt<From>::cx) {
        return static_cast<To>(v);
    } else {
        static_assert(!vt<To>::cx, "real to complex is not a spectrum type");
        return static_cast<To>(v.x * v.x + v.y * v.y);
    }
}

// Storage -> math conversion for volume reads (FP32 math on FP16/BF16 complex storage).
template<class Math, class Storage> __device__ __forceinline__ Math load_as(const Storage* p) {
    if constexpr (__hip_internal::is_same<Storage, __half2>::value) {
        const float2 f = __half22float2(*p);
        return f;
    } else if constexpr (__hip_internal::is_same<Storage, __hip_bfloat162>::value) {
        const __hip_bfloat162 h = *p;
        return make_float2(__bfloat162float(h.x), __bfloat162float(h.y));
    } else {
        return vcast<Math>(*p);
    }
}

// The value `v` of lane `lane`, for every lane of the wave (float or double).
template<class C> __device__ __forceinline__ C bcast(C v, i32 lane) {
    if constexpr (sizeof(C) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
    } else {
        const long long b = __double_as_longlong(v);
        const i32 lo = __builtin_amdgcn_readlane(static_cast<i32>(b & 0xffffffffLL), lane);
        const i32 hi = __builtin_amdgcn_readlane(static_cast<i32>(b >> 32), lane);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    }
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves: shifts by 1, 2, 4, 8 inside each row of 16
// lanes (lanes whose source falls outside the row add 0), then the last lane of rows 0 and 2 is added to rows 1
// and 3 (row_bcast:15), and lane 31 to rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ i32 wave_inclusive_sum(i32 x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
    return x;
}

// ---------------------------------------------------------------------------------------------------------
// Index <-> frequency (fft/core/Frequency.hpp:80-87, :115-122).
// ---------------------------------------------------------------------------------------------------------
__device__ __forceinline__ i32 index2frequency(bool centered, i32 index, i32 size) {
    if (centered) return index - size / 2;
    return index < (size + 1) / 2 ? index : index - size;
}
__device__ __forceinline__ i32 frequency2index(bool centered, i32 frequency, i32 size) {
    if (centered) return frequency + size / 2;
    return frequency < 0 ? frequency + size : frequency;
}

// ---------------------------------------------------------------------------------------------------------
// Transforms. A transform is read from device memory (batched) or from the kernel arguments (broadcast).
// ---------------------------------------------------------------------------------------------------------
template<class C> struct vec3 { C z, y, x; };
template<class C> struct vec2 { C y, x; };

template<class C>
struct XformArg {
    const void* data;   // device pointer when batched
    i32 kind;           // NOA_HIP_MAT33_F32, ...
    i32 batched;
    double single[9];   // broadcast value (Mat33 row-major, Quaternion zyxw, Mat22 row-major)
};

template<class C> __device__ __forceinline__ vec3<C> matvec3(const C m[9], vec3<C> v) {
    vec3<C> o;
    o.z = (m[0] * v.z + m[1] * v.y) + m[2] * v.x;
    o.y = (m[3] * v.z + m[4] * v.y) + m[5] * v.x;
    o.x = (m[6] * v.z + m[7] * v.y) + m[8] * v.x;
    return o;
}

template<class C> __device__ __forceinline__ vec3<C> quat_rotate(C qz, C qy, C qx, C qw, vec3<C> v) {
    const C a0 = qw * v.x + qy * v.z - qz * v.y;
    const C a1 = qw * v.y + qz * v.x - qx * v.z;
    const C a2 = qw * v.z + qx * v.y - qy * v.x;
    const C two = 2;
    vec3<C> o;
    o.z = v.z + two * (qx * a1 - qy * a0);
    o.y = v.y + two * (qz * a0 - qx * a2);
    o.x = v.x + two * (qy * a2 - qz * a1);
    return o;
}

// A rotation resolved for one batch index: either 9 matrix coefficients or a quaternion.
template<class C>
struct Rot {
    C m[9];
    bool quat;
    __device__ __forceinline__ vec3<C> apply(vec3<C> v) const {
        return quat ? quat_rotate<C>(m[0], m[1], m[2], m[3], v) : matvec3<C>(m, v);
    }
};

// The 3x3 matrix of a rotation (for quaternions: Quaternion::to_matrix, Quaternion.hpp:69-94); only used to
// bound and enumerate (plane geometry), never for the exact per-sample expressions.
template<class C>
__device__ __forceinline__ void rot_matrix(const Rot<C>& r, C m[9]) {
    if (!r.quat) {
#pragma unroll
        for (int i = 0; i < 9; ++i) m[i] = r.m[i];
        return;
    }
    const C z = r.m[0], y = r.m[1], x = r.m[2], w = r.m[3];
    const C tx = 2 * x, ty = 2 * y, tz = 2 * z;
    const C twx = tx * w, twy = ty * w, twz = tz * w;
    const C txx = tx * x, txy = ty * x, txz = tz * x, tyy = ty * y, tyz = tz * y, tzz = tz * z;
    m[0] = 1 - (txx + tyy); m[1] = tyz + twx; m[2] = txz - twy;
    m[3] = tyz - twx; m[4] = 1 - (txx + tzz); m[5] = txy + twz;
    m[6] = txz + twy; m[7] = txy - twz; m[8] = 1 - (tyy + tzz);
}

template<class C>
__device__ __forceinline__ Rot<C> load_rotation(const XformArg<C>& a, i64 b) {
    Rot<C> r;
    r.quat = a.kind == NOA_HIP_QUAT_F32 || a.kind == NOA_HIP_QUAT_F64;
    if (!a.batched) {
#pragma unroll
        for (int i = 0; i < 9; ++i) r.m[i] = static_cast<C>(a.single[i]);
        return r;
    }
    const unsigned char* base = static_cast<const unsigned char*>(a.data);
    switch (a.kind) {
        case NOA_HIP_MAT33_F32: {
            const float* p = reinterpret_cast<const float*>(base + b * 48);   // alignas(16) Mat33<f32>
#pragma unroll
            for (int i = 0; i < 9; ++i) r.m[i] = static_cast<C>(p[i]);
            break;
        }
        case NOA_HIP_MAT33_F64: {
            const double* p = reinterpret_cast<const double*>(base + b * 80);  // alignas(16) Mat33<f64>
#pragma unroll
            for (int i = 0; i < 9; ++i) r.m[i] = static_cast<C>(p[i]);
            break;
        }
        case NOA_HIP_QUAT_F32: {
            const float4 q = reinterpret_cast<const float4*>(base)[b];
            r.m[0] = static_cast<C>(q.x); r.m[1] = static_cast<C>(q.y);
            r.m[2] = static_cast<C>(q.z); r.m[3] = static_cast<C>(q.w);
            break;
        }
        default: {  // NOA_HIP_QUAT_F64
            const double* p = reinterpret_cast<const double*>(base + b * 32);
#pragma unroll
            for (int i = 0; i < 4; ++i) r.m[i] = static_cast<C>(p[i]);
            break;
        }
    }
    return r;
}

template<class C>
struct Scale {
    C m[4];
    bool identity;   // Empty: no arithmetic at all
    __device__ __forceinline__ vec2<C> apply(vec2<C> v) const {
        if (identity) return v;
        vec2<C> o;
        o.y = m[0] * v.y + m[1] * v.x;
        o.x = m[2] * v.y + m[3] * v.x;
        return o;
    }
};

template<class C>
__device__ __forceinline__ Scale<C> load_scaling(const XformArg<C>& a, i64 b) {
    Scale<C> s;
    s.identity = a.kind == NOA_HIP_XFORM_NONE;
    if (s.identity) return s;
    if (!a.batched) {
#pragma unroll
        for (int i = 0; i < 4; ++i) s.m[i] = static_cast<C>(a.single[i]);
        return s;
    }
    if (a.kind == NOA_HIP_MAT22_F32) {
        const float4 q = reinterpret_cast<const float4*>(a.data)[b];
        s.m[0] = static_cast<C>(q.x); s.m[1] = static_cast<C>(q.y);
        s.m[2] = static_cast<C>(q.z); s.m[3] = static_cast<C>(q.w);
    } else {
        const double* p = reinterpret_cast<const double*>(a.data) + b * 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) s.m[i] = static_cast<C>(p[i]);
    }
    return s;
}

// fourier_slice2grid (FourierProject.hpp:41-68). GENERAL = scaling or EWS present.
template<bool GENERAL, class C>
__device__ __forceinline__ vec3<C> slice2grid(vec2<C> f, const Scale<C>& s, const Rot<C>& r, C ews0, C ews1) {
    vec3<C> f3;
    if constexpr (GENERAL) {
        f = s.apply(f);
        f3.z = (ews0 * f.y) * f.y + (ews1 * f.x) * f.x;
    } else {
        f3.z = 0;
    }
    f3.y = f.y;
    f3.x = f.x;
    return r.apply(f3);
}

// fourier_grid2slice (FourierProject.hpp:71-95). Returns fz; f2 receives the 2D frequency.
template<bool GENERAL, class C>
__device__ __forceinline__ C grid2slice(vec3<C> f, const Scale<C>& s, const Rot<C>& r, C ews0, C ews1, vec2<C>& f2) {
    const vec3<C> g = r.apply(f);
    f2.y = g.y;
    f2.x = g.x;
    C fz = g.z;
    if constexpr (GENERAL) {
        fz -= (ews0 * f2.y) * f2.y + (ews1 * f2.x) * f2.x;
        f2 = s.apply(f2);
    }
    return fz;
}

template<class C> __device__ __forceinline__ C dot3(vec3<C> a) { return (a.z * a.z + a.y * a.y) + a.x * a.x; }

// windowed_sinc (FourierProject.hpp:99-110).
template<class C>
__device__ __forceinline__ C windowed_sinc(C fftfreq, C fftfreq_sinc, C fftfreq_blackman) {
    const C PI = static_cast<C>(3.141592653589793238462643383279502884);
    fftfreq *= PI;
    const C xs = fftfreq / fftfreq_sinc;
    const C s = xs == 0 ? C(1) : sin(xs) / xs;
    const C bc = fftfreq / fftfreq_blackman;
    const C blackman = static_cast<C>(0.42) + static_cast<C>(0.5) * cos(bc) + static_cast<C>(0.08) * cos(2 * bc);
    return s * blackman;
}

// ---------------------------------------------------------------------------------------------------------
// Spectrum interpolation (Interpolation.hpp:435-526), rfft along x: NEAREST and LINEAR (early flip), and the
// windows of 4-8 taps (per-tap flip).
// ---------------------------------------------------------------------------------------------------------
struct Bounds3 { i32 lo_z, hi_z, lo_y, hi_y, hi_x; i32 nz, ny; };

__device__ __forceinline__ Bounds3 make_bounds3(i32 nz, i32 ny, i32 nx) {
    return {-nz / 2, (nz - 1) / 2, -ny / 2, (ny - 1) / 2, nx / 2, nz, ny};
}

// Template value of INTERP for the windows larger than 2 (CUBIC, CUBIC_BSPLINE, LANCZOS4/6/8): one instantiation,
// the method itself is a run-time argument.
constexpr int INTERP_WINDOW = NOA_HIP_INTERP_CUBIC;

// Interp::window_size (Interp.hpp:55-77) of a method (the _FAST spellings erased).
__host__ __device__ __forceinline__ int window_size(int interp) {
    switch (interp) {
        case NOA_HIP_INTERP_NEAREST: case NOA_HIP_INTERP_LINEAR: return 2;
        case NOA_HIP_INTERP_CUBIC: case NOA_HIP_INTERP_CUBIC_BSPLINE: case NOA_HIP_INTERP_LANCZOS4: return 4;
        case NOA_HIP_INTERP_LANCZOS6: return 6;
        default: return 8;
    }
}

// interpolation_weights (Interpolation.hpp:81-141) for one dimension of a window of 4, 6 or 8, in the value's real
// precision R: CUBIC (Keys, A = -0.5), CUBIC_BSPLINE, LANCZOS (OpenCV's one-sincos form, normalised).
template<class R>
__device__ __forceinline__ void window_weights(int interp, R f, R (&w)[8]) {
    if (interp == NOA_HIP_INTERP_CUBIC) {
        const R A = static_cast<R>(-0.5);
        w[0] = ((A * (f + 1) - 5 * A) * (f + 1) + 8 * A) * (f + 1) - 4 * A;
        w[1] = ((A + 2) * f - (A + 3)) * f * f + 1;
        w[2] = ((A + 2) * (1 - f) - (A + 3)) * (1 - f) * (1 - f) + 1;
        w[3] = 1 - w[0] - w[1] - w[2];
    } else if (interp == NOA_HIP_INTERP_CUBIC_BSPLINE) {
        const R one_sixth = static_cast<R>(1) / static_cast<R>(6);
        const R two_third = static_cast<R>(2) / static_cast<R>(3);
        const R one_minus = 1 - f;
        const R one_squared = one_minus * one_minus;
        const R squared = f * f;
        w[0] = one_sixth * one_squared * one_minus;
        w[1] = two_third - static_cast<R>(0.5) * squared * (2 - f);
        w[2] = two_third - static_cast<R>(0.5) * one_squared * (2 - one_minus);
        w[3] = one_sixth * squared * f;
    } else {
        const int size = window_size(interp);
        const int center = size / 2 - 1;
        const R PI = static_cast<R>(3.1415926535897932384626433832795);
        const R s45 = static_cast<R>(0.7071067811865475);
        const R cs0[8] = {1, -s45, 0, s45, -1, s45, 0, -s45};
        const R cs1[8] = {0, -s45, 1, -s45, 0, s45, -1, s45};
        const R eps = sizeof(R) == 4 ? static_cast<R>(1.1920928955078125e-07) : static_cast<R>(2.220446049250313e-16);
        for (int i = 0; i < 8; ++i) w[i] = 0;
        R sum = 0;
        if (f <= eps) {
            w[center] = 1;
            sum = 1;
        } else if (f >= 1 - eps) {
            w[center + 1] = 1;
            sum = 1;
        } else {
            const R x0 = -(f + static_cast<R>(center));
            const R y0 = x0 * PI * static_cast<R>(0.25);
            R sn, cn;
            if constexpr (sizeof(R) == 4) sincosf(y0, &sn, &cn);
            else ::sincos(y0, &sn, &cn);
            for (int i = 0; i < size; ++i) {
                const R fij = -(f + static_cast<R>(center) - static_cast<R>(i));
                const R yj = fij * PI * static_cast<R>(0.25);
                w[i] = (cs0[i] * sn + cs1[i] * cn) / (yj * yj);
                sum += w[i];
            }
        }
        const R inv = 1 / sum;
        for (int i = 0; i < size; ++i) w[i] *= inv;
    }
}

// interpolate_spectrum for windows larger than 2 (Interpolation.hpp:435-526): the flip is decided per tap (a tap at
// x < 0 reads the conjugate of the mirrored element), taps outside the bounds contribute 0; x innermost.
template<class Math, class Storage, class C>
__device__ Math interp3_window(int interp, const Storage* __restrict__ vol, i32 sz, i32 sy, i32 sx, const Bounds3& bd,
                               bool centered, C fz, C fy, C fx) {
    using R = typename vt<Math>::real;
    const int size = window_size(interp);
    const i32 start = -(size - 1) / 2;
    const C flz = floor(fz), fly = floor(fy), flx = floor(fx);
    const i32 iz = static_cast<i32>(flz) + start, iy = static_cast<i32>(fly) + start, ix = static_cast<i32>(flx) + start;
    R wz[8], wy[8], wx[8];
    window_weights<R>(interp, static_cast<R>(fz - flz), wz);
    window_weights<R>(interp, static_cast<R>(fy - fly), wy);
    window_weights<R>(interp, static_cast<R>(fx - flx), wx);
    Math value = vzero<Math>();
    for (int z = 0; z < size; ++z) {
        Math vz = vzero<Math>();
        for (int y = 0; y < size; ++y) {
            Math vy = vzero<Math>();
            for (int x = 0; x < size; ++x) {
                i32 tz = iz + z, ty = iy + y, tx = ix + x;
                R conj = 1;
                if (tx < 0) { tz = -tz; ty = -ty; tx = -tx; conj = -1; }
                if (tz < bd.lo_z || tz > bd.hi_z || ty < bd.lo_y || ty > bd.hi_y || tx > bd.hi_x) continue;
                Math v = load_as<Math>(vol + (frequency2index(centered, tz, bd.nz) * sz +
                                              frequency2index(centered, ty, bd.ny) * sy + tx * sx));
                vimag_mul(v, conj);
                vy = vadd(vy, vmul(v, wx[x]));
            }
            vz = vadd(vz, vmul(vy, wy[y]));
        }
        value = vadd(value, vmul(vz, wz[z]));
    }
    return value;
}

template<class Math, class C>
__device__ Math interp2_window(int interp, const Math* __restrict__ base, i32 sy, i32 sx, i32 ny, i32 nx,
                               bool centered, C fy, C fx, bool is_value, Math cval) {
    using R = typename vt<Math>::real;
    const int size = window_size(interp);
    const i32 start = -(size - 1) / 2;
    const i32 lo_y = -ny / 2, hi_y = (ny - 1) / 2, hi_x = nx / 2;
    const C fly = floor(fy), flx = floor(fx);
    const i32 iy = static_cast<i32>(fly) + start, ix = static_cast<i32>(flx) + start;
    R wy[8], wx[8];
    window_weights<R>(interp, static_cast<R>(fy - fly), wy);
    window_weights<R>(interp, static_cast<R>(fx - flx), wx);
    Math value = vzero<Math>();
    for (int y = 0; y < size; ++y) {
        Math vy = vzero<Math>();
        for (int x = 0; x < size; ++x) {
            i32 ty = iy + y, tx = ix + x;
            R conj = 1;
            if (tx < 0) { ty = -ty; tx = -tx; conj = -1; }
            if (ty < lo_y || ty > hi_y || tx > hi_x) continue;
            Math v = is_value ? cval : base[frequency2index(centered, ty, ny) * sy + tx * sx];
            vimag_mul(v, conj);
            vy = vadd(vy, vmul(v, wx[x]));
        }
        value = vadd(value, vmul(vy, wy[y]));
    }
    return value;
}

// 3D: `vol` with strides (sz, sy, sx); Math = float2/float/double2/double, Storage = stored element type.
// INTERP_WINDOW: the window method `mode` at run time.
template<int INTERP, class Math, class Storage, class C>
__device__ __forceinline__ Math interp3(const Storage* __restrict__ vol, i32 sz, i32 sy, i32 sx, const Bounds3& bd,
                                       bool centered, C fz, C fy, C fx, int mode = 0) {
    if constexpr (INTERP == INTERP_WINDOW) {
        return interp3_window<Math>(mode, vol, sz, sy, sx, bd, centered, fz, fy, fx);
    }
    using R = typename vt<Math>::real;
    R conjugate = 1;
    if (fx < 0) { fz *= -1; fy *= -1; fx *= -1; conjugate = -1; }
    Math value = vzero<Math>();
    if constexpr (INTERP == NOA_HIP_INTERP_NEAREST) {
        const i32 z = static_cast<i32>(round(fz)), y = static_cast<i32>(round(fy)), x = static_cast<i32>(round(fx));
        if (z >= bd.lo_z && z <= bd.hi_z && y >= bd.lo_y && y <= bd.hi_y && x >= 0 && x <= bd.hi_x) {
            const i32 iz = frequency2index(centered, z, bd.nz), iy = frequency2index(centered, y, bd.ny);
            value = vadd(value, vmul(load_as<Math>(vol + (iz * sz + iy * sy + x * sx)), R(1)));
        }
        return value;   // no conjugation after the flip for NEAREST (Interpolation.hpp:479-484)
    } else {
        const C flz = floor(fz), fly = floor(fy), flx = floor(fx);
        const i32 z0 = static_cast<i32>(flz), y0 = static_cast<i32>(fly), x0 = static_cast<i32>(flx);
        const R tz = static_cast<R>(fz - flz), ty = static_cast<R>(fy - fly), tx = static_cast<R>(fx - flx);
        const R wz[2] = {1 - tz, tz}, wy[2] = {1 - ty, ty}, wx[2] = {1 - tx, tx};
        bool okz[2], oky[2], okx[2];
        i32 oz[2], oy[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            okz[i] = z0 + i >= bd.lo_z && z0 + i <= bd.hi_z;
            oky[i] = y0 + i >= bd.lo_y && y0 + i <= bd.hi_y;
            okx[i] = x0 + i >= 0 && x0 + i <= bd.hi_x;
            oz[i] = frequency2index(centered, z0 + i, bd.nz) * sz;
            oy[i] = frequency2index(centered, y0 + i, bd.ny) * sy;
        }
#pragma unroll
        for (int z = 0; z < 2; ++z) {
            Math vz = vzero<Math>();
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                Math vy = vzero<Math>();
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    if (okz[z] && oky[y] && okx[x])
                        vy = vadd(vy, vmul(load_as<Math>(vol + (oz[z] + oy[y] + (x0 + x) * sx)), wx[x]));
                }
                vz = vadd(vz, vmul(vy, wy[y]));
            }
            value = vadd(value, vmul(vz, wz[z]));
        }
        vimag_mul(value, conjugate);
        return value;
    }
}

// Element loader of the two-phase interpolation: a flat pointer.
template<class S>
struct GlobalLoad {
    const S* __restrict__ base;
    template<class Math> __device__ __forceinline__ Math load(u32 elem) const { return load_as<Math>(base + elem); }
};

// The same 3D interpolation split in two phases, so that a lane can have the loads of several samples in flight:
// fetch3 resolves the flip, the fractions and the tap offsets and issues every load (an out-of-bounds tap reads
// element 0 and is zeroed), combine3 sums the taps in noa's order (x innermost, then y, then z) and conjugates.
// `valid == false` fetches nothing useful (all taps zero). The result equals interp3's bit for bit.
template<class Math, class C>
struct Taps3 {
    Math v[2][2][2];      // raw loads; zeroed in combine3 where `ok` is clear (the loads stay in flight until then)
    u32 ok;               // bit (z*4 + y*2 + x): tap inside the bounds
    typename vt<Math>::real tz, ty, tx, conj;
};

// Element offsets of a volume stored row-major with strides (sz, sy, sx).
struct LinearLayout {
    i32 sz, sy, sx;
    __device__ __forceinline__ i32 z(i32 i) const { return i * sz; }
    __device__ __forceinline__ i32 y(i32 i) const { return i * sy; }
    __device__ __forceinline__ i32 x(i32 i) const { return i * sx; }
};

// Element offsets of a volume stored in blocks of 32 elements (4 x 2 x 4 along x, y, z), blocks row-major: a
// 128-byte line of c32 values holds a 4x2x2 neighbourhood, so the eight taps of a trilinear gather fall in one to
// four lines whatever the slice's orientation (row-major storage spends a whole line on the two x taps of a plane
// seen edge-on along x). Measured at 256^3 against 4x2x2, 2x2x4, 2x4x2, 4x4x1, 8x2x1, 4x4x2, 2x4x4 and 4x4x4
// blocks: 4x2x4 is the fastest by 2-3 %.
#ifndef NOA_HIP_BRICK_LOG2
#define NOA_HIP_BRICK_LOG2 2, 1, 2
#endif
struct BrickLayout {
    // log2 of the block extent along x, y, z (4 x 2 x 4: 32 c32 values, two 128-byte lines).
    static constexpr i32 LOG2[3] = {NOA_HIP_BRICK_LOG2};
    static constexpr i32 LX = LOG2[0], LY = LOG2[1], LZ = LOG2[2];
    i32 bz, by;   // element strides of one block step along z and along y
    __device__ __forceinline__ i32 z(i32 i) const { return (i >> LZ) * bz + ((i & ((1 << LZ) - 1)) << (LX + LY)); }
    __device__ __forceinline__ i32 y(i32 i) const { return (i >> LY) * by + ((i & ((1 << LY) - 1)) << LX); }
    __device__ __forceinline__ i32 x(i32 i) const { return ((i >> LX) << (LX + LY + LZ)) + (i & ((1 << LX) - 1)); }
    // Strides and element count of the block copy of a stored (nz, ny, wr) volume.
    static __host__ __device__ void strides(i64 nz, i64 ny, i64 wr, i64& bz_, i64& by_, i64& count) {
        const i64 bx = (wr + (1 << LX) - 1) >> LX;
        by_ = bx << (LX + LY + LZ);
        bz_ = ((ny + (1 << LY) - 1) >> LY) * by_;
        count = ((nz + (1 << LZ) - 1) >> LZ) * bz_;
    }
};

template<int INTERP, class Math, class Loader, class Layout, class C>
__device__ __forceinline__ Taps3<Math, C> fetch3(const Loader& vol, const Layout& L, const Bounds3& bd, bool centered,
                                                 C fz, C fy, C fx, bool valid);

template<int INTERP, class Math, class Loader, class C>
__device__ __forceinline__ Taps3<Math, C> fetch3(const Loader& vol, i32 sz, i32 sy, i32 sx,
                                                 const Bounds3& bd, bool centered, C fz, C fy, C fx, bool valid) {
    return fetch3<INTERP, Math>(vol, LinearLayout{sz, sy, sx}, bd, centered, fz, fy, fx, valid);
}

template<int INTERP, class Math, class Loader, class Layout, class C>
__device__ __forceinline__ Taps3<Math, C> fetch3(const Loader& vol, const Layout& L, const Bounds3& bd, bool centered,
                                                 C fz, C fy, C fx, bool valid) {
    using R = typename vt<Math>::real;
    Taps3<Math, C> t;
    t.conj = 1;
    if (fx < 0) { fz *= -1; fy *= -1; fx *= -1; t.conj = -1; }
    if constexpr (INTERP == NOA_HIP_INTERP_NEAREST) {
        const i32 z = static_cast<i32>(round(fz)), y = static_cast<i32>(round(fy)), x = static_cast<i32>(round(fx));
        const bool ok = valid && z >= bd.lo_z && z <= bd.hi_z && y >= bd.lo_y && y <= bd.hi_y && x >= 0 && x <= bd.hi_x;
        const u32 off = ok ? static_cast<u32>(L.z(frequency2index(centered, z, bd.nz)) +
                                              L.y(frequency2index(centered, y, bd.ny)) + L.x(x)) : 0u;
        t.v[0][0][0] = vol.template load<Math>(off);
        t.ok = ok ? 1u : 0u;
        t.tz = t.ty = t.tx = 0;
        t.conj = 1;   // no conjugation after the flip for NEAREST (Interpolation.hpp:479-484)
    } else {
        const C flz = floor(fz), fly = floor(fy), flx = floor(fx);
        const i32 z0 = static_cast<i32>(flz), y0 = static_cast<i32>(fly), x0 = static_cast<i32>(flx);
        t.tz = static_cast<R>(fz - flz);
        t.ty = static_cast<R>(fy - fly);
        t.tx = static_cast<R>(fx - flx);
        bool okz[2], oky[2], okx[2];
        i32 oz[2], oy[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            okz[i] = valid && z0 + i >= bd.lo_z && z0 + i <= bd.hi_z;
            oky[i] = y0 + i >= bd.lo_y && y0 + i <= bd.hi_y;
            okx[i] = x0 + i >= 0 && x0 + i <= bd.hi_x;
            oz[i] = L.z(frequency2index(centered, z0 + i, bd.nz));
            oy[i] = L.y(frequency2index(centered, y0 + i, bd.ny));
        }
        t.ok = 0;
#pragma unroll
        for (int z = 0; z < 2; ++z)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const bool ok = okz[z] && oky[y] && okx[x];
                    const u32 off = ok ? static_cast<u32>(oz[z] + oy[y] + L.x(x0 + x)) : 0u;
                    t.v[z][y][x] = vol.template load<Math>(off);
                    t.ok |= ok ? (1u << (z * 4 + y * 2 + x)) : 0u;
                }
    }
    return t;
}

template<int INTERP, class Math, class C>
__device__ __forceinline__ Math combine3(const Taps3<Math, C>& t) {
    using R = typename vt<Math>::real;
    if constexpr (INTERP == NOA_HIP_INTERP_NEAREST) {
        return vadd(vzero<Math>(), vmul(t.ok ? t.v[0][0][0] : vzero<Math>(), R(1)));
    } else {
        const R wz[2] = {1 - t.tz, t.tz}, wy[2] = {1 - t.ty, t.ty}, wx[2] = {1 - t.tx, t.tx};
        Math value = vzero<Math>();
#pragma unroll
        for (int z = 0; z < 2; ++z) {
            Math vz = vzero<Math>();
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                Math vy = vzero<Math>();
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const Math v = (t.ok >> (z * 4 + y * 2 + x)) & 1u ? t.v[z][y][x] : vzero<Math>();
                    vy = vadd(vy, vmul(v, wx[x]));
                }
                vz = vadd(vz, vmul(vy, wy[y]));
            }
            value = vadd(value, vmul(vz, wz[z]));
        }
        vimag_mul(value, t.conj);
        return value;
    }
}

// 2D: slice `b` of a stack with strides (sb, sy, sx); returns the interpolated value.
// A constant input (noa's AccessorValue) is read through the same code with `is_value` set.
template<int INTERP, class Math, class C>
__device__ __forceinline__ Math interp2(const Math* __restrict__ base, i32 sy, i32 sx, i32 ny, i32 nx,
                                       bool centered, C fy, C fx, bool is_value = false, Math cval = Math{},
                                       int mode = 0) {
    if constexpr (INTERP == INTERP_WINDOW) {
        return interp2_window<Math>(mode, base, sy, sx, ny, nx, centered, fy, fx, is_value, cval);
    }
    using R = typename vt<Math>::real;
    R conjugate = 1;
    if (fx < 0) { fy *= -1; fx *= -1; conjugate = -1; }
    const i32 lo_y = -ny / 2, hi_y = (ny - 1) / 2, hi_x = nx / 2;
    Math value = vzero<Math>();
    if constexpr (INTERP == NOA_HIP_INTERP_NEAREST) {
        const i32 y = static_cast<i32>(round(fy)), x = static_cast<i32>(round(fx));
        if (y >= lo_y && y <= hi_y && x >= 0 && x <= hi_x)
            value = vadd(value, vmul(is_value ? cval : base[frequency2index(centered, y, ny) * sy + x * sx], R(1)));
        return value;
    } else {
        const C fly = floor(fy), flx = floor(fx);
        const i32 y0 = static_cast<i32>(fly), x0 = static_cast<i32>(flx);
        const R ty = static_cast<R>(fy - fly), tx = static_cast<R>(fx - flx);
        const R wy[2] = {1 - ty, ty}, wx[2] = {1 - tx, tx};
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            Math vy = vzero<Math>();
            const bool oky = y0 + y >= lo_y && y0 + y <= hi_y;
            const i32 oy = frequency2index(centered, y0 + y, ny) * sy;
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                if (oky && x0 + x >= 0 && x0 + x <= hi_x)
                    vy = vadd(vy, vmul(is_value ? cval : base[oy + (x0 + x) * sx], wx[x]));
            }
            value = vadd(value, vmul(vy, wy[y]));
        }
        vimag_mul(value, conjugate);
        return value;
    }
}

// Phase shift of one rfft element (signal/PhaseShift.hpp:113-131 with nf::phase_shift, fft/core/Frequency.hpp:
// 228-248): fftfreq = frequency * (1 / shape) in the shift precision R, no shift beyond the cutoff, otherwise
// exp(-2 pi i (shift . fftfreq)) evaluated in the value precision T. mode 2: a shift of half the shape, the exact
// +-1 pattern of PhaseShiftHalf (:18-69); mode 1: no shift.
template<class V, class R>
__device__ __forceinline__ V phase_shift_at(int mode, R sy, R sx, i32 fy, i32 fx, R norm_y, R norm_x, R cutoff_sqd) {
    using T = typename vt<V>::real;
    V phase;
    phase.x = T(1);
    phase.y = T(0);
    if (mode == 2) {
        const i32 ay = fy % 2, ax = fx % 2;
        phase.x = static_cast<T>((1 - 2 * (ay < 0 ? -ay : ay)) * (1 - 2 * (ax < 0 ? -ax : ax)));
    } else if (mode == 0) {
        const R fz = R(0) * R(1);
        const R fyn = static_cast<R>(fy) * norm_y, fxn = static_cast<R>(fx) * norm_x;
        if ((fz * fz + fyn * fyn) + fxn * fxn <= cutoff_sqd) {
            const R two_pi = R(2) * static_cast<R>(3.141592653589793238462643383279502884);
            const R factor = -((two_pi * sy) * fyn + (two_pi * sx) * fxn);
            T sn, cs;
            if constexpr (sizeof(T) == 4) sincosf(static_cast<T>(factor), &sn, &cs);
            else ::sincos(static_cast<T>(factor), &sn, &cs);
            phase.x = cs;
            phase.y = sn;
        }
    }
    return phase;
}

// noa's Complex multiplication order (base/Complex.hpp:195-201).
template<class V> __device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// Device-wide float atomics. Outputs are coarse-grained hipMalloc memory: the hardware float atomic is exact.
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float2* p, float2 v) {
    unsafeAtomicAdd(&p->x, v.x);
    unsafeAtomicAdd(&p->y, v.y);
}
__device__ __forceinline__ void atomic_add(double2* p, double2 v) {
    unsafeAtomicAdd(&p->x, v.x);
    unsafeAtomicAdd(&p->y, v.y);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block i runs on XCD i % 8). Remap the hardware block
// index so that each XCD processes one contiguous range of logical blocks, which keeps neighbouring tiles (and
// their volume neighbourhoods) in the same L2.
__device__ __forceinline__ u32 xcd_swizzle(u32 bid, u32 nblocks) {
    constexpr u32 NXCD = 8;
    const u32 q = nblocks / NXCD, r = nblocks % NXCD;
    const u32 xcd = bid % NXCD, local = bid / NXCD;
    return xcd * q + (xcd < r ? xcd : r) + local;
}

} // namespace noa_hip
